"""Batches on the GPU: many views over several resident scenes, one launch per stage, optionally a hipGraph.

Bar: every view of a batch is bit-identical to the same view rendered alone (image, stats, and its slice of the
integer artefacts: keys with the view in bits 16-23, the depth order, the tile-sorted pairs with tile ids offset
by the view's tile base); the single-view renders are themselves held to the oracle by test_frame_gpu.py.
"""
import sys

import numpy as np
import pytest

from splatlight_amd import native as N
from splatlight_amd import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_scene(ctx, sc):
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
    rng = np.random.default_rng(sc.count)
    col = np.concatenate([sc.gaussians["albedo"] * rng.uniform(0.5, 2.0, (sc.count, 1)), np.ones((sc.count, 1))], 1)
    scene.write_colour(col.astype(np.float32))
    return scene


def render_alone(ctx, scene, frame, w, h, fmt):
    out = ctx.alloc(w * h * 16)
    try:
        N.frame_encode(ctx, scene, frame, w, h, out, fmt)
        st = N.frame_finish(ctx)
        img = ctx.read(out, w * h * (16 if fmt == N.SL_OUT_RGBA32F else 8), np.uint8)
        keys = N.frame_read(ctx, N.ART_KEYS)
        order = N.frame_read(ctx, N.ART_ORDER)
        ptile = N.frame_read(ctx, N.ART_PAIR_TILES)
        pslot = N.frame_read(ctx, N.ART_PAIR_SLOTS)
    finally:
        ctx.free(out)
    return img, st, keys, order, ptile, pslot


@pytest.fixture(scope="module")
def world(ctx):
    scs = [scenes.cornell(n=20_000, seed=1), scenes.hill(n=30_000, seed=2), scenes.cornell(n=5_000, seed=3)]
    handles = [make_scene(ctx, sc) for sc in scs]
    yield scs, handles
    for h in handles:
        h.destroy()


VIEWS = [  # (scene, width, height, orbit radius, azimuth, elevation, format)
    (0, 320, 240, 3.5, -1.57, 0.2, N.SL_OUT_RGBA32F),
    (1, 256, 256, 5.0, 0.3, 0.4, N.SL_OUT_RGBA16F),
    (0, 200, 150, 2.5, -0.9, 0.1, N.SL_OUT_RGBA32F),   # ragged tiles
    (2, 640, 480, 4.0, -2.0, 0.3, N.SL_OUT_RGBA32F),
    (1, 64, 48, 6.0, 2.5, 0.6, N.SL_OUT_RGBA16F),
]


def frames_for(scs, views, shift=0.0):
    out = []
    for si, w, h, r, az, el, fmt in views:
        sc = scs[si]
        out.append(sc.frame_bytes(width=w, height=h, eye=scenes.orbit_eye(sc, r, az + shift, el)))
    return out


def run_batch(ctx, handles, views, frames, graph):
    outs = [ctx.alloc(w * h * 16) for _, w, h, *_ in views]
    try:
        N.batch_encode(ctx, [(handles[si], fb, w, h, o, fmt)
                             for (si, w, h, r, az, el, fmt), fb, o in zip(views, frames, outs)], graph=graph)
        stats = N.batch_finish(ctx, len(views))
        imgs = [ctx.read(o, w * h * (16 if fmt == N.SL_OUT_RGBA32F else 8), np.uint8)
                for (si, w, h, r, az, el, fmt), o in zip(views, outs)]
        arts = {k: N.frame_read(ctx, k) for k in (N.ART_KEYS, N.ART_ORDER, N.ART_PAIR_TILES, N.ART_PAIR_SLOTS)}
    finally:
        for o in outs:
            ctx.free(o)
    return imgs, stats, arts


def check_against_alone(ctx, scs, handles, views, frames, imgs, stats, arts):
    slot_base, tile_base, rank0, pair0 = 0, 0, 0, 0
    for v, ((si, w, h, r, az, el, fmt), fb) in enumerate(zip(views, frames)):
        img, st, keys, order, ptile, pslot = render_alone(ctx, handles[si], fb, w, h, fmt)
        assert np.array_equal(imgs[v], img), f"view {v}: image differs from the view alone"
        assert (stats[v].visible, stats[v].pairs, stats[v].slots, stats[v].tiles) == (st.visible, st.pairs, st.slots,
                                                                                       st.tiles), f"view {v}"
        n = st.slots
        bk = arts[N.ART_KEYS][slot_base:slot_base + n]
        valid = keys != 0xFFFFFFFF
        assert np.array_equal(bk[valid], keys[valid] | (v << 16)) and np.all(bk[~valid] == 0xFFFFFFFF)
        # the view's ranks in the batch order are its own order shifted to its slot base
        bo = arts[N.ART_ORDER][rank0:rank0 + st.visible]
        assert np.array_equal(bo, order[:st.visible] + slot_base)
        bt = arts[N.ART_PAIR_TILES][pair0:pair0 + st.pairs]
        bs = arts[N.ART_PAIR_SLOTS][pair0:pair0 + st.pairs]
        assert np.array_equal(bt, ptile + tile_base) and np.array_equal(bs, pslot + slot_base)
        log(f"  view {v}: {w}x{h} visible={st.visible} pairs={st.pairs} identical to the view alone")
        slot_base += (n + 255) // 256 * 256
        tile_base += st.tiles
        rank0 += st.visible
        pair0 += st.pairs


def test_batch_matches_views_alone(ctx, world):
    scs, handles = world
    frames = frames_for(scs, VIEWS)
    imgs, stats, arts = run_batch(ctx, handles, VIEWS, frames, graph=False)
    check_against_alone(ctx, scs, handles, VIEWS, frames, imgs, stats, arts)


def test_batch_graph_replays(ctx, world):
    """Captured once, replayed with moved cameras, then re-captured for a new layout; each result is the
    views-alone result."""
    scs, handles = world
    for shift in (0.0, 0.7):
        frames = frames_for(scs, VIEWS, shift)
        imgs, stats, arts = run_batch(ctx, handles, VIEWS, frames, graph=True)
        log(f"graph replay, camera shift {shift}")
        check_against_alone(ctx, scs, handles, VIEWS, frames, imgs, stats, arts)
    views2 = VIEWS[:3] + [(2, 128, 96, 3.0, 1.0, 0.5, N.SL_OUT_RGBA32F)]
    frames = frames_for(scs, views2, 0.3)
    imgs, stats, arts = run_batch(ctx, handles, views2, frames, graph=True)
    log("graph re-captured for a new layout")
    check_against_alone(ctx, scs, handles, views2, frames, imgs, stats, arts)


def test_batch_refuses_bad_views(ctx, world):
    scs, handles = world
    fb = scs[0].frame_bytes(width=64, height=48)
    out = ctx.alloc(64 * 48 * 16)
    try:
        with pytest.raises(N.SplatlightError):
            N.batch_encode(ctx, [(handles[0], fb, 64, 48, out, 0), (handles[0], fb[:100], 64, 48, out, 0)])
        with pytest.raises(N.SplatlightError):
            N.batch_encode(ctx, [(handles[0], fb, 32, 48, out, 0)])       # size != viewport
        N.batch_encode(ctx, [(handles[0], fb, 64, 48, out, 0)])          # still usable
        N.batch_finish(ctx, 1)
    finally:
        ctx.free(out)


def test_every_tuning_gives_the_same_results(ctx, world):
    """The launch-shape candidates of the autotuner change no result: images and pair lists bit-identical."""
    from splatlight_amd import autotune
    scs, handles = world
    frames = frames_for(scs, VIEWS, 0.2)
    ref_imgs, ref_stats, ref_arts = run_batch(ctx, handles, VIEWS, frames, graph=False)
    refs = {16: (ref_stats, ref_arts)}
    try:
        N.set_tuning(ctx, 256, 16, 4, tile_px=32)        # the 32 px tiling's own pair lists (eager, 8-bit digits)
        _, st32, art32 = run_batch(ctx, handles, VIEWS, frames, graph=False)
        refs[32] = (st32, art32)
        for cand in autotune.CANDIDATES:
            N.set_tuning(ctx, **cand)
            imgs, stats, arts = run_batch(ctx, handles, VIEWS, frames, graph=True)
            rs, ra = refs[cand["tile_px"]]
            assert all(np.array_equal(a, b) for a, b in zip(imgs, ref_imgs)), cand
            assert [s.pairs for s in stats] == [s.pairs for s in rs], cand
            for k in ra:
                assert np.array_equal(arts[k], ra[k]), (cand, k)
        log(f"  {len(autotune.CANDIDATES)} tunings: images identical across tile sizes and digit widths, "
            f"artefacts identical per tile size")
    finally:
        N.set_tuning(ctx, 256, 16, 4)


def test_autotune_picks_and_caches(ctx, world, tmp_path):
    from splatlight_amd import autotune
    scs, handles = world
    frames = frames_for(scs, VIEWS, 0.4)
    outs = [ctx.alloc(w * h * 16) for _, w, h, *_ in VIEWS]
    try:
        views = [(handles[si], fb, w, h, o, fmt) for (si, w, h, r, az, el, fmt), fb, o in zip(VIEWS, frames, outs)]
        cache = str(tmp_path / "tuning.json")
        best, table = autotune.tune(ctx, views, repeats=2, cache=cache)
        assert len(table) == len(autotune.CANDIDATES) and best in autotune.CANDIDATES
        assert N.get_tuning(ctx) == best
        N.set_tuning(ctx, 256, 16, 4)
        again, _ = autotune.tune(ctx, views, cache=cache)               # from the cache, no timing
        assert again == best and N.get_tuning(ctx) == best
        log(f"  autotune: best {best}, {min(ms for _, ms in table):.3f} ms .. {max(ms for _, ms in table):.3f} ms")
    finally:
        N.set_tuning(ctx, 256, 16, 4)
        for o in outs:
            ctx.free(o)
