"""The no-alloc frame (sl_frame_encode_noalloc) against the ordinary frame (sl_frame_encode), through the C ABI.

Both run the same kernels over the same inputs; only the workspace differs (the device's fixed scratch array
instead of the context's own allocation), so every artefact and the image must be bit-identical. Two no-alloc calls
in a row, the second over a smaller scene and target, check that nothing of the first call's contents (which the
array keeps) leaks into the second. Its error paths: a frame whose slots and tiles cannot fit is refused before
anything is launched; a frame whose pairs overflow the rest of the array is reported by sl_frame_finish with the
counts, without the growth-and-re-run the ordinary frame does.
"""
import sys

import numpy as np
import pytest

from splatlight_amd import native as N
from splatlight_amd import scenes

pytestmark = pytest.mark.gpu

ARTS = {"keys": N.ART_KEYS, "order": N.ART_ORDER, "splats": N.ART_SPLATS, "tiles": N.ART_PAIR_TILES,
        "slots": N.ART_PAIR_SLOTS, "ranges": N.ART_RANGES}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def colour_of(sc, seed):
    rng = np.random.default_rng(seed)
    c = np.concatenate([sc.gaussians["albedo"] * rng.uniform(0.5, 3.0, (sc.count, 1)), np.ones((sc.count, 1))], 1)
    return c.astype(np.float32)


def run(ctx, scene, sc, fmt, noalloc):
    W, H = sc.width, sc.height
    px = 16 if fmt == N.SL_OUT_RGBA32F else 8
    out = ctx.alloc(W * H * px)
    try:
        enc = N.frame_encode_noalloc if noalloc else N.frame_encode
        enc(ctx, scene, sc.frame_bytes(), W, H, out, fmt)
        st = N.frame_finish(ctx)
        img = ctx.read(out, W * H * px, np.uint8)
        art = {k: N.frame_read(ctx, w) for k, w in ARTS.items()}
    finally:
        ctx.free(out)
    return st, img, art


@pytest.fixture(scope="module")
def nctx():
    c = N.Context(0)
    yield c
    c.close()


def test_noalloc_matches_frame_encode_twice_second_smaller(nctx):
    ctx = nctx
    cases = [
        ("cornell 200k 640x480 rgba16f", scenes.cornell(n=200_000, seed=0, width=640, height=480), N.SL_OUT_RGBA16F),
        ("tiny 3k 160x120 rgba32f", scenes.tiny(n=3000, seed=5, width=160, height=120), N.SL_OUT_RGBA32F),
    ]
    handles = []
    got = []
    try:
        for name, sc, fmt in cases:   # both no-alloc calls first, back to back
            log(f"no-alloc frame: {name}")
            scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
            scene.write_colour(colour_of(sc, 1))
            handles.append(scene)
            got.append(run(ctx, scene, sc, fmt, noalloc=True))
        for (name, sc, fmt), scene, (st, img, art) in zip(cases, handles, got):
            log(f"ordinary frame: {name}")
            st0, img0, art0 = run(ctx, scene, sc, fmt, noalloc=False)
            log(f"  visible={st.visible} pairs={st.pairs} (ordinary: {st0.visible}, {st0.pairs})")
            assert st.pairs > 0 and st.visible > 0
            assert (st.slots, st.visible, st.pairs, st.overflow, st.tiles) == (
                st0.slots, st0.visible, st0.pairs, st0.overflow, st0.tiles)
            for k in ARTS:
                assert art[k].shape == art0[k].shape, k
                if k == "splats":   # words of invalid slots are undefined
                    valid = art0["keys"] != 0xFFFFFFFF
                    np.testing.assert_array_equal(art[k].reshape(-1, 12)[valid], art0[k].reshape(-1, 12)[valid])
                else:
                    np.testing.assert_array_equal(art[k], art0[k], err_msg=k)
            np.testing.assert_array_equal(img, img0)
    finally:
        for s in handles:
            s.destroy()


def test_noalloc_refuses_a_frame_that_cannot_fit(nctx):
    ctx = nctx
    # ~80 B per slot before any pair: 900k slots cannot fit 64 MiB; refused on the host, nothing launched
    sc = scenes.tiny(n=900_000, seed=2, width=320, height=240)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
    out = ctx.alloc(sc.width * sc.height * 16)
    try:
        with pytest.raises(N.SplatlightError) as e:
            N.frame_encode_noalloc(ctx, scene, sc.frame_bytes(), sc.width, sc.height, out)
        assert e.value.status == N.SL_ERR_CAPACITY
        assert "more than the scratch" in str(e.value)
        assert ctx.L.sl_ctx_status(ctx.h) == N.SL_OK   # a refusal does not poison the context
    finally:
        ctx.free(out)
        scene.destroy()


def test_noalloc_reports_pair_overflow(nctx):
    ctx = nctx
    # 600k large splats at 1080p: the slots fit, their pairs cannot; the kernels drop what does not fit (as the
    # ordinary frame's first attempt does) and sl_frame_finish reports it instead of growing
    sc = scenes.tiny(n=600_000, seed=3, width=1920, height=1080, scale_mu=-3.5)   # 31.6M pairs
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
    scene.write_colour(colour_of(sc, 2))
    out = ctx.alloc(sc.width * sc.height * 16)
    try:
        N.frame_encode_noalloc(ctx, scene, sc.frame_bytes(), sc.width, sc.height, out)
        s = N.FrameStats()
        rc = ctx.L.sl_frame_finish(ctx.h, N.C.byref(s))
        log(f"  overflow frame: status={rc} pairs={s.pairs} overflow={s.overflow}")
        assert rc == N.SL_ERR_CAPACITY
        assert s.overflow > 0 and s.pairs > 0
        assert ctx.L.sl_ctx_status(ctx.h) == N.SL_OK
        # the ordinary frame renders it (growing its own workspace)
        N.frame_encode(ctx, scene, sc.frame_bytes(), sc.width, sc.height, out)
        st = N.frame_finish(ctx)
        assert st.pairs == s.pairs + s.overflow
    finally:
        ctx.free(out)
        scene.destroy()
