"""The counting scatter (shaders/scatter.wgsl:9-18, the reference's non-default ?sort=counting) on the GPU.

Its order is not deterministic among equal keys (one atomic per slot), so the bars are the order's invariants
against the oracle: the same keys and visible count, the visible slots first with ascending keys and only invalid
slots after them, every key's group the same set of slots as the stable order's; every tile's pair list the same set of
splats as the oracle's, in ascending key; and the image within 1/1024 and >= 55 dB of the oracle's blend in the
device's own order.
"""
import sys

import numpy as np
import pytest

import oracle
from splatlight_amd import native as N
from splatlight_amd import records as R
from splatlight_amd import scenes
from test_frame_gpu import lit_colour, psnr

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.parametrize("n,W,H,seed", [(3000, 160, 120, 2), (200_000, 640, 480, 0)])
def test_counting_scatter_invariants(ctx, n, W, H, seed):
    sc = scenes.tiny(n=n, seed=seed, width=W, height=H) if n < 10000 else scenes.cornell(n=n, seed=seed, width=W, height=H)
    col = lit_colour(sc, seed)
    fb = sc.frame_bytes()
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
    scene.write_colour(col)
    out = ctx.alloc(W * H * 16)
    try:
        N.frame_encode(ctx, scene, fb, W, H, out, counting=True)
        st = N.frame_finish(ctx)
        img = ctx.read(out, W * H * 16, np.float32).reshape(H, W, 4)
        keys = N.frame_read(ctx, N.ART_KEYS)
        order = N.frame_read(ctx, N.ART_ORDER)
        splats = N.frame_read(ctx, N.ART_SPLATS).reshape(-1, 12)
        tiles = N.frame_read(ctx, N.ART_PAIR_TILES)
        slots = N.frame_read(ctx, N.ART_PAIR_SLOTS)
    finally:
        ctx.free(out)
        scene.destroy()
    o_splats, o_keys, o_vis = oracle.project(fb, sc, colour=col)
    _, _, o_order, _ = oracle.sort_keys(o_keys)
    assert st.visible == o_vis
    np.testing.assert_array_equal(keys, o_keys)
    ns = o_keys.shape[0]
    # the visible slots first, keys ascending; after them invalid slots only (the frame's own and the padding of
    # the slot space to 256, in any order: they share the invalid bucket)
    vis_set = np.flatnonzero(o_keys != R.INVALID_KEY)
    assert np.array_equal(np.sort(order[:o_vis]), vis_set)
    tail = order[o_vis:ns]
    assert np.all((tail >= ns) | (o_keys[np.minimum(tail, ns - 1)] == R.INVALID_KEY))
    order = order[:o_vis]
    k = keys[order].astype(np.int64)
    assert np.all(np.diff(k) >= 0)
    ties = int(np.count_nonzero(np.diff(k[:o_vis]) == 0))
    starts = np.flatnonzero(np.r_[True, np.diff(k[:o_vis]) != 0])
    ends = np.r_[starts[1:], o_vis]
    for a, b in zip(starts, ends):                                               # each key's group, as a set
        assert np.array_equal(np.sort(order[a:b]), np.sort(o_order[a:b]))
    start, vals = oracle.tile_lists(o_splats, o_order, o_vis, W, H)
    assert st.pairs == vals.shape[0]
    cnt = np.diff(start).astype(np.int64)
    np.testing.assert_array_equal(tiles, np.repeat(np.arange(cnt.size), cnt))
    for t in np.flatnonzero(cnt):                                                # each tile's list
        a, b = int(start[t]), int(start[t + 1])
        assert np.array_equal(np.sort(slots[a:b]), np.sort(vals[a:b])), t
        assert np.all(np.diff(keys[slots[a:b]].astype(np.int64)) >= 0), t
    ref = oracle.composite(fb, o_splats, order, o_vis, W, H)                     # the oracle's blend in this order
    err = float(np.abs(img - ref).max())
    log(f"  counting scatter: {o_vis} visible, {ties} tied neighbours, {st.pairs} pairs; image max|err| {err:.3g} "
        f"psnr {psnr(img, ref):.1f} dB")
    assert err <= 1 / 1024 and psnr(img, ref) >= 55


def test_counting_scatter_refused_for_a_batch(ctx):
    sc = scenes.tiny(n=500, seed=1, width=64, height=48)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
    out = ctx.alloc(64 * 48 * 16)
    try:
        fb = sc.frame_bytes()
        assert _batch_counting(ctx, scene, fb, out) == 5              # SL_ERR_UNSUPPORTED, nothing launched
        assert ctx.L.sl_ctx_status(ctx.h) == 0
    finally:
        ctx.free(out)
        scene.destroy()


def _batch_counting(ctx, scene, fb, out):
    import ctypes as C
    arr = (N.View * 2)()
    keep = (C.c_uint8 * len(fb)).from_buffer_copy(fb)
    for k in range(2):
        arr[k] = N.View(scene.h, C.cast(keep, C.c_void_p), len(fb), 64, 48, 0, out)
    return ctx.L.sl_batch_encode(ctx.h, arr, 2, N.SL_FRAME_COUNTING_SORT)
