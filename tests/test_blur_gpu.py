"""Motion blur of moving parts (shaders/project.wgsl:138-149, 208-210, 242-268) on the GPU against the oracle: the
Cornell box with its tall box moving (Part.flags bit 1, a previous pose turned and shifted), at shutter fractions
0.5 and 1 with the coverage cap on and off. Bars as for the sharp frame: splat words, keys, order and pair lists
bit-exact (the cap's exp and log correctly rounded on both sides), the image within 1/1024 and >= 55 dB.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from splatlight_amd import native as N
from splatlight_amd import scenes
from test_frame_gpu import check_integer_artefacts, lit_colour, psnr, render

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def moving_box(n=30_000, seed=4):
    sc = scenes.cornell(n=n, seed=seed, width=320, height=240)
    p = 2                                                       # the tall box
    sc.parts["flags"][p] = 2
    c, s = math.cos(0.25), math.sin(0.25)
    sc.parts["xf_prev"][p] = [[c, -s, 0, 0.3], [s, c, 0, -0.1], [0, 0, 1, 0.05]]
    return sc


@pytest.mark.parametrize("shutter,cover", [(0.5, True), (0.5, False), (1.0, True)])
def test_motion_blur_matches_oracle(ctx, shutter, cover):
    sc = moving_box()
    col = lit_colour(sc, 4)
    fb = sc.frame_bytes(shutter=shutter, blur_coverage=cover)
    st, img, art = render(ctx, sc, col, frame=fb)
    o_splats, o_keys, o_vis = oracle.project(fb, sc, colour=col)
    _, _, o_order, _ = oracle.sort_keys(o_keys)
    sharp, _, _ = oracle.project(sc.frame_bytes(), sc, colour=col)
    moved = int(np.count_nonzero(np.any(sharp != o_splats, axis=1)))
    check_integer_artefacts(sc, st, art, o_keys, o_order, o_vis, o_splats)
    ref = oracle.composite(fb, o_splats, o_order, o_vis, sc.width, sc.height)
    err = float(np.abs(img - ref).max())
    log(f"  blur shutter {shutter} cover {cover}: {moved} splats blurred, visible {st.visible}, pairs {st.pairs}, "
        f"image max|err| {err:.3g} psnr {psnr(img, ref):.1f} dB")
    assert moved > 1000
    assert err <= 1 / 1024 and psnr(img, ref) >= 55


def test_motion_blur_in_a_batch_equals_alone(ctx):
    sc = moving_box(n=8000, seed=5)
    col = lit_colour(sc, 5)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
    scene.write_colour(col)
    W, H = 160, 120
    fbs = [sc.frame_bytes(width=W, height=H, shutter=0.5), sc.frame_bytes(width=W, height=H)]
    outs = [ctx.alloc(W * H * 16) for _ in fbs]
    try:
        for rep in range(2):
            N.batch_encode(ctx, [(scene, fb, W, H, o, 0) for fb, o in zip(fbs, outs)], graph=True)
            N.batch_finish(ctx, 2)
        got = [ctx.read(o, W * H * 16, np.uint8) for o in outs]
        for fb, o, g in zip(fbs, outs, got):
            N.frame_encode(ctx, scene, fb, W, H, o)
            N.frame_finish(ctx)
            assert np.array_equal(ctx.read(o, W * H * 16, np.uint8), g)
        assert not np.array_equal(got[0], got[1])                  # the blurred view differs from the sharp one
    finally:
        for o in outs:
            ctx.free(o)
        scene.destroy()
