"""GPU parity on real visibility grids and real SGGX matrices: the cuts of the reference's shipped assets
(golden/asset_cut.npz, built into scenes by cuts.py), through the C ABI against the oracle.

Every lighting pass that reads a V grid or S_hat runs here on baked data (V varying along both axes, S_hat
rotated and anisotropic): iso_init's V weighting, the emitter map's and the sun maps' optical depth (dirm of
S_hat), both shading kernels' V-grid lookups (the one-wave cut walk, and the deep-tree walk a chain-shaped emitter
tree needs), env_direct's V and A, the bounce's leaf weights A V and its transfer. Bars as in the other GPU tests:
map sums bit-exact, iso_norm 2e-5, T 1e-4, colours 1e-3 relative, the lit frame 1/1024 and 55 dB, the bounce's
passes as in test_bounce_gpu.py.
"""
import math
import sys

import numpy as np
import pytest

import cuts
import oracle
from splatlight_amd import native as N
from test_bounce_gpu import check_frames
from test_lighting_gpu import check_lighting

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.parametrize("name", ["hill", "cornell"])
def test_cut_lighting(ctx, name):
    sc, nd, vg = cuts.load(name)
    log(f"lighting parity on the {name} cut: {sc.count} Gaussians, baked V grids and S_hat")
    scene, _, _ = check_lighting(ctx, sc, sun_res=512, nd=nd, vg=vg)
    scene.destroy()


def test_cut_lighting_deep_emitter_tree(ctx):
    """24 lanterns in a chain-shaped tree (every node splits off one leaf): the cut walk needs a 25-entry stack,
    so the shading runs its deep-walk kernel instead of the one-wave walk."""
    sc, nd, vg = cuts.load("hill")
    from splatlight_amd import scenes
    rng = np.random.default_rng(23)
    mu = sc.gaussians["mu"][:3000].astype(np.float64)
    lo, hi = mu.min(0), mu.max(0)
    emu = np.column_stack([rng.uniform(lo[0], hi[0], 24), rng.uniform(lo[1], hi[1], 24),
                           rng.uniform(lo[2] + 0.3, hi[2] + 1.0, 24)])
    sc.emitters = scenes._emitter_records(emu, np.repeat(rng.uniform(1.0, 20.0, 24)[:, None], 3, 1),
                                          (0.03, 0.03, 0.03))
    log("lighting parity on the hill cut with a chain emitter tree (deep-walk shading kernel)")
    scene, _, _ = check_lighting(ctx, sc, sun_res=256, nd=nd, vg=vg, emit_split=lambda n: 1)
    scene.destroy()


def test_cut_lit_frame(ctx):
    """The Cornell cut lit by its own lamp (setAssetEmitters), drawn on the device against the oracle's frame."""
    sc, nd, vg = cuts.load("cornell", width=256, height=256)
    scene, inp, o_col = check_lighting(ctx, sc, nd=nd, vg=vg)
    W, H = sc.width, sc.height
    out = ctx.alloc(W * H * 16)
    try:
        fb = sc.frame_bytes()
        N.frame_encode(ctx, scene, fb, W, H, out)
        st = N.frame_finish(ctx)
        img = ctx.read(out, W * H * 16, np.float32).reshape(H, W, 4)
    finally:
        ctx.free(out)
        scene.destroy()
    o_splats, o_keys, o_vis = oracle.project(fb, sc, colour=o_col)
    _, _, o_order, _ = oracle.sort_keys(o_keys)
    assert st.visible == o_vis
    ref = oracle.composite(fb, o_splats, o_order, o_vis, W, H)
    err = float(np.abs(img - ref).max())
    mse = float(np.mean((img.astype(np.float64) - ref) ** 2))
    psnr = math.inf if mse == 0 else 10 * math.log10(1.0 / mse)
    log(f"  lit cut frame: visible {st.visible}, max|err| {err:.3g}, psnr {psnr:.1f} dB, mean rgb {ref[..., :3].mean():.3f}")
    assert err <= 1 / 1024 and psnr >= 55


@pytest.mark.parametrize("name", ["hill", "cornell"])
def test_cut_bounce(ctx, name):
    sc, nd, vg = cuts.load(name)
    log(f"bounce parity on the {name} cut: baked V grids and S_hat in the leaves (tree built over them)")
    scene, _ = check_frames(ctx, sc, frames=2, leaves=nd, vg=vg)
    scene.destroy()
