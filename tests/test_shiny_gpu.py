"""GPU parity of shiny receivers (material flags bit 3: mirror flakes mixed into the M39 transfer) against the
oracle, through the C ABI: shade_emit with the lights as points and as shapes (the ASG phase, EmitParams.n5 bit 3),
both cut-walk kernels, and env_direct's highlight of a distant light of angular width; the lit frame drawn from them.

Scenes: the hill cut of the reference's shipped asset with its own materials (grass and leaves shiny, k = 0.3), and
a synthetic hill whose materials sweep roughness 0.1-1 (1: fibre-like flakes, the point phase) and k 0.2-1. Bars as
for the matte lighting: map sums bit-exact, colours within 1e-3 relative, the frame within 1/1024 and 55 dB.
"""
import math
import sys

import numpy as np
import pytest

import cuts
import oracle
from splatlight_amd import lighting as L
from splatlight_amd import native as N
from splatlight_amd import records as R
from splatlight_amd import scenes
from test_lighting_gpu import check_lighting, make

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def shiny_materials(n=8):
    m = np.zeros(n, R.MATERIAL_DTYPE)
    for i in range(n):
        r = (0.1, 0.3, 0.6, 1.0)[i % 4]
        m["flake"][i] = (1.0, r * r, r * r)
        m["phase_samples"][i] = 8
        m["vfloor"][i] = 0.02
        m["flags"][i] = R.MAT_SPECULAR if i < n - 1 else 0          # the last one matte
        m["pad0"][i] = R.pack2x16float((0.2, 0.5, 1.0)[i % 3], 0.0)
    return m


def shiny_hill(n=30_000, seed=6):
    sc = scenes.hill(n=n, seed=seed, width=320, height=180)
    sc.materials = shiny_materials()
    sc.gaussians["flags"] = np.random.default_rng(seed).integers(0, 8, sc.count).astype(np.uint32)
    return sc


def test_shiny_hill_cut_own_materials(ctx):
    sc, nd, vg = cuts.load("hill", matte=False)
    assert np.any((sc.materials["flags"] & R.MAT_SPECULAR) != 0)
    log("shiny parity: hill cut with its own shiny grass and leaf materials, sun 512^2")
    scene, _, _ = check_lighting(ctx, sc, sun_res=512, nd=nd, vg=vg)
    scene.destroy()


def test_shiny_sweep_lights_as_shapes(ctx):
    log("shiny parity: hill 30k, roughness 0.1-1 and k 0.2-1, lights as shapes (ASG), sun 256^2")
    sc = shiny_hill()
    scene, _, _ = check_lighting(ctx, sc, sun_res=256)
    scene.destroy()


def test_shiny_sweep_lights_as_points(ctx, monkeypatch):
    log("shiny parity: the same with EmitParams.n5 bit 3 clear (lights as points)")
    sc = shiny_hill(seed=7)
    real = R.emit_params

    def no_asg(**kw):
        kw["n5"] = 0
        return real(**kw)
    monkeypatch.setattr(R, "emit_params", no_asg)
    scene, _, _ = check_lighting(ctx, sc, sun_res=256)
    scene.destroy()


def test_shiny_deep_walk_kernel(ctx):
    log("shiny parity: chain emitter tree (deep-walk kernel) on the shiny sweep")
    sc = shiny_hill(n=12_000, seed=8)
    scene, _, _ = check_lighting(ctx, sc, sun_res=256, emit_split=lambda n: 1)
    scene.destroy()


def test_shiny_lit_frame(ctx):
    sc = shiny_hill(n=20_000, seed=9)
    scene, inp, o_col = check_lighting(ctx, sc, sun_res=256)
    W, H = sc.width, sc.height
    out = ctx.alloc(W * H * 16)
    try:
        fb = sc.frame_bytes()
        N.frame_encode(ctx, scene, fb, W, H, out)
        N.frame_finish(ctx)
        img = ctx.read(out, W * H * 16, np.float32).reshape(H, W, 4)
    finally:
        ctx.free(out)
        scene.destroy()
    o_splats, o_keys, o_vis = oracle.project(fb, sc, colour=o_col)
    _, _, o_order, _ = oracle.sort_keys(o_keys)
    ref = oracle.composite(fb, o_splats, o_order, o_vis, W, H)
    err = float(np.abs(img - ref).max())
    mse = float(np.mean((img.astype(np.float64) - ref) ** 2))
    psnr = math.inf if mse == 0 else 10 * math.log10(1.0 / mse)
    log(f"  shiny lit frame: max|err| {err:.3g} psnr {psnr:.1f} dB")
    assert err <= 1 / 1024 and psnr >= 55


def test_shiny_refusals(ctx):
    """The spec cut (EmitParams.n5 bit 4) and the bounce toward shiny receivers are refused; the context stays usable."""
    from splatlight_amd import bounce as BN
    from splatlight_amd import tree
    sc = shiny_hill(n=3000, seed=10)
    scene, nd, vg = make(ctx, sc)
    try:
        a = L.lighting_inputs(sc)
        EP = bytearray(a["EP"])
        np.frombuffer(EP, np.uint32)[11] |= 16
        with pytest.raises(N.SplatlightError) as e:
            scene.light(a["U"], bytes(EP), a["emit_nodes"], a["LS"])
        assert e.value.status == 5 and "spec cut" in str(e.value)
        t = tree.scene_tree(sc)
        with pytest.raises(N.SplatlightError) as e:
            BN.Bounce(scene, sc, t)
        assert e.value.status == 5 and "shiny" in str(e.value)
        scene.light(a["U"], a["EP"], a["emit_nodes"], a["LS"])        # still usable
    finally:
        scene.destroy()
