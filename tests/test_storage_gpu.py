"""Reduced-precision attribute storage (FP16 / BF16) on the GPU.

A scene stored in half or bfloat16 holds quat, scale and albedo rounded to nearest even; every kernel decodes them
exactly and computes in f32. Bars: against the oracle run on the same rounded records, everything is held to the
f32 bars (integer artefacts and splat words bit-exact, image <= 1/1024 and >= 55 dB, opacity-map sums bit-exact,
lit colour <= 1e-3 relative); depth keys of splats both storages keep equal the f32 scene's (they depend on the
means only). The error the
rounding itself brings is stated per channel against the float64 oracle on the original f32 records.
"""
import dataclasses
import math
import sys

import numpy as np
import pytest

import oracle
from splatlight_amd import lighting as L
from splatlight_amd import native as N
from splatlight_amd import records as R
from splatlight_amd import scenes

pytestmark = pytest.mark.gpu
STORAGES = [(R.STORAGE_F16, "f16"), (R.STORAGE_BF16, "bf16")]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def psnr(a, b):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return math.inf if mse == 0 else 10 * math.log10(1.0 / mse)


def frame(ctx, scene, sc, fb):
    W, H = sc.width, sc.height
    out = ctx.alloc(W * H * 16)
    try:
        N.frame_encode(ctx, scene, fb, W, H, out)
        st = N.frame_finish(ctx)
        img = ctx.read(out, W * H * 16, np.float32).reshape(H, W, 4)
        keys = N.frame_read(ctx, N.ART_KEYS)
        order = N.frame_read(ctx, N.ART_ORDER)
        splats = N.frame_read(ctx, N.ART_SPLATS).reshape(-1, 12)
    finally:
        ctx.free(out)
    return st, img, keys, order, splats


@pytest.mark.parametrize("storage,name", STORAGES)
def test_frame_with_reduced_storage(ctx, storage, name):
    sc = scenes.hill(n=40_000, seed=9, width=320, height=180)
    fb = sc.frame_bytes()
    rng = np.random.default_rng(3)
    colour = np.concatenate([rng.uniform(0.0, 2.0, (sc.count, 3)), np.ones((sc.count, 1))], 1).astype(np.float32)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, storage=storage)
    ref32 = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
    try:
        scene.write_colour(colour)
        ref32.write_colour(colour)
        st, img, keys, order, splats = frame(ctx, scene, sc, fb)
        _, _, keys32, _, _ = frame(ctx, ref32, sc, fb)
        assert scene.device_bytes < ref32.device_bytes
    finally:
        scene.destroy()
        ref32.destroy()
    # a depth key depends on the mean only; which splats are culled depends on the rounded scales too
    both = (keys != R.INVALID_KEY) & (keys32 != R.INVALID_KEY)
    np.testing.assert_array_equal(keys[both], keys32[both])
    log(f"{name}: {int(both.sum())} splats kept by both storages have equal keys; "
        f"{int(((keys == R.INVALID_KEY) != (keys32 == R.INVALID_KEY)).sum())} kept by only one")
    q = dataclasses.replace(sc, gaussians=R.dequantized(sc.gaussians, storage))
    o_splats, o_keys, o_vis = oracle.project(fb, q, colour=colour)
    _, _, o_order, _ = oracle.sort_keys(o_keys)
    assert st.visible == o_vis
    np.testing.assert_array_equal(keys, o_keys)
    np.testing.assert_array_equal(order, o_order)
    valid = o_keys != R.INVALID_KEY
    np.testing.assert_array_equal(splats[valid], o_splats[valid])
    ref = oracle.composite(fb, o_splats, o_order, o_vis, sc.width, sc.height)
    err = float(np.abs(img - ref).max())
    log(f"{name}: vs oracle on the stored values max|err| {err:.3g}, psnr {psnr(img, ref):.1f} dB")
    assert err <= 1 / 1024 and psnr(img, ref) >= 55
    # the rounding's own effect: against the float64 oracle on the original f32 records, per channel
    sp64, valid64 = oracle.project(fb, sc, colour=colour, f64=True)
    k32 = oracle.project(fb, sc, colour=colour)[1]
    _, _, order32, _ = oracle.sort_keys(k32)
    ref64 = oracle.composite_f64(fb, sp64, order32, int((k32 != R.INVALID_KEY).sum()), sc.width, sc.height)
    errs = [float(np.abs(img[..., c] - ref64[..., c]).max()) for c in range(4)]
    ps = [psnr(img[..., c], ref64[..., c]) for c in range(4)]
    log(f"{name}: vs float64 oracle on f32 records, per channel r/g/b/a max|err| "
        + " ".join(f"{e:.3g}" for e in errs) + ", psnr " + " ".join(f"{p:.1f}" for p in ps) + " dB")
    assert min(ps) >= 40.0


@pytest.mark.parametrize("storage,name", STORAGES)
def test_lighting_with_reduced_storage(ctx, storage, name):
    sc = scenes.hill(n=20_000, seed=4, width=320, height=180)
    nd = scenes.leaf_nodes(sc.gaussians)
    vg = scenes.vgrid_open(sc.count)
    inp = L.lighting_inputs(sc, sun_res=256)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials, nodes=nd, vgrid=vg,
                    storage=storage)
    try:
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        acc = scene.read(N.SCENE_LSM_ACC, np.uint32)
        sacc = scene.read(N.SCENE_SUN_ACC, np.uint32)
        col = scene.read(N.SCENE_COLOUR).reshape(-1, 4)
    finally:
        scene.destroy()
    q = R.dequantized(sc.gaussians, storage)
    nq = scenes.leaf_nodes(sc.gaussians)          # the leaf records are the caller's f32 nodes either way
    o_iso = oracle.iso_init(q, nq, vg)
    o_acc, _, o_T = oracle.lsm(inp["U"], inp["LS"], q, nq, sc.parts)
    assert int(np.count_nonzero(acc != o_acc.ravel())) == 0
    o_col = oracle.shade_emit(inp["U"], inp["EP"], inp["emit_nodes"], q, nq, vg, sc.parts, o_T, sc.materials, o_iso)
    o_sacc, _, o_envT = oracle.sun(inp["U"], inp["SM"], inp["desc"], q, nq, sc.parts)
    assert int(np.count_nonzero(sacc != o_sacc.ravel())) == 0
    o_col = oracle.env_direct(inp["U"], inp["ENV"], q, nq, vg, sc.parts, sc.materials, o_iso, o_envT, o_col)
    rel = float(np.max(np.abs(col[:, :3] - o_col[:, :3]) / np.maximum(np.abs(o_col[:, :3]), 1e-3)))
    log(f"{name}: lighting maps bit-exact, colour rel err {rel:.2e}")
    assert rel < 1e-3
