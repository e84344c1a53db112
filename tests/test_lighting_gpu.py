"""GPU parity of the relighting against the CPU oracle, through the C ABI.

Bars: the opacity maps' fixed-point sums bit-exact (integer atomics of correctly rounded per-texel values);
iso_norm, transmittances and lit colours within float32 rounding of the oracle (transcendentals differ by an ulp
or two between the device and host maths libraries); the lit frame within 1/1024 and >= 55 dB of the oracle's.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from splatlight_amd import lighting as L
from splatlight_amd import native as N
from splatlight_amd import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make(ctx, sc, nd=None, vg=None):
    """The scene on the device with leaf Nodes and V grids: given (baked ones), else derived (S_hat = I, V = 1)."""
    nd = scenes.leaf_nodes(sc.gaussians) if nd is None else nd
    vg = scenes.vgrid_open(sc.count) if vg is None else vg
    em = sc.emitters if sc.emitters.shape[0] else None
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=em, materials=sc.materials, nodes=nd, vgrid=vg)
    return scene, nd, vg


def rel_err(a, b, floor):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def check_lighting(ctx, sc, sun_res=None, nd=None, vg=None, emit_split=None):
    scene, nd, vg = make(ctx, sc, nd, vg)
    inp = L.lighting_inputs(sc, sun_res=sun_res, emit_split=emit_split)
    try:
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        iso = scene.read(N.SCENE_ISO_NORM)
        o_iso = oracle.iso_init(sc.gaussians, nd, vg)
        e_iso = rel_err(iso, o_iso, 1e-30)
        log(f"  iso_norm rel err {e_iso:.2e}")
        assert e_iso < 2e-5
        acc = scene.read(N.SCENE_LSM_ACC, np.uint32)
        o_acc, o_mip, o_T = oracle.lsm(inp["U"], inp["LS"], sc.gaussians, nd, sc.parts)
        mism = int(np.count_nonzero(acc != o_acc.ravel()))
        log(f"  emitter map: {o_acc.size} words, {mism} differ; texels hit {int(np.count_nonzero(o_acc[..., 0]))}")
        assert mism == 0
        mip = scene.read(N.SCENE_LSM_MIP).reshape(o_mip.shape)
        log(f"  emitter mips: {int(np.count_nonzero(mip != o_mip))} of {mip.size} differ, max {np.abs(mip - o_mip).max():.2e}")
        assert np.abs(mip - o_mip).max() < 2e-5
        T = scene.read(N.SCENE_LAMP_T)
        log(f"  lamp T: max|err| {np.abs(T - o_T).max():.2e}, shadowed {(o_T < 0.5).mean():.3f}")
        assert np.abs(T - o_T).max() < 1e-4
        o_col = oracle.shade_emit(inp["U"], inp["EP"], inp["emit_nodes"], sc.gaussians, nd, vg, sc.parts, o_T,
                                  sc.materials, o_iso)
        if inp["SM"] is not None:
            s_acc = scene.read(N.SCENE_SUN_ACC, np.uint32)
            o_sacc, o_smip, o_envT = oracle.sun(inp["U"], inp["SM"], inp["desc"], sc.gaussians, nd, sc.parts)
            smism = int(np.count_nonzero(s_acc != o_sacc.ravel()))
            log(f"  sun maps: {o_sacc.size} words, {smism} differ; texels hit {int(np.count_nonzero(o_sacc[..., 0]))}")
            assert smism == 0
            smip = scene.read(N.SCENE_SUN_MIP).reshape(o_smip.shape)
            log(f"  sun mips: {int(np.count_nonzero(smip != o_smip))} of {smip.size} differ, max {np.abs(smip - o_smip).max():.2e}")
            assert np.abs(smip - o_smip).max() < 2e-5
            envT = scene.read(N.SCENE_ENV_T).reshape(-1, 4)
            log(f"  env T: max|err| {np.abs(envT - o_envT).max():.2e}, shadowed {(o_envT[:, 0] < 0.5).mean():.3f}")
            assert np.abs(envT - o_envT).max() < 1e-4
            o_col = oracle.env_direct(inp["U"], inp["ENV"], sc.gaussians, nd, vg, sc.parts, sc.materials, o_iso,
                                      o_envT, o_col)
        col = scene.read(N.SCENE_COLOUR).reshape(-1, 4)
        e_col = rel_err(col[:, :3], o_col[:, :3], 1e-3)
        log(f"  colour rel err {e_col:.2e} (max colour {o_col[:, :3].max():.3g})")
        assert e_col < 1e-3
        return scene, inp, o_col
    except Exception:
        scene.destroy()
        raise


def test_cornell_lighting(ctx):
    log("lighting parity: cornell n=20000")
    sc = scenes.cornell(n=20_000, seed=3, width=320, height=240)
    scene, _, _ = check_lighting(ctx, sc)
    scene.destroy()


def test_hill_lighting_with_sun(ctx):
    log("lighting parity: hill n=30000, sun map 256^2")
    sc = scenes.hill(n=30_000, seed=2, width=320, height=180)
    scene, _, _ = check_lighting(ctx, sc, sun_res=256)
    scene.destroy()


def test_map_culling_stress(ctx):
    """Needle-thin Gaussians at every angle and Gaussians whose 3-sigma box is clamped to the map's maximum half-size:
    the cases where the map sums' ellipse masks (which tiles a record may touch, which of a tile's 64-texel units),
    their per-record lane layout (tall boxes take 4x16 columns, wide ones 16x4 bands, squarish ones 8x8 quads) and
    the box clamp matter most. The maps must still equal the oracle's bit for bit."""
    log("lighting parity: hill n=24000 with needles and clamped giants, sun map 512^2")
    sc = scenes.hill(n=24_000, seed=11, width=320, height=180)
    g = sc.gaussians
    rng = np.random.default_rng(5)
    idx = rng.permutation(np.flatnonzero(g["part_surface"] == 0))      # the ground part
    needles, giants = idx[: len(idx) // 3], idx[len(idx) // 3: 2 * len(idx) // 3]
    g["scale"][needles] = np.sort(np.column_stack([np.full(len(needles), 0.002), np.full(len(needles), 0.004),
                                                   rng.uniform(0.2, 0.6, len(needles))]), axis=1)
    g["scale"][giants] = np.sort(rng.uniform(0.3, 1.5, (len(giants), 3)), axis=1)
    scene, _, _ = check_lighting(ctx, sc, sun_res=512)
    scene.destroy()


def test_lit_frame_against_oracle(ctx):
    """Relight then draw on the device; the oracle relights and draws on the host."""
    sc = scenes.cornell(n=20_000, seed=5, width=320, height=240)
    scene, inp, o_col = check_lighting(ctx, sc)
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
    log(f"  lit frame: max|err| {err:.3g} psnr {psnr:.1f} dB, mean rgb {ref[..., :3].mean():.3f}")
    assert err <= 1 / 1024 and psnr >= 55


def test_relit_frame_repeats_bit_exact(ctx):
    """Two fully relit frames of one scene in one process: the sun-map chain on the context's second stream, the
    frame's geometry on the third beside the lighting. Maps, transmittances, colours and the image are the same bit
    for bit both times (integer map sums are order-independent; nothing depends on which stream finishes first)."""
    log("relit frame twice: hill n=120000, sun map 1024^2, 640x360 rgba16f")
    sc = scenes.hill(n=120_000, seed=4, width=640, height=360)
    scene, nd, vg = make(ctx, sc)
    W, H = sc.width, sc.height
    out = ctx.alloc(W * H * 8)
    runs = []
    try:
        fb = sc.frame_bytes()
        for _ in range(2):
            inp = L.lighting_inputs(sc, sun_res=1024)
            scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
            N.frame_encode(ctx, scene, fb, W, H, out, N.SL_OUT_RGBA16F)
            st = N.frame_finish(ctx)
            runs.append({
                "lsm": scene.read(N.SCENE_LSM_ACC, np.uint32), "sun": scene.read(N.SCENE_SUN_ACC, np.uint32),
                "lamp_T": scene.read(N.SCENE_LAMP_T), "env_T": scene.read(N.SCENE_ENV_T),
                "colour": scene.read(N.SCENE_COLOUR), "image": ctx.read(out, W * H * 8, np.uint16),
                "pairs": st.pairs, "visible": st.visible})
    finally:
        ctx.free(out)
        scene.destroy()
    a, b = runs
    log(f"  visible {a['visible']}, pairs {a['pairs']}, sun texels hit {int(np.count_nonzero(a['sun']))}")
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k
    assert np.count_nonzero(a["sun"]) > 0 and np.count_nonzero(a["image"]) > 0


def test_maps_kept_when_nothing_moved(ctx):
    sc = scenes.cornell(n=5000, seed=1)
    scene, nd, vg = make(ctx, sc)
    try:
        a = L.lighting_inputs(sc, dirty_all=True)
        scene.light(a["U"], a["EP"], a["emit_nodes"], a["LS"])
        acc1 = scene.read(N.SCENE_LSM_ACC, np.uint32)
        col1 = scene.read(N.SCENE_COLOUR)
        b = L.lighting_inputs(sc, dirty_all=False)          # nothing dirty: the maps hold, lookups rerun
        scene.light(b["U"], b["EP"], b["emit_nodes"], b["LS"])
        np.testing.assert_array_equal(scene.read(N.SCENE_LSM_ACC, np.uint32), acc1)
        np.testing.assert_array_equal(scene.read(N.SCENE_COLOUR), col1)
    finally:
        scene.destroy()


def test_lighting_refuses_unsupported(ctx):
    sc = scenes.cornell(n=1000, seed=1)
    scene, nd, vg = make(ctx, sc)
    try:
        a = L.lighting_inputs(sc)
        U = bytearray(a["U"])
        np.frombuffer(U, np.uint32)[15] |= 16                   # reflection split
        with pytest.raises(N.SplatlightError) as e:
            scene.light(bytes(U), a["EP"], a["emit_nodes"], a["LS"])
        assert e.value.status == 5
        scene.light(a["U"], a["EP"], a["emit_nodes"], a["LS"])  # still usable
    finally:
        scene.destroy()
