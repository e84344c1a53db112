"""GPU parity of the real-asset path through the C ABI: the screen-space LOD cut and its node colours
(runtime/wgsl/lod.wgsl, sl_scene_set_lod + RenderFrame.lod), the frame drawn from the cut, and the display fill
(runtime/wgsl/fill.wgsl, sl_fill_encode), on whole subtrees of the reference's shipped hill and Cornell trees
(tests/golden/lod_cut.npz) with their own parts and materials.

Bars: lod_select's flag per node and lod_colour's node colours bit-exact against the oracle's level-by-level passes;
keys, depth order, splat words and pair lists bit-exact; the image within 1/1024 and >= 55 dB; the fill's colours
bit-exact. Thresholds 0.5, 2 (the reference's default) and 8 px, projected-area and radius rules, two cameras, the
lit and flat views, and the frame encoded after a lighting encode (its geometry on the second stream, the node
colours after the wait).
"""
import math
import sys

import numpy as np
import pytest

import lodcut
import oracle
from splatlight_amd import lighting as L
from splatlight_amd import lod as LOD
from splatlight_amd import native as N
from splatlight_amd import records as R
from splatlight_amd import renderer as RR
from splatlight_amd import scenes
from test_frame_gpu import check_integer_artefacts, psnr

pytestmark = pytest.mark.gpu
DRAW, UNDER = 2, 4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lit_inputs(sc, seed):
    rng = np.random.default_rng(seed)
    colour = np.concatenate([rng.uniform(0, 1.5, (sc.count, 3)), np.ones((sc.count, 1))], 1).astype(np.float32)
    lo_b = rng.uniform(0, 0.3, (sc.count, 12)).astype(np.float32)
    return colour, lo_b


def setup(ctx, name, size_rule=1, storage=N.SL_STORAGE_F32):
    sc, nodes, kids, _, emit = lodcut.load(name)
    rec = LOD.node_gaussians(nodes, kids, sc.gaussians)["records"]
    scene = N.Scene(ctx, sc.gaussians, sc.parts, materials=sc.materials, storage=storage)
    scene.set_lod(nodes, kids, rec, emit_part=emit, size_rule=size_rule)
    return sc, nodes, kids, emit, rec, scene


def render_lod(ctx, scene, sc, fb):
    W, H = sc.width, sc.height
    out = ctx.alloc(W * H * 16)
    try:
        N.frame_encode(ctx, scene, fb, W, H, out)
        st = N.frame_finish(ctx)
        img = ctx.read(out, W * H * 16, np.float32).reshape(H, W, 4)
        art = {"keys": N.frame_read(ctx, N.ART_KEYS), "order": N.frame_read(ctx, N.ART_ORDER),
               "splats": N.frame_read(ctx, N.ART_SPLATS).reshape(-1, 12),
               "tiles": N.frame_read(ctx, N.ART_PAIR_TILES), "slots": N.frame_read(ctx, N.ART_PAIR_SLOTS),
               "ranges": N.frame_read(ctx, N.ART_RANGES).reshape(-1, 2),
               "flags": N.frame_read(ctx, N.ART_LOD_FLAGS)}
        if int(np.frombuffer(fb, np.uint32)[79]) == R.VIEW_LIT:
            art["ncol"] = N.frame_read(ctx, N.ART_LOD_COLOUR, np.float32).reshape(-1, 4)
    finally:
        ctx.free(out)
    return st, img, art


def check_lod_frame(sc, nodes, kids, emit, rec, fb, colour, lo_b, st, img, art, size_rule=1, what=""):
    G = sc.count
    f = np.frombuffer(fb, np.float32)
    view = f[0:16]
    p = oracle.lod_params(view, float(f[180]), float(f[74]), float(f[76]), emit, 0, size_rule)
    flag = oracle.lod_select(p, nodes, kids, sc.parts, G)
    np.testing.assert_array_equal(art["flags"], flag, err_msg=f"{what}: lod flags")
    lit = int(np.frombuffer(fb, np.uint32)[79]) == R.VIEW_LIT
    gain = float(f[44])
    ncol = oracle.lod_colour(nodes, kids, G, flag, colour, lo_b, gain)
    need = (flag[G:] == DRAW) | (flag[G:] == UNDER)
    if lit:
        np.testing.assert_array_equal(art["ncol"][need], ncol[need], err_msg=f"{what}: node colours")
    o_splats, o_keys, o_vis = oracle.project_lod(fb, sc, flag, rec, ncol, colour=colour, lo_b=lo_b)
    _, _, o_order, _ = oracle.sort_keys(o_keys)
    check_integer_artefacts(sc, st, art, o_keys, o_order, o_vis, o_splats)
    ref = oracle.composite(fb, o_splats, o_order, o_vis, sc.width, sc.height)
    err = float(np.abs(img - ref).max())
    ps = psnr(img, ref)
    log(f"  {what}: {int((flag[:G] == DRAW).sum())} leaves + {int((flag[G:] == DRAW).sum())} nodes drawn, visible "
        f"{st.visible}, pairs {st.pairs}; flags, colours, keys, order, splats, pairs bit-exact; max|err| {err:.3g} "
        f"psnr {ps:.1f} dB")
    assert err <= 1 / 1024 and ps >= 55


@pytest.mark.parametrize("name", ["hill", "cornell"])
def test_lod_cut_frames(ctx, name):
    sc, nodes, kids, emit, rec, scene = setup(ctx, name)
    try:
        colour, lo_b = lit_inputs(sc, 1)
        scene.write_colour(colour)
        scene.write_bounce(lo_b)
        nn = nodes.shape[0] - sc.count
        eyes = [None, tuple(np.asarray(sc.target) + (np.asarray(sc.eye) - np.asarray(sc.target)) * 0.45)]
        for ei, eye in enumerate(eyes):
            for thr in (0.5, 2.0, 8.0):
                fb = lodcut.lod_frame(sc, thr, nn, eye=eye)
                st, img, art = render_lod(ctx, scene, sc, fb)
                check_lod_frame(sc, nodes, kids, emit, rec, fb, colour, lo_b, st, img, art,
                                what=f"{name} eye {ei} threshold {thr}")
    finally:
        scene.destroy()


def test_lod_radius_rule_and_flat_view(ctx):
    sc, nodes, kids, emit, rec, scene = setup(ctx, "hill", size_rule=0)
    try:
        colour, lo_b = lit_inputs(sc, 2)
        scene.write_colour(colour)
        scene.write_bounce(lo_b)
        nn = nodes.shape[0] - sc.count
        fb = lodcut.lod_frame(sc, 2.0, nn)
        st, img, art = render_lod(ctx, scene, sc, fb)
        check_lod_frame(sc, nodes, kids, emit, rec, fb, colour, lo_b, st, img, art, size_rule=0, what="radius rule")
        fb = lodcut.lod_frame(sc, 2.0, nn, view_mode=0)
        st, img, art = render_lod(ctx, scene, sc, fb)
        check_lod_frame(sc, nodes, kids, emit, rec, fb, colour, lo_b, st, img, art, size_rule=0, what="flat view")
    finally:
        scene.destroy()


def test_lod_frame_after_lighting(ctx):
    """Encoded after a lighting encode, the frame's geometry (and the cut) runs on the second stream and the node
    colours after the wait for the lighting: the same artefacts."""
    other = scenes.hill(n=3000, seed=4, width=160, height=90)
    h = N.Scene(ctx, other.gaussians, other.parts, emitters=other.emitters, materials=other.materials,
                nodes=scenes.leaf_nodes(other.gaussians), vgrid=scenes.vgrid_open(other.count))
    sc, nodes, kids, emit, rec, scene = setup(ctx, "cornell")
    try:
        colour, lo_b = lit_inputs(sc, 3)
        scene.write_colour(colour)
        scene.write_bounce(lo_b)
        inp = L.lighting_inputs(other, dirty_all=True, sun_res=128)
        h.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        fb = lodcut.lod_frame(sc, 2.0, nodes.shape[0] - sc.count)
        st, img, art = render_lod(ctx, scene, sc, fb)
        check_lod_frame(sc, nodes, kids, emit, rec, fb, colour, lo_b, st, img, art, what="after a lighting encode")
    finally:
        scene.destroy()
        h.destroy()


def test_lod_reduced_storage(ctx):
    sc, nodes, kids, emit, rec, scene = setup(ctx, "hill", storage=N.SL_STORAGE_F16)
    try:
        colour, lo_b = lit_inputs(sc, 5)
        scene.write_colour(colour)
        scene.write_bounce(lo_b)
        fb = lodcut.lod_frame(sc, 2.0, nodes.shape[0] - sc.count)
        st, img, art = render_lod(ctx, scene, sc, fb)
        deq = scenes.Scene(sc.name, R.dequantized(sc.gaussians, R.STORAGE_F16), sc.parts, sc.materials, sc.emitters,
                           sc.eye, sc.target, sc.fov_deg, sc.width, sc.height, sc.far)
        check_lod_frame(deq, nodes, kids, emit, R.dequantized(rec, R.STORAGE_F16), fb, colour, lo_b, st, img, art,
                        what="f16 storage")
    finally:
        scene.destroy()


def test_display_fill(ctx):
    sc, nodes, kids, emit, rec, scene = setup(ctx, "cornell")
    try:
        fill = lodcut.fill("cornell")
        colour, lo_b = lit_inputs(sc, 6)
        scene.write_colour(colour)
        scene.write_bounce(lo_b)
        scene.set_fill(fill)
        scene.fill()
        got = scene.read(N.SCENE_COLOUR).reshape(-1, 4)
        o_col, o_lb = oracle.display_fill(fill, sc.count, colour, lo_b)
        np.testing.assert_array_equal(got, o_col)
        # the bounce toward the eye through the frame it feeds (gain 1, every leaf drawn)
        fb = sc.frame_bytes(bounce_gain=1.0)
        W, H = sc.width, sc.height
        out = ctx.alloc(W * H * 16)
        try:
            N.frame_encode(ctx, scene, fb, W, H, out)
            st = N.frame_finish(ctx)
            splats = N.frame_read(ctx, N.ART_SPLATS).reshape(-1, 12)
        finally:
            ctx.free(out)
        o_splats, o_keys, _ = oracle.project(fb, sc, colour=o_col, lo_b=o_lb)
        ok = o_keys != R.INVALID_KEY
        np.testing.assert_array_equal(splats[ok], o_splats[ok])
        log(f"  fill: {int((np.diff(fill[:sc.count + 1]) > 0).sum())} targets; colours and the frame's splat "
            f"colours (colour + bounce) bit-exact")
    finally:
        scene.destroy()


def test_lod_renderer_surface(ctx):
    """SplatRenderer(device, format, shaders, asset, lodNodes) as src/renderer.js:38 takes it."""
    sc, nodes, kids, _, emit = lodcut.load("hill")
    lod_nodes = LOD.node_gaussians(nodes, kids, sc.gaussians)
    r = RR.SplatRenderer(ctx, "rgba16float", None, {"gaussians": sc.gaussians, "parts": sc.parts,
                                                     "materials": sc.materials, "nodes": nodes, "kids": kids,
                                                     "emitPart": emit}, lod_nodes)
    try:
        colour, lo_b = lit_inputs(sc, 7)
        r.bindColour(colour, lo_b)
        r.resize(sc.width, sc.height)
        assert r.nodeBase == sc.count + 1 and r.lodNodeCount == nodes.shape[0] - sc.count
        st = r.encode(lodcut.lod_frame(sc, 2.0, r.lodNodeCount), 3)
        assert st.visible > 0
    finally:
        r.destroy()


def test_lod_refusals(ctx):
    sc, nodes, kids, emit, rec, scene = setup(ctx, "cornell")
    plain = N.Scene(ctx, sc.gaussians, sc.parts, materials=sc.materials)
    W, H = sc.width, sc.height
    out = ctx.alloc(W * H * 8)
    nn = nodes.shape[0] - sc.count
    try:
        fb = lodcut.lod_frame(sc, 2.0, nn)
        with pytest.raises(N.SplatlightError) as e:
            N.frame_encode(ctx, plain, fb, W, H, out, N.SL_OUT_RGBA16F)
        assert e.value.status == 5 and "sl_scene_set_lod" in str(e.value)
        with pytest.raises(N.SplatlightError) as e:
            N.frame_encode(ctx, scene, lodcut.lod_frame(sc, 2.0, nn - 1), W, H, out, N.SL_OUT_RGBA16F)
        assert e.value.status == 1
        with pytest.raises(N.SplatlightError) as e:
            N.batch_encode(ctx, [(scene, fb, W, H, out, N.SL_OUT_RGBA16F)] * 2)
        assert e.value.status == 5
        with pytest.raises(N.SplatlightError) as e:
            N.frame_encode_noalloc(ctx, scene, fb, W, H, out, N.SL_OUT_RGBA16F)
        assert e.value.status == 5
        with pytest.raises(N.SplatlightError) as e:
            scene.set_lod(nodes, kids, rec, emit_part=emit)
        assert e.value.status == 1
        bad = nodes.copy()
        bad["depth"][kids[0]] += 1
        with pytest.raises(N.SplatlightError) as e:
            plain.set_lod(bad, kids, rec)
        assert e.value.status == 2
        f = lodcut.fill("cornell").copy()
        G = sc.count
        f[G + 1] = np.nonzero(np.diff(f[:G + 1]) > 0)[0][0]     # a source that is itself a target
        with pytest.raises(N.SplatlightError) as e:
            plain.set_fill(f)
        assert e.value.status == 2
        N.frame_encode(ctx, scene, fb, W, H, out, N.SL_OUT_RGBA16F)   # still usable
        N.frame_finish(ctx)
    finally:
        ctx.free(out)
        plain.destroy()
        scene.destroy()
