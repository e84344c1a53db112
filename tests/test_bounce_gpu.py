"""GPU parity of the sky and bounce through probes against the CPU oracle, through the C ABI.

Each pass is checked on the device's own inputs for that pass (so a failure names the pass), then the chain end
to end from the lighting's outputs. Bars: node SH2 halves bit-exact (the same IEEE operations on the same sums,
rounded to half the same way); node sums, probe gathers, probe SH2, transfers and the bounce within float32
rounding of the oracle (exp, erfc and the V-grid lookup's asin / atan2 differ by an ulp or two between the
device and host libraries: relative 1e-4; the transfer's SH2, whose higher bands cancel, 1e-5 of its largest
coefficient); a probe ray's first-hit part and flags may differ only where a cull
decision sits within rounding of its threshold (at most 0.2 % of rays).
"""
import sys

import numpy as np
import pytest

import oracle
from splatlight_amd import bounce as BN
from splatlight_amd import lighting as L
from splatlight_amd import native as N
from splatlight_amd import scenes, tree

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel(a, b, floor):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


def mixed_arity(t, G):
    """The scene tree with nodes of three or more kids (every 7th internal node takes its first internal child's
    kids in that child's place; the child is left unreferenced, out of every part) and single-kid nodes (every 5th
    wraps its first leaf child in a new node holding that leaf's record): the node passes' general path. Depths are
    recounted from the scene root, which stays the last node."""
    N, root = t.count, t.scene_root
    lo, hi = t.nodes["child_lo"], t.nodes["child_hi"]
    kids = {i: [int(k) for k in t.kids[lo[i]:hi[i]]] for i in range(G, N)}
    root_kids = kids.pop(root)
    orphan, wraps = set(), []
    for n, p in enumerate(i for i in range(G, N) if i != root):
        if p in orphan:
            continue
        ks = kids[p]
        if n % 7 == 0 and ks and ks[0] >= G and ks[0] not in orphan:
            orphan.add(ks[0])
            kids[p] = kids[ks[0]] + ks[1:]
        elif n % 5 == 0:
            for j, k in enumerate(ks):
                if k < G:
                    y = N - 1 + len(wraps)                # before the scene root, which moves to the end
                    wraps.append(k)
                    kids[y] = [k]
                    ks[j] = y
                    break
    W = len(wraps)
    new_root = N - 1 + W
    nodes = np.concatenate([t.nodes[:N - 1], t.nodes[np.array(wraps, np.int64)], t.nodes[N - 1:]]) if W else t.nodes.copy()
    kids[new_root] = root_kids
    for q in orphan:
        nodes["part"][q] = tree.NO_PART
    flat = []
    for i in range(nodes.shape[0]):
        ks = kids.get(i, []) if i >= G else []
        nodes["child_lo"][i], nodes["child_hi"][i] = len(flat), len(flat) + len(ks)
        flat += ks
    depth = np.zeros(nodes.shape[0], np.int64)
    stack = [(new_root, 0)]
    while stack:
        i, d = stack.pop()
        depth[i] = d
        stack += [(k, d + 1) for k in kids.get(i, []) if i >= G]
    nodes["depth"] = depth.astype(nodes["depth"].dtype)
    roots = list(t.roots)                              # part roots keep their ids
    log(f"    tree reshaped: {len(orphan)} nodes spliced into their parents, {W} single-kid wrappers, "
        f"kids per node up to {max(len(v) for k, v in kids.items() if k != new_root)}")
    return tree.Tree(nodes, np.array(flat, np.uint32), roots, new_root, int(depth[:G].max()))


def setup(ctx, sc, shape=None, leaves=None, vg=None):
    """The scene, its tree (over the given leaf Nodes, else the derived ones) and V grids (given, else V = 1)."""
    if leaves is None:
        t = tree.scene_tree(sc)
    else:
        t = tree.build_tree(sc.gaussians, sc.parts, leaves)
        for p, r in enumerate(t.roots):
            if r >= 0:
                sc.parts["root"][p] = r
    if shape is not None:
        t = shape(t, sc.count)
    vg = scenes.vgrid_open(sc.count) if vg is None else vg
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters if sc.emitters.shape[0] else None,
                    materials=sc.materials, nodes=t.nodes[:sc.count], vgrid=vg)
    b = BN.Bounce(scene, sc, t)
    return scene, t, vg, b


def read_state(scene, n_probes, G, N_nodes):
    return {
        "c0": scene.bounce_read(N.SL_BOUNCE_C0).reshape(G, 4),
        "lo_dir": scene.bounce_read(N.SL_BOUNCE_LO_DIR).reshape(G, 4),
        "lo_b": scene.bounce_read(N.SL_BOUNCE_LO_B).reshape(G, 3, 4),
        "node_acc": scene.bounce_read(N.SL_BOUNCE_NODE_ACC).reshape(N_nodes, 14, 4),
        "node_sh": scene.bounce_read(N.SL_BOUNCE_NODE_SH, np.uint32).reshape(N_nodes, 14),
        "probe_data": scene.bounce_read(N.SL_BOUNCE_PROBE_DATA).reshape(n_probes, 128, 2, 4),
        "probe_sh": scene.bounce_read(N.SL_BOUNCE_PROBE_SH).reshape(n_probes, 9, 4),
        "xfer": scene.bounce_read(N.SL_BOUNCE_XFER_SH).reshape(G, 3, 4),
    }


def check_frames(ctx, sc, frames=2, moved_mask=0, shape=None, leaves=None, vg=None):
    scene, t, vg, b = setup(ctx, sc, shape, leaves, vg)
    G, Nn, NP = sc.count, t.count, b.n_probes
    cells = oracle.cells_table(*BN.cell_table())
    up, levels = BN.up_levels(t.nodes, G, t.max_depth)
    prev = {"lo_dir": np.zeros((G, 4), np.float32), "lo_b": np.zeros((G, 3, 4), np.float32),
            "probe_data": np.zeros((NP, 128, 2, 4), np.float32), "xfer": np.zeros((G, 3, 4), np.float32)}
    o_st = oracle.bounce_state(G, Nn, NP)
    try:
        for f in range(frames):
            inp = L.lighting_inputs(sc, frame=f)
            scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
            run_xfer = b.force_near or moved_mask != 0
            PP = b.encode(moved_mask=moved_mask)
            ctx.sync()
            s = read_state(scene, NP, G, Nn)
            U, ENV = inp["U"], inp["ENV"]
            iso = scene.read(N.SCENE_ISO_NORM)
            env_T = scene.read(N.SCENE_ENV_T).reshape(G, 4) if ENV is not None else None
            log(f"  frame {f}: G {G}, nodes {Nn}, probes {NP}")
            # c0: the shading's emitter light before a Gaussian's own emission (the lit colour before env_direct
            # with the materials' emissive bit cleared)
            no_emit = sc.materials.copy()
            no_emit["flags"] &= ~np.uint32(1)
            o_col = oracle.shade_emit(U, inp["EP"], inp["emit_nodes"], sc.gaussians, t.nodes[:G], vg, sc.parts,
                                      scene.read(N.SCENE_LAMP_T), no_emit, iso)
            e = rel(s["c0"][:, :3], o_col[:, :3], 1e-3)
            log(f"    c0 rel {e:.2e}")
            assert e < 1e-3
            # lo_dir from the device's c0 and env T (a quarter refreshed, the rest kept)
            o = oracle.bounce_lo_dir(U, ENV, sc.gaussians, t.nodes, vg, sc.parts, sc.materials, iso, env_T, s["c0"],
                                     prev["lo_dir"])
            e = rel(s["lo_dir"], o, 1e-4 * max(float(np.abs(o).max()), 1e-12))
            log(f"    lo_dir rel {e:.2e} (refreshed {int(np.count_nonzero(np.any(s['lo_dir'] != prev['lo_dir'], 1)))})")
            assert e < 1e-4
            # node sums from the device's lo_dir and last frame's lo_b
            o = oracle.node_leaf(U, PP, sc.gaussians, t.nodes, vg, sc.parts, s["lo_dir"], prev["lo_b"],
                                 np.zeros((Nn, 14, 4), np.float32))
            w = scene.bounce_read(N.SL_BOUNCE_LEAF_W)
            if w.size:
                # one refit launch: the leaves' cached weights against node_leaf's, then the internal sums and every
                # node's SH2 bit for bit from the device's leaf weights and light
                w = w.reshape(G, 14)
                e = rel(w, o[:G, :, 3], 1e-4 * max(float(o[:G, :, 3].max()), 1e-12))
                log(f"    leaf weights rel {e:.2e}")
                assert e < 1e-4
                Lg = s["lo_dir"][:, :3] + prev["lo_b"][:, 1, :3]
                leaf = np.zeros((Nn, 14, 4), np.float32)
                leaf[:G, :, :3] = Lg[:, None, :] * w[:, :, None]
                leaf[:G, :, 3] = w
                o = oracle.node_up(U, PP, t.nodes, t.kids, up, levels, t.max_depth, leaf)
                mism = int(np.count_nonzero(s["node_acc"][G:] != o[G:]))
                log(f"    internal node sums: {mism} of {o[G:].size} differ")
                assert mism == 0
                acc_for_sh = o
            else:
                o = oracle.node_up(U, PP, t.nodes, t.kids, up, levels, t.max_depth, o)
                for c in range(4):
                    e = rel(s["node_acc"][..., c], o[..., c], 1e-4 * max(float(np.abs(o[..., c]).max()), 1e-12))
                    log(f"    node_acc[{c}] rel {e:.2e}")
                    assert e < 1e-4
                acc_for_sh = s["node_acc"]
            o = oracle.node_sh2(PP, acc_for_sh, Nn)
            mism = int(np.count_nonzero(o != s["node_sh"]))
            log(f"    node_sh: {mism} of {o.size} words differ")
            assert mism == 0
            # the gather from the device's node SH2 and last frame's records, then validity
            o = oracle.probe_gather(U, PP, ENV, t.nodes, vg, sc.parts, t.kids, s["node_sh"], cells, prev["probe_data"])
            o = oracle.probe_valid(U, PP, sc.parts, o).reshape(NP, 128, 2, 4)
            d = s["probe_data"]
            fl = max(float(np.abs(o[:, :, 0, :3]).max()), 1e-6)
            bad_rgb = np.any(np.abs(d[:, :, 0, :3] - o[:, :, 0, :3]) > 1e-4 * np.maximum(np.abs(o[:, :, 0, :3]), 1e-3 * fl), -1)
            bad_m = np.any(np.abs(d[:, :, 1, :2] - o[:, :, 1, :2]) > 1e-4 * np.maximum(np.abs(o[:, :, 1, :2]), 1e-3), -1)
            bad_f = (d[:, :, 0, 3] != o[:, :, 0, 3]) | (d[:, :, 1, 3] != o[:, :, 1, 3])
            bad = bad_rgb | bad_m | bad_f
            log(f"    probe rays: {int(bad.sum())} of {bad.size} differ (rgb {int(bad_rgb.sum())}, moments "
                f"{int(bad_m.sum())}, flags {int(bad_f.sum())}); valid {float(o[:, 0, 1, 2].mean()):.3f}, "
                f"mean rgb {o[:, :, 0, :3].mean((0, 1))}")
            assert bad.sum() <= 0.002 * bad.size
            assert np.array_equal(d[:, 0, 1, 2], o[:, 0, 1, 2]) or bad.sum() > 0
            o = oracle.probe_sh2(PP, s["probe_data"], cells, NP)
            e = rel(s["probe_sh"], o, 1e-4 * max(float(np.abs(o).max()), 1e-12))
            log(f"    probe_sh rel {e:.2e}")
            assert e < 1e-4
            if run_xfer:
                # the higher bands are sums of terms of both signs (near zero for a near-isotropic Gaussian): their
                # error is measured against the largest coefficient, the scale of the terms
                o = oracle.iso_xfer(U, PP, sc.gaussians, t.nodes, vg, sc.parts, iso, cells, prev["xfer"])
                e = float(np.abs(s["xfer"] - o).max()) / max(float(np.abs(o).max()), 1e-12)
                log(f"    xfer abs err / max {e:.2e}")
                assert e < 1e-5
            o = oracle.bounce_sh(U, PP, sc.gaussians, t.nodes, sc.parts, cells, s["probe_data"], s["probe_sh"],
                                 s["xfer"], prev["lo_b"])
            e = rel(s["lo_b"], o, 1e-4 * max(float(np.abs(o).max()), 1e-12))
            log(f"    lo_b rel {e:.2e} (mean eye {o[:, 0, :3].mean(0)})")
            assert e < 1e-4
            # end to end from the lighting's outputs (the device's c0, env T and iso_norm)
            o_st = oracle.bounce_frame(o_st, U, PP, ENV, sc.gaussians, t.nodes, vg, sc.parts, sc.materials, iso,
                                       env_T, s["c0"], t.kids, up, levels, t.max_depth, cells, NP, run_xfer=run_xfer)
            e = rel(s["lo_b"], o_st["lo_b"], 1e-3 * max(float(np.abs(o_st["lo_b"]).max()), 1e-12))
            log(f"    end to end lo_b rel {e:.2e}")
            assert e < 1e-3
            prev = {k: s[k].copy() for k in ("lo_dir", "lo_b", "probe_data", "xfer")}
        return scene, b
    except Exception:
        scene.destroy()
        raise


def test_bounce_hill_terrain(ctx):
    """Terrain columns over the hill, probe volumes on the trunk and canopy, sky and sun."""
    log("bounce: hill 20k")
    scene, _ = check_frames(ctx, scenes.hill(n=20_000, shadow_res=512))
    scene.destroy()


def test_bounce_cornell_room(ctx):
    """The room lattice in the Cornell box: emitter light only (no environment)."""
    log("bounce: cornell 20k")
    scene, _ = check_frames(ctx, scenes.cornell(n=20_000))
    scene.destroy()


def test_bounce_tree_of_mixed_arity(ctx):
    """A tree whose nodes have one, two, three or more kids (the format allows any count): node sums and SH2 bit
    for bit, and every pass to its bar, as on the binary trees."""
    log("bounce: hill 12k, tree with single-kid and many-kid nodes")
    scene, _ = check_frames(ctx, scenes.hill(n=12_000, shadow_res=256), shape=mixed_arity)
    scene.destroy()


def test_bounce_movable_part(ctx):
    """A movable part that moved this frame (Part.flags bit 0 clear, bit 1 set, its previous pose shifted): its 8
    part probes gather without it and serve its Gaussians alone, and every frame rebuilds the transfer of the
    Gaussians on it or within its bound (plus near_r) of either pose."""
    log("bounce: cornell 20k, the tall box movable and moved")
    sc = scenes.cornell(n=20_000)
    sc.parts["flags"][2] = 2
    sc.parts["xf_prev"][2][0][3] = 0.15
    scene, b = check_frames(ctx, sc, moved_mask=1 << 2)
    scene.destroy()


def test_bounce_light_still_and_frame(ctx):
    """lightStill re-reads the settled probes (the same bounce); the frame's colours add gain x bounce."""
    log("bounce: light_still + frame, hill 8k")
    sc = scenes.hill(n=8_000, width=320, height=240, shadow_res=256)
    scene, t, vg, b = setup(ctx, sc)
    try:
        inp = L.lighting_inputs(sc, frame=0)
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        b.encode()
        ctx.sync()
        lob = scene.bounce_read(N.SL_BOUNCE_LO_B)
        b.encode(light_still=True)
        ctx.sync()
        assert np.array_equal(scene.bounce_read(N.SL_BOUNCE_LO_B), lob)
        W, H = sc.width, sc.height
        out = ctx.alloc(W * H * 16)
        imgs = []
        for gain in (0.0, 1.0):
            N.frame_encode(ctx, scene, sc.frame_bytes(bounce_gain=gain), W, H, out)
            N.frame_finish(ctx)
            imgs.append(ctx.read(out, W * H * 16, np.float32).reshape(H, W, 4))
        ctx.free(out)
        diff = float(np.abs(imgs[1][..., :3] - imgs[0][..., :3]).max())
        log(f"    image change with bounce gain 1: max {diff:.3g}")
        assert diff > 1e-3 and np.all(imgs[1][..., :3] >= imgs[0][..., :3] - 1e-6)
    finally:
        scene.destroy()


def test_bounce_sky_closed_form(ctx):
    """An open sky over faint isotropic Gaussians: every probe sees the sky, so each Gaussian's bounce is its albedo
    times the sky's mean radiance A (the M39 transfer of an unoccluded isotropic Gaussian under a sky A + B z)."""
    from test_bounce_oracle import sky_scene
    sc, A = sky_scene(400)
    scene, t, vg, b = setup(ctx, sc)
    try:
        inp = L.lighting_inputs(sc, frame=0)
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        b.encode()
        ctx.sync()
        lob = scene.bounce_read(N.SL_BOUNCE_LO_B).reshape(sc.count, 3, 4)
        want = sc.gaussians["albedo"] * A
        e = rel(lob[1:, 0, :3], want[1:], 1e-6)
        log(f"    sky closed form: rel {e:.2e}")
        assert e < 1e-3
    finally:
        scene.destroy()
