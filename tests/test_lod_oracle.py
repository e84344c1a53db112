"""CPU pins of the real-asset path's host code and oracle: the asset loader (src/asset.js), the LOD node records
(src/lod.js nodeGaussians), lod_select / lod_colour (runtime/wgsl/lod.wgsl) and display_fill (runtime/wgsl/fill.wgsl),
on whole subtrees of the reference's shipped hill and Cornell trees (tests/golden/lod_cut.npz).

* loader: the reference's layout checks and messages, the buffer size check, and the fingerprint against a
  word-by-word restatement of its 32-bit arithmetic;
* node records: R diag(s^2) R^T equals the inverse of the baked C^-1, scales ascending, unit quaternions with a
  proper rotation, the albedo is the mass-weighted mean over the node's leaves, the material the first leaf's;
* the cut: every leaf of a part covered exactly once (drawn, or under exactly one drawn node whose ancestors up to
  the part root all open); threshold 0 draws every leaf and a huge one the part roots; the emitters' part opens to
  its leaves; with the radius rule the set of open nodes shrinks as the threshold grows; a node's size by the
  projected-area rule against J V R C R^T V^T J^T in float64;
* node colour: the mass-weighted mean of the children's shown colours, recomputed in float64 over the tree;
* fill: a sparse matrix product in float64.
"""
import json
import math
import os

import numpy as np
import pytest

import lodcut
import oracle
from splatlight_amd import asset as A
from splatlight_amd import lod as LOD
from splatlight_amd import records as R

NONE, OPEN, DRAW, UNDER = 0, 1, 2, 4


# ---- loader ---------------------------------------------------------------------------------------------------------

def write_asset(tmp, name="hill"):
    """The cut written as a gs-runtime directory with the shipped header's buffer descriptions."""
    r = lodcut.raw(name)
    h = json.loads(r["header"].tobytes().decode())
    bufs = {"gaussians": r["gaussians"], "parts": r["parts"], "materials": r["materials"], "nodes": r["nodes"],
            "kids": r["kids"], "vgrid": np.zeros((r["gaussians"].shape[0], 32), np.uint32),
            "lamps": np.zeros((1, 24), np.uint32)}
    for k, a in bufs.items():
        a = np.ascontiguousarray(a)
        h["buffers"][k]["count"] = a.nbytes // h["buffers"][k]["stride"]
        (tmp / h["buffers"][k]["file"]).write_bytes(a.tobytes())
    parts = []
    for k in range(r["parts"].shape[0]):
        parts.append({"index": k, "name": f"p{k}", "root": int(r["parts"][k, 28]),
                      "gaussians": [int(r["parts"][k, 29]), int(r["parts"][k, 30])]})
    h["parts"] = parts
    (tmp / "asset.json").write_text(json.dumps(h))
    return h


def test_load_asset_cut(tmp_path):
    h = write_asset(tmp_path)
    a = A.load_asset(str(tmp_path))
    r = lodcut.raw("hill")
    assert a["count"] == r["gaussians"].shape[0] and a["partCount"] == 3
    assert np.array_equal(a["gaussians"].view(np.uint32).reshape(-1, 16), r["gaussians"])
    assert a["nodes"].shape[0] == r["nodes"].shape[0] and np.array_equal(a["kids"], r["kids"])
    assert a["fingerprint"]["parts"] == [p["gaussians"] for p in h["parts"]]


def imul_fingerprint(words):
    """The fingerprint's loop as src/asset.js:84-86 writes it, with Math.imul's 32-bit wrap spelled out."""
    def imul(a, b):
        return ((a & 0xFFFFFFFF) * (b & 0xFFFFFFFF)) & 0xFFFFFFFF
    h = 0
    for i, w in enumerate(words):
        h = (h + imul(int(w), (imul(i, 2654435761) + 1) & 0xFFFFFFFF)) & 0xFFFFFFFF
    return h


def test_fingerprint_arithmetic():
    rng = np.random.default_rng(0)
    for n in (1, 7, 1000):
        w = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        hdr = {"buffers": {"gaussians": {"count": n}}, "parts": []}
        assert A.fingerprint(hdr, w.tobytes())["hash"] == imul_fingerprint(w)
    w = np.full(4096, 0xFFFFFFFF, np.uint32)           # every product and partial sum wraps
    assert A.fingerprint({"buffers": {"gaussians": {"count": 1}}, "parts": []}, w.tobytes())["hash"] == \
        imul_fingerprint(w)


@pytest.mark.parametrize("edit,msg", [
    (lambda h: h.update(format="x"), "Unknown asset format x"),
    (lambda h: h.update(version="1.0"), "Unsupported asset version 1.0"),
    (lambda h: h["buffers"].pop("kids"), "Asset has no kids buffer"),
    (lambda h: h["buffers"]["parts"].update(stride=96), "parts stride 96, renderer expects 128"),
    (lambda h: h["buffers"]["gaussians"]["fields"]["albedo"].update(offset=44), "gaussians.albedo at 44, renderer expects 48"),
    (lambda h: h["buffers"]["nodes"].update(count=0), "Bad nodes count 0"),
    (lambda h: h["buffers"]["vgrid"].update(count=(1 << 24) + 1), "Bad vgrid count 16777217"),
])
def test_check_header_refusals(tmp_path, edit, msg):
    h = write_asset(tmp_path)
    edit(h)
    with pytest.raises(A.AssetError, match=msg):
        A.check_header(h)


def test_check_buffer_size(tmp_path):
    h = write_asset(tmp_path)
    with pytest.raises(A.AssetError, match=r"kids.bin is 8 B, expected \d+ x 4 B"):
        A.check_buffer(h, "kids", b"\0" * 8)


# ---- node records -----------------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["hill", "cornell"])
def test_node_gaussians(name):
    sc, nodes, kids, _, _ = lodcut.load(name)
    G = sc.count
    rec = LOD.node_gaussians(nodes, kids, sc.gaussians)["records"]
    assert rec.shape[0] == nodes.shape[0] - G
    q = rec["quat"].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    s = rec["scale"].astype(np.float64)
    assert np.all(np.diff(s, axis=1) >= -1e-7 * s[:, 1:])
    x, y, z, w = q.T
    Rm = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], 1),
                   np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], 1),
                   np.stack([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], 1)], 2)   # columns
    assert np.all(np.linalg.det(Rm) > 0.999)
    C = Rm @ (s[:, :, None] ** 2 * np.transpose(Rm, (0, 2, 1)))
    cd, co = nodes["cinv_diag"][G:].astype(np.float64), nodes["cinv_off"][G:].astype(np.float64)
    Ci = np.zeros((cd.shape[0], 3, 3))
    Ci[:, [0, 1, 2], [0, 1, 2]] = cd
    Ci[:, 0, 1] = Ci[:, 1, 0] = co[:, 0]
    Ci[:, 0, 2] = Ci[:, 2, 0] = co[:, 1]
    Ci[:, 1, 2] = Ci[:, 2, 1] = co[:, 2]
    Cref = np.linalg.inv(Ci)
    err = np.abs(C - Cref).max((1, 2)) / np.abs(Cref).max((1, 2))
    assert err.max() < 2e-4, err.max()
    assert np.array_equal(rec["part_surface"], nodes["part"][G:] & 0xFFFF)
    # albedo: mass-weighted over the subtree's leaves (node masses as baked on the leaves); material: first leaf's
    cdl, col = nodes["cinv_diag"][:G].astype(np.float64), nodes["cinv_off"][:G].astype(np.float64)
    det = (cdl[:, 0] * (cdl[:, 1] * cdl[:, 2] - col[:, 2] ** 2) - col[:, 0] * (col[:, 0] * cdl[:, 2] - col[:, 2] * col[:, 1])
           + col[:, 1] * (col[:, 0] * col[:, 2] - cdl[:, 1] * col[:, 1]))
    m_leaf = nodes["kappa"][:G] / np.sqrt(det)
    rng = np.random.default_rng(1)
    for n in rng.choice(np.arange(G, nodes.shape[0]), 60, replace=False):
        leaves, stack, first = [], [int(n)], None
        while stack:
            k = stack.pop()
            if k < G:
                leaves.append(k)
                continue
            ch = kids[nodes["child_lo"][k]:nodes["child_hi"][k]]
            stack.extend(int(c) for c in ch[::-1])
        first = leaves[0]
        want = (m_leaf[leaves, None] * sc.gaussians["albedo"][leaves]).sum(0) / m_leaf[leaves].sum()
        np.testing.assert_allclose(rec["albedo"][n - G], want, rtol=1e-5, atol=1e-7)
        assert rec["flags"][n - G] == sc.gaussians["flags"][first] & 0xFFFF


# ---- the cut ------------------------------------------------------------------------------------------------------

def params(sc, threshold, emit, size_rule=1, eye=None):
    v = R.look_at(eye or sc.eye, sc.target, (0.0, 0.0, 1.0))
    focal = sc.height / (2 * math.tan(math.radians(sc.fov_deg) / 2))
    return oracle.lod_params(v, threshold, focal, 0.02, emit, 0, size_rule), v, focal


def check_cover(sc, nodes, kids, flag, emit):
    """Each leaf drawn or under exactly one drawn node, the path above it open up to its part root."""
    G, N = sc.count, nodes.shape[0]
    up = np.full(N, -1)
    for n in range(G, N):
        up[kids[nodes["child_lo"][n]:nodes["child_hi"][n]]] = n
    roots = set(int(r) for r in sc.parts["root"])
    for g in range(G):
        drawn, a = [], g
        while a >= 0:
            if flag[a] == DRAW:
                drawn.append(a)
            if a in roots:
                break
            a = up[a]
        assert len(drawn) == 1, (g, drawn)
        a = up[drawn[0]] if drawn[0] not in roots else -1
        while a >= 0:
            assert flag[a] == OPEN, (g, a, flag[a])
            if a in roots:
                break
            a = up[a]
        if (sc.gaussians["part_surface"][g] & 0xFFFF) == emit:
            assert flag[g] == DRAW


@pytest.mark.parametrize("name", ["hill", "cornell"])
def test_cut_covers_every_leaf_once(name):
    sc, nodes, kids, _, emit = lodcut.load(name)
    for thr in (0.5, 2.0, 8.0, 64.0):
        p, _, _ = params(sc, thr, emit)
        flag = oracle.lod_select(p, nodes, kids, sc.parts, sc.count)
        check_cover(sc, nodes, kids, flag, emit)
        assert set(np.unique(flag)) <= {NONE, OPEN, DRAW, UNDER}


@pytest.mark.parametrize("name", ["hill", "cornell"])
def test_cut_limits(name):
    sc, nodes, kids, _, emit = lodcut.load(name)
    G = sc.count
    p, _, _ = params(sc, 0.0, emit)
    flag = oracle.lod_select(p, nodes, kids, sc.parts, G)
    assert np.all(flag[:G] == DRAW) and not np.any(flag[G:] == DRAW)
    p, _, _ = params(sc, 1e30, emit)
    flag = oracle.lod_select(p, nodes, kids, sc.parts, G)
    for k, r in enumerate(sc.parts["root"]):
        if k == emit:
            leaves = np.arange(sc.parts["g_lo"][k], sc.parts["g_hi"][k])
            assert np.all(flag[leaves] == DRAW) and flag[r] == OPEN
        else:
            assert flag[r] == DRAW
    assert (emit != 0xFFFFFFFF) == (name == "cornell")


@pytest.mark.parametrize("name", ["hill", "cornell"])
def test_cut_monotone_radius_rule(name):
    sc, nodes, kids, _, emit = lodcut.load(name)
    prev = None
    for thr in (0.25, 1.0, 2.0, 4.0, 16.0, 128.0):
        p, _, _ = params(sc, thr, emit, size_rule=0)
        opened = set(np.nonzero(oracle.lod_select(p, nodes, kids, sc.parts, sc.count) == OPEN)[0].tolist())
        if prev is not None:
            assert opened <= prev
        prev = opened


def test_node_size_projected_area():
    sc, nodes, kids, _, emit = lodcut.load("hill")
    G = sc.count
    p, view, focal = params(sc, 2.0, emit)
    V = np.asarray(view, np.float64).reshape(4, 4).T          # column-major words -> matrix
    rng = np.random.default_rng(2)
    checked = 0
    for n in rng.choice(np.arange(G, nodes.shape[0]), 200, replace=False):
        P = int(nodes["part"][n])
        xf = sc.parts["xf"][P].astype(np.float64).reshape(3, 4)
        w = xf[:, :3] @ nodes["mean"][n].astype(np.float64) + xf[:, 3]
        pv = V[:3, :3] @ w + V[:3, 3]
        z, Rr = -pv[2], float(nodes["radius"][n])
        size = (2 / 3) * focal * Rr / max(z - Rr, 0.02)
        if z - Rr > 0.02:
            cd, co = nodes["cinv_diag"][n].astype(np.float64), nodes["cinv_off"][n].astype(np.float64)
            Ci = np.array([[cd[0], co[0], co[1]], [co[0], cd[1], co[2]], [co[1], co[2], cd[2]]])
            M = V[:3, :3] @ xf[:, :3]
            Cv = M @ np.linalg.inv(Ci) @ M.T
            J = np.array([[focal / z, 0, focal * pv[0] / z ** 2], [0, focal / z, focal * pv[1] / z ** 2]])
            S = J @ Cv @ J.T
            ev = np.linalg.eigvalsh(S)
            size *= math.sqrt(math.sqrt(max(ev[0], 0) / ev[1]))    # sqrt(short / long) of the footprint
            checked += 1
        got = oracle.lod_node_size(p, nodes, sc.parts, int(n))
        np.testing.assert_allclose(got, size, rtol=2e-3)
    assert checked > 100


def test_node_colour_mass_weighted():
    sc, nodes, kids, _, emit = lodcut.load("cornell")
    G, N = sc.count, nodes.shape[0]
    rng = np.random.default_rng(3)
    colour = rng.uniform(0, 2, (G, 4)).astype(np.float32)
    lo_b = rng.uniform(0, 0.5, (G, 12)).astype(np.float32)
    gain = 0.8
    p, _, _ = params(sc, 6.0, emit)
    flag = oracle.lod_select(p, nodes, kids, sc.parts, G)
    got = oracle.lod_colour(nodes, kids, G, flag, colour, lo_b, gain)
    mass = np.array([oracle.node_mass(nodes, n) for n in range(N)], np.float64)
    shown = np.zeros((N, 3))
    shown[:G] = colour[:, :3].astype(np.float64) + gain * lo_b[:, :3].astype(np.float64)
    order = np.argsort(-nodes["depth"][G:].astype(np.int64), kind="stable") + G
    for n in order:
        ch = kids[nodes["child_lo"][n]:nodes["child_hi"][n]]
        shown[n] = (mass[ch, None] * shown[ch]).sum(0) / mass[ch].sum()
    need = np.nonzero((flag[G:] == DRAW) | (flag[G:] == UNDER))[0]
    assert need.size > 50
    np.testing.assert_allclose(got[need, :3], shown[G + need], rtol=2e-5)
    assert np.all(got[need, 3] == 1.0)


def test_display_fill_matrix():
    sc, _, _, _, _ = lodcut.load("cornell")
    G = sc.count
    f = lodcut.fill("cornell")
    A.check_fill(f, G)
    rng = np.random.default_rng(4)
    colour = rng.uniform(0, 2, (G, 4)).astype(np.float32)
    lo_b = rng.uniform(0, 1, (G, 12)).astype(np.float32)
    col, lb = oracle.display_fill(f, G, colour, lo_b)
    off, pairs = f[:G + 1], f[G + 1:].reshape(-1, 2)
    W = np.zeros((G, G))
    for g in range(G):
        for j, w in pairs[off[g]:off[g + 1]]:
            W[g, j] += np.frombuffer(np.uint32(w).tobytes(), np.float32)[0]
    tgt = np.diff(off) > 0
    assert tgt.sum() > 100
    np.testing.assert_allclose(col[tgt, :3], (W @ colour[:, :3].astype(np.float64))[tgt], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(lb[tgt, :3], (W @ lo_b[:, :3].astype(np.float64))[tgt], rtol=1e-5, atol=1e-6)
    assert np.all(lb[tgt, 3] == 0) and np.array_equal(col[tgt, 3], colour[tgt, 3])
    assert np.array_equal(col[~tgt], colour[~tgt]) and np.array_equal(lb[~tgt], lo_b[~tgt])
    np.testing.assert_allclose(W[tgt].sum(1), 1.0, rtol=1e-4)       # the baked weights are a mean


def test_check_fill_refusals():
    G = 4
    good = np.array([0, 1, 1, 1, 1, 2, 0], np.uint32)          # target 0 <- source 2 (weight bits 0)
    A.check_fill(good, G)
    with pytest.raises(A.AssetError, match="overlap"):
        A.check_fill(np.array([0, 1, 1, 1, 1, 0, 0], np.uint32), G)
    with pytest.raises(A.AssetError, match="out of range"):
        A.check_fill(np.array([0, 1, 1, 1, 1, 9, 0], np.uint32), G)
    with pytest.raises(A.AssetError, match="ascending"):
        A.check_fill(np.array([0, 1, 0, 1, 1, 2, 0], np.uint32), G)


def test_project_lod_slots():
    """With every leaf drawn the leaf slots are the plain projection's; undrawn node slots stay invalid, drawn ones
    get their record's splat."""
    sc, nodes, kids, _, emit = lodcut.load("hill")
    G = sc.count
    rec = LOD.node_gaussians(nodes, kids, sc.gaussians)["records"]
    colour = np.concatenate([sc.gaussians["albedo"], np.ones((G, 1))], 1).astype(np.float32)
    nn = nodes.shape[0] - G
    fb = lodcut.lod_frame(sc, 2.0, nn)
    p, _, _ = params(sc, 0.0, emit)
    flag = oracle.lod_select(p, nodes, kids, sc.parts, G)
    ncol = np.zeros((nn, 4), np.float32)
    s1, k1, v1 = oracle.project_lod(fb, sc, flag, rec, ncol, colour=colour)
    s0, k0, v0 = oracle.project(sc.frame_bytes(), sc, colour=colour)
    assert np.array_equal(k1[:G + 1], k0[:G + 1]) and v1 == v0 and np.all(k1[G + 1:] == 0xFFFFFFFF)
    p, _, _ = params(sc, 4.0, emit)
    flag = oracle.lod_select(p, nodes, kids, sc.parts, G)
    ncol = oracle.lod_colour(nodes, kids, G, flag, colour, None, 0.0)
    s2, k2, v2 = oracle.project_lod(fb, sc, flag, rec, ncol, colour=colour)
    drawn_nodes = np.nonzero(flag[G:] == DRAW)[0]
    assert drawn_nodes.size > 0 and np.all(k2[:G][flag[:G] != DRAW] == 0xFFFFFFFF)
    vis = drawn_nodes[k2[G + 1 + drawn_nodes] != 0xFFFFFFFF]
    assert vis.size > 0
    np.testing.assert_array_equal(s2[G + 1 + vis, 8:11].view(np.float32), ncol[vis, :3])
