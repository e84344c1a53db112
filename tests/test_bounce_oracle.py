"""Pins of the bounce oracle (oracle/bounce.c) and of the host side (splatlight_amd/bounce.py, tree.py).

Each check is fixed by something outside the code it checks: numpy's pseudo-inverse for the D14 fit, the shipped
hill_lite tree for the moment-matching rule, numpy's float16 for the half packing, a sky closed form (an open sky
over faint isotropic Gaussians gives each one its albedo times the sky's mean radiance, and probe SH2 equal to the
sky's projection), brute-force subtree sums for node_up, an exact SH2 function for node_sh2, an f64 brute force of
the probe ray composite when a ray meets at most 16 nodes, and conservation of optical depth by the hit list's fold.
"""
import math
import os

import numpy as np
import pytest

import oracle
from splatlight_amd import bounce as BN
from splatlight_amd import lighting as L
from splatlight_amd import records as R
from splatlight_amd import scenes, tree

HERE = os.path.dirname(os.path.abspath(__file__))


# ---- host side: D14 fit, cell table, layout ---------------------------------------------------------------------

def test_pinv_d14_matches_numpy():
    P = np.array(BN.pinv_d14())
    A = np.array([BN.sh2(d) for d in BN.D14])
    assert P.shape == (9, 14)
    assert np.abs(P - np.linalg.pinv(A)).max() < 1e-12
    # fits any SH2 function on D14 exactly
    c = np.random.default_rng(0).normal(size=9)
    assert np.abs(P @ (A @ c) - c).max() < 1e-12


def test_cell_table():
    d, om = BN.cell_table()
    assert abs(float(om.astype(np.float64).sum()) - 4 * math.pi) < 1e-5
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
    assert np.all(np.diff(d[::16, 2]) < 0)               # rows from the zenith down
    assert abs(float(d[0, 2]) - math.cos(math.pi / 16)) < 1e-6


def test_terrain_layout_ground():
    sc = scenes.hill(n=20_000)
    lay = BN.probe_layout(sc)
    assert lay["terrain"] and lay["dims"] == [16, 16, 5]
    mu = sc.gaussians["mu"][sc.parts["g_lo"][0]:sc.parts["g_hi"][0]].astype(np.float64)
    lo = mu.min(0)
    step = lay["step"]
    for (i, j) in ((0, 0), (7, 9), (15, 15)):
        m = (np.floor((mu[:, 0] - sc.gaussians["mu"][:, 0].min()) / step[0]) == i) & \
            (np.floor((mu[:, 1] - sc.gaussians["mu"][:, 1].min()) / step[1]) == j)
        if m.any():
            assert lay["ground"][i * 16 + j] == np.float32(mu[m, 2].max())
    assert lo[2] < 0.0
    # probe volumes on the trunk and canopy (static, 1 m < radius < a quarter of the half-diagonal)
    assert [v["part"] for v in lay["vols"]] == [1, 2]


def test_room_layout_positions():
    sc = scenes.cornell(n=4_000)
    t = tree.scene_tree(sc)
    lay = BN.probe_layout(sc)
    PP = BN.frame_params(BN.probe_params(lay, t.count, 3), 0, True, True)
    U = R.frame_uniforms(gaussian_count=sc.count, part_count=3, eye=(0, 0, 0), material_count=8)
    pos = oracle.probe_pos(U, PP, t.nodes, sc.parts, 512)
    ijk = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    want = np.array([-0.875, -0.875, 0.125]) + 0.25 * ijk
    assert np.abs(pos - want).max() < 1e-6


# ---- tree ----------------------------------------------------------------------------------------------------------

def test_tree_rule_matches_shipped_tree():
    """The moment-matching rule reproduces the shipped hill_lite tree's internal nodes from their children."""
    z = np.load(os.path.join(HERE, "golden", "tree_hill_lite.npz"))
    par = z["parents"].view(R.NODE_DTYPE).ravel()
    ch = z["children"].view(R.NODE_DTYPE).ravel()
    grp = z["group"]
    m, mu, C, rad, sd, so = tree.combine(*tree.node_moments(ch), grp)
    rec = tree.pack_internal(m, mu, C, rad, sd, so)
    scale = np.maximum(np.abs(par["mean"]).max(1, keepdims=True), 1.0)
    assert np.abs(rec["mean"] - par["mean"]).max() / 1.0 < 1e-4 * scale.max()
    assert np.max(np.abs(rec["radius"] - par["radius"]) / par["radius"]) < 1e-4
    assert np.max(np.abs(rec["kappa"] - par["kappa"]) / par["kappa"]) < 1e-3
    assert np.max(np.abs(rec["smax"] - par["smax"]) / par["smax"]) < 1e-4
    ci = np.concatenate([par["cinv_diag"], par["cinv_off"]], 1)
    cr = np.concatenate([rec["cinv_diag"], rec["cinv_off"]], 1)
    assert np.max(np.abs(cr - ci) / np.abs(ci).max(1, keepdims=True)) < 2e-3
    assert np.abs(rec["sn_diag"] - par["sn_diag"]).max() < 1e-5
    assert np.abs(rec["sn_off"] - par["sn_off"]).max() < 1e-5


def test_tree_structure():
    sc = scenes.tiny(n=3000)
    t = tree.scene_tree(sc)
    G, nd = sc.count, t.nodes
    seen = np.zeros(t.count, np.int64)
    stack = [t.scene_root]
    while stack:
        n = stack.pop()
        seen[n] += 1
        if n >= G:
            stack.extend(int(c) for c in t.kids[nd["child_lo"][n]:nd["child_hi"][n]])
    assert np.all(seen == 1)
    assert nd["depth"][t.scene_root] == 0 and nd["part"][t.scene_root] == tree.NO_PART
    for p, r in enumerate(t.roots):
        assert sc.parts["root"][p] == r and nd["depth"][r] == 1 and nd["part"][r] == p
    # masses add up: the scene root holds the leaves' total mass
    lm = tree.leaf_moments(sc.gaussians, nd[:G])[0]
    rm = tree.node_moments(nd[t.scene_root:t.scene_root + 1])[0][0]
    assert abs(rm - lm.sum()) / lm.sum() < 1e-4
    # every internal node's sphere contains its children's
    for n in range(G, t.count):
        c = t.kids[nd["child_lo"][n]:nd["child_hi"][n]]
        reach = np.linalg.norm(nd["mean"][c].astype(np.float64) - nd["mean"][n], axis=1) + nd["radius"][c]
        assert np.all(reach <= nd["radius"][n] * (1 + 1e-5) + 1e-6)


# ---- half packing --------------------------------------------------------------------------------------------

def test_f16_bits_match_numpy():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-9, 6, 2000),
                        [0.0, -0.0, 65504.0, 65520.0, 1e6, -1e6, 6e-8, 3e-8, 2.98e-8, np.inf, -np.inf]]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    got = np.array([oracle.f16_bits(float(v)) for v in x], np.uint32)
    assert np.array_equal(got, want.astype(np.uint32))


# ---- node sums ----------------------------------------------------------------------------------------------------

def _small_tree(n=1500, seed=5):
    sc = scenes.tiny(n=n, seed=seed)
    t = tree.scene_tree(sc)
    lay = BN.probe_layout(sc)
    P = sc.parts.shape[0]
    PP = BN.frame_params(BN.probe_params(lay, t.count, P), 0, True, True)
    U = R.frame_uniforms(gaussian_count=sc.count, part_count=P, eye=(0, 0, 0), material_count=8)
    return sc, t, PP, U, lay


def test_node_up_is_the_subtree_sum():
    sc, t, PP, U, _ = _small_tree()
    G, Nn = sc.count, t.count
    rng = np.random.default_rng(1)
    acc = np.zeros((Nn, 14, 4), np.float32)
    acc[:G] = rng.uniform(0, 1, (G, 14, 4))
    up, levels = BN.up_levels(t.nodes, G, t.max_depth)
    got = oracle.node_up(U, PP, t.nodes, t.kids, up, levels, t.max_depth, acc)
    # brute force: each internal node below the part roots' depth holds the f64 sum of its leaves
    want = acc.astype(np.float64)
    order = np.argsort(-t.nodes["depth"][G:], kind="stable") + G
    for n in order:
        c = t.kids[t.nodes["child_lo"][n]:t.nodes["child_hi"][n]]
        want[n] = want[c].sum(0)
    internal = np.arange(G, Nn)
    summed = internal[(t.nodes["depth"][internal] >= 1) & (t.nodes["depth"][internal] <= t.max_depth - 1) &
                      (t.nodes["part"][internal] != tree.NO_PART)]
    assert summed.size == Nn - G - 1
    assert np.max(np.abs(got[summed] - want[summed]) / np.maximum(want[summed], 1e-3)) < 1e-5
    assert np.all(got[t.scene_root] == 0)                 # depth 0 is not summed (node_up runs depths >= 1)


def test_node_sh2_fits_sh2_functions():
    """Light that is an SH2 function of direction on D14 comes back as its coefficients (to half precision)."""
    sc, t, PP, U, _ = _small_tree(n=200)
    Nn = t.count
    rng = np.random.default_rng(2)
    coef = rng.normal(size=(Nn, 9, 3))
    A = np.array([BN.sh2(d) for d in np.array(BN.D14, np.float32).astype(np.float64)])
    Lk = np.einsum("km,nmc->nkc", A, coef)
    w = rng.uniform(0.5, 2.0, (Nn, 14))
    acc = np.zeros((Nn, 14, 4), np.float32)
    acc[..., :3] = Lk * w[..., None]
    acc[..., 3] = w
    sh = oracle.node_sh2(PP, acc, Nn)
    h = sh.view(np.uint16).reshape(Nn, 28)[:, :27].view(np.float16).astype(np.float64).reshape(Nn, 9, 3)
    assert np.max(np.abs(h - coef)) < 2e-3 * np.abs(coef).max()


# ---- probe gather -------------------------------------------------------------------------------------------------

def _seg_geom64(mean, cinv, P, w, t0, t1):
    d = mean - P
    cw = cinv @ w
    a = cw @ w
    m = (cw @ d) / a
    e = d - m * w
    q = max(e @ cinv @ e, 0.0)
    s = math.sqrt(0.5 * a)
    return math.exp(-0.5 * q) * math.sqrt(math.pi / (2 * a)) * 0.5 * (math.erfc(s * (t0 - m)) - math.erfc(s * (t1 - m))) * 2


def _leaf_hits(t, sc, o, w, t1):
    """Every leaf a ray meets (bounding sphere) with its optical depth and composite depth, in f64."""
    nd = t.nodes
    out = []
    for g in range(sc.count):
        mean = nd["mean"][g].astype(np.float64)
        d = mean - o
        tt = min(max(d @ w, 0.0), t1)
        e = d - tt * w
        if e @ e > float(nd["radius"][g]) ** 2:
            continue
        cd, co = nd["cinv_diag"][g].astype(np.float64), nd["cinv_off"][g].astype(np.float64)
        ci = np.array([[cd[0], co[0], co[1]], [co[0], cd[1], co[2]], [co[1], co[2], cd[2]]])
        tau = float(nd["kappa"][g]) * 1.0 * _seg_geom64(mean, ci, o, w, 0.0, t1)
        if tau <= 1e-6:
            continue
        cw = ci @ w
        tp = min(max((cw @ d) / (cw @ w), 0.0), t1)
        out.append((tp, tau, g))
    return out


def test_gather_matches_f64_brute_force():
    """theta huge opens every node: the gather composites the leaves a ray meets, nearest first. Rays that meet at
    most 16 leaves match an f64 composite of all of them; for every ray, what escapes (the sky term) is
    exp(-sum of all optical depths), which the hit list's fold conserves."""
    rng = np.random.default_rng(11)
    n = 400
    g = np.zeros(n, R.GAUSSIAN_DTYPE)
    g["mu"] = rng.uniform([-1, -1, 0.2], [1, 1, 1.8], (n, 3))
    g["scale"] = np.sort(rng.uniform(0.03, 0.12, (n, 3)), 1)
    q = rng.normal(size=(n, 4))
    g["quat"] = q / np.linalg.norm(q, axis=1, keepdims=True)
    g["kappa"] = rng.uniform(2.0, 20.0, n)
    g["albedo"] = 0.5
    parts = R.identity_part(0, n, bound=(0, 0, 1, 2.0), flags=1)
    sc = scenes.Scene("room", g, parts, scenes._materials(), scenes._emitter_records(np.zeros((1, 3)), np.zeros((1, 3)), (0.01,) * 3),
                      (0, -3, 1), (0, 0, 1), 40.0, 64, 48, 50.0)
    t = tree.scene_tree(sc)
    hdr = {"parts": [{"name": "p", "mobility": "static", "gaussians": [0, n]}], "environment": True, "probes": {},
           "room": {"x": [-1.5, 1.5], "y": [-1.5, 1.5], "z": [0.0, 2.0]}}
    lay = BN.probe_layout(sc, hdr)
    base = BN.probe_params(lay, t.count, 1)
    u = base.view(np.uint32)
    base.view(np.float32)[9] = 1e30                      # theta: open every internal node
    PP = BN.frame_params(base, 0, True, True)
    U = R.frame_uniforms(gaussian_count=n, part_count=1, eye=(0, 0, 0), material_count=8)
    ENV = L.env_bytes([], np.vstack([[1.0 / 0.282095] * 3, np.zeros((15, 3))]))   # sky radiance 1 everywhere
    # node light: band 0 only, radiance r_n = 0.282095 h0 along every direction
    rad = rng.uniform(0.1, 2.0, t.count).astype(np.float32)
    h = np.zeros((t.count, 28), np.float32)
    h[:, 0] = rad / 0.282095
    h[:, 1] = h[:, 0] * 0.5
    h[:, 2] = h[:, 0] * 0.25
    node_sh = np.array([[oracle.f16_bits(float(a)) | (oracle.f16_bits(float(b)) << 16) for a, b in zip(r[0::2], r[1::2])]
                        for r in h], np.uint32)
    radf = (node_sh[:, 0] & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64) * 0.282095
    rad3 = np.stack([radf, (node_sh[:, 0] >> 16).astype(np.uint16).view(np.float16).astype(np.float64) * 0.282095,
                     (node_sh[:, 1] & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64) * 0.282095], 1)
    cells = oracle.cells_table(*BN.cell_table())
    NP = BN.probe_count(lay, 1)
    probes = [0, 73, 200, 511]
    pd0 = np.zeros((NP, 128, 2, 4), np.float32)
    pd = oracle.probe_gather(U, PP, ENV, t.nodes, scenes.vgrid_open(n), sc.parts, t.kids, node_sh, cells, pd0)
    pd = pd.reshape(NP, 128, 2, 4)
    pos = oracle.probe_pos(U, PP, t.nodes, sc.parts, NP)
    tmax = float(lay["tmax"])
    checked = folded = 0
    for p in probes:
        for c in range(0, 128, 3):
            o = pos[p].astype(np.float64)
            w = cells[c, :3].astype(np.float64)
            hits = sorted(_leaf_hits(t, sc, o, w, tmax), key=lambda x: x[0])
            T_all = math.exp(-sum(min(h_[1], 64.0) for h_ in hits))
            got = pd[p, c, 0, :3]
            if len(hits) <= 16:
                T, acc, m1 = 1.0, np.zeros(3), 0.0
                for tp, tau, gid in hits:
                    al = 1 - math.exp(-tau)
                    acc += T * al * rad3[gid]
                    m1 += T * al * tp
                    T *= 1 - al
                acc += T * 1.0
                m1 += T * tmax
                assert np.allclose(got, acc, rtol=2e-4, atol=2e-5), (p, c, got, acc)
                assert abs(pd[p, c, 1, 0] - m1) <= 2e-4 * max(m1, 1.0)
                checked += 1
            else:
                folded += 1
            # what escapes: the sky term T_final x 1, the rest of acc is node light >= 0
            sky_left = float(pd[p, c, 0, 0]) - 0.0
            assert sky_left >= T_all * (1 - 1e-3) - 1e-6
    assert checked > 50


def sky_scene(n=400, seed=4):
    """Faint isotropic Gaussians (no ray meets an optical depth above 1e-6) under a sky A + B z + C x and no lights;
    returns the scene and A. Their bounce is albedo x A."""
    rng = np.random.default_rng(seed)
    g = np.zeros(n, R.GAUSSIAN_DTYPE)
    g["mu"] = rng.uniform([-3, -3, 0.0], [3, 3, 2.0], (n, 3))
    g["scale"] = 0.05
    g["quat"] = [0.0, 0.0, 0.0, 1.0]
    g["kappa"] = 1e-7
    g["albedo"] = rng.uniform(0.2, 0.9, (n, 3))
    parts = R.identity_part(0, n, bound=(0, 0, 1, 5.0), flags=1)
    em = scenes._emitter_records(np.array([[0.0, 0.0, 50.0]]), np.zeros((1, 3)), (0.01, 0.01, 0.01))
    A = np.array([0.4, 0.5, 0.7])
    sky = np.zeros((16, 3))
    sky[0] = A / 0.282095
    sky[2] = np.array([0.1, 0.15, 0.2]) / 0.488603
    sky[3] = np.array([0.05, 0.0, -0.05]) / 0.488603
    sc = scenes.Scene("sky", g, parts, scenes._materials(), em, (0, -8, 2), (0, 0, 1), 40.0, 64, 48, 50.0, sky=sky)
    return sc, A


def test_sky_closed_form():
    sc, A = sky_scene()
    t = tree.scene_tree(sc)
    G = sc.count
    lay = BN.probe_layout(sc)
    P = 1
    NP = BN.probe_count(lay, P)
    PP = BN.frame_params(BN.probe_params(lay, t.count, P), 0, True, True)
    inp = L.lighting_inputs(sc)
    vg = scenes.vgrid_open(G)
    iso = oracle.iso_init(sc.gaussians, t.nodes[:G], vg)
    up, levels = BN.up_levels(t.nodes, G, t.max_depth)
    cells = oracle.cells_table(*BN.cell_table())
    st = oracle.bounce_state(G, t.count, NP)
    st = oracle.bounce_frame(st, inp["U"], PP, inp["ENV"], sc.gaussians, t.nodes, vg, sc.parts, sc.materials, iso,
                             None, np.zeros((G, 4), np.float32), t.kids, up, levels, t.max_depth, cells, NP)
    # every probe ray escapes and sees the sky
    pd = st["probe_data"].reshape(NP, 128, 2, 4)
    lat = lay["dims"][0] * lay["dims"][1] * lay["dims"][2]
    d = cells[:, :3].astype(np.float64)
    sky = A[None, :] + d[:, 2:3] * np.array([0.1, 0.15, 0.2]) + d[:, 0:1] * np.array([0.05, 0.0, -0.05])
    assert np.abs(pd[:lat, :, 0, :3] - sky[None]).max() < 1e-5
    # probe SH2 = the sky's projection: A sqrt(4 pi) on band 0, B and C x 0.488603 x 4 pi / 3 on z and x
    psh = st["probe_sh"][:lat, :, :3]
    want = np.zeros((9, 3))
    want[0] = A * 0.282095 * 4 * math.pi
    want[2] = np.array([0.1, 0.15, 0.2]) * 0.488603 * 4 * math.pi / 3
    want[3] = np.array([0.05, 0.0, -0.05]) * 0.488603 * 4 * math.pi / 3
    assert np.abs(psh - want[None]).max() < 0.02 * np.abs(want).max()
    assert np.abs(psh[:, 0] - want[0]).max() < 1e-4 * want[0].max()
    # each Gaussian's bounce: albedo x A (Gaussian 0 has no transfer: the near-field placeholder reads zero there)
    lob = st["lo_b"]
    # (to the 128-cell quadrature's second-order cross term, band 2 of the transfer against band 2 of the light:
    # 3e-4 relative; a wrong basis index or a dropped term is off by 0.1 or more)
    want = sc.gaussians["albedo"][1:] * A
    assert np.abs(lob[1:, 0, :3] - want).max() < 1e-3 * want.max()
    assert np.array_equal(lob[1:, 0], lob[1:, 1])
    assert np.all(lob[0] == 0)
