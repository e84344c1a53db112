"""Pins for the CPU oracle (runs without a GPU).

The reference has no tests or golden vectors for this path, so each oracle function is pinned by what the
mathematics fixes: closed forms, brute force on small inputs, and invariants, chosen so that a dropped term, a
wrong sign or index, or a transposed operand fails one of them.
"""
import math

import numpy as np
import pytest

import oracle
from splatlight_amd import records as R
from splatlight_amd import scenes

ALPHA_MIN = np.float32(1.0) / np.float32(255.0)


def one_gaussian_scene(mu, scale, quat=(0.0, 0.0, 0.0, 1.0), kappa=50.0, albedo=(0.5, 0.25, 0.125),
                       eye=(0.0, -3.0, 0.0), target=(0.0, 0.0, 0.0), W=64, H=48, fov=40.0, extra=()):
    g = np.zeros(1 + len(extra), R.GAUSSIAN_DTYPE)
    g["mu"][0], g["scale"][0], g["quat"][0], g["kappa"][0], g["albedo"][0] = mu, scale, quat, kappa, albedo
    for k, e in enumerate(extra):
        g["mu"][k + 1], g["scale"][k + 1], g["quat"][k + 1], g["kappa"][k + 1], g["albedo"][k + 1] = e
    parts = R.identity_part(0, g.shape[0])
    s = scenes.Scene("one", g, parts, scenes._materials(), np.zeros(0, R.GAUSSIAN_DTYPE), eye, target, fov, W, H, 50.0)
    return s


# ---- depth key (shaders/project.wgsl:366-367) -------------------------------------------------------------------

def test_depth_key_closed_forms():
    near, far = np.float32(0.02), np.float32(50.0)
    assert oracle.depth_key(near, near, far) == 0xFFFE          # t = 0 -> 65535 clamped to 0xFFFE
    assert oracle.depth_key(far, near, far) == 0                # t = 1 -> 65535 - 65535
    assert oracle.depth_key(np.float32(100.0), near, far) == 0  # clamped t
    mid = float(np.float32(math.sqrt(float(near) * float(far))))
    assert oracle.depth_key(mid, near, far) in (32767, 32768)   # t = 1/2 -> 65535 - u32(32768.0)


def test_depth_key_against_float64_formula():
    rng = np.random.default_rng(3)
    near, far = np.float32(0.02), np.float32(50.0)
    d = rng.uniform(0.02, 50.0, 20000).astype(np.float32)
    got = np.array([oracle.depth_key(float(x), float(near), float(far)) for x in d])
    t = np.log(d.astype(np.float64) / float(near)) / math.log(float(far) / float(near))
    s = np.clip(t, 0, 1) * 65535.0 + 0.5
    want = np.minimum(65535 - np.floor(s).astype(np.int64), 0xFFFE)
    # float32 rounding may only move a key across a boundary the exact value sits next to
    near_boundary = np.abs(s - np.round(s)) < 1e-2
    assert np.all((got == want) | near_boundary)
    assert np.mean(got == want) > 0.995
    assert np.all(np.diff(got[np.argsort(d)].astype(np.int64)) <= 0)   # monotone: farther -> smaller key


# ---- project (shaders/project.wgsl:84-358) ------------------------------------------------------------------

def _splat_fields(sp):
    f = sp.view(np.float32)
    return f


def test_project_isotropic_on_axis_closed_form():
    W, H, fov, s, kappa = 64, 48, 40.0, 0.05, 30.0
    sc = one_gaussian_scene((0.0, 0.0, 0.0), (s, s, s), kappa=kappa, W=W, H=H, fov=fov)
    fb = sc.frame_bytes()
    sp, keys, vis = oracle.project(fb, sc, colour=np.array([[1, 2, 3, 1]], np.float32))
    assert vis == 1 and keys[0] != R.INVALID_KEY and keys[1] == R.INVALID_KEY
    f = _splat_fields(sp)[0]
    focal = H / (2 * math.tan(math.radians(fov) / 2))
    z = 3.0
    a = (focal * s / z) ** 2
    lp = 0.3
    tau0 = kappa * math.sqrt(2 * math.pi) * s * a / (a + lp)
    tau_min = -math.log(1 - 1 / 255)
    r = math.sqrt(2 * math.log(tau0 / tau_min))
    ax = math.sqrt(a + lp) * r
    np.testing.assert_allclose(f[0:2], [0.0, 0.0], atol=1e-6)
    np.testing.assert_allclose(abs(f[2]) * W / 2, ax, rtol=1e-5)     # axis0 along x (a == c: e0 = (1, 0))
    np.testing.assert_allclose(abs(f[5]) * H / 2, ax, rtol=1e-5)
    np.testing.assert_allclose(f[6], r, rtol=1e-5)
    np.testing.assert_allclose(f[7], tau0, rtol=1e-5)
    np.testing.assert_allclose(f[8:11], [1, 2, 3])                 # colour[i] (lit view, palette 1)
    np.testing.assert_allclose(sp[0, 11] & ~np.uint32(3), np.float32(z).view(np.uint32) & ~np.uint32(3))


def _ewa_cov_numpy(sc, i, fb):
    """Sigma_2D = J W R S^2 R^T W^T J^T + lowpass I by explicit float64 matrix products."""
    g = sc.gaussians[i]
    f = np.frombuffer(fb, np.float32)
    V = f[0:16].astype(np.float64).reshape(4, 4).T          # column-major -> row-major
    x, y, z, w = g["quat"].astype(np.float64)
    Rq = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                   [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                   [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    S = np.diag(g["scale"].astype(np.float64) ** 2)
    Wm = V[:3, :3]
    t = V @ np.append(g["mu"].astype(np.float64), 1.0)
    cov = Wm @ Rq @ S @ Rq.T @ Wm.T
    fx = float(f[74])
    depth = -t[2]
    J = np.array([[fx / depth, 0, fx * t[0] / depth ** 2], [0, fx / depth, fx * t[1] / depth ** 2]])
    return J @ cov @ J.T + float(f[83]) * np.eye(2), t


def test_project_rotated_matches_explicit_ewa():
    rng = np.random.default_rng(11)
    extra = []
    for _ in range(40):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        s = np.sort(np.exp(rng.normal(-2.5, 0.5, 3)))
        extra.append((rng.uniform(-0.4, 0.4, 3), s, q, 40.0, (0.3, 0.3, 0.3)))
    sc = one_gaussian_scene((0.0, 0.0, 0.0), (0.05, 0.08, 0.1), extra=extra)
    fb = sc.frame_bytes()
    sp64, valid = oracle.project(fb, sc, colour=np.ones((sc.count, 4), np.float32), f64=True)
    sp32, keys, vis = oracle.project(fb, sc, colour=np.ones((sc.count, 4), np.float32))
    W, H = sc.width, sc.height
    checked = 0
    for i in range(sc.count):
        if not valid[i]:
            continue
        cov2, t = _ewa_cov_numpy(sc, i, fb)
        a0 = sp64[i, 2:4] * np.array([W / 2, H / 2])
        a1 = sp64[i, 4:6] * np.array([W / 2, H / 2])
        r = sp64[i, 6]
        # the quad's axes are the covariance's eigenvectors scaled by sqrt(eigenvalue) * r
        recon = (np.outer(a0, a0) + np.outer(a1, a1)) / r ** 2
        np.testing.assert_allclose(recon, cov2, rtol=1e-9, atol=1e-9)
        assert abs(a0 @ a1) < 1e-9 * (a0 @ a0 + 1e-30)
        # float32 restatement within float rounding of the float64 one
        f32 = sp32[i].view(np.float32)[:11]
        np.testing.assert_allclose(f32, sp64[i, :11], rtol=2e-4, atol=2e-6)
        checked += 1
    assert checked >= 30


# ---- sort artefacts -----------------------------------------------------------------------------------------------

def test_sort_keys_matches_stable_argsort():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 0xFFFF, 50000).astype(np.uint32)
    keys[rng.random(50000) < 0.2] = R.INVALID_KEY
    keys[:1000] = 1234                                        # many ties
    hist, off, order, vis = oracle.sort_keys(keys)
    low = (keys & 0xFFFF).astype(np.int64)
    want = np.argsort(low, kind="stable")
    np.testing.assert_array_equal(order, want)
    valid = keys != R.INVALID_KEY
    assert vis == valid.sum()
    np.testing.assert_array_equal(hist, np.bincount(keys[valid], minlength=65536))
    np.testing.assert_array_equal(off, np.concatenate([[0], np.cumsum(hist)[:-1]]))


# ---- tiles --------------------------------------------------------------------------------------------------------

def test_tile_lists_brute_force():
    sc = scenes.tiny(n=300, width=80, height=56)
    fb = sc.frame_bytes()
    sp, keys, vis = oracle.project(fb, sc, colour=np.ones((sc.count, 4), np.float32))
    _, _, order, _ = oracle.sort_keys(keys)
    W, H = sc.width, sc.height
    start, vals = oracle.tile_lists(sp, order, vis, W, H)
    tw, th = (W + 15) // 16, (H + 15) // 16
    f = sp.view(np.float32)
    for t in range(tw * th):
        tx, ty = t % tw, t // tw
        lst = []
        for r in range(vis):
            i = order[r]
            cx, cy = f[i, 0], f[i, 1]
            ex, ey = abs(f[i, 2]) + abs(f[i, 4]), abs(f[i, 3]) + abs(f[i, 5])
            # pixel centres of the tile inside the box?
            pxs = (np.arange(tx * 16, min(tx * 16 + 16, W)) + 0.5) * 2 / W - 1
            pys = 1 - (np.arange(ty * 16, min(ty * 16 + 16, H)) + 0.5) * 2 / H
            if np.any(np.abs(pxs - cx) <= ex) and np.any(np.abs(pys - cy) <= ey):
                lst.append(i)
        got = vals[start[t]:start[t + 1]]
        assert list(got) == lst, t


# ---- composite ----------------------------------------------------------------------------------------------------

def test_composite_single_splat_closed_form():
    sc = one_gaussian_scene((0.0, 0.0, 0.0), (0.1, 0.15, 0.2), quat=(0.2, 0.3, 0.1, 0.927), kappa=20.0)
    fb = sc.frame_bytes()
    col = np.array([[0.5, 0.25, 2.0, 1.0]], np.float32)
    sp64, valid = oracle.project(fb, sc, colour=col, f64=True)
    img = oracle.composite_f64(fb, sp64, np.array([0, 1], np.uint32), 1, sc.width, sc.height)
    cov2, t = _ewa_cov_numpy(sc, 0, fb)
    f = np.frombuffer(fb, np.float32)
    fx = float(f[74])
    c = np.array([fx * t[0] / -t[2], fx * t[1] / -t[2]])        # pixels, y up, origin at the image centre
    inv = np.linalg.inv(cov2)
    tau0, r = sp64[0, 7], sp64[0, 6]
    W, H = sc.width, sc.height
    for py in range(H):
        for px in range(W):
            p = np.array([px + 0.5 - W / 2, H / 2 - (py + 0.5)])
            m = (p - c) @ inv @ (p - c)
            alpha = 1 - math.exp(-tau0 * math.exp(-0.5 * m))
            if m > r * r or alpha < 1 / 255:
                # outside the cut ellipse the splat adds nothing (the quad may still reach, alpha < 1/255)
                assert img[py, px, 3] == 0 or m <= 2 * r * r
                continue
            np.testing.assert_allclose(img[py, px, 3], alpha, rtol=1e-9)
            np.testing.assert_allclose(img[py, px, :3], col[0, :3] * alpha, rtol=1e-6)


def test_composite_stack_closed_form_and_fp16_target():
    # N identical-footprint splats at increasing depth: colour sum_i c_i a_i prod_{j nearer} (1 - a_j)
    extra = [((0.0, 0.05 * k, 0.0), (0.2, 0.2, 0.2), (0, 0, 0, 1), 3.0, (0.1 * k, 0.5, 1.0 - 0.1 * k)) for k in range(1, 6)]
    sc = one_gaussian_scene((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), kappa=3.0, albedo=(0.9, 0.1, 0.1), extra=extra)
    fb = sc.frame_bytes()
    col = np.concatenate([sc.gaussians["albedo"], np.ones((sc.count, 1))], 1).astype(np.float32)
    sp, keys, vis = oracle.project(fb, sc, colour=col)
    _, _, order, _ = oracle.sort_keys(keys)
    sp64, _ = oracle.project(fb, sc, colour=col, f64=True)
    img32 = oracle.composite(fb, sp, order, vis, sc.width, sc.height, mode=0)
    img64 = oracle.composite_f64(fb, sp64, order, vis, sc.width, sc.height)
    img16 = oracle.composite(fb, sp, order, vis, sc.width, sc.height, mode=2)
    py, px = sc.height // 2, sc.width // 2
    # back to front = farther first: the camera looks along +y, so larger y is farther
    alphas, cols = [], []
    for r in range(vis):
        i = order[r]
        f = sp64[i]
        a0, a1 = f[2:4], f[4:6]
        p = np.array([(px + 0.5) * 2 / sc.width - 1, 1 - (py + 0.5) * 2 / sc.height]) - f[0:2]
        M = np.column_stack([a0, a1])
        u, v = np.linalg.solve(M, p)
        e = math.exp(-0.5 * (u * u + v * v) * f[6] ** 2)
        a = 1 - math.exp(-f[7] * e)
        alphas.append(a)
        cols.append(f[8:11])
    Tn, C = 1.0, np.zeros(3)
    for a, c in zip(alphas[::-1], cols[::-1]):               # front to back
        C += Tn * a * np.asarray(c)
        Tn *= 1 - a
    np.testing.assert_allclose(img64[py, px, :3], C, rtol=1e-9)
    np.testing.assert_allclose(img64[py, px, 3], 1 - Tn, rtol=1e-9)
    assert np.abs(img32 - img64).max() < 1e-5
    # the fp16 target rounds after every blend: every stored value is a float16 value
    assert np.all(img16.astype(np.float16).astype(np.float32) == img16)
    assert np.abs(img16 - img64).max() < 2e-2


def test_fp16_rounding_matches_numpy():
    # one splat, fp16 target: each pixel is half(src) exactly
    sc = one_gaussian_scene((0.0, 0.0, 0.0), (0.1, 0.1, 0.1), kappa=10.0)
    fb = sc.frame_bytes()
    col = np.array([[1.37, 0.211, 7.3, 1.0]], np.float32)
    sp, keys, vis = oracle.project(fb, sc, colour=col)
    _, _, order, _ = oracle.sort_keys(keys)
    a = oracle.composite(fb, sp, order, vis, sc.width, sc.height, mode=0)
    b = oracle.composite(fb, sp, order, vis, sc.width, sc.height, mode=2)
    np.testing.assert_array_equal(b, a.astype(np.float16).astype(np.float32))
