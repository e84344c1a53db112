"""Screen-space LOD, host side (src/lod.js:1-196): the tree's internal nodes as Gaussian records.

node_gaussians follows nodeGaussians (src/lod.js:50-110) without the node V grids (only lit nodes read them): each
internal node becomes a 64 B Gaussian record with the node's mean and kappa, its covariance (the inverse of the
baked C^-1) as an ascending-scale eigen-decomposition (Jacobi, src/lod.js:9-26) and the rotation's quaternion
(src/lod.js:29-41), its part, the mass-weighted albedo of its leaves and the material of its first leaf. The
records go to sl_scene_set_lod with the tree; the frame then draws the cut lod_select picks (RenderFrame.lod).
"""
from __future__ import annotations

import math

import numpy as np

from . import records as R


def eig3(A):
    """Symmetric 3x3 eigen-decomposition by Jacobi sweeps (src/lod.js:9-26): (values ascending, vectors as
    columns: vectors[k] is the k-th)."""
    a = [list(map(float, r)) for r in A]
    V = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    for _ in range(20):
        off = sum(a[p][q] * a[p][q] for p in range(3) for q in range(p + 1, 3))
        if off < 1e-30:
            break
        for p in range(3):
            for q in range(p + 1, 3):
                if abs(a[p][q]) < 1e-30:
                    continue
                th = 0.5 * math.atan2(2 * a[p][q], a[q][q] - a[p][p])
                c, s = math.cos(th), math.sin(th)
                for k in range(3):
                    x, y = a[k][p], a[k][q]
                    a[k][p], a[k][q] = c * x - s * y, s * x + c * y
                for k in range(3):
                    x, y = a[p][k], a[q][k]
                    a[p][k], a[q][k] = c * x - s * y, s * x + c * y
                for k in range(3):
                    x, y = V[k][p], V[k][q]
                    V[k][p], V[k][q] = c * x - s * y, s * x + c * y
    idx = sorted(range(3), key=lambda i: a[i][i])
    return [a[i][i] for i in idx], [[V[0][i], V[1][i], V[2][i]] for i in idx]


def quat_from_columns(c):
    """Rotation with columns c[0], c[1], c[2] -> quaternion (x, y, z, w) (src/lod.js:29-41)."""
    m00, m11, m22 = c[0][0], c[1][1], c[2][2]
    m10, m20, m01, m21, m02, m12 = c[0][1], c[0][2], c[1][0], c[1][2], c[2][0], c[2][1]
    tr = m00 + m11 + m22
    if tr > 0:
        s = 2 * math.sqrt(tr + 1)
        w, x, y, z = s / 4, (m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s
    elif m00 > m11 and m00 > m22:
        s = 2 * math.sqrt(1 + m00 - m11 - m22)
        x, w, y, z = s / 4, (m21 - m12) / s, (m01 + m10) / s, (m02 + m20) / s
    elif m11 > m22:
        s = 2 * math.sqrt(1 + m11 - m00 - m22)
        y, w, x, z = s / 4, (m02 - m20) / s, (m01 + m10) / s, (m12 + m21) / s
    else:
        s = 2 * math.sqrt(1 + m22 - m00 - m11)
        z, w, x, y = s / 4, (m10 - m01) / s, (m02 + m20) / s, (m12 + m21) / s
    l = math.hypot(x, y, z, w) or 1.0
    return [x / l, y / l, z / l, w / l]


def _cov(nd):
    """(C, det C^-1) of a node record: C the inverse of the baked C^-1 by cofactors (src/lod.js:62-70)."""
    a, b = nd["cinv_diag"].astype(np.float64), nd["cinv_off"].astype(np.float64)
    M = [[a[0], b[0], b[1]], [b[0], a[1], b[2]], [b[1], b[2], a[2]]]
    det = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
           + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]))
    inv = [[M[1][1] * M[2][2] - M[1][2] * M[2][1], M[0][2] * M[2][1] - M[0][1] * M[2][2], M[0][1] * M[1][2] - M[0][2] * M[1][1]],
           [M[1][2] * M[2][0] - M[1][0] * M[2][2], M[0][0] * M[2][2] - M[0][2] * M[2][0], M[0][2] * M[1][0] - M[0][0] * M[1][2]],
           [M[1][0] * M[2][1] - M[1][1] * M[2][0], M[0][1] * M[2][0] - M[0][0] * M[2][1], M[0][0] * M[1][1] - M[0][1] * M[1][0]]]
    return [[v / det for v in r] for r in inv], det


def node_gaussians(nodes: np.ndarray, kids: np.ndarray, gaussians: np.ndarray) -> dict:
    """src/lod.js:50-110 without V grids: {"records": (N - G) Gaussian records, "N", "G", "maxDepth"}. nodes: all N
    node records (leaves [0, G) first), kids: the child index list, gaussians: the G leaf records."""
    nodes = np.asarray(nodes).view(R.NODE_DTYPE).reshape(-1)
    gaussians = np.asarray(gaussians).view(R.GAUSSIAN_DTYPE).reshape(-1)
    kids = np.asarray(kids, np.uint32)
    N, G = nodes.shape[0], gaussians.shape[0]
    mass = np.zeros(N)
    alb = np.zeros((N, 3))
    mat = np.zeros(N, np.uint32)
    # leaves: mass kappa / sqrt(det C^-1), their albedo and material
    cd, co = nodes["cinv_diag"][:G].astype(np.float64), nodes["cinv_off"][:G].astype(np.float64)
    det = (cd[:, 0] * (cd[:, 1] * cd[:, 2] - co[:, 2] * co[:, 2]) - co[:, 0] * (co[:, 0] * cd[:, 2] - co[:, 2] * co[:, 1])
           + co[:, 1] * (co[:, 0] * co[:, 2] - cd[:, 1] * co[:, 1]))
    mass[:G] = nodes["kappa"][:G] / np.sqrt(np.maximum(det, 1e-30))
    alb[:G] = gaussians["albedo"]
    mat[:G] = gaussians["flags"] & 0xFFFF
    depth = nodes["depth"]
    max_depth = int(depth.max()) if N else 0
    order = np.arange(G, N)
    order = order[np.argsort(-depth[G:].astype(np.int64), kind="stable")]   # deepest first: children before parents
    out = np.zeros(max(1, N - G), R.GAUSSIAN_DTYPE)
    for n in order:
        lo, hi = int(nodes["child_lo"][n]), int(nodes["child_hi"][n])
        ch = kids[lo:hi]
        m = float(mass[ch].sum())
        mass[n] = m
        alb[n] = (mass[ch, None] * alb[ch]).sum(0) / max(m, 1e-30)
        mat[n] = mat[ch[0]] if hi > lo else 0
        C, _ = _cov(nodes[n])
        vals, vecs = eig3(C)
        cols = [list(v) for v in vecs]
        d = (cols[0][0] * (cols[1][1] * cols[2][2] - cols[1][2] * cols[2][1])
             - cols[1][0] * (cols[0][1] * cols[2][2] - cols[0][2] * cols[2][1])
             + cols[2][0] * (cols[0][1] * cols[1][2] - cols[0][2] * cols[1][1]))
        if d < 0:
            cols[2] = [-v for v in cols[2]]
        r = n - G
        out["mu"][r] = nodes["mean"][n]
        out["kappa"][r] = nodes["kappa"][n]
        out["quat"][r] = quat_from_columns(cols)
        out["scale"][r] = [math.sqrt(max(v, 1e-12)) for v in vals]
        out["part_surface"][r] = nodes["part"][n] & 0xFFFF
        out["albedo"][r] = alb[n]
        out["flags"][r] = mat[n]
    return {"records": out[:N - G], "N": N, "G": G, "maxDepth": max_depth}


def lod_words(threshold_px: float, count: int, n_emit: int, n_nodes: int):
    """RenderFrame.lod (words 180-182, src/main.js:548): the threshold's f32 bits, the first node slot and the node
    count. The reference places the nodes after its 16,384 emitter slots; the first node slot here follows this
    scene's emitters, with the nodes in the same order after them (the depth order's ties go by slot)."""
    return float(threshold_px), count + 1 + n_emit, n_nodes
