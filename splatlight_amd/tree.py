"""Scene trees for the probe bounce: binary trees over each part's leaves plus a scene root.

The probe gather walks a tree of Gaussian nodes per part (runtime/wgsl/bounce.wgsl:1317-1373) and the node passes
sum the leaves' outgoing light up that tree (bounce.wgsl:1166-1209). The reference receives its trees baked with its
assets (nodes.bin + kids.bin; the builder, runtime/bintree.py, is not shipped); the synthetic scenes carry leaves
only, so this module builds the tree the way the shipped ones are laid out:

* nodes [0, G) are the leaves (one per Gaussian), internal nodes follow, the scene root is the last node;
* each part's subtree is binary, its root at depth 1 with ``part`` = the part index; the scene root has depth 0,
  ``part`` = 0xFFFFFFFF and the part roots as its kids (hill_lite_bin/asset.json: scene_root, parts[].root);
* an internal node's children are kids[child_lo:child_hi]; leaves have child_lo = child_hi = 0;
* internal nodes are moment-matched over their children with mass m = kappa sqrt(det C): mean and covariance are
  mass-weighted (covariance plus the spread of the children's means), kappa = m / sqrt(det C) conserves the mass,
  smax = sqrt of C's largest eigenvalue, radius = the larger of 3 smax (a leaf's rule) and the largest
  |mean_k - mean| + radius_k over the children, and the SGGX matrix S_hat (sn) is the mass-weighted mean of the
  children's.

That last rule is checked against the shipped hill_lite_bin tree (a fixture under tests/golden/). The split itself
is a median split of the leaves along the longest axis of their centres' box, level by level.
"""
from __future__ import annotations

import numpy as np

from . import records as R

NO_PART = 0xFFFFFFFF


def _sym(cd, co):
    """(n, 3) diagonal and (n, 3) off-diagonal (xy, xz, yz) -> (n, 3, 3) symmetric matrices."""
    m = np.empty(cd.shape[:-1] + (3, 3))
    m[..., 0, 0], m[..., 1, 1], m[..., 2, 2] = cd[..., 0], cd[..., 1], cd[..., 2]
    m[..., 0, 1] = m[..., 1, 0] = co[..., 0]
    m[..., 0, 2] = m[..., 2, 0] = co[..., 1]
    m[..., 1, 2] = m[..., 2, 1] = co[..., 2]
    return m


def _unsym(m):
    return (np.stack([m[..., 0, 0], m[..., 1, 1], m[..., 2, 2]], -1),
            np.stack([m[..., 0, 1], m[..., 0, 2], m[..., 1, 2]], -1))


def node_moments(nodes: np.ndarray):
    """Float64 moments of stored Node records: (mass, mean, covariance, radius, sn_diag, sn_off)."""
    cinv = _sym(nodes["cinv_diag"].astype(np.float64), nodes["cinv_off"].astype(np.float64))
    cov = np.linalg.inv(cinv)
    mass = nodes["kappa"].astype(np.float64) * np.sqrt(np.maximum(np.linalg.det(cov), 0.0))
    return (mass, nodes["mean"].astype(np.float64), cov, nodes["radius"].astype(np.float64),
            nodes["sn_diag"].astype(np.float64), nodes["sn_off"].astype(np.float64))


def combine(mass, mean, cov, radius, sn_d, sn_o, group):
    """Moment-match children into parents. Arrays are per child; group[i] is child i's parent (0..P-1, every
    parent with at least one child). Returns per parent (mass, mean, cov, radius, sn_d, sn_o)."""
    P = int(group.max()) + 1
    m = np.bincount(group, mass, P)
    w = mass / np.maximum(m[group], 1e-300)
    mu = np.stack([np.bincount(group, w * mean[:, k], P) for k in range(3)], 1)
    d = mean - mu[group]
    spread = cov + d[:, :, None] * d[:, None, :]
    C = np.stack([np.stack([np.bincount(group, w * spread[:, i, j], P) for j in range(3)], 1) for i in range(3)], 1)
    sd = np.stack([np.bincount(group, w * sn_d[:, k], P) for k in range(3)], 1)
    so = np.stack([np.bincount(group, w * sn_o[:, k], P) for k in range(3)], 1)
    reach = np.linalg.norm(d, axis=1) + radius
    rad = np.full(P, -np.inf)
    np.maximum.at(rad, group, reach)
    rad = np.maximum(rad, 3.0 * np.sqrt(np.maximum(np.linalg.eigvalsh(C)[:, -1], 0.0)))   # at least 3 sigma, as a leaf
    return m, mu, C, rad, sd, so


def pack_internal(mass, mean, cov, radius, sn_d, sn_o) -> np.ndarray:
    """Internal Node records (part, children and depth left zero) from float64 moments."""
    n = mass.shape[0]
    out = np.zeros(n, R.NODE_DTYPE)
    det = np.maximum(np.linalg.det(cov), 1e-300)
    cinv = np.linalg.inv(cov)
    out["mean"] = mean
    out["radius"] = radius
    out["cinv_diag"], out["cinv_off"] = _unsym(cinv)
    out["kappa"] = mass / np.sqrt(det)
    out["smax"] = np.sqrt(np.maximum(np.linalg.eigvalsh(cov)[:, -1], 0.0))
    out["sn_diag"], out["sn_off"] = sn_d, sn_o
    return out


def leaf_moments(g: np.ndarray, leaves: np.ndarray):
    """Float64 moments of the leaves from the Gaussian records (mass kappa s0 s1 s2, covariance R diag(s^2) R^T)."""
    q = g["quat"].astype(np.float64)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    Rm = np.empty((q.shape[0], 3, 3))
    Rm[:, 0, 0] = 1 - 2 * (y * y + z * z)
    Rm[:, 1, 0] = 2 * (x * y + w * z)
    Rm[:, 2, 0] = 2 * (x * z - w * y)
    Rm[:, 0, 1] = 2 * (x * y - w * z)
    Rm[:, 1, 1] = 1 - 2 * (x * x + z * z)
    Rm[:, 2, 1] = 2 * (y * z + w * x)
    Rm[:, 0, 2] = 2 * (x * z + w * y)
    Rm[:, 1, 2] = 2 * (y * z - w * x)
    Rm[:, 2, 2] = 1 - 2 * (x * x + y * y)
    s = g["scale"].astype(np.float64)
    cov = np.einsum("nij,nj,nkj->nik", Rm, s * s, Rm)
    mass = g["kappa"].astype(np.float64) * s[:, 0] * s[:, 1] * s[:, 2]
    return (mass, g["mu"].astype(np.float64), cov, leaves["radius"].astype(np.float64),
            leaves["sn_diag"].astype(np.float64), leaves["sn_off"].astype(np.float64))


class Tree:
    """nodes (N x 96 B records, leaves first), kids (u32), per-part roots, the scene root and max_depth (the
    deepest leaf's depth: node_up sums depths max_depth - 1 .. 1, src/lighting.js:763-772, 1091)."""

    def __init__(self, nodes, kids, roots, scene_root, max_depth):
        self.nodes, self.kids, self.roots, self.scene_root, self.max_depth = nodes, kids, roots, scene_root, max_depth

    @property
    def count(self) -> int:
        return int(self.nodes.shape[0])


def build_tree(g: np.ndarray, parts: np.ndarray, leaves: np.ndarray) -> Tree:
    """Binary trees over each part's Gaussians [g_lo, g_hi) plus a scene root (see the module docstring).
    leaves: the G leaf Node records (their depth is set here)."""
    G = g.shape[0]
    mu = g["mu"].astype(np.float64)
    perm = np.arange(G, dtype=np.int64)
    # open segments [lo, hi) of perm with their node id, depth and part; children recorded per internal node
    seg_lo, seg_hi, seg_node, seg_depth, seg_part = [], [], [], [], []
    child_of = []            # (parent internal index, child node id) pair arrays, in order
    int_depth, int_part = [], []
    roots = []
    for p in range(parts.shape[0]):
        lo, hi = int(parts["g_lo"][p]), int(parts["g_hi"][p])
        if hi <= lo:
            roots.append(None)
            continue
        if hi - lo == 1:
            roots.append(lo)
            continue
        k = len(int_depth)
        int_depth.append(1)
        int_part.append(p)
        roots.append(G + k)
        seg_lo.append(lo); seg_hi.append(hi); seg_node.append(k); seg_depth.append(1); seg_part.append(p)
    leaf_depth = np.zeros(G, np.int64)
    for p in range(parts.shape[0]):
        if roots[p] is not None and roots[p] < G:
            leaf_depth[roots[p]] = 1
    lo = np.array(seg_lo, np.int64)
    hi = np.array(seg_hi, np.int64)
    node = np.array(seg_node, np.int64)
    depth = np.array(seg_depth, np.int64)
    part = np.array(seg_part, np.int64)
    while lo.size:
        size = hi - lo
        pos = np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)]) if lo.size < 64 else \
            np.repeat(lo, size) + (np.arange(size.sum()) - np.repeat(np.cumsum(size) - size, size))
        sid = np.repeat(np.arange(lo.size), size)
        pts = mu[perm[pos]]
        bmin = np.stack([np.minimum.reduceat(pts[:, k], np.cumsum(size) - size) for k in range(3)], 1)
        bmax = np.stack([np.maximum.reduceat(pts[:, k], np.cumsum(size) - size) for k in range(3)], 1)
        axis = np.argmax(bmax - bmin, 1)
        key = pts[np.arange(pos.size), axis[sid]]
        order = np.lexsort((perm[pos], key, sid))
        perm[pos] = perm[pos][order]
        mid = lo + size // 2
        nlo, nhi, nnode, ndep, npart = [], [], [], [], []
        for a, b in ((lo, mid), (mid, hi)):
            one = (b - a) == 1
            # single leaves: the child is the Gaussian
            single = np.nonzero(one)[0]
            gid = perm[a[single]]
            child_of.append(np.stack([node[single], gid], 1))
            leaf_depth[gid] = depth[single] + 1
            many = np.nonzero(~one)[0]
            base = len(int_depth)
            child_of.append(np.stack([node[many], G + base + np.arange(many.size)], 1))
            int_depth.extend((depth[many] + 1).tolist())
            int_part.extend(part[many].tolist())
            nlo.append(a[many]); nhi.append(b[many]); nnode.append(base + np.arange(many.size))
            ndep.append(depth[many] + 1); npart.append(part[many])
        lo, hi = np.concatenate(nlo), np.concatenate(nhi)
        node, depth, part = np.concatenate(nnode), np.concatenate(ndep), np.concatenate(npart)
    n_int = len(int_depth)
    N = G + n_int + 1
    scene_root = N - 1
    # kids: each internal node's children contiguous, in the order they were recorded; then the scene root's
    pairs = np.concatenate(child_of).astype(np.int64).reshape(-1, 2) if child_of else np.zeros((0, 2), np.int64)
    order = np.argsort(pairs[:, 0], kind="stable")
    pairs = pairs[order]
    counts = np.bincount(pairs[:, 0], minlength=n_int)
    first = np.cumsum(counts) - counts
    part_roots = [r for r in roots if r is not None]
    kids = np.concatenate([pairs[:, 1], np.array(part_roots, np.int64)]).astype(np.uint32)
    # moments bottom-up: internal nodes deepest first
    lm = leaf_moments(g, leaves)
    mass = np.zeros(N); mean = np.zeros((N, 3)); cov = np.zeros((N, 3, 3)); rad = np.zeros(N)
    snd = np.zeros((N, 3)); sno = np.zeros((N, 3))
    mass[:G], mean[:G], cov[:G], rad[:G], snd[:G], sno[:G] = lm
    idep = np.array(int_depth, np.int64)
    par_of_pair = pairs[:, 0]
    for d in range(int(idep.max()) if n_int else 0, 0, -1):
        at = np.nonzero(idep == d)[0]
        sel = np.isin(par_of_pair, at)
        ch = pairs[sel, 1]
        remap = np.full(n_int, -1, np.int64)
        remap[at] = np.arange(at.size)
        grp = remap[par_of_pair[sel]]
        m, mu_, C, r, sd, so = combine(mass[ch], mean[ch], cov[ch], rad[ch], snd[ch], sno[ch], grp)
        ids = G + at
        mass[ids], mean[ids], cov[ids], rad[ids], snd[ids], sno[ids] = m, mu_, C, r, sd, so
    out = np.zeros(N, R.NODE_DTYPE)
    out[:G] = leaves
    out["depth"][:G] = leaf_depth
    if n_int:
        ids = G + np.arange(n_int)
        rec = pack_internal(mass[ids], mean[ids], cov[ids], rad[ids], snd[ids], sno[ids])
        rec["part"] = np.array(int_part, np.uint32)
        rec["child_lo"] = first.astype(np.uint32)
        rec["child_hi"] = (first + counts).astype(np.uint32)
        rec["depth"] = idep.astype(np.uint32)
        out[G:N - 1] = rec
    pr = np.array(part_roots, np.int64)
    m, mu_, C, r, sd, so = combine(mass[pr], mean[pr], cov[pr], rad[pr], snd[pr], sno[pr], np.zeros(pr.size, np.int64))
    root = pack_internal(m, mu_, C, r, sd, so)
    root["part"] = NO_PART
    root["child_lo"] = pairs.shape[0]
    root["child_hi"] = pairs.shape[0] + pr.size
    root["depth"] = 0
    out[N - 1] = root[0]
    return Tree(out, kids, [(-1 if r is None else int(r)) for r in roots], scene_root, int(leaf_depth.max()))


def scene_tree(sc) -> Tree:
    """The tree of a synthetic scene (scenes.Scene), with parts[].root pointing at each part's subtree."""
    from .scenes import leaf_nodes
    t = build_tree(sc.gaussians, sc.parts, leaf_nodes(sc.gaussians))
    for p, r in enumerate(t.roots):
        if r >= 0:
            sc.parts["root"][p] = r
    return t
