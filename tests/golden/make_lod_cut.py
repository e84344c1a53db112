"""Write lod_cut.npz: whole subtrees of the reference's two shipped trees (runtime/assets/hill_lite_bin and
cornell_sdf_room_plain_bin), for the screen-space LOD and display-fill tests. Reads the reference's data files only
(raw little-endian records, asset.json and onepath.json/.bin); the tests never read the reference.

For each asset a few subtrees are cut whole: from a part's root, the descent follows the child with more leaves until
the subtree holds at most `target` leaves. hill: a subtree of the ground, the grass and the tree (about 2,000
leaves each); cornell: a subtree of the floor and of the short box, and the whole lamp part (the asset's emitter
part, 320 leaves). Each subtree becomes a part of the cut (its part record with root and Gaussian range moved to
the cut's indices). The cut is a self-contained tree: leaves first (by part, then asset order), internal nodes after
(asset order), the kids list rebuilt in the same child order, node and Gaussian part words renumbered, depths kept.
cornell also keeps the display-fill rows (onepath.bin `fill`) of the cut's leaves whose sources all lie in the cut,
with the source indices renumbered. The asset.json headers are stored as text (for the loader tests).
"""
import json
import os

import numpy as np

SRC = "/root/reference/runtime/assets"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lod_cut.npz")


def load(name):
    d = os.path.join(SRC, name)
    h = json.load(open(os.path.join(d, "asset.json")))
    G = h["gaussian_count"]
    g = np.fromfile(os.path.join(d, "gaussians.bin"), dtype="<u4").reshape(G, 16)
    nodes = np.fromfile(os.path.join(d, "nodes.bin"), dtype="<u4").reshape(-1, 24)
    kids = np.fromfile(os.path.join(d, "kids.bin"), dtype="<u4")
    parts = np.fromfile(os.path.join(d, "parts.bin"), dtype="<u4").reshape(-1, 32)
    mats = np.fromfile(os.path.join(d, "materials.bin"), dtype="<u4").reshape(-1, 8)
    return h, g, nodes, kids, parts, mats


def leaf_counts(nodes, kids, G):
    N = nodes.shape[0]
    cnt = np.zeros(N, np.int64)
    cnt[:G] = 1
    order = np.argsort(-nodes[G:, 21].astype(np.int64), kind="stable") + G   # deepest first
    for n in order:
        lo, hi = nodes[n, 19], nodes[n, 20]
        cnt[n] = cnt[kids[lo:hi]].sum()
    return cnt


def pick(nodes, kids, cnt, root, target):
    n = root
    while cnt[n] > target:
        ch = kids[nodes[n, 19]:nodes[n, 20]]
        n = int(ch[np.argmax(cnt[ch])])
    return n


def subtree(nodes, kids, G, r):
    out, stack = [], [r]
    while stack:
        n = stack.pop()
        out.append(n)
        if n >= G:
            stack.extend(int(c) for c in kids[nodes[n, 19]:nodes[n, 20]])
    return out


def cut(name, picks, fill_meta=None):
    h, g, nodes, kids, parts, mats = load(name)
    G = g.shape[0]
    cnt = leaf_counts(nodes, kids, G)
    roots, src_parts = [], []
    for p, target in picks:
        roots.append(pick(nodes, kids, cnt, int(parts[p, 28]), target))
        src_parts.append(p)
    leaves, inner = [], []
    for k, r in enumerate(roots):
        s = subtree(nodes, kids, G, r)
        leaves += [(k, n) for n in sorted(x for x in s if x < G)]
        inner += [(k, n) for n in s if n >= G]
    inner.sort(key=lambda t: t[1])
    new = {}
    for i, (_, n) in enumerate(leaves):
        new[n] = i
    for i, (_, n) in enumerate(inner):
        new[n] = len(leaves) + i
    Gc, Nc = len(leaves), len(leaves) + len(inner)
    nd = np.zeros((Nc, 24), np.uint32)
    gs = np.zeros((Gc, 16), np.uint32)
    part_of = {}
    for k, n in leaves + inner:
        part_of[n] = k
    for i, (k, n) in enumerate(leaves):
        gs[i] = g[n]
        gs[i, 11] = (g[n, 11] & 0xFFFF0000) | k
    kid_list = []
    for k, n in leaves + inner:
        row = nodes[n].copy()
        row[15] = k
        if n >= G:
            ch = kids[nodes[n, 19]:nodes[n, 20]]
            row[19] = len(kid_list)
            kid_list += [new[int(c)] for c in ch]
            row[20] = len(kid_list)
        nd[new[n]] = row
    pr = np.zeros((len(roots), 32), np.uint32)
    for k, (r, p) in enumerate(zip(roots, src_parts)):
        pr[k] = parts[p]
        idx = [i for i, (kk, _) in enumerate(leaves) if kk == k]
        pr[k, 28], pr[k, 29], pr[k, 30] = new[r], idx[0], idx[-1] + 1
    out = {"gaussians": gs, "nodes": nd, "kids": np.array(kid_list, np.uint32), "parts": pr, "materials": mats,
           "src_leaves": np.array([n for _, n in leaves], np.int64), "src_parts": np.array(src_parts, np.int64),
           "camera": np.array([*h["camera"]["eye"], *h["camera"]["look_at"], h["camera"]["fov_deg"]], np.float64),
           "header": np.frombuffer(json.dumps(h).encode(), np.uint8)}
    if fill_meta:
        d = os.path.join(SRC, name)
        meta = json.load(open(os.path.join(d, "onepath.json")))
        with open(os.path.join(d, "onepath.bin"), "rb") as f:
            f.seek(meta["fill"]["offset"])
            fill = np.frombuffer(f.read(meta["fill"]["bytes"]), "<u4")
        off = fill[:G + 1]
        pairs = fill[G + 1:].reshape(-1, 2)
        offs, rows = [0], []
        for i, (_, n) in enumerate(leaves):
            prs = pairs[off[n]:off[n + 1]]
            if len(prs) and all(int(j) in new and new[int(j)] < Gc for j in prs[:, 0]):
                rows += [(new[int(j)], int(w)) for j, w in prs]
            offs.append(len(rows))
        out["fill"] = np.concatenate([np.array(offs, np.uint32), np.array(rows, np.uint32).reshape(-1)])
        out["fill_rows_full"] = np.array([int(off[G]), int((np.diff(off) > 0).sum())], np.int64)
    return out


def main():
    hill = cut("hill_lite_bin", [(0, 2000), (1, 2000), (2, 2500)])
    cornell = cut("cornell_sdf_room_plain_bin", [(0, 1500), (5, 1500), (7, 400)], fill_meta=True)
    np.savez_compressed(OUT, **{f"hill_{k}": v for k, v in hill.items()}, **{f"cornell_{k}": v for k, v in cornell.items()})


if __name__ == "__main__":
    main()
