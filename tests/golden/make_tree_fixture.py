"""Write tree_hill_lite.npz: internal nodes of the shipped hill_lite_bin tree (runtime/assets/hill_lite_bin,
nodes.bin + kids.bin, raw 96 B Node records and u32 kids) with their children's records, as inputs and expected
outputs for the moment-matching rule of splatlight_amd/tree.py. Reads the reference's data files only."""
import json
import os
import sys

import numpy as np

SRC = "/root/reference/runtime/assets/hill_lite_bin"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tree_hill_lite.npz")


def main(n=256, seed=0):
    h = json.load(open(os.path.join(SRC, "asset.json")))
    raw = np.fromfile(os.path.join(SRC, "nodes.bin"), dtype="<u4").reshape(-1, 24)
    kids = np.fromfile(os.path.join(SRC, "kids.bin"), dtype="<u4")
    G = h["gaussian_count"]
    rng = np.random.default_rng(seed)
    cand = np.arange(G, raw.shape[0] - 1)                      # internal nodes, not the scene root
    depth = raw[cand, 21]
    pick = []
    for d in np.unique(depth):                                 # every depth, then random ones
        pick.append(rng.choice(cand[depth == d], size=min(4, int((depth == d).sum())), replace=False))
    pick = np.unique(np.concatenate(pick + [rng.choice(cand, n, replace=False)]))
    lo, hi = raw[pick, 19], raw[pick, 20]
    ch = np.concatenate([kids[a:b] for a, b in zip(lo, hi)])
    group = np.repeat(np.arange(pick.size), hi - lo)
    np.savez_compressed(OUT, parents=raw[pick], children=raw[ch], group=group.astype(np.int64),
                        parent_ids=pick.astype(np.int64), child_ids=ch.astype(np.int64))
    print(OUT, pick.size, "parents", ch.size, "children")


if __name__ == "__main__":
    sys.exit(main())
