"""Write asset_cut.npz: small cuts of the reference's two shipped assets (runtime/assets/hill_lite_bin and
runtime/assets/cornell_sdf_room_plain_bin), with everything the relighting reads per Gaussian as the assets store
it: the 64 B Gaussian records, their baked leaf Node records (96 B: C^-1, kappa and the SGGX matrix S_hat), their
baked 8x16 visibility grids (128 B), the part rows and the material table, plus the asset's own camera, emitter
part, room box and environment (sun and sky SH). Reads the reference's data files only (raw little-endian records and
asset.json); the tests build their scenes from this file and never read the reference.

hill: the ground (part 0), the grass (part 1) and the tree (part 2), the 1,000 Gaussians of each nearest a point
near the tree's foot, and 300 of a far part (part 3); cornell: 350 random Gaussians of each of the room's seven
surface parts and the whole lamp (part 7, the asset's emitter part). Within a part the cut keeps the asset's order.
"""
import json
import os
import sys

import numpy as np

SRC = "/root/reference/runtime/assets"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "asset_cut.npz")


def load(name):
    d = os.path.join(SRC, name)
    h = json.load(open(os.path.join(d, "asset.json")))
    G = h["gaussian_count"]
    g = np.fromfile(os.path.join(d, "gaussians.bin"), dtype="<u4").reshape(G, 16)
    nodes = np.fromfile(os.path.join(d, "nodes.bin"), dtype="<u4").reshape(-1, 24)[:G]   # leaves first
    vg = np.fromfile(os.path.join(d, "vgrid.bin"), dtype=np.uint8).reshape(G, 128)
    parts = np.fromfile(os.path.join(d, "parts.bin"), dtype="<u4").reshape(-1, 32)
    mats = np.fromfile(os.path.join(d, "materials.bin"), dtype="<u4").reshape(-1, 8)
    return h, g, nodes, vg, parts, mats


def cut(prefix, name, picks):
    """picks: [(part, count, point or None)]: the count Gaussians of the part nearest the point (None: random)."""
    h, g, nodes, vg, parts, mats = load(name)
    rng = np.random.default_rng(0)
    mu = g[:, 0:3].view(np.float32).astype(np.float64)
    ids, part_of = [], []
    for k, (p, n, pt) in enumerate(picks):
        lo, hi = int(parts[p, 29]), int(parts[p, 30])
        cand = np.arange(lo, hi)
        if n >= cand.size:
            sel = cand
        elif pt is None:
            sel = np.sort(rng.choice(cand, n, replace=False))
        else:
            d = np.linalg.norm(mu[cand] - np.asarray(pt), axis=1)
            sel = np.sort(cand[np.argsort(d, kind="stable")[:n]])
        ids.append(sel)
        part_of.append(np.full(sel.size, k, np.int64))
    ids = np.concatenate(ids)
    cam = h["camera"]
    out = {
        prefix + "gaussians": g[ids], prefix + "nodes": nodes[ids], prefix + "vgrid": vg[ids],
        prefix + "part_of": np.concatenate(part_of), prefix + "src_index": ids.astype(np.int64),
        prefix + "parts": parts[[p for p, _, _ in picks]], prefix + "part_ids": np.array([p for p, _, _ in picks]),
        prefix + "materials": mats,
        prefix + "camera": np.array([*cam["eye"], *cam["look_at"], cam["fov_deg"]], np.float64),
    }
    env = h.get("environment")
    if env:
        sun = env["lights"][0]
        d = np.asarray(sun["dir"], np.float64)
        out[prefix + "sun"] = np.array([*(d / np.linalg.norm(d)), *sun["irradiance_rgb"],
                                        sun.get("angular_std_deg", 0.27)], np.float64)
        out[prefix + "sky_sh"] = np.asarray(env["sky"]["sh_rgb"][:16], np.float64)
    if h.get("room"):
        out[prefix + "room"] = np.array([*h["room"]["x"], *h["room"]["y"], *h["room"]["z"]], np.float64)
    em = h.get("emitters")
    if em:
        out[prefix + "emitter"] = np.array([[p for p, _, _ in picks].index(em["part"]), *em["emission"]], np.float64)
    return out


def main():
    arrs = {}
    foot = (0.5, -3.0, 2.5)
    arrs.update(cut("hill_", "hill_lite_bin", [(0, 1000, foot), (1, 1000, foot), (2, 1000, (0.5, -2.0, 4.5)),
                                               (3, 300, (19.2, 3.6, 18.8))]))
    arrs.update(cut("cornell_", "cornell_sdf_room_plain_bin", [(p, 350, None) for p in range(7)] + [(7, 320, None)]))
    np.savez_compressed(OUT, **arrs)
    print(OUT, os.path.getsize(OUT), "bytes;", {k: v.shape for k, v in arrs.items() if k.endswith("gaussians")})


if __name__ == "__main__":
    sys.exit(main())
