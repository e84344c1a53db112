"""gs-runtime asset directories made from the LOD fixture's cuts of the shipped hill and Cornell assets
(tests/golden/lod_cut.npz), for the JavaScript host tests: asset.json is the shipped header with the cut's buffer
counts, its part list (names, mobility, Gaussian ranges) and the emitters' part renumbered; the buffers are the
cut's records, its tree, open V grids (every cell 255), one disabled lamp record, and for Cornell the cut's display
fill as onepath.json / onepath.bin."""
import json
import os

import numpy as np

import lodcut


def write(name, out_dir, matte=False):
    """Writes the `name` cut ("hill" or "cornell") as an asset directory; returns (header, raw dict of arrays).
    matte: clear the specular flag (Material.flags bit 3) of every material (the hill's grass and leaf), for the
    probe bounce, whose specular receivers are not on this path."""
    r = lodcut.raw(name)
    h = json.loads(r["header"].tobytes().decode())
    g, nodes, kids = r["gaussians"], r["nodes"], r["kids"]
    parts, mats, src = r["parts"], r["materials"].copy(), [int(p) for p in r["src_parts"]]
    if matte:
        mats[:, 5] &= ~np.uint32(8)
    G, N = g.shape[0], nodes.shape[0]
    vgrid = np.full((G, 128), 255, np.uint8)
    lamps = np.zeros((1, 24), np.uint32)
    bufs = {"gaussians": g, "parts": parts, "lamps": lamps, "materials": mats, "nodes": nodes, "kids": kids,
            "vgrid": vgrid}
    for k, v in bufs.items():
        h["buffers"][k]["count"] = int(v.shape[0])
    hp = []
    for k, p in enumerate(src):
        e = dict(h["parts"][p])
        e["gaussians"] = [int(parts[k, 29]), int(parts[k, 30])]
        hp.append(e)
    h["parts"] = hp
    h["gaussian_count"], h["node_count"], h["part_count"] = G, N, len(src)
    if h.get("emitters"):
        h["emitters"] = dict(h["emitters"], part=src.index(int(h["emitters"]["part"])),
                             gaussians=int(parts[src.index(int(h["emitters"]["part"])), 30] -
                                           parts[src.index(int(h["emitters"]["part"])), 29]))
    h.pop("blockers", None)
    os.makedirs(out_dir, exist_ok=True)
    for k, v in bufs.items():
        with open(os.path.join(out_dir, h["buffers"][k]["file"]), "wb") as f:
            f.write(np.ascontiguousarray(v).tobytes())
    fill = r.get("fill")
    if fill is not None:
        with open(os.path.join(out_dir, "onepath.bin"), "wb") as f:
            f.write(fill.astype("<u4").tobytes())
        with open(os.path.join(out_dir, "onepath.json"), "w") as f:
            json.dump({"G": G, "fill": {"offset": 0, "bytes": int(fill.size * 4)}}, f)
    with open(os.path.join(out_dir, "asset.json"), "w") as f:
        json.dump(h, f)
    return h, dict(bufs, fill=fill)
