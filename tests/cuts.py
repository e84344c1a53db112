"""Scenes made of the cuts of the reference's shipped assets in golden/asset_cut.npz (written by
golden/make_asset_cut.py): the assets' own Gaussian records, baked leaf Nodes (SGGX S_hat) and visibility grids,
parts and materials, so the relighting runs on real V grids and real S_hat instead of V = 1 and S_hat = I.

* hill: ground, grass and tree near the tree's foot plus a far part, lit by the asset's own sun and sky SH
  (asset.json environment) and 8 lanterns placed over the cut (the asset has no emitter part).
* cornell: the room's seven surfaces and its lamp (part 7), lit by the lamp itself as setAssetEmitters does
  (src/lighting.js:690-721): the lamp's Gaussians are the emitter set (emission from asset.json, one-sided as the
  lamp's material says), drawn with the scene and skipped by the emitter map; the asset's room box.

The cut's parts keep the shipped part rows (poses, bounds, flags) with their Gaussian ranges renumbered; each
Gaussian's part index (part_surface low 16 bits, and its Node's part) is renumbered with them. matte=True clears the
specular flag (Material.flags bit 3) of the hill's grass and leaf materials.
"""
import math
import os

import numpy as np

from splatlight_amd import lighting as L
from splatlight_amd import records as R
from splatlight_amd import scenes

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "asset_cut.npz")


def load(name, width=320, height=240, matte=True):
    """(scenes.Scene, leaf Node records, V grids) of the `name` cut ("hill" or "cornell")."""
    d = np.load(FIXTURE)
    p = name + "_"
    g = np.ascontiguousarray(d[p + "gaussians"]).view(R.GAUSSIAN_DTYPE).reshape(-1).copy()
    nd = np.ascontiguousarray(d[p + "nodes"]).view(R.NODE_DTYPE).reshape(-1).copy()
    vg = np.ascontiguousarray(d[p + "vgrid"]).copy()
    part_of = d[p + "part_of"]
    parts = np.ascontiguousarray(d[p + "parts"]).view(R.PART_DTYPE).reshape(-1).copy()
    mats = np.ascontiguousarray(d[p + "materials"]).view(R.MATERIAL_DTYPE).reshape(-1).copy()
    if matte:
        mats["flags"] &= ~np.uint32(R.MAT_SPECULAR)
    counts = np.bincount(part_of, minlength=parts.shape[0])
    lo = np.concatenate([[0], np.cumsum(counts)[:-1]])
    parts["g_lo"] = lo
    parts["g_hi"] = lo + counts
    parts["root"] = 0
    g["part_surface"] = (g["part_surface"] & np.uint32(0xFFFF0000)) | part_of.astype(np.uint32)
    nd["part"] = part_of.astype(np.uint32)
    cam = d[p + "camera"]
    mu = g["mu"].astype(np.float64)
    far = max(50.0, 1.5 * float(np.linalg.norm(mu.max(0) - mu.min(0))))
    if name == "hill":
        near = mu[part_of < 3]
        target = tuple(float(v) for v in near.mean(0))
        eye = (target[0] + 3.0, target[1] - 5.5, target[2] + 2.0)
        sun = d[p + "sun"]
        rng = np.random.default_rng(17)
        lo3, hi3 = near.min(0), near.max(0)
        emu = np.column_stack([rng.uniform(lo3[0], hi3[0], 8), rng.uniform(lo3[1], hi3[1], 8),
                               rng.uniform(lo3[2] + 0.3, hi3[2] + 1.0, 8)])
        le = rng.uniform(1.0, 20.0, 8)
        em = scenes._emitter_records(emu, np.repeat(le[:, None], 3, 1), (0.03, 0.03, 0.03))
        sc = scenes.Scene("hill_cut", g, parts, mats, em, eye, target, 40.0, width, height, far,
                          sun={"dir": tuple(float(v) for v in sun[:3]), "E": tuple(float(v) for v in sun[3:6]),
                               "std": math.radians(float(sun[6])), "res": 512},
                          sky=d[p + "sky_sh"].copy())
    else:
        em_part, emission = int(d[p + "emitter"][0]), [float(v) for v in d[p + "emitter"][1:4]]
        lamp_mat = int(g["flags"][parts["g_lo"][em_part]])
        es = L.asset_emitters(g, int(parts["g_lo"][em_part]), int(parts["g_hi"][em_part]), emission)
        es["part"] = em_part
        rm = d[p + "room"]
        sc = scenes.Scene("cornell_cut", g, parts, mats, np.zeros(0, R.GAUSSIAN_DTYPE), tuple(cam[:3]),
                          tuple(cam[3:6]), float(cam[6]), width, height, far,
                          one_sided=bool(mats["flags"][lamp_mat] & R.MAT_ONE_SIDED), emit_set=es,
                          room={"x": [rm[0], rm[1]], "y": [rm[2], rm[3]], "z": [rm[4], rm[5]]})
    return sc, nd, vg
