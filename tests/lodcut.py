"""Scenes from the LOD fixture (tests/golden/lod_cut.npz): whole subtrees of the reference's shipped hill and Cornell
trees with their Gaussians, parts, materials and camera, plus the Cornell cut's display-fill rows."""
import json
import math
import os

import numpy as np

from splatlight_amd import records as R
from splatlight_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
_cache = {}


def raw(name):
    if "d" not in _cache:
        _cache["d"] = np.load(os.path.join(HERE, "golden", "lod_cut.npz"))
    d = _cache["d"]
    return {k[len(name) + 1:]: d[k] for k in d.files if k.startswith(name + "_")}


def load(name, width=640, height=360):
    """(Scene, nodes NODE_DTYPE [N], kids u32, header dict, emit part)."""
    r = raw(name)
    g = np.frombuffer(r["gaussians"].tobytes(), R.GAUSSIAN_DTYPE).copy()
    parts = np.frombuffer(r["parts"].tobytes(), R.PART_DTYPE).copy()
    mats = np.frombuffer(r["materials"].tobytes(), R.MATERIAL_DTYPE).copy()
    nodes = np.frombuffer(r["nodes"].tobytes(), R.NODE_DTYPE).copy()
    header = json.loads(r["header"].tobytes().decode())
    cam = r["camera"]
    mu = g["mu"].astype(np.float64)
    far = max(50.0, 1.5 * float(np.linalg.norm(mu.max(0) - mu.min(0))))
    sc = scenes.Scene(name, g, parts, mats, np.zeros(0, R.GAUSSIAN_DTYPE), tuple(cam[:3]), tuple(cam[3:6]),
                      float(cam[6]), width, height, far)
    emit = 0xFFFFFFFF
    if header.get("emitters"):
        src = int(header["emitters"]["part"])
        hits = np.nonzero(r["src_parts"] == src)[0]
        emit = int(hits[0]) if hits.size else 0xFFFFFFFF
    return sc, nodes, r["kids"].copy(), header, emit


def fill(name):
    return raw(name).get("fill")


def lod_frame(sc, threshold, n_nodes, eye=None, view_mode=R.VIEW_LIT, bounce_gain=1.0):
    v = R.look_at(eye or sc.eye, sc.target, (0.0, 0.0, 1.0))
    return R.render_frame(view=v, width=sc.width, height=sc.height, fovy=math.radians(sc.fov_deg), far=sc.far,
                          count=sc.count, view_mode=view_mode, bounce_gain=bounce_gain,
                          lod=(threshold, sc.count + 1, n_nodes))
