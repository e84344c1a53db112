"""Sky and bounce through probes: the host side of Lighting.initBounce / encodeBounce (src/lighting.js:170-316,
723-840, 1020-1103) for the default path (no sender LOD, no near field, every probe ray and Gaussian refreshed
each frame, probe gathers taken as they are).

The device side (sl_bounce_init / sl_bounce_encode) runs runtime/wgsl/bounce.wgsl's passes; this module keeps the
reference's parameter logic: the probe layout (room lattice or terrain columns, part probes, static volumes), the
14 fixed directions and their SH2 least-squares fit, and the ProbeParams bytes.
"""
from __future__ import annotations

import math

import numpy as np

from . import records as R

NCELL = 128
PPK = 8                        # probes per movable part (src/lighting.js:176)
PROBES_AXIS = 8                # room lattice per axis (src/lighting.js:177)
PROBE_THETA = 4.0
TMAX = 6.0
NEAR_R = 0.3
FULL_ROUNDS = 8
HIT_MAX = 16                   # probe_gather's hit list (runtime/wgsl/bounce.wgsl:21)
LOB = 3                        # lo_b vec4s per Gaussian (runtime/wgsl/common.wgsl:54-57)
PROBE_PARAMS_BYTES = 832 + 16 + 64 + 4096 + 528 + 4096      # src/lighting.js:173
PART_STATIC = 1                # Part.flags bit 0: static (src/lighting.js:298)
PART_MOVED = 2                 # Part.flags bit 1: moved this frame (bounce.wgsl iso_xfer)

# D14: 6 axes + 8 diagonals; the vertical axes lean 1e-4 off the pole (src/lighting.js:186-187)
D14 = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1e-4, 0.0, 1.0], [-1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [1e-4, 0.0, -1.0]] + \
    [[x / math.sqrt(3), y / math.sqrt(3), z / math.sqrt(3)] for x in (1, -1) for y in (1, -1) for z in (1, -1)]


def sh2(d):
    """Real SH bands 0-2 (src/lighting.js:188-189)."""
    x, y, z = d
    return [0.282095, 0.488603 * y, 0.488603 * z, 0.488603 * x, 1.092548 * x * y, 1.092548 * y * z,
            0.315392 * (3 * z * z - 1), 1.092548 * x * z, 0.546274 * (x * x - y * y)]


def pinv_d14():
    """The 9 x 14 least-squares SH2 fit of values on D14 (src/lighting.js:190-206): Gauss-Jordan with partial
    pivoting on [A^T A | I], then (A^T A)^-1 A^T, in double."""
    A = [sh2(d) for d in D14]
    M = [[sum(r[i] * r[j] for r in A) for j in range(9)] + [1.0 if i == j else 0.0 for j in range(9)]
         for i in range(9)]
    for c in range(9):
        piv = c
        for r in range(c + 1, 9):
            if abs(M[r][c]) > abs(M[piv][c]):
                piv = r
        M[c], M[piv] = M[piv], M[c]
        d = M[c][c]
        M[c] = [v / d for v in M[c]]
        for r in range(9):
            if r != c:
                f = M[r][c]
                M[r] = [a - f * b for a, b in zip(M[r], M[c])]
    inv = [row[9:] for row in M]
    return [[sum(inv[m][j] * row[j] for j in range(9)) for row in A] for m in range(9)]


def cell_table():
    """The 128 V-grid cells' centre directions and solid angles (bounce.wgsl cell_dir / cell_omega, 171-197):
    8 rows of elevation from the zenith x 16 columns of azimuth from -pi. Computed in double and rounded to float
    once (WGSL's f32 cos / sin are implementation-defined); the device uses this table."""
    dirs = np.zeros((NCELL, 3), np.float32)
    om = np.zeros(NCELL, np.float32)
    for c in range(NCELL):
        i, j = c // 16, c % 16
        el = 0.5 * math.pi - (i + 0.5) * math.pi / 8.0
        az = -math.pi + (j + 0.5) * 2.0 * math.pi / 16.0
        dirs[c] = [math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)]
        top = 0.5 * math.pi - i * math.pi / 8.0
        om[c] = (2.0 * math.pi / 16.0) * (math.sin(top) - math.sin(top - math.pi / 8.0))
    return dirs, om


def scene_header(sc) -> dict:
    """The asset.json fields probeLayout reads, for a synthetic scene: a room (the Cornell box) or bounds (the
    Gaussian centres' box), the parts' mobility (static) and whether it has an environment (sun or sky)."""
    parts = [{"name": f"part{p}", "mobility": "static" if sc.parts["flags"][p] & PART_STATIC else "movable",
              "gaussians": [int(sc.parts["g_lo"][p]), int(sc.parts["g_hi"][p])]} for p in range(sc.parts.shape[0])]
    h = {"parts": parts, "environment": sc.sun is not None or getattr(sc, "sky", None) is not None,
         "probes": {}}
    if getattr(sc, "room", None) is not None:
        h["room"] = dict(sc.room)
    elif sc.name == "cornell":
        h["room"] = {"x": [-1.0, 1.0], "y": [-1.0, 1.0], "z": [0.0, 2.0]}
    else:
        mu = sc.gaussians["mu"].astype(np.float64)
        h["bounds"] = {"lo": mu.min(0).tolist(), "hi": mu.max(0).tolist()}
    return h


def probe_layout(sc, header=None) -> dict:
    """probeLayout (src/lighting.js:220-290): the room lattice, or terrain columns with ground heights from the
    terrain part (the static part with the widest footprint) plus probe volumes on static objects."""
    h = header or scene_header(sc)
    pr = h.get("probes", {})
    terrain = pr.get("layout") == "terrain" or (not h.get("room") and bool(h.get("bounds")) and pr.get("layout") != "room")
    if h.get("room"):
        box = h["room"]
    elif h.get("bounds"):
        b = h["bounds"]
        box = {"x": [b["lo"][0], b["hi"][0]], "y": [b["lo"][1], b["hi"][1]], "z": [b["lo"][2], b["hi"][2]]}
    else:
        raise ValueError("asset.json has neither room nor bounds (probe lattice)")
    ax = [box[k] for k in ("x", "y", "z")]
    tmax = max(TMAX, math.hypot(*[b - a for a, b in ax])) if h.get("environment") else TMAX
    if not terrain:
        n = PROBES_AXIS
        cen = [(a + b) / 2 for a, b in ax]
        half = [(b - a) / 2 - 0.125 for a, b in ax]
        lo = [c - hh for c, hh in zip(cen, half)]
        step = [2 * v / (n - 1) for v in half]
        return {"terrain": False, "dims": [n, n, n], "lo": lo, "step": step,
                "levels": [lo[2] + k * step[2] for k in range(n)], "ground": np.zeros(n * n, np.float32),
                "tmax": tmax, "vols": []}
    dims = list(pr.get("dims", [16, 16, 5]))
    if dims[0] * dims[1] > 1024 or dims[2] > 16:
        raise ValueError("probe terrain layout: at most 1024 columns and 16 levels")
    levels = pr.get("heights_m", [0.3, 1, 3, 8, 20])
    dims[2] = len(levels)
    step = [(ax[k][1] - ax[k][0]) / dims[k] for k in (0, 1)]
    lo = [ax[0][0] + step[0] / 2, ax[1][0] + step[1] / 2, 0.0]
    gf = sc.gaussians["mu"]
    tp = []
    best, best_a = -1, -1.0
    for i, p in enumerate(h["parts"]):
        if p["mobility"] == "movable":
            continue
        g0, g1 = p["gaussians"]
        idx = np.arange(g0, g1, max(1, (g1 - g0) >> 12))
        if idx.size == 0:
            continue
        x, y = gf[idx, 0].astype(np.float64), gf[idx, 1].astype(np.float64)
        a = (x.max() - x.min()) * (y.max() - y.min())
        if a > best_a:
            best_a, best = a, i
    tp = [best] if best >= 0 else []
    ground = np.full(dims[0] * dims[1], -np.inf, np.float32)
    for i in tp:
        g0, g1 = h["parts"][i]["gaussians"]
        m = gf[g0:g1]
        ci = np.floor((m[:, 0].astype(np.float64) - ax[0][0]) / step[0]).astype(np.int64)
        cj = np.floor((m[:, 1].astype(np.float64) - ax[1][0]) / step[1]).astype(np.int64)
        ok = (ci >= 0) & (cj >= 0) & (ci < dims[0]) & (cj < dims[1])
        np.maximum.at(ground, ci[ok] * dims[1] + cj[ok], m[ok, 2].astype(np.float32))
    for _ in range(dims[0] + dims[1]):
        if not np.isinf(ground).any():
            break
        prev = ground.copy()
        for i in range(dims[0]):
            for j in range(dims[1]):
                if prev[i * dims[1] + j] != -np.inf:
                    continue
                s, c = 0.0, 0
                for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                    a, b = i + di, j + dj
                    if a < 0 or b < 0 or a >= dims[0] or b >= dims[1] or prev[a * dims[1] + b] == -np.inf:
                        continue
                    s += float(prev[a * dims[1] + b])
                    c += 1
                if c:
                    ground[i * dims[1] + j] = s / c
    ground[np.isinf(ground)] = ax[2][0]
    return {"terrain": True, "dims": dims, "lo": lo, "step": [*step, 0.0], "levels": list(levels), "ground": ground,
            "tmax": tmax, "vols": probe_volumes(sc, h, tp, ax)}


def probe_volumes(sc, h, terrain_parts, ax) -> list:
    """probeVolumes (src/lighting.js:295-316): every static part that is not terrain and whose bounding radius is
    1 m to a quarter of the scene's half-diagonal gets a grid of about 1.2 m spacing, 2-6 per axis."""
    pr = h.get("probes", {})
    if pr.get("volumes"):
        return [{"part": v["part"], "dims": v["dims"]} for v in pr["volumes"] if v["part"] < 32]
    half_diag = 0.5 * math.hypot(*[b - a for a, b in ax])
    out = []
    for q in range(min(sc.parts.shape[0], 32)):
        if q in terrain_parts or h["parts"][q]["mobility"] == "movable":
            continue
        Rb = float(sc.parts["bound"][q][3])
        if Rb < 1 or Rb > 0.25 * half_diag:
            continue
        n = max(2, min(6, round(Rb / 1.2)))
        out.append({"part": q, "dims": [n, n, n]})
    return out


def probe_count(layout, part_count) -> int:
    vol = sum(v["dims"][0] * v["dims"][1] * v["dims"][2] for v in layout["vols"])
    d = layout["dims"]
    return d[0] * d[1] * d[2] + PPK * part_count + vol


def probe_params(layout, node_count, part_count) -> np.ndarray:
    """ProbeParams words as initBounce writes them (src/lighting.js:776-797); the per-frame words (alpha,
    force_near, subset, phase) are set by frame_params."""
    u = np.zeros(PROBE_PARAMS_BYTES // 4, np.uint32)
    f = u.view(np.float32)
    f[0:3] = layout["lo"]
    u[3] = layout["dims"][0]
    f[4:7] = layout["step"]
    u[7] = node_count
    f[8:12] = [0.5 * math.hypot(*layout["step"]), PROBE_THETA, layout["tmax"], NEAR_R]
    u[208:212] = [*layout["dims"], 1 if layout["terrain"] else 0]
    f[212:212 + len(layout["levels"])] = layout["levels"]
    f[228:228 + len(layout["ground"])] = layout["ground"]
    off = PPK * part_count
    vol = 0
    for v in layout["vols"]:
        u[1252 + 4 * v["part"]: 1256 + 4 * v["part"]] = [off, *v["dims"]]
        off += v["dims"][0] * v["dims"][1] * v["dims"][2]
        vol += v["dims"][0] * v["dims"][1] * v["dims"][2]
    u[1380] = PPK * part_count
    u[1381] = 0
    f[12:16] = [1.0, 0.0, 0.02, 0.95]
    u[16] = 1
    u[18] = PPK * part_count + vol
    u[19] = 1
    for k, d in enumerate(D14):
        f[20 + 4 * k: 23 + 4 * k] = d
    f[76:76 + 126] = np.array(pinv_d14(), np.float64).ravel()
    return u


def frame_params(base: np.ndarray, round_: int, first: bool, force_near: bool, subset: int = 1) -> bytes:
    """The words encodeBounce rewrites each frame (src/lighting.js:1024-1034): alpha (1: each gather taken as is),
    force_near, subset and phase (every ray while round < FULL_ROUNDS), the near-field reflection words off."""
    u = base.copy()
    f = u.view(np.float32)
    f[12] = 1.0
    u[13] = 1 if force_near else 0
    u[205] = 0
    sub = 1 if round_ < FULL_ROUNDS else subset
    u[16] = sub
    u[17] = round_ % sub
    u[204] = 0
    u[206] = 8
    u[207] = 0
    return u.tobytes()


def up_levels(nodes: np.ndarray, gaussian_count: int, max_depth: int):
    """node_up's lists (src/lighting.js:762-772): the internal nodes of each depth <= max_depth that belong to a part,
    concatenated by depth, with (depth, offset, count) per depth."""
    d = nodes["depth"][gaussian_count:].astype(np.int64)
    ok = (nodes["part"][gaussian_count:] != 0xFFFFFFFF) & (d <= max_depth)
    ids = np.nonzero(ok)[0] + gaussian_count
    dd = d[ok]
    order = np.argsort(dd, kind="stable")
    ids, dd = ids[order], dd[order]
    counts = np.bincount(dd, minlength=max_depth + 1)[:max_depth + 1]
    offs = np.cumsum(counts) - counts
    return ids.astype(np.uint32), [(k, int(offs[k]), int(counts[k])) for k in range(max_depth + 1)]


class Bounce:
    """Lighting.initBounce / encodeBounce (src/lighting.js:723-840, 1020-1103) over the HIP library, default path:
    the probe layout and ProbeParams on the host, every pass on the device. The scene's tree comes from
    tree.scene_tree (which also points each part's root at its subtree); encode() follows a lighting encode of the
    same frame."""

    def __init__(self, scene_native, sc, tree):
        self.scene = scene_native
        self.sc = sc
        self.tree = tree
        self.layout = probe_layout(sc)
        P = sc.parts.shape[0]
        self.n_probes = probe_count(self.layout, P)
        self.base = probe_params(self.layout, tree.count, P)
        scene_native.write_parts(sc.parts)
        scene_native.bounce_init(tree.nodes, tree.kids, tree.max_depth, frame_params(self.base, 0, True, True))
        self.round = 0
        self.force_near = True
        self.first = True

    def params(self) -> bytes:
        return frame_params(self.base, self.round, self.first, self.force_near)

    def encode(self, light_still=False, moved_mask=0, timers=False):
        """One encodeBounce: light_still skips to the eye-dependent read; the transfer is rebuilt on the first
        frame (force_near) or when a part moved (moved_mask)."""
        from . import native as N
        light_still = light_still and not self.force_near and not self.first
        pp = self.params()
        iso_stale = self.force_near or moved_mask != 0
        flags = (N.SL_BOUNCE_XFER if iso_stale else 0) | (N.SL_BOUNCE_LIGHT_STILL if light_still else 0) | \
            (N.SL_BOUNCE_TIMERS if timers else 0)
        self.round += 1
        self.force_near = False
        self.first = False
        self.scene.bounce_encode(pp, flags)
        return pp
