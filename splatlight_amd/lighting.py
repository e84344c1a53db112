"""Host side of the relighting: the parameter logic of src/lighting.js, src/light_shadow.js, src/sun_map.js and
src/environment.js, producing the same uniform byte layouts their writers produce. The GPU work behind them is
libsplatlight_hip.so (sl_lighting_iso_init, sl_lighting_encode).

Everything here runs on the host in float64, as the JavaScript does, and is stored as float32 where the
reference stores it.
"""
from __future__ import annotations

import math

import numpy as np

from . import records as R

MAX_EMITTERS = 16384
EMIT_THETA = 4.0                      # src/lighting.js:53
EMIT_EDGE_SOFT = 0.2                  # src/lighting.js:578
LS_RES = 512                          # src/light_shadow.js:18
SUN_CASCADES = (8.0, 32.0)            # src/sun_map.js:13
LS_PARAMS_BYTES = 336
SUN_PARAMS_BYTES = 432
SUN_LAYER_DESC_BYTES = 32
ENV_BYTES = 416


def _ein3() -> float:
    """Ein(3) by the midpoint sum of src/lighting.js:55-56."""
    acc = 0.0
    for i in range(3000):
        x = (i + 0.5) * 0.001
        acc += (1 - math.exp(-x)) / x * 0.001
    return acc


EIN3 = _ein3()
ASG_WIDEN = EIN3 / (1 - math.exp(-3))  # src/lighting.js:67


def _sym_det(m):
    a, d, f, b, c, e = m
    return a * (d * f - e * e) - b * (b * f - e * c) + c * (b * e - d * c)


def _sym_inv(m):
    a, d, f, b, c, e = m
    det = _sym_det(m)
    return [(d * f - e * e) / det, (a * f - c * c) / det, (a * d - b * b) / det,
            (c * e - b * f) / det, (b * e - c * d) / det, (b * c - a * e) / det]


def _sym_quad(m, w):
    a, d, f, b, c, e = m
    x, y, z = w
    return a * x * x + d * y * y + f * z * z + 2 * (b * x * y + c * x * z + e * y * z)


def _sym_axes(m):
    """Eigenvectors of a symmetric 3x3 by 12 Jacobi sweeps (src/lighting.js:79-93)."""
    A = [[m[0], m[3], m[4]], [m[3], m[1], m[5]], [m[4], m[5], m[2]]]
    V = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    for _ in range(12):
        for p, q in ((0, 1), (0, 2), (1, 2)):
            if abs(A[p][q]) < 1e-30:
                continue
            th = 0.5 * math.atan2(2 * A[p][q], A[q][q] - A[p][p])
            c, s = math.cos(th), math.sin(th)
            for k in range(3):
                x, y = A[k][p], A[k][q]
                A[k][p] = c * x - s * y
                A[k][q] = s * x + c * y
            for k in range(3):
                x, y = A[p][k], A[q][k]
                A[p][k] = c * x - s * y
                A[q][k] = s * x + c * y
            for k in range(3):
                x, y = V[k][p], V[k][q]
                V[k][p] = c * x - s * y
                V[k][q] = s * x + c * y
    return [[V[0][k], V[1][k], V[2][k]] for k in range(3)]


def _lum(c):
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


def build_emitter_tree(rec: np.ndarray, split_at=None):
    """buildEmitterTree (src/lighting.js:99-169) over emitter records (GAUSSIAN_DTYPE or float32 [n, 16]).

    split_at: None, the reference's median split (len >> 1); else a function of a node's leaf count giving how many
    of its leaves, in order along the longest axis, go to its first child (1 gives a chain as deep as the count: the
    trees a host that adds emitters one at a time would build, which need the shading's deep-walk stack).
    Returns (nodes float32 [2n-1, 16] with u32 words at 11 and 15, root index, cov float32 [2n-1, 8]).
    """
    r = np.ascontiguousarray(rec).view(np.float32).reshape(-1, 16).astype(np.float64)
    n = r.shape[0]
    if not n >= 1:
        raise ValueError(f"emitter tree needs at least one emitter (got {n})")
    f = np.zeros(((2 * n - 1), 16), np.float32)
    u = f.view(np.uint32)
    cov = np.zeros(((2 * n - 1), 8), np.float32)
    info = {}
    nxt = [0]

    def leaf(i):
        o = r[i]
        mu = [o[0], o[1], o[2]]
        x, y, z, w = o[4], o[5], o[6], o[7]
        Rm = [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
              [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
              [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]]
        s2 = [o[8] ** 2, o[9] ** 2, o[10] ** 2]

        def c(a, k):
            return Rm[a][0] * Rm[k][0] * s2[0] + Rm[a][1] * Rm[k][1] * s2[1] + Rm[a][2] * Rm[k][2] * s2[2]
        C = [c(0, 0), c(1, 1), c(2, 2), c(0, 1), c(0, 2), c(1, 2)]
        g = (2 * math.pi * EIN3) ** 2 * _sym_det(C)
        M = [v * g for v in _sym_inv(C)]
        mass = o[3] * (2 * math.pi) ** 1.5 * math.sqrt(_sym_det(C))
        return {"mu": mu, "C": C, "M": M, "mass": mass, "Cs": [v * ASG_WIDEN for v in C],
                "Le": [o[12], o[13], o[14]], "radius": 3 * max(o[8], o[9], o[10]), "lo": 0xFFFFFFFF, "hi": i}

    def write(nd):
        i = nxt[0]
        nxt[0] += 1
        f[i, 0:11] = [*nd["mu"], nd["radius"], nd["M"][0], nd["M"][1], nd["M"][2], 0.0, nd["M"][3], nd["M"][4],
                      nd["M"][5]]
        u[i, 11] = nd["lo"]
        f[i, 12:15] = nd["Le"]
        u[i, 15] = nd["hi"]
        Cs = nd["Cs"]
        cov[i] = [Cs[0], Cs[1], Cs[2], 0.0, Cs[3], Cs[4], Cs[5], 0.0]
        info[i] = nd
        return i

    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))

    def build(idx):
        if len(idx) == 1:
            return write(leaf(idx[0]))
        lo = [min(r[i][k] for i in idx) for k in range(3)]
        hi = [max(r[i][k] for i in idx) for k in range(3)]
        ax = 0
        for k in range(3):
            if hi[k] - lo[k] > hi[ax] - lo[ax]:
                ax = k
        srt = sorted(idx, key=lambda i: r[i][ax])
        half = len(srt) >> 1 if split_at is None else min(max(int(split_at(len(srt))), 1), len(srt) - 1)
        ia = build(srt[:half])
        ib = build(srt[half:])
        kids = [info[ia], info[ib]]
        m = kids[0]["mass"] + kids[1]["mass"]
        mu = [sum(ch["mass"] * ch["mu"][k] for ch in kids) / m for k in range(3)]
        C = [sum(ch["mass"] * (ch["C"][j] + (ch["mu"][a] - mu[a]) * (ch["mu"][b] - mu[b])) for ch in kids) / m
             for j, (a, b) in enumerate(pairs)]
        Cs = [sum(ch["mass"] * (ch["Cs"][j] + (ch["mu"][a] - mu[a]) * (ch["mu"][b] - mu[b])) for ch in kids) / m
              for j, (a, b) in enumerate(pairs)]
        axes = _sym_axes(C)

        def along(ch, w):
            return _lum(ch["Le"]) * math.sqrt(max(_sym_quad(ch["M"], w), 0.0))
        tot = [sum(along(ch, w) for ch in kids) for w in axes]
        wsum = [sum(math.sqrt(max(_sym_quad(ch["M"], w), 0.0)) for w in axes) for ch in kids]
        Le = [sum(ch["Le"][c] * ws for ch, ws in zip(kids, wsum)) / sum(wsum) for c in range(3)]
        L = max(_lum(Le), 1e-12)
        M = [0.0] * 6
        for k, w in enumerate(axes):
            lam = (tot[k] / L) ** 2
            for j, (a, b) in enumerate(pairs):
                M[j] += lam * w[a] * w[b]
        radius = max(math.hypot(*[ch["mu"][k] - mu[k] for k in range(3)]) + ch["radius"] for ch in kids)
        return write({"mu": mu, "C": C, "Cs": Cs, "M": M, "mass": m, "Le": Le, "radius": radius, "lo": ia, "hi": ib})

    root = build(list(range(n)))
    return f, root, cov


def pack_emit_nodes(nodes: np.ndarray, cov: np.ndarray) -> np.ndarray:
    """packEmitNodes (src/lighting.js:61-65): 24 floats (96 B) per node."""
    out = np.zeros((nodes.shape[0], 24), np.float32)
    out[:, :16] = nodes
    out[:, 16:] = cov
    return out


def one_sided_normal(rec0) -> list:
    """The first emitter's short axis R x (src/lighting.js:703-705)."""
    x, y, z, w = [float(v) for v in rec0["quat"]]
    return [1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)]


def asset_emitters(gaussians: np.ndarray, g0: int, g1: int, emission, center=None) -> dict:
    """setAssetEmitters (src/lighting.js:690-721): an asset's emissive part [g0, g1) as the light set. Its records
    are moved so their rest centroid sits at `center` (None: where it is) and carry Le = `emission` in the albedo
    words; the emitting normal is the first record's short axis; EP.scale is the overlap correction, the area of the
    grid of emitter centres (extent along the first record's y and z axes, each stretched by cols / (cols - 1) over
    its distinct columns at 1/1000 of the span) over the emitters' summed footprint 2 pi s_y s_z Ein(3). Returns
    {"records", "normal", "scale"}; the part itself is EP.skip_part."""
    src = gaussians[g0:g1]
    n = int(src.shape[0])
    rec = src.copy()
    mu = src["mu"].astype(np.float64)
    c0 = mu.sum(0) / n
    if center is not None:
        rec["mu"] = (mu + (np.asarray(center, np.float64) - c0)).astype(np.float32)
    rec["albedo"] = np.asarray(emission, np.float32)
    x, y, z, w = [float(v) for v in rec["quat"][0]]
    normal = [1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)]
    axes = ([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
            [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)])
    rm = rec["mu"].astype(np.float64)
    extent = []
    for a in axes:
        vals = a[0] * rm[:, 0] + a[1] * rm[:, 1] + a[2] * rm[:, 2]
        lo, hi = float(vals.min()), float(vals.max())
        cols = len({int(math.floor((v - lo) / (hi - lo + 1e-9) * 1000 + 0.5)) for v in vals.tolist()})   # Math.round
        extent.append((hi - lo) * cols / max(cols - 1, 1))
    sc = rec["scale"].astype(np.float64)
    foot = 0.0
    for i in range(n):
        foot += 2 * math.pi * sc[i, 1] * sc[i, 2] * EIN3
    return {"records": rec, "normal": normal, "scale": extent[0] * extent[1] / foot}


def eig3(m):
    """Ascending eigenvalues of a symmetric 3x3 [xx yy zz xy xz yz] (src/light_shadow.js:169-180)."""
    a, d, f, b, c, e = m
    p1 = b * b + c * c + e * e
    q = (a + d + f) / 3
    p2 = (a - q) ** 2 + (d - q) ** 2 + (f - q) ** 2 + 2 * p1
    p = math.sqrt(p2 / 6)
    if p < 1e-30:
        return [q, q, q]
    B = [(a - q) / p, (d - q) / p, (f - q) / p, b / p, c / p, e / p]
    rr = (B[0] * (B[1] * B[2] - B[5] * B[5]) - B[3] * (B[3] * B[2] - B[5] * B[4]) +
          B[4] * (B[3] * B[5] - B[1] * B[4])) / 2
    phi = math.acos(min(1.0, max(-1.0, rr))) / 3
    e1 = q + 2 * p * math.cos(phi)
    e3 = q + 2 * p * math.cos(phi + 2 * math.pi / 3)
    return [e3, 3 * q - e1 - e3, e1]


def ls_levels(res: int):
    levels, off, k = [], 0, res
    while k >= 1:
        levels.append((off, k))
        off += k * k
        k >>= 1
    return levels, off


def ls_params(tree_nodes, root, cov, normal, one_sided, part_count, dirty, skip_part=0xFFFFFFFF, res=LS_RES):
    """LSParams (336 B) as LightShadow.writeParams packs it (src/light_shadow.js:113-140)."""
    while res > 64 and math.ceil(part_count * res * res / 64) > 65535:
        res >>= 1
    levels, layer_texels = ls_levels(res)
    buf = np.zeros(84, np.uint32)
    f = buf.view(np.float32)
    c = [float(v) for v in cov[root]]
    center = [float(v) for v in tree_nodes[root, :3]]
    n = [float(v) for v in normal]
    nl = math.hypot(*n)
    n = [v / nl for v in n]
    a = [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0]
    t1 = [n[1] * a[2] - n[2] * a[1], n[2] * a[0] - n[0] * a[2], n[0] * a[1] - n[1] * a[0]]
    tl = math.hypot(*t1)
    t1 = [v / tl for v in t1]
    t2 = [n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]]
    ev = eig3([c[0], c[1], c[2], c[4], c[5], c[6]])
    sigma = math.sqrt(0.5 * (ev[1] + ev[2]))
    f[0:3] = center
    buf[3] = res
    f[4:7] = n
    buf[7] = len(levels)
    f[8:11] = t1
    buf[11] = layer_texels
    f[12:15] = t2
    buf[15] = dirty
    f[16] = sigma
    buf[17] = skip_part
    buf[18] = 1 if one_sided else 0
    buf[19] = part_count
    f[20], buf[21], buf[22], f[23] = 0.25, 32, 1024, 8.0            # method B's fields, kept as the reference's
    f[24:27] = c[0:3]
    buf[27] = 1 << 17
    f[28:31] = c[4:7]
    buf[31] = 256
    buf[32], f[33], f[34], buf[35] = 10, 0.25, math.log(3.5 / 0.25) / 10, 2
    for k, (o, _) in enumerate(levels):
        buf[36 + k] = o
    f[52:84] = 1.0
    return buf.tobytes()


def env_bytes(lights, sh=None) -> bytes:
    """Env (416 B) as Environment.write packs it (src/environment.js:111-153): lights = [(dir, std, irradiance,
    shadows)] with dir the unit direction toward the light."""
    buf = np.zeros(ENV_BYTES // 4, np.uint32)
    f = buf.view(np.float32)
    if sh is not None:
        for k, c in enumerate(sh[:16]):
            f[4 * k: 4 * k + 3] = c
    for k, (d, std, irr, shadows) in enumerate(lights[:4]):
        f[64 + 8 * k: 68 + 8 * k] = [*d, std]
        f[68 + 8 * k: 72 + 8 * k] = [*irr, 1.0 if shadows else 0.0]
    buf[96] = len(lights[:4])
    buf[97] = 1 if sh is not None and len(sh) else 0
    return buf.tobytes()


def _basis_k(w):
    """common.wgsl basis_k / sun_map.js basisK: a unit vector perpendicular to w."""
    a = [1.0, 0.0, 0.0] if abs(w[2]) >= 0.999 else [0.0, 0.0, 1.0]
    c = [a[1] * w[2] - a[2] * w[1], a[2] * w[0] - a[0] * w[2], a[0] * w[1] - a[1] * w[0]]
    l = math.hypot(*c)
    return [v / l for v in c]


def sun_plan(parts: np.ndarray, lights, eye, fwd, res=512, cap_bytes=None, static=True):
    """SunMap's layer plan and SunParams (src/sun_map.js:24-159) for every layer dirty.

    lights: [(dir, std, casts_shadows)]. cap_bytes: the reference halves res until each array fits 120 MB; None
    keeps the requested resolution (the 2048^2 hill map needs the cap lifted). Returns (SunParams bytes, LayerDesc
    bytes, n_layers, res).
    """
    P = min(parts.shape[0], 32)
    K = min(len(lights), 4)
    bound = [[float(v) for v in parts["bound"][p]] for p in range(P)]
    ext = max(math.hypot(b[0], b[1], b[2]) + b[3] for b in bound)
    sig_max = max(L[1] for L in lights[:K])
    margin = 3 * sig_max * 2 * ext
    desc = []
    part_range = []
    for p in range(P):
        first = len(desc)
        big = bound[p][3] > SUN_CASCADES[0] and static
        if parts["g_hi"][p] > parts["g_lo"][p]:
            for k in range(K):
                if big:
                    for c, h in enumerate(SUN_CASCADES):
                        if h < bound[p][3] * 1.5:
                            desc.append({"p": p, "k": k, "kind": 1, "half": h})
                desc.append({"p": p, "k": k, "kind": 0})
        part_range.append((first, len(desc) - first))
    n_layers = len(desc)
    if cap_bytes is not None:
        while res > 64 and (n_layers * res * res * 12 > cap_bytes or n_layers * res * res * 16 * 4 / 3 > cap_bytes):
            res >>= 1
    levels, layer_texels = ls_levels(res)
    du = np.zeros((n_layers, 8), np.uint32)
    df = du.view(np.float32)
    for l, d in enumerate(desc):
        du[l, 0:4] = [d["p"], d["k"], d["kind"], 1]
        if d["kind"] == 1:
            L = lights[d["k"]][0]
            t1 = _basis_k(L)
            t2 = [L[1] * t1[2] - L[2] * t1[1], L[2] * t1[0] - L[0] * t1[2], L[0] * t1[1] - L[1] * t1[0]]
            h = d["half"]
            at = [eye[i] + fwd[i] * h * 0.75 for i in range(3)]
            df[l, 4:7] = [sum(a * b for a, b in zip(at, t1)), sum(a * b for a, b in zip(at, t2)), h]
    sp = np.zeros(SUN_PARAMS_BYTES // 4, np.uint32)
    sf = sp.view(np.float32)
    sp[0:8] = [n_layers, res, len(levels), layer_texels, K, P, 0, 0]
    sf[8:11] = [margin, 0.0, 0.0]
    sp[11] = 0
    for k in range(4):
        if k < K:
            d, std, cast = lights[k]
            sf[12 + 4 * k: 16 + 4 * k] = [*d, std if cast else -1.0]
        else:
            sf[12 + 4 * k: 16 + 4 * k] = [0.0, 0.0, 1.0, -1.0]
    for p, (a, cnt) in enumerate(part_range):
        sp[28 + 2 * p] = a
        sp[29 + 2 * p] = cnt
    for k, (o, _) in enumerate(levels):
        sp[92 + k] = o
    return sp.tobytes(), du.tobytes(), n_layers, res


# ---- the per-scene lighting setup the host hands to sl_lighting_encode ----------------------------------------

def lighting_inputs(sc, eye=None, fwd=None, dirty_all=True, sun_res=None, frame=0, emit_split=None):
    """All uniform bytes for one lit frame of a synthetic scene (scenes.Scene): FrameUniforms, EmitParams, the
    emitter tree (emit_split: build_emitter_tree's split_at), LSParams, Env, SunParams and the sun layer plan. Maps
    are marked dirty when dirty_all."""
    eye = tuple(eye or sc.eye)
    if fwd is None:
        d = np.asarray(sc.target, np.float64) - np.asarray(eye, np.float64)
        fwd = tuple(d / np.linalg.norm(d))
    U = R.frame_uniforms(frame=frame, gaussian_count=sc.count, part_count=sc.parts.shape[0],
                         theta=8.0, alpha_min=1 / 16, alpha_part=0.5, eye=eye,
                         material_count=sc.materials.shape[0], iso=1 | 2)
    es = getattr(sc, "emit_set", None)
    if es is None:                  # the scene's own emitter records, drawn after its Gaussians, EP.scale 1
        recs, scale, skip = sc.emitters, 1.0, 0xFFFFFFFF
        normal = one_sided_normal(sc.emitters[0]) if sc.one_sided else [0.0, 0.0, 1.0]
    else:                           # an asset's emissive part (asset_emitters): drawn with the scene, skipped
        recs, scale, skip = es["records"], es["scale"], es["part"]
        normal = es["normal"]
    nodes, root, cov = build_emitter_tree(recs, split_at=emit_split)
    n = recs.shape[0]
    EP = R.emit_params(count=n, root=root, scale=scale, theta=EMIT_THETA, mode=1, node_count=nodes.shape[0],
                       skip_part=skip, one_sided=1 if sc.one_sided else 0, normal=normal, n5=1 << 3,
                       edge_soft=EMIT_EDGE_SOFT)
    P = sc.parts.shape[0]
    LS = ls_params(nodes, root, cov, normal, sc.one_sided, P, ((1 << P) - 1) if dirty_all else 0, skip_part=skip)
    out = {"U": U, "EP": EP, "emit_nodes": pack_emit_nodes(nodes, cov), "LS": LS, "ENV": None, "SM": None,
           "desc": None}
    if sc.sun is not None:
        d = sc.sun["dir"]
        out["ENV"] = env_bytes([(d, sc.sun["std"], sc.sun["E"], True)], getattr(sc, "sky", None))
        SM, desc, nl, res = sun_plan(sc.parts, [(d, sc.sun["std"], True)], eye, fwd,
                                     res=sun_res or sc.sun.get("res", 512))
        if not dirty_all:
            du = np.frombuffer(desc, np.uint32).reshape(-1, 8).copy()
            du[:, 3] = 0
            desc = du.tobytes()
        out["SM"], out["desc"] = SM, desc
    elif getattr(sc, "sky", None) is not None:
        out["ENV"] = env_bytes([], sc.sky)          # a sky and no distant lights
    return out


class Lighting:
    """Lighting (src/lighting.js:318-1275) over the HIP library, for the default static M39 matte path: the host
    keeps the reference's parameter logic; encode() runs the opacity maps of what moved, shade_emit and the
    distant lights' direct light on the device."""

    def __init__(self, device, scene_native, sc, sun_res=None):
        self.device = device
        self.scene = scene_native
        self.sc = sc
        self.sun_res = sun_res
        self.frame = 0
        self.first = True

    def encode(self, eye=None, changed=True, timers=False):
        inp = lighting_inputs(self.sc, eye=eye, dirty_all=self.first or changed, sun_res=self.sun_res,
                              frame=self.frame)
        self.frame += 1
        self.scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"],
                         timers=timers)
        self.first = False
