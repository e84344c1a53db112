"""CPU oracle for the splat frame path: TEST INFRASTRUCTURE ONLY.

A ctypes wrapper around liboracle.so (plain C restatement of the reference's WGSL/JS, see oracle.h). Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this, as the checker; the product
path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, u32 = C.c_void_p, C.c_uint32
        L.or_depth_key.restype = u32
        L.or_depth_key.argtypes = [C.c_float, C.c_float, C.c_float]
        L.or_project_f32.restype = u32
        L.or_project_f32.argtypes = [vp, vp, vp, vp, vp, vp, u32, vp, vp]
        L.or_project_f64.restype = None
        L.or_project_f64.argtypes = [vp, vp, vp, vp, vp, vp, u32, vp, vp]
        L.or_sort_keys.restype = u32
        L.or_sort_keys.argtypes = [vp, u32, vp, vp, vp]
        L.or_tile_box.restype = u32
        L.or_tile_box.argtypes = [vp, u32, u32, u32, vp]
        L.or_tile_lists.restype = C.c_uint64
        L.or_tile_lists.argtypes = [vp, vp, u32, u32, u32, u32, vp, vp]
        L.or_composite.restype = None
        L.or_composite.argtypes = [vp, vp, vp, u32, u32, u32, C.c_int, vp]
        L.or_composite_rows.restype = None
        L.or_composite_rows.argtypes = [vp, vp, vp, u32, u32, u32, C.c_int, u32, vp]
        L.or_composite_f64.restype = None
        L.or_composite_f64.argtypes = [vp, vp, vp, u32, u32, u32, vp]
        f32 = C.c_float
        L.or_exp_map_n.restype = None
        L.or_exp_map_n.argtypes = [vp, vp, u32]
        L.or_erfc_pos.restype = f32
        L.or_erfc_pos.argtypes = [f32]
        L.or_ein.restype = f32
        L.or_threads.restype = C.c_uint32
        L.or_phase_specular.restype = f32
        L.or_asg_phase.restype = f32
        L.or_ein.argtypes = [f32]
        L.or_footprint.restype = f32
        L.or_footprint.argtypes = [vp, f32, f32, f32, f32]
        L.or_vgrid_lookup.restype = f32
        L.or_vgrid_lookup.argtypes = [vp, f32, f32, f32]
        L.or_iso_init.restype = None
        L.or_iso_init.argtypes = [vp, vp, vp, u32, vp]
        L.or_shade_emit.restype = None
        L.or_shade_emit.argtypes = [vp] * 11
        for name, n in (("or_lsm_splat", 6), ("or_lsm_resolve", 3), ("or_lsm_lookup", 6), ("or_sun_frame", 4),
                        ("or_sun_splat", 8), ("or_sun_resolve", 4), ("or_sun_lookup", 8), ("or_env_direct", 10)):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [vp] * n
        L.or_f16_bits.restype = u32
        L.or_f16_bits.argtypes = [f32]
        for name, n in (("or_bounce_lo_dir", 11), ("or_node_leaf", 9), ("or_node_sh2", 3), ("or_probe_valid", 4),
                        ("or_probe_sh2", 4), ("or_iso_xfer", 9), ("or_bounce_sh", 10)):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [vp] * n
        L.or_node_up.restype = None
        L.or_node_up.argtypes = [vp, vp, vp, vp, vp, u32, u32, u32, vp]
        L.or_probe_gather.restype = None
        L.or_probe_gather.argtypes = [vp] * 10 + [u32, u32]
        L.or_probe_weights.restype = None
        L.or_probe_weights.argtypes = [vp] * 6 + [f32, f32, f32, u32, vp, vp]
        L.or_probe_pos.restype = None
        L.or_probe_pos.argtypes = [vp, vp, vp, vp, u32, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def depth_key(depth: float, near: float, far: float) -> int:
    return int(lib().or_depth_key(depth, near, far))


def project(frame: bytes, scene, colour=None, lo_b=None, f64=False):
    """Returns (splats u32 [slots, 12], keys u32 [slots], visible) or, f64, (splats f64 [slots, 12], valid)."""
    fb = np.frombuffer(frame, np.uint8).copy()
    g = np.ascontiguousarray(scene.gaussians)
    p = np.ascontiguousarray(scene.parts)
    e = np.ascontiguousarray(scene.emitters) if scene.emitters is not None else np.zeros(0, g.dtype)
    n_emit = e.shape[0]
    nslots = g.shape[0] + 1 + n_emit
    col = None if colour is None else np.ascontiguousarray(colour, np.float32).reshape(-1, 4)
    lb = None if lo_b is None else np.ascontiguousarray(lo_b, np.float32).reshape(-1, 12)
    if f64:
        out = np.zeros((nslots, 12), np.float64)
        valid = np.zeros(nslots, np.uint8)
        lib().or_project_f64(_p(fb), _p(g), _p(p), _p(col), _p(lb), _p(e), n_emit, _p(out), _p(valid))
        return out, valid.astype(bool)
    splats = np.zeros((nslots, 12), np.uint32)
    keys = np.zeros(nslots, np.uint32)
    vis = lib().or_project_f32(_p(fb), _p(g), _p(p), _p(col), _p(lb), _p(e), n_emit, _p(splats), _p(keys))
    return splats, keys, int(vis)


def sort_keys(keys: np.ndarray):
    """(hist[65536], offsets[65536], order[n], visible)."""
    k = np.ascontiguousarray(keys, np.uint32)
    hist = np.zeros(65536, np.uint32)
    off = np.zeros(65536, np.uint32)
    order = np.zeros(k.shape[0], np.uint32)
    vis = lib().or_sort_keys(_p(k), k.shape[0], _p(hist), _p(off), _p(order))
    return hist, off, order, int(vis)


def tile_box(splat_row: np.ndarray, W: int, H: int, tile_px: int = 16):
    box = np.zeros(4, np.uint32)
    n = lib().or_tile_box(_p(np.ascontiguousarray(splat_row, np.uint32)), W, H, tile_px, _p(box))
    return int(n), box


def tile_lists(splats: np.ndarray, order: np.ndarray, visible: int, W: int, H: int, tile_px: int = 16):
    """(tile_start[ntiles + 1] u64, vals[P] u32)."""
    s = np.ascontiguousarray(splats, np.uint32)
    o = np.ascontiguousarray(order, np.uint32)
    nt = ((W + tile_px - 1) // tile_px) * ((H + tile_px - 1) // tile_px)
    P = int(lib().or_tile_lists(_p(s), _p(o), visible, W, H, tile_px, None, None))
    start = np.zeros(nt + 1, np.uint64)
    vals = np.zeros(max(P, 1), np.uint32)
    lib().or_tile_lists(_p(s), _p(o), visible, W, H, tile_px, _p(start), _p(vals))
    return start, vals[:P]


def composite(frame: bytes, splats, order, visible, W, H, mode=0, row_step=1):
    """mode 0: float32 back-to-front; 2: float16 target after every blend. Returns [H, W, 4] float32.
    row_step > 1 blends only the rows py % row_step == 0 (a timing sample)."""
    out = np.zeros((H, W, 4), np.float32)
    fb = np.frombuffer(frame, np.uint8).copy()
    lib().or_composite_rows(_p(fb), _p(np.ascontiguousarray(splats, np.uint32)),
                            _p(np.ascontiguousarray(order, np.uint32)), visible, W, H, mode, row_step, _p(out))
    return out


def composite_f64(frame: bytes, splats64, order, visible, W, H):
    out = np.zeros((H, W, 4), np.float64)
    fb = np.frombuffer(frame, np.uint8).copy()
    lib().or_composite_f64(_p(fb), _p(np.ascontiguousarray(splats64, np.float64)),
                           _p(np.ascontiguousarray(order, np.uint32)), visible, W, H, _p(out))
    return out


# ---- relighting -------------------------------------------------------------------------------------------------

def _b(x):
    return np.frombuffer(x, np.uint8).copy() if isinstance(x, (bytes, bytearray)) else np.ascontiguousarray(x)


def exp_map(x):
    """The opacity-map splats' e^x (float32 in, float32 out)."""
    xs = np.ascontiguousarray(x, np.float32).ravel()
    y = np.zeros_like(xs)
    lib().or_exp_map_n(_p(xs), _p(y), xs.size)
    return y


def erfc_pos(z):
    return float(lib().or_erfc_pos(z))


def _f3(v):
    return np.ascontiguousarray(v, np.float32).ravel()


def phase_specular(Sd, So, wi, wo):
    """The specular flake phase (material_ext.wgsl:27-37); wo one direction or [n, 3]."""
    a = [_f3(Sd), _f3(So), _f3(wi)]
    w = _f3(wo)
    if w.size == 3:
        return float(lib().or_phase_specular(*[_p(x) for x in a], _p(w)))
    out = np.zeros(w.size // 3, np.float32)
    lib().or_phase_specular_n(*[_p(x) for x in a], _p(w), out.size, _p(out))
    return out


def onb(n):
    out = np.zeros(6, np.float32)
    lib().or_onb(_p(_f3(n)), _p(out))
    return out.reshape(2, 3)


def asg_phase(wi, lb, sl, n0, t1, t2, flake, Sd, So, wo):
    """The specular phase over a light ASG (material_ext.wgsl:128-181); lb [2, 3], sl (xx, xy, yy)."""
    args = [_f3(x) for x in (wi, lb, sl, n0, t1, t2, flake, Sd, So, wo)]
    return float(lib().or_asg_phase(*[_p(x) for x in args]))


def threads() -> int:
    """Threads the oracle's parallel loops run on."""
    return int(lib().or_threads())


def ein(x):
    return float(lib().or_ein(x))


def footprint(node_row, detC, w):
    return float(lib().or_footprint(_p(np.ascontiguousarray(node_row)), detC, *[float(v) for v in w]))


def vgrid_lookup(vg128, w):
    return float(lib().or_vgrid_lookup(_p(np.ascontiguousarray(vg128, np.uint8)), *[float(v) for v in w]))


def iso_init(gaussians, nodes, vgrid):
    n = gaussians.shape[0]
    out = np.zeros(n, np.float32)
    lib().or_iso_init(_p(_b(gaussians)), _p(_b(nodes)), _p(np.ascontiguousarray(vgrid, np.uint8)), n, _p(out))
    return out


def shade_emit(U, EP, emit_nodes, gaussians, nodes, vgrid, parts, lamp_T, materials, iso_norm):
    n = gaussians.shape[0]
    col = np.zeros((n, 4), np.float32)
    lt = None if lamp_T is None else np.ascontiguousarray(lamp_T, np.float32)
    lib().or_shade_emit(_p(_b(U)), _p(_b(EP)), _p(np.ascontiguousarray(emit_nodes, np.float32)), _p(_b(gaussians)),
                        _p(_b(nodes)), _p(np.ascontiguousarray(vgrid, np.uint8)), _p(_b(parts)), _p(lt),
                        _p(_b(materials)), _p(np.ascontiguousarray(iso_norm, np.float32)), _p(col))
    return col


def lsm(U, LS, gaussians, nodes, parts):
    """(acc u32 [P, res^2, 3], mip f32 [P, layer_texels, 4], T f32 [G])."""
    w = np.frombuffer(LS, np.uint32)
    res, lt, pc = int(w[3]), int(w[11]), int(w[19])
    acc = np.zeros((pc, res * res, 3), np.uint32)
    mip = np.zeros((pc, lt, 4), np.float32)
    T = np.zeros(gaussians.shape[0], np.float32)
    L = lib()
    L.or_lsm_splat(_p(_b(U)), _p(_b(LS)), _p(_b(gaussians)), _p(_b(nodes)), _p(_b(parts)), _p(acc))
    L.or_lsm_resolve(_p(_b(LS)), _p(acc), _p(mip))
    L.or_lsm_lookup(_p(_b(U)), _p(_b(LS)), _p(_b(gaussians)), _p(_b(parts)), _p(mip), _p(T))
    return acc, mip, T


def sun(U, SM, desc, gaussians, nodes, parts):
    """(acc u32 [L, res^2, 3], mip f32 [L, layer_texels, 4], env_T f32 [G, 4])."""
    w = np.frombuffer(SM, np.uint32)
    nl, res, lt = int(w[0]), int(w[1]), int(w[3])
    layers = np.zeros((nl, 12), np.float32)
    acc = np.zeros((nl, res * res, 3), np.uint32)
    mip = np.zeros((nl, lt, 4), np.float32)
    T = np.zeros((gaussians.shape[0], 4), np.float32)
    L = lib()
    L.or_sun_frame(_p(_b(SM)), _p(_b(desc)), _p(_b(parts)), _p(layers))
    L.or_sun_splat(_p(_b(U)), _p(_b(SM)), _p(_b(desc)), _p(layers), _p(_b(gaussians)), _p(_b(nodes)), _p(_b(parts)),
                   _p(acc))
    L.or_sun_resolve(_p(_b(SM)), _p(_b(desc)), _p(acc), _p(mip))
    L.or_sun_lookup(_p(_b(U)), _p(_b(SM)), _p(_b(desc)), _p(layers), _p(_b(gaussians)), _p(_b(parts)), _p(mip), _p(T))
    return acc, mip, T


def lsm_resolve(LS, acc):
    """The emitter map's resolve and mips alone (per texel), from fixed-point sums."""
    w = np.frombuffer(LS, np.uint32)
    mip = np.zeros((int(w[19]), int(w[11]), 4), np.float32)
    lib().or_lsm_resolve(_p(_b(LS)), _p(np.ascontiguousarray(acc, np.uint32)), _p(mip))
    return mip


def sun_resolve(SM, desc, acc):
    """The sun maps' resolve and mips alone (per texel), from fixed-point sums."""
    w = np.frombuffer(SM, np.uint32)
    mip = np.zeros((int(w[0]), int(w[3]), 4), np.float32)
    lib().or_sun_resolve(_p(_b(SM)), _p(_b(desc)), _p(np.ascontiguousarray(acc, np.uint32)), _p(mip))
    return mip


def env_direct(U, ENV, gaussians, nodes, vgrid, parts, materials, iso_norm, env_T, colour):
    col = np.ascontiguousarray(colour, np.float32).copy()
    lib().or_env_direct(_p(_b(U)), _p(_b(ENV)), _p(_b(gaussians)), _p(_b(nodes)), _p(np.ascontiguousarray(vgrid, np.uint8)),
                        _p(_b(parts)), _p(_b(materials)), _p(np.ascontiguousarray(iso_norm, np.float32)),
                        _p(np.ascontiguousarray(env_T, np.float32)), _p(col))
    return col


# ---- sky and bounce through probes ------------------------------------------------------------------------------

def f16_bits(x: float) -> int:
    return int(lib().or_f16_bits(x))


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def cells_table(dirs, omega):
    """The 128 x 4 cell table (direction, solid angle) the bounce passes take."""
    t = np.zeros((128, 4), np.float32)
    t[:, :3] = dirs
    t[:, 3] = omega
    return t


def bounce_lo_dir(U, ENV, gaussians, nodes, vgrid, parts, materials, iso_norm, env_T, c0, lo_dir):
    out = _f(lo_dir).copy()
    lib().or_bounce_lo_dir(_p(_b(U)), None if ENV is None else _p(_b(ENV)), _p(_b(gaussians)), _p(_b(nodes)),
                           _p(np.ascontiguousarray(vgrid, np.uint8)), _p(_b(parts)), _p(_b(materials)),
                           _p(_f(iso_norm)), None if env_T is None else _p(_f(env_T)), _p(_f(c0)), _p(out))
    return out


def node_leaf(U, PP, gaussians, nodes, vgrid, parts, lo_dir, lo_b, node_acc):
    out = _f(node_acc).copy()
    lib().or_node_leaf(_p(_b(U)), _p(_b(PP)), _p(_b(gaussians)), _p(_b(nodes)), _p(np.ascontiguousarray(vgrid, np.uint8)),
                       _p(_b(parts)), _p(_f(lo_dir)), _p(_f(lo_b)), _p(out))
    return out


def node_up(U, PP, nodes, kids, up_nodes, levels, max_depth, node_acc):
    """node_up for depths max_depth - 1 .. 1 (levels: (depth, offset, count) per depth)."""
    out = _f(node_acc).copy()
    k = np.ascontiguousarray(kids, np.uint32)
    un = np.ascontiguousarray(up_nodes, np.uint32)
    for d in range(max_depth - 1, 0, -1):
        dep, off, cnt = levels[d]
        lib().or_node_up(_p(_b(U)), _p(_b(PP)), _p(_b(nodes)), _p(k), _p(un), dep, off, cnt, _p(out))
    return out


def node_sh2(PP, node_acc, node_count):
    out = np.zeros((node_count, 14), np.uint32)
    lib().or_node_sh2(_p(_b(PP)), _p(_f(node_acc)), _p(out))
    return out


def probe_gather(U, PP, ENV, nodes, vgrid, parts, kids, node_sh, cells, probe_data, id0=0, id1=None):
    out = _f(probe_data).copy()
    n = out.size // 8
    lib().or_probe_gather(_p(_b(U)), _p(_b(PP)), None if ENV is None else _p(_b(ENV)), _p(_b(nodes)),
                          _p(np.ascontiguousarray(vgrid, np.uint8)), _p(_b(parts)), _p(np.ascontiguousarray(kids, np.uint32)),
                          _p(np.ascontiguousarray(node_sh, np.uint32)), _p(_f(cells)), _p(out), id0, n if id1 is None else id1)
    return out


def probe_valid(U, PP, parts, probe_data):
    out = _f(probe_data).copy()
    lib().or_probe_valid(_p(_b(U)), _p(_b(PP)), _p(_b(parts)), _p(out))
    return out


def probe_sh2(PP, probe_data, cells, n_probes):
    out = np.zeros((n_probes, 9, 4), np.float32)
    lib().or_probe_sh2(_p(_b(PP)), _p(_f(probe_data)), _p(_f(cells)), _p(out))
    return out


def iso_xfer(U, PP, gaussians, nodes, vgrid, parts, iso_norm, cells, xfer):
    out = _f(xfer).copy()
    lib().or_iso_xfer(_p(_b(U)), _p(_b(PP)), _p(_b(gaussians)), _p(_b(nodes)), _p(np.ascontiguousarray(vgrid, np.uint8)),
                      _p(_b(parts)), _p(_f(iso_norm)), _p(_f(cells)), _p(out))
    return out


def probe_weights(U, PP, nodes, parts, cells, probe_data, P, own):
    pid = np.zeros(8, np.uint32)
    w = np.zeros(8, np.float32)
    lib().or_probe_weights(_p(_b(U)), _p(_b(PP)), _p(_b(nodes)), _p(_b(parts)), _p(_f(cells)), _p(_f(probe_data)),
                           float(P[0]), float(P[1]), float(P[2]), own, _p(pid), _p(w))
    return pid, w


def bounce_sh(U, PP, gaussians, nodes, parts, cells, probe_data, probe_sh, xfer, lo_b):
    out = _f(lo_b).copy()
    lib().or_bounce_sh(_p(_b(U)), _p(_b(PP)), _p(_b(gaussians)), _p(_b(nodes)), _p(_b(parts)), _p(_f(cells)),
                       _p(_f(probe_data)), _p(_f(probe_sh)), _p(_f(xfer)), _p(out))
    return out


def probe_pos(U, PP, nodes, parts, count):
    out = np.zeros((count, 3), np.float32)
    lib().or_probe_pos(_p(_b(U)), _p(_b(PP)), _p(_b(nodes)), _p(_b(parts)), count, _p(out))
    return out


def bounce_frame(st: dict, U, PP, ENV, gaussians, nodes, vgrid, parts, materials, iso_norm, env_T, c0, kids,
                 up_nodes, levels, max_depth, cells, n_probes, run_xfer=True):
    """One encodeBounce (src/lighting.js:1020-1103, default path) on the CPU. st holds the state carried across
    frames (lo_dir, lo_b, node_acc, probe_data, xfer; zeros at first) and is updated in place."""
    G = gaussians.shape[0]
    N = nodes.shape[0]
    if run_xfer:
        st["xfer"] = iso_xfer(U, PP, gaussians, nodes, vgrid, parts, iso_norm, cells, st["xfer"])
    st["lo_dir"] = bounce_lo_dir(U, ENV, gaussians, nodes, vgrid, parts, materials, iso_norm, env_T, c0, st["lo_dir"])
    st["node_acc"] = node_leaf(U, PP, gaussians, nodes, vgrid, parts, st["lo_dir"], st["lo_b"], st["node_acc"])
    st["node_acc"] = node_up(U, PP, nodes, kids, up_nodes, levels, max_depth, st["node_acc"])
    st["node_sh"] = node_sh2(PP, st["node_acc"], N)
    st["probe_data"] = probe_gather(U, PP, ENV, nodes, vgrid, parts, kids, st["node_sh"], cells, st["probe_data"])
    st["probe_data"] = probe_valid(U, PP, parts, st["probe_data"])
    st["probe_sh"] = probe_sh2(PP, st["probe_data"], cells, n_probes)
    st["lo_b"] = bounce_sh(U, PP, gaussians, nodes, parts, cells, st["probe_data"], st["probe_sh"], st["xfer"], st["lo_b"])
    return st


def bounce_state(G, N, n_probes):
    return {"lo_dir": np.zeros((G, 4), np.float32), "lo_b": np.zeros((G, 3, 4), np.float32),
            "node_acc": np.zeros((N, 14, 4), np.float32), "probe_data": np.zeros((n_probes, 128, 2, 4), np.float32),
            "xfer": np.zeros((G, 3, 4), np.float32)}


# ---- display (display.c) ------------------------------------------------------------------------------------------

def display(frame: bytes, hdr, W, H, env=None, bg=None):
    """The display pass over an HDR target [H, W, 4]: (encoded colour [H, W, 4] float32, RGBA8 [H, W, 4] uint8)."""
    L = lib()
    fb = np.frombuffer(frame, np.uint8).copy()
    h = np.ascontiguousarray(hdr, np.float32).reshape(H, W, 4)
    out = np.zeros((H, W, 4), np.float32)
    by = np.zeros((H, W, 4), np.uint8)
    e = None if env is None else _b(env)
    g = None if bg is None else _b(bg)
    L.or_display.restype = None
    L.or_display.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p]
    L.or_display(_p(fb), _p(h), W, H, _p(e), _p(g), _p(out), _p(by))
    return out, by


def tone_map(x, mode: int):
    xs = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    y = np.zeros_like(xs)
    lib().or_tone_map_n(_p(xs), C.c_uint32(xs.shape[0]), C.c_uint32(mode), _p(y))
    return y


def srgb_encode(x):
    xs = np.ascontiguousarray(x, np.float32).ravel()
    y = np.zeros_like(xs)
    lib().or_srgb_encode_n(_p(xs), C.c_uint32(xs.size), _p(y))
    return y


def heat(t):
    ts = np.ascontiguousarray(t, np.float32).ravel()
    y = np.zeros((ts.size, 3), np.float32)
    lib().or_heat_n(_p(ts), C.c_uint32(ts.size), _p(y))
    return y


def sky(env, bg, w, which: int):
    """which 0: the lighting sky (Env SH), 1: the background sky (SkyBg), 2: the lights' discs; w [n, 3]."""
    ws = np.ascontiguousarray(w, np.float32).reshape(-1, 3)
    y = np.zeros_like(ws)
    e = None if env is None else _b(env)
    g = None if bg is None else _b(bg)
    lib().or_sky_n(_p(e), _p(g), _p(ws), C.c_uint32(ws.shape[0]), C.c_uint32(which), _p(y))
    return y


def view_dir(frame: bytes, pxy):
    p = np.ascontiguousarray(pxy, np.float32).reshape(-1, 2)
    w = np.zeros((p.shape[0], 3), np.float32)
    lib().or_view_dir_n(_p(np.frombuffer(frame, np.uint8).copy()), _p(p), C.c_uint32(p.shape[0]), _p(w))
    return w


# ---- screen-space LOD and display fill (lod.c) ---------------------------------------------------------------------

class LodParams(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("threshold", C.c_float), ("focal", C.c_float), ("near_", C.c_float),
                ("emit_part", C.c_uint32), ("open_parts", C.c_uint32), ("size_rule", C.c_uint32)]


class LodView(C.Structure):
    _fields_ = [("flag", C.c_void_p), ("records", C.c_void_p), ("colour", C.c_void_p), ("n_nodes", C.c_uint32)]


def lod_params(view, threshold, focal, near=0.02, emit_part=0xFFFFFFFF, open_parts=0, size_rule=1):
    p = LodParams()
    p.view[:] = [float(x) for x in np.asarray(view, np.float32).ravel()]
    p.threshold, p.focal, p.near_ = threshold, focal, near
    p.emit_part, p.open_parts, p.size_rule = emit_part, open_parts, size_rule
    return p


def lod_select(params: LodParams, nodes, kids, parts, G: int):
    """lod_select's flags per node (0 none, 1 open, 2 draw, 4 under)."""
    nd = _b(nodes)
    N = nd.nbytes // 96
    flag = np.zeros(N, np.uint32)
    L = lib()
    L.or_lod_select.restype = None
    L.or_lod_select.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_uint32, C.c_void_p]
    L.or_lod_select(C.byref(params), _p(nd), _p(np.ascontiguousarray(kids, np.uint32)), _p(_b(parts)), N, G,
                    _p(flag))
    return flag


def node_mass(nodes, n: int) -> float:
    L = lib()
    L.or_node_mass.restype = C.c_float
    L.or_node_mass.argtypes = [C.c_void_p, C.c_uint32]
    return float(L.or_node_mass(_p(_b(nodes)), n))


def lod_colour(nodes, kids, G: int, flag, colour, lo_b, gain: float):
    """Node colours (vec4 per internal node) of the drawn nodes and their subtrees; zeros elsewhere."""
    nd = _b(nodes)
    N = nd.nbytes // 96
    out = np.zeros((max(N - G, 1), 4), np.float32)
    col = np.ascontiguousarray(colour, np.float32).reshape(-1, 4)
    lb = np.zeros((G, 12), np.float32) if lo_b is None else np.ascontiguousarray(lo_b, np.float32).reshape(-1, 12)
    L = lib()
    L.or_lod_colour.restype = None
    L.or_lod_colour.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_float, C.c_void_p]
    L.or_lod_colour(_p(nd), _p(np.ascontiguousarray(kids, np.uint32)), N, G,
                    _p(np.ascontiguousarray(flag, np.uint32)), _p(col), _p(lb), gain, _p(out))
    return out[:N - G]


def display_fill(fill, G: int, colour, lo_b):
    """display_fill on copies of colour (vec4 per Gaussian) and lo_b (3 vec4 per Gaussian)."""
    col = np.array(colour, np.float32).reshape(-1, 4)
    lb = np.array(lo_b, np.float32).reshape(-1, 12)
    L = lib()
    L.or_display_fill.restype = None
    L.or_display_fill.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    L.or_display_fill(_p(np.ascontiguousarray(fill, np.uint32)), G, _p(col), _p(lb))
    return col, lb


def project_lod(frame: bytes, scene, flag, records, node_colour, colour=None, lo_b=None):
    """or_project_f32 with the LOD's cut: (splats u32 [slots, 12], keys u32 [slots], visible)."""
    fb = np.frombuffer(frame, np.uint8).copy()
    g = np.ascontiguousarray(scene.gaussians)
    p = np.ascontiguousarray(scene.parts)
    e = np.ascontiguousarray(scene.emitters) if scene.emitters is not None else np.zeros(0, g.dtype)
    n_emit = e.shape[0]
    rec = np.ascontiguousarray(records)
    nn = rec.nbytes // 64
    ncol = np.ascontiguousarray(node_colour, np.float32).reshape(-1, 4)
    fl = np.ascontiguousarray(flag, np.uint32)
    lv = LodView(_p(fl), _p(rec), _p(ncol), nn)
    nslots = g.shape[0] + 1 + n_emit + nn
    col = None if colour is None else np.ascontiguousarray(colour, np.float32).reshape(-1, 4)
    lb = None if lo_b is None else np.ascontiguousarray(lo_b, np.float32).reshape(-1, 12)
    splats = np.zeros((nslots, 12), np.uint32)
    keys = np.zeros(nslots, np.uint32)
    L = lib()
    L.or_project_lod_f32.restype = C.c_uint32
    L.or_project_lod_f32.argtypes = [C.c_void_p] * 6 + [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    vis = L.or_project_lod_f32(_p(fb), _p(g), _p(p), _p(col), _p(lb), _p(e), n_emit, C.byref(lv), _p(splats),
                               _p(keys))
    return splats, keys, int(vis)


def lod_node_size(params: LodParams, nodes, parts, n: int) -> float:
    """lod_select's on-screen size (px) of internal node n (before the threshold)."""
    L = lib()
    L.or_lod_node_size.restype = C.c_float
    L.or_lod_node_size.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
    return float(L.or_lod_node_size(C.byref(params), _p(_b(nodes)), _p(_b(parts)), n))
