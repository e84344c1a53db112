"""ctypes binding of libsplatlight_hip.so (include/splatlight.h).

This is the binding a Python host uses; the JavaScript host's is the N-API addon in splatlight_amd/node/. The
library is the only compute path: if it is missing or fails to load this module raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplatlight_hip.so")

SL_OK = 0
SL_ERR_CAPACITY = 7
STATUS_NAMES = {0: "SL_OK", 1: "SL_ERR_INVALID", 2: "SL_ERR_LAYOUT", 3: "SL_ERR_DEVICE", 4: "SL_ERR_POISONED",
                5: "SL_ERR_UNSUPPORTED", 6: "SL_ERR_BUSY", 7: "SL_ERR_CAPACITY"}
SL_STORAGE_F32, SL_STORAGE_F16, SL_STORAGE_BF16 = 0, 1, 2
SL_OUT_RGBA32F, SL_OUT_RGBA16F = 0, 1
SL_FRAME_TIMERS = 1
SL_FRAME_COUNTING_SORT = 4
SL_DISPLAY_RGBA8, SL_DISPLAY_BGRA8 = 0, 1
SL_DISPLAY_EXACT = 1
ART_KEYS, ART_ORDER, ART_SPLATS, ART_PAIR_TILES, ART_PAIR_SLOTS, ART_RANGES = 1, 2, 3, 4, 5, 6
ART_LOD_FLAGS, ART_LOD_COLOUR = 7, 8

# Every symbol include/splatlight.h declares (checked by the CPU tests).
EXPORTS = [
    "sl_session_acquire", "sl_session_release", "sl_session_error",
    "sl_ctx_create", "sl_ctx_destroy", "sl_last_error", "sl_ctx_status", "sl_sync",
    "sl_buf_alloc", "sl_buf_free", "sl_buf_write", "sl_buf_read",
    "sl_scene_create", "sl_scene_destroy", "sl_scene_write_parts", "sl_scene_write_colour",
    "sl_scene_write_bounce", "sl_scene_device_bytes", "sl_scene_write_gaussians", "sl_scene_write_materials",
    "sl_frame_encode", "sl_frame_finish", "sl_frame_read", "sl_timers_read", "sl_frame_encode_noalloc",
    "sl_lighting_iso_init", "sl_lighting_encode", "sl_scene_read", "sl_lighting_stats_read",
    "sl_batch_encode", "sl_batch_finish", "sl_ctx_set_tuning", "sl_ctx_get_tuning",
    "sl_server_create", "sl_server_submit", "sl_server_poll", "sl_server_info", "sl_server_error",
    "sl_server_destroy", "sl_server_fail_next", "sl_bounce_init", "sl_bounce_encode", "sl_bounce_read",
    "sl_display_encode", "sl_scene_set_lod", "sl_scene_set_fill", "sl_fill_encode",
]
SL_BATCH_GRAPH = 2
SCENE_COLOUR, SCENE_LAMP_T, SCENE_ENV_T, SCENE_ISO_NORM = 1, 2, 3, 4
SCENE_LSM_ACC, SCENE_LSM_MIP, SCENE_SUN_ACC, SCENE_SUN_MIP = 5, 6, 7, 8
SL_LIGHT_TIMERS = 1


class LightingStats(C.Structure):
    _fields_ = [(f"{m}_{k}", C.c_uint32) for m in ("lsm", "sun") for k in ("records", "pairs", "tiles", "tiles_used")]


class SceneDesc(C.Structure):
    _fields_ = [
        ("gaussians", C.c_void_p), ("gaussian_count", C.c_uint32), ("gaussian_stride", C.c_uint32),
        ("parts", C.c_void_p), ("part_count", C.c_uint32), ("part_stride", C.c_uint32),
        ("materials", C.c_void_p), ("material_count", C.c_uint32), ("material_stride", C.c_uint32),
        ("nodes", C.c_void_p), ("node_count", C.c_uint32), ("node_stride", C.c_uint32),
        ("vgrid", C.c_void_p), ("vgrid_count", C.c_uint32), ("vgrid_stride", C.c_uint32),
        ("emitters", C.c_void_p), ("emitter_count", C.c_uint32),
        ("storage", C.c_uint32),
    ]


class FrameOpts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("out_format", C.c_uint32), ("flags", C.c_uint32)]


class DisplayOpts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("hdr_format", C.c_uint32), ("out_format", C.c_uint32),
                ("flags", C.c_uint32)]


class LodTree(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint32), ("node_stride", C.c_uint32),
                ("kids", C.c_void_p), ("kid_count", C.c_uint32),
                ("node_gaussians", C.c_void_p), ("node_gaussian_count", C.c_uint32),
                ("emit_part", C.c_uint32), ("size_rule", C.c_uint32)]


class FrameStats(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("visible", C.c_uint32), ("pairs", C.c_uint32), ("overflow", C.c_uint32),
                ("tiles", C.c_uint32), ("pad", C.c_uint32 * 3)]


class View(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("frame_bytes", C.c_void_p), ("frame_len", C.c_size_t), ("width", C.c_uint32),
                ("height", C.c_uint32), ("out_format", C.c_uint32), ("out_device", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("composite_batch", C.c_uint32), ("radix_items", C.c_uint32), ("grid_mul", C.c_uint32),
                ("radix_bits", C.c_uint32), ("tile_px", C.c_uint32), ("pad", C.c_uint32)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("lanes", C.c_uint32), ("max_batch_views", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("batch_views", C.c_uint32), ("stats", FrameStats),
                ("submit_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double)]


class LightingDesc(C.Structure):
    _fields_ = [("frame_uniforms", C.c_void_p), ("emit_params", C.c_void_p), ("emit_nodes", C.c_void_p),
                ("emit_node_count", C.c_uint32), ("ls_params", C.c_void_p), ("env", C.c_void_p),
                ("sun_params", C.c_void_p), ("sun_layers", C.c_void_p), ("sun_layer_count", C.c_uint32),
                ("flags", C.c_uint32)]


class BounceTree(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint32), ("node_stride", C.c_uint32),
                ("kids", C.c_void_p), ("kid_count", C.c_uint32), ("max_depth", C.c_uint32)]


SL_BOUNCE_TIMERS, SL_BOUNCE_XFER, SL_BOUNCE_LIGHT_STILL = 1, 2, 4
(SL_BOUNCE_LO_DIR, SL_BOUNCE_LO_B, SL_BOUNCE_NODE_ACC, SL_BOUNCE_NODE_SH, SL_BOUNCE_PROBE_DATA, SL_BOUNCE_PROBE_SH,
 SL_BOUNCE_XFER_SH, SL_BOUNCE_C0, SL_BOUNCE_LEAF_W) = range(1, 10)


class SplatlightError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                          f"g.build()'` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, u32, i32, sz = C.c_void_p, C.c_uint32, C.c_int, C.c_size_t
    sig = {
        "sl_session_acquire": (i32, [C.c_char_p, C.POINTER(C.c_int64)]),
        "sl_session_release": (i32, [C.c_int64]),
        "sl_session_error": (C.c_char_p, []),
        "sl_ctx_create": (i32, [i32, C.POINTER(vp)]),
        "sl_ctx_destroy": (None, [vp]),
        "sl_last_error": (C.c_char_p, [vp]),
        "sl_ctx_status": (i32, [vp]),
        "sl_sync": (i32, [vp]),
        "sl_buf_alloc": (i32, [vp, sz, C.POINTER(vp)]),
        "sl_buf_free": (i32, [vp, vp]),
        "sl_buf_write": (i32, [vp, vp, sz, vp, sz]),
        "sl_buf_read": (i32, [vp, vp, sz, vp, sz]),
        "sl_scene_create": (i32, [vp, C.POINTER(SceneDesc), C.POINTER(vp)]),
        "sl_scene_destroy": (None, [vp, vp]),
        "sl_scene_write_parts": (i32, [vp, vp, vp, sz]),
        "sl_scene_write_colour": (i32, [vp, vp, vp, sz]),
        "sl_scene_write_bounce": (i32, [vp, vp, vp, sz]),
        "sl_scene_write_gaussians": (i32, [vp, vp, vp, sz]),
        "sl_scene_write_materials": (i32, [vp, vp, vp, sz]),
        "sl_scene_device_bytes": (sz, [vp]),
        "sl_frame_encode": (i32, [vp, vp, vp, sz, C.POINTER(FrameOpts), vp]),
        "sl_frame_encode_noalloc": (i32, [vp, vp, vp, sz, C.POINTER(FrameOpts), vp]),
        "sl_frame_finish": (i32, [vp, C.POINTER(FrameStats)]),
        "sl_display_encode": (i32, [vp, vp, sz, vp, sz, vp, sz, C.POINTER(DisplayOpts), vp, vp]),
        "sl_scene_set_lod": (i32, [vp, vp, C.POINTER(LodTree)]),
        "sl_scene_set_fill": (i32, [vp, vp, vp, sz]),
        "sl_fill_encode": (i32, [vp, vp]),
        "sl_frame_read": (i32, [vp, u32, vp, sz, C.POINTER(sz)]),
        "sl_timers_read": (i32, [vp, C.POINTER(C.c_char_p), C.POINTER(C.c_float), u32, C.POINTER(u32)]),
        "sl_lighting_iso_init": (i32, [vp, vp]),
        "sl_lighting_encode": (i32, [vp, vp, C.POINTER(LightingDesc)]),
        "sl_scene_read": (i32, [vp, vp, u32, vp, sz, C.POINTER(sz)]),
        "sl_lighting_stats_read": (i32, [vp, vp, C.POINTER(LightingStats)]),
        "sl_batch_encode": (i32, [vp, C.POINTER(View), u32, u32]),
        "sl_ctx_set_tuning": (i32, [vp, C.POINTER(Tuning)]),
        "sl_server_create": (i32, [vp, C.POINTER(ServerOpts), C.POINTER(vp)]),
        "sl_server_submit": (i32, [vp, C.POINTER(View), C.POINTER(C.c_uint64)]),
        "sl_server_poll": (i32, [vp, C.POINTER(RequestResult), u32, C.POINTER(u32), u32]),
        "sl_server_info": (i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_uint64)]),
        "sl_server_error": (C.c_char_p, [vp]),
        "sl_server_destroy": (None, [vp]),
        "sl_server_fail_next": (i32, [vp]),
        "sl_ctx_get_tuning": (i32, [vp, C.POINTER(Tuning)]),
        "sl_batch_finish": (i32, [vp, C.POINTER(FrameStats), u32]),
        "sl_bounce_init": (i32, [vp, vp, C.POINTER(BounceTree), vp, sz]),
        "sl_bounce_encode": (i32, [vp, vp, vp, sz, u32]),
        "sl_bounce_read": (i32, [vp, vp, u32, vp, sz, C.POINTER(sz)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """One device and one HIP stream (sl_context)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        st = self.L.sl_ctx_create(device, C.byref(h))
        if st != SL_OK:
            raise SplatlightError(st, f"sl_ctx_create({device}) failed (no usable HIP device?)")
        self.h = h

    def check(self, st: int):
        if st != SL_OK:
            raise SplatlightError(st, (self.L.sl_last_error(self.h) or b"").decode())

    def close(self):
        if self.h:
            self.L.sl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.L.sl_sync(self.h))

    # buffers
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self.check(self.L.sl_buf_alloc(self.h, nbytes, C.byref(p)))
        return p.value

    def free(self, dev: int):
        self.check(self.L.sl_buf_free(self.h, C.c_void_p(dev)))

    def write(self, dev: int, host: np.ndarray, offset: int = 0):
        host = np.ascontiguousarray(host)
        self.check(self.L.sl_buf_write(self.h, C.c_void_p(dev), offset, _ptr(host), host.nbytes))

    def read(self, dev: int, nbytes: int, dtype=np.uint8, offset: int = 0) -> np.ndarray:
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        self.check(self.L.sl_buf_read(self.h, C.c_void_p(dev), offset, _ptr(out), out.nbytes))
        return out


class Scene:
    """Device-resident scene records (sl_scene)."""

    def __init__(self, ctx: Context, gaussians: np.ndarray, parts: np.ndarray, emitters: np.ndarray | None = None,
                 materials: np.ndarray | None = None, nodes: np.ndarray | None = None,
                 vgrid: np.ndarray | None = None, storage: int = SL_STORAGE_F32):
        self.ctx = ctx
        self._keep = [np.ascontiguousarray(a) for a in (gaussians, parts)]
        g, p = self._keep
        d = SceneDesc()
        d.gaussians, d.gaussian_count, d.gaussian_stride = _ptr(g), g.shape[0], g.dtype.itemsize
        d.parts, d.part_count, d.part_stride = _ptr(p), p.shape[0], p.dtype.itemsize
        if materials is not None:
            m = np.ascontiguousarray(materials)
            self._keep.append(m)
            d.materials, d.material_count, d.material_stride = _ptr(m), m.shape[0], m.dtype.itemsize
        if nodes is not None:
            n = np.ascontiguousarray(nodes)
            self._keep.append(n)
            d.nodes, d.node_count, d.node_stride = _ptr(n), n.shape[0], n.dtype.itemsize
        if vgrid is not None:
            v = np.ascontiguousarray(vgrid)
            self._keep.append(v)
            d.vgrid, d.vgrid_count, d.vgrid_stride = _ptr(v), v.shape[0], v.reshape(v.shape[0], -1).shape[1] * v.itemsize
        if emitters is not None and emitters.shape[0] > 0:
            e = np.ascontiguousarray(emitters)
            self._keep.append(e)
            d.emitters, d.emitter_count = _ptr(e), e.shape[0]
        d.storage = storage
        h = C.c_void_p()
        ctx.check(ctx.L.sl_scene_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.count = g.shape[0]
        self.n_emit = 0 if emitters is None else emitters.shape[0]
        self._keep = None

    def destroy(self):
        if self.h:
            self.ctx.L.sl_scene_destroy(self.ctx.h, self.h)
            self.h = None

    def write_parts(self, parts: np.ndarray):
        p = np.ascontiguousarray(parts)
        self.ctx.check(self.ctx.L.sl_scene_write_parts(self.ctx.h, self.h, _ptr(p), p.nbytes))

    def write_colour(self, rgba: np.ndarray):
        c = np.ascontiguousarray(rgba, dtype=np.float32).reshape(-1, 4)
        self.ctx.check(self.ctx.L.sl_scene_write_colour(self.ctx.h, self.h, _ptr(c), c.shape[0]))

    def write_gaussians(self, gaussians: np.ndarray):
        """sl_scene_write_gaussians: the records rewritten in place (same count)."""
        g = np.ascontiguousarray(gaussians)
        self.ctx.check(self.ctx.L.sl_scene_write_gaussians(self.ctx.h, self.h, _ptr(g), g.nbytes))

    def write_materials(self, materials: np.ndarray):
        """sl_scene_write_materials: the material table rewritten (1 to 256 records of 32 B)."""
        m = np.ascontiguousarray(materials)
        self.ctx.check(self.ctx.L.sl_scene_write_materials(self.ctx.h, self.h, _ptr(m), m.nbytes))

    def write_bounce(self, lo_b: np.ndarray | None):
        if lo_b is None:
            self.ctx.check(self.ctx.L.sl_scene_write_bounce(self.ctx.h, self.h, None, 0))
            return
        b = np.ascontiguousarray(lo_b, dtype=np.float32).reshape(-1, 12)
        self.ctx.check(self.ctx.L.sl_scene_write_bounce(self.ctx.h, self.h, _ptr(b), b.shape[0]))

    def set_lod(self, nodes: np.ndarray, kids: np.ndarray, node_gaussians: np.ndarray,
                emit_part: int = 0xFFFFFFFF, size_rule: int = 1):
        """sl_scene_set_lod: the asset's tree (all nodes, leaves first; kids) and its internal nodes as Gaussian
        records (lod.node_gaussians)."""
        nd = np.ascontiguousarray(nodes)
        kd = np.ascontiguousarray(kids, np.uint32)
        ng = np.ascontiguousarray(node_gaussians)
        t = LodTree(_ptr(nd), nd.shape[0], nd.nbytes // max(nd.shape[0], 1), _ptr(kd), kd.size, _ptr(ng),
                    ng.nbytes // 64, emit_part, size_rule)
        self.ctx.check(self.ctx.L.sl_scene_set_lod(self.ctx.h, self.h, C.byref(t)))

    def set_fill(self, fill: np.ndarray):
        """sl_scene_set_fill: the display fill table (offsets [count + 1], then (source, weight bits) pairs)."""
        f = np.ascontiguousarray(fill, np.uint32)
        self.ctx.check(self.ctx.L.sl_scene_set_fill(self.ctx.h, self.h, _ptr(f), f.size))

    def fill(self):
        """sl_fill_encode: the display fill on the lit colour and the bounce toward the eye."""
        self.ctx.check(self.ctx.L.sl_fill_encode(self.ctx.h, self.h))

    @property
    def device_bytes(self) -> int:
        return int(self.ctx.L.sl_scene_device_bytes(self.h))

    def iso_init(self):
        self.ctx.check(self.ctx.L.sl_lighting_iso_init(self.ctx.h, self.h))

    def light(self, frame_uniforms: bytes, emit_params: bytes, emit_nodes: np.ndarray, ls_params: bytes | None = None,
              env: bytes | None = None, sun_params: bytes | None = None, sun_layers: bytes | None = None,
              timers: bool = False):
        """sl_lighting_encode: the lighting passes for this scene, leaving the lit colour on the device."""
        keep = []

        def buf(b):
            if b is None:
                return None
            a = (C.c_uint8 * len(b)).from_buffer_copy(b)
            keep.append(a)
            return C.cast(a, C.c_void_p)
        en = np.ascontiguousarray(emit_nodes, np.float32)
        d = LightingDesc()
        d.frame_uniforms = buf(frame_uniforms)
        d.emit_params = buf(emit_params)
        d.emit_nodes = en.ctypes.data_as(C.c_void_p)
        d.emit_node_count = en.nbytes // 96
        d.ls_params = buf(ls_params)
        d.env = buf(env)
        d.sun_params = buf(sun_params)
        d.sun_layers = buf(sun_layers)
        d.sun_layer_count = 0 if sun_layers is None else len(sun_layers) // 32
        d.flags = SL_LIGHT_TIMERS if timers else 0
        self.ctx.check(self.ctx.L.sl_lighting_encode(self.ctx.h, self.h, C.byref(d)))

    def lighting_stats(self) -> dict:
        """Records and (record, tile) pairs of the last relighting's opacity-map splats."""
        st = LightingStats()
        self.ctx.check(self.ctx.L.sl_lighting_stats_read(self.ctx.h, self.h, C.byref(st)))
        return {name: int(getattr(st, name)) for name, _ in LightingStats._fields_}

    def bounce_init(self, nodes: np.ndarray, kids: np.ndarray, max_depth: int, probe_params: bytes):
        """sl_bounce_init: the scene's tree (leaves first) and the probe layout's ProbeParams."""
        nd = np.ascontiguousarray(nodes)
        kd = np.ascontiguousarray(kids, np.uint32)
        t = BounceTree(nd.ctypes.data_as(C.c_void_p), nd.nbytes // 96, 96, kd.ctypes.data_as(C.c_void_p), kd.size,
                       max_depth)
        pp = (C.c_uint8 * len(probe_params)).from_buffer_copy(probe_params)
        self.ctx.check(self.ctx.L.sl_bounce_init(self.ctx.h, self.h, C.byref(t), C.cast(pp, C.c_void_p), len(probe_params)))

    def bounce_encode(self, probe_params: bytes, flags: int = 0):
        """sl_bounce_encode: one encodeBounce after this frame's sl_lighting_encode."""
        pp = (C.c_uint8 * len(probe_params)).from_buffer_copy(probe_params)
        self.ctx.check(self.ctx.L.sl_bounce_encode(self.ctx.h, self.h, C.cast(pp, C.c_void_p), len(probe_params), flags))

    def bounce_read(self, which: int, dtype=np.float32) -> np.ndarray:
        n = C.c_size_t()
        self.ctx.check(self.ctx.L.sl_bounce_read(self.ctx.h, self.h, which, None, 0, C.byref(n)))
        out = np.empty(n.value // np.dtype(dtype).itemsize, dtype)
        self.ctx.check(self.ctx.L.sl_bounce_read(self.ctx.h, self.h, which, _ptr(out), out.nbytes, C.byref(n)))
        return out

    def read(self, which: int, dtype=np.float32) -> np.ndarray:
        n = C.c_size_t()
        self.ctx.check(self.ctx.L.sl_scene_read(self.ctx.h, self.h, which, None, 0, C.byref(n)))
        out = np.empty(n.value // np.dtype(dtype).itemsize, dtype)
        self.ctx.check(self.ctx.L.sl_scene_read(self.ctx.h, self.h, which, _ptr(out), out.nbytes, C.byref(n)))
        return out


def frame_encode(ctx: Context, scene: Scene, frame: bytes, width: int, height: int, out_dev: int,
                 out_format: int = SL_OUT_RGBA32F, timers: bool = False, counting: bool = False):
    """sl_frame_encode; counting: the depth order by the counting scatter instead of the stable radix sort."""
    fb = (C.c_uint8 * len(frame)).from_buffer_copy(frame)
    o = FrameOpts(width, height, out_format, (SL_FRAME_TIMERS if timers else 0) | (SL_FRAME_COUNTING_SORT if counting else 0))
    ctx.check(ctx.L.sl_frame_encode(ctx.h, scene.h, fb, len(frame), C.byref(o), C.c_void_p(out_dev)))


def display_encode(ctx: Context, frame: bytes, width: int, height: int, hdr_dev: int, out_dev: int,
                   hdr_format: int = SL_OUT_RGBA32F, out_format: int = SL_DISPLAY_RGBA8, env: bytes | None = None,
                   sky_bg: bytes | None = None, exact: bool = False):
    """sl_display_encode: the HDR target (hdr_format) to width * height * 4 display bytes, enqueued on the context's
    stream; env / sky_bg: the Env (416 B) and SkyBg (1520 B) uniforms or None."""
    fb = (C.c_uint8 * len(frame)).from_buffer_copy(frame)
    eb = None if env is None else (C.c_uint8 * len(env)).from_buffer_copy(env)
    bb = None if sky_bg is None else (C.c_uint8 * len(sky_bg)).from_buffer_copy(sky_bg)
    o = DisplayOpts(width, height, hdr_format, out_format, SL_DISPLAY_EXACT if exact else 0)
    ctx.check(ctx.L.sl_display_encode(ctx.h, fb, len(frame), eb, 0 if env is None else len(env), bb,
                                      0 if sky_bg is None else len(sky_bg), C.byref(o), C.c_void_p(hdr_dev),
                                      C.c_void_p(out_dev)))


# Size of the device array the no-alloc frame lays its workspace over (SL_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 64 << 20


def frame_encode_noalloc(ctx: Context, scene: Scene, frame: bytes, width: int, height: int, out_dev: int,
                         out_format: int = SL_OUT_RGBA32F, timers: bool = False):
    """sl_frame_encode_noalloc: the same frame with its workspace in the device's fixed scratch array."""
    fb = (C.c_uint8 * len(frame)).from_buffer_copy(frame)
    o = FrameOpts(width, height, out_format, SL_FRAME_TIMERS if timers else 0)
    ctx.check(ctx.L.sl_frame_encode_noalloc(ctx.h, scene.h, fb, len(frame), C.byref(o), C.c_void_p(out_dev)))


def batch_encode(ctx: Context, views, graph: bool = False, timers: bool = False):
    """sl_batch_encode: views = [(scene, frame_bytes, width, height, out_dev, out_format)], one launch per stage."""
    arr = (View * len(views))()
    keep = []
    for k, (scene, frame, w, h, out_dev, fmt) in enumerate(views):
        fb = (C.c_uint8 * len(frame)).from_buffer_copy(frame)
        keep.append(fb)
        arr[k] = View(scene.h, C.cast(fb, C.c_void_p), len(frame), w, h, fmt, out_dev)
    flags = (SL_BATCH_GRAPH if graph else 0) | (SL_FRAME_TIMERS if timers else 0)
    ctx.check(ctx.L.sl_batch_encode(ctx.h, arr, len(views), flags))


def batch_finish(ctx: Context, n: int) -> list:
    st = (FrameStats * n)()
    ctx.check(ctx.L.sl_batch_finish(ctx.h, st, n))
    return [st[k] for k in range(n)]


class Server:
    """sl_server_*: continuous batching of render requests over resident scenes."""

    def __init__(self, ctx: Context, hbm_budget: int, lanes: int = 2, max_batch_views: int = 32):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(ctx.L.sl_server_create(ctx.h, C.byref(ServerOpts(hbm_budget, lanes, max_batch_views)),
                                         C.byref(self.h)))

    def submit(self, scene: "Scene", frame: bytes, width: int, height: int, out_dev: int,
               out_format: int = SL_OUT_RGBA32F) -> int:
        fb = (C.c_uint8 * len(frame)).from_buffer_copy(frame)
        v = View(scene.h, C.cast(fb, C.c_void_p), len(frame), width, height, out_format, out_dev)
        rid = C.c_uint64()
        st = self.ctx.L.sl_server_submit(self.h, C.byref(v), C.byref(rid))
        if st != 0:
            raise SplatlightError(st, "sl_server_submit refused the request")
        return int(rid.value)

    def poll(self, wait_us: int = 0, cap: int = 256) -> list:
        out = (RequestResult * cap)()
        n = C.c_uint32()
        self.ctx.L.sl_server_poll(self.h, out, cap, C.byref(n), wait_us)
        return [out[k] for k in range(n.value)]

    def info(self):
        p, r, b = C.c_uint32(), C.c_uint32(), C.c_uint64()
        self.ctx.L.sl_server_info(self.h, C.byref(p), C.byref(r), C.byref(b))
        return int(p.value), int(r.value), int(b.value)

    def error(self) -> str:
        return (self.ctx.L.sl_server_error(self.h) or b"").decode()

    def fail_next(self):
        """sl_server_fail_next: the next completion check of a busy lane reports a device error (host side)."""
        self.ctx.check(self.ctx.L.sl_server_fail_next(self.h))

    def close(self):
        if self.h:
            self.ctx.L.sl_server_destroy(self.h)
            self.h = C.c_void_p()


def set_tuning(ctx: Context, composite_batch: int, radix_items: int, grid_mul: int, radix_bits: int = 8,
               tile_px: int = 16):
    """sl_ctx_set_tuning: composite LDS batch depth, onesweep keys per lane, persistent sort workgroups per CU, the
    tile sort's digit width (8 or 11 bits) and the binning tile (16 or 32 px; either is composited in 16x16-pixel
    workgroups)."""
    ctx.check(ctx.L.sl_ctx_set_tuning(ctx.h, C.byref(Tuning(composite_batch, radix_items, grid_mul, radix_bits,
                                                             tile_px, 0))))


def get_tuning(ctx: Context) -> dict:
    t = Tuning()
    ctx.check(ctx.L.sl_ctx_get_tuning(ctx.h, C.byref(t)))
    return {"composite_batch": t.composite_batch, "radix_items": t.radix_items, "grid_mul": t.grid_mul,
            "radix_bits": t.radix_bits, "tile_px": t.tile_px}


def frame_finish(ctx: Context) -> FrameStats:
    s = FrameStats()
    ctx.check(ctx.L.sl_frame_finish(ctx.h, C.byref(s)))
    return s


def frame_read(ctx: Context, which: int, dtype=np.uint32) -> np.ndarray:
    n = C.c_size_t()
    ctx.check(ctx.L.sl_frame_read(ctx.h, which, None, 0, C.byref(n)))
    out = np.empty(n.value // np.dtype(dtype).itemsize, dtype)
    ctx.check(ctx.L.sl_frame_read(ctx.h, which, _ptr(out), out.nbytes, C.byref(n)))
    return out


def timers(ctx: Context) -> dict:
    names = (C.c_char_p * 32)()
    ms = (C.c_float * 32)()
    n = C.c_uint32()
    ctx.check(ctx.L.sl_timers_read(ctx.h, names, ms, 32, C.byref(n)))
    return {names[k].decode(): float(ms[k]) for k in range(n.value)}
