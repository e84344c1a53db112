"""Byte layouts of the "gs-runtime" records and the per-frame uniforms.

These are the external contract of the splat frame path: the host hands these
exact bytes to the library, which validates them and keeps its own device
layouts (SoA) internally.

Layouts follow:
  * Gaussian, 64 B       shaders/common.wgsl:6-14, src/asset.js:9
  * Part, 128 B          shaders/common.wgsl:17-25, src/asset.js:10
  * Node, 96 B           runtime/wgsl/common.wgsl:12-19
  * Material, 32 B       runtime/wgsl/common.wgsl:59-65, src/asset.js:12
  * RenderFrame, 1200 B  shaders/common.wgsl:44-77, written by src/main.js:526-556
  * FrameUniforms, 64 B  runtime/wgsl/common.wgsl:39-47, src/lighting.js:1229-1238
  * EmitParams, 64 B     runtime/wgsl/emissive.wgsl:34-53, src/lighting.js:657-673
  * EmitNode, 96 B       runtime/wgsl/emissive.wgsl:58-65, src/lighting.js:61-65
"""
from __future__ import annotations

import math
import numpy as np

GAUSSIAN_DTYPE = np.dtype({
    "names": ["mu", "kappa", "quat", "scale", "part_surface", "albedo", "flags"],
    "formats": [("<f4", 3), "<f4", ("<f4", 4), ("<f4", 3), "<u4", ("<f4", 3), "<u4"],
    "offsets": [0, 12, 16, 32, 44, 48, 60],
    "itemsize": 64,
})

PART_DTYPE = np.dtype({
    "names": ["xf", "xf_prev", "bound", "root", "g_lo", "g_hi", "flags"],
    "formats": [("<f4", (3, 4)), ("<f4", (3, 4)), ("<f4", 4), "<u4", "<u4", "<u4", "<u4"],
    "offsets": [0, 48, 96, 112, 116, 120, 124],
    "itemsize": 128,
})

NODE_DTYPE = np.dtype({
    "names": ["mean", "radius", "cinv_diag", "kappa", "cinv_off", "smax", "sn_diag", "part", "sn_off",
              "child_lo", "child_hi", "depth", "pad0", "pad1"],
    "formats": [("<f4", 3), "<f4", ("<f4", 3), "<f4", ("<f4", 3), "<f4", ("<f4", 3), "<u4", ("<f4", 3),
                "<u4", "<u4", "<u4", "<u4", "<u4"],
    "offsets": [0, 12, 16, 28, 32, 44, 48, 60, 64, 76, 80, 84, 88, 92],
    "itemsize": 96,
})

MATERIAL_DTYPE = np.dtype({
    "names": ["flake", "phase_samples", "vfloor", "flags", "pad0", "pad1"],
    "formats": [("<f4", 3), "<u4", "<f4", "<u4", "<u4", "<u4"],
    "offsets": [0, 12, 16, 20, 24, 28],
    "itemsize": 32,
})

EMIT_NODE_DTYPE = np.dtype({
    "names": ["mean", "radius", "m_diag", "tag", "m_off", "child_lo", "Le", "child_hi", "c_diag", "nrm",
              "c_off", "part"],
    "formats": [("<f4", 3), "<f4", ("<f4", 3), "<u4", ("<f4", 3), "<u4", ("<f4", 3), "<u4", ("<f4", 3),
                "<u4", ("<f4", 3), "<u4"],
    "offsets": [0, 12, 16, 28, 32, 44, 48, 60, 64, 76, 80, 92],
    "itemsize": 96,
})

VGRID_BYTES = 128
FRAME_BYTES = 1200          # src/renderer.js:9
SPLAT_BYTES = 48            # src/renderer.js:10
NUM_BUCKETS = 65536         # shaders/common.wgsl:37
INVALID_KEY = 0xFFFFFFFF    # shaders/common.wgsl:38
MAX_PARTS = 32              # shaders/common.wgsl:39
EMITTER_MATERIAL = 254      # shaders/common.wgsl:41
MAX_EMITTERS = 16384        # src/renderer.js:17
VIEW_LIT = 3
# RenderFrame.tone.x: the display's tone curve (shaders/display.wgsl:140-152; AgX by default, src/main.js:198)
TONE_CLAMP, TONE_AGX, TONE_ACES, TONE_REINHARD, TONE_NEUTRAL, TONE_FILMIC = range(6)

# Material.flags bits (runtime/wgsl/common.wgsl:63)
MAT_EMISSIVE = 1
MAT_PROBE_ONLY = 2
MAT_ONE_SIDED = 4
MAT_SPECULAR = 8


def f32_bits(x: float) -> int:
    return int(np.array([x], dtype="<f4").view("<u4")[0])


def pack2x16float(a: float, b: float) -> int:
    """WGSL pack2x16float: two IEEE halves, round to nearest even."""
    h = np.array([a, b], dtype="<f2").view("<u2").astype(np.uint32)
    return int(h[0] | (h[1] << 16))


def unpack2x16float(u: int) -> tuple[float, float]:
    h = np.array([u & 0xFFFF, u >> 16], dtype="<u2").view("<f2").astype(np.float64)
    return float(h[0]), float(h[1])


# ---- camera (src/math.js:36-57, src/camera.js:32-43) -------------------------------------------------------

def look_at(eye, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """Column-major 4x4 view matrix as Float32Array(16) (src/math.js:36-46), computed in float64 like JS."""
    eye = np.asarray(eye, np.float64)
    target = np.asarray(target, np.float64)
    up = np.asarray(up, np.float64)

    def norm(a):
        return a / math.hypot(*a)
    z = norm(eye - target)
    x = norm(np.cross(up, z))
    y = np.cross(z, x)
    m = np.array([x[0], y[0], z[0], 0,
                  x[1], y[1], z[1], 0,
                  x[2], y[2], z[2], 0,
                  -x.dot(eye), -y.dot(eye), -z.dot(eye), 1], np.float64)
    return m.astype(np.float32)


def perspective(fovy, aspect, near, far) -> np.ndarray:
    """WebGPU clip-space projection (src/math.js:50-57)."""
    f = 1.0 / math.tan(fovy / 2)
    m = np.zeros(16, np.float64)
    m[0] = f / aspect
    m[5] = f
    m[10] = far / (near - far)
    m[11] = -1
    m[14] = near * far / (near - far)
    return m.astype(np.float32)


def render_frame(*, view, width, height, fovy, near=0.02, far=50.0, count, view_mode=VIEW_LIT,
                 part_material=None, palette=None, bounce_gain=1.0, emitter_count=0, albedo_scale=1.0,
                 lowpass=0.3, exposure=1.0, shutter=0.0, blur_coverage=True, tone=TONE_AGX, lod=None) -> bytes:
    """Pack a RenderFrame (1200 B) the way writeFrame does (src/main.js:526-556, 470-475).

    shutter: the motion blur's shutter fraction (palette[3].z; moving parts blur over it, 0 = off); blur_coverage:
    the blurred splats' coverage cap (palette[3].w, ?blurcov); tone: the display pass's tone curve (tone.x, word
    260, src/main.js:553); lod: (threshold px, first node slot, internal nodes) for the screen-space LOD (words
    180-182, src/main.js:548; None or threshold 0 = off). Explode, refraction, particles and the
    reflection split are left off (zero).
    """
    buf = np.zeros(FRAME_BYTES // 4, np.uint32)
    f32 = buf.view(np.float32)
    f32[0:16] = np.asarray(view, np.float32)
    f32[16:32] = perspective(fovy, width / height, near, far)
    pal = palette if palette is not None else [(1.0, 1.0, 1.0)] * 3
    for i in range(3):
        f32[32 + 4 * i: 36 + 4 * i] = [*pal[i], 1.0]
    f32[44:48] = [bounce_gain, float(emitter_count), shutter, 1.0 if blur_coverage else 0.0]   # palette[3]
    pm = part_material if part_material is not None else [0] * MAX_PARTS
    for i, m in enumerate(pm):
        buf[(48 if i < 16 else 164 - 16) + i] = m
    for k in range(4):
        f32[84 + 4 * k: 88 + 4 * k] = [1.0, 1.0, 1.0, 0.0]          # palette_ext neutral
    f32[64:68] = [0.0, 0.0, 0.0, 0.0]                                # lamp (w = 0: no flake colour)
    f32[68:72] = [4.0, 4.0, 4.0, 0.0]                                # lamp_color, w = 0: no reflection split
    focal = height / (2 * math.tan(fovy / 2))
    f32[72:78] = [width, height, focal, focal, near, far]
    buf[78] = count
    buf[79] = view_mode
    f32[80:84] = [albedo_scale, 0.01, exposure, lowpass]
    buf[260] = tone
    if lod is not None and lod[0] > 0:
        f32[180] = lod[0]
        buf[181], buf[182] = int(lod[1]), int(lod[2])
    return buf.tobytes()


def frame_uniforms(*, frame=0, gaussian_count, part_count, lamp_count=0, theta=0.0, alpha_min=1.0 / 255.0,
                   alpha_part=0.0, eye, material_count, iso=1 | 2, stratum_step=0, light_samples_axis=0) -> bytes:
    """FrameUniforms (64 B) as writeUniform packs it (src/lighting.js:1229-1238). stratum_step and
    light_samples_axis (the reference writes 7 and the asset's light_samples_axis) are read by no kernel on this
    path."""
    buf = np.zeros(16, np.uint32)
    f = buf.view(np.float32)
    buf[0] = frame
    buf[1] = gaussian_count
    buf[2] = part_count
    buf[3] = lamp_count
    f[4] = theta
    f[5] = alpha_min
    f[6] = alpha_part
    buf[7] = 1
    f[8:11] = eye
    buf[11] = stratum_step
    buf[12] = light_samples_axis
    buf[13] = material_count
    buf[14] = 0
    buf[15] = iso
    return buf.tobytes()


def emit_params(*, count, root, scale, theta=4.0, mode=1, node_count, skip_part=0xFFFFFFFF, one_sided=0,
                normal=(0.0, 0.0, 1.0), n5=0, edge_soft=0.2, lights=0) -> bytes:
    """EmitParams (64 B) as writeEmitParams packs it (src/lighting.js:657-673)."""
    buf = np.zeros(16, np.uint32)
    f = buf.view(np.float32)
    buf[0] = count
    buf[1] = root
    f[2] = scale
    f[3] = theta
    buf[4] = mode
    buf[5] = node_count
    buf[6] = skip_part
    buf[7] = one_sided
    f[8:11] = normal
    buf[11] = n5
    f[12] = edge_soft
    buf[15] = lights
    return buf.tobytes()


def identity_part(g_lo: int, g_hi: int, bound=(0.0, 0.0, 0.0, 1.0), flags=0) -> np.ndarray:
    p = np.zeros(1, PART_DTYPE)
    p["xf"][0] = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]]
    p["xf_prev"][0] = p["xf"][0]
    p["bound"][0] = bound
    p["g_lo"][0] = g_lo
    p["g_hi"][0] = g_hi
    p["flags"][0] = flags
    return p


STORAGE_F32, STORAGE_F16, STORAGE_BF16 = 0, 1, 2


def _bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 -> float32, rounding to nearest even (NaN kept)."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    nan = (b & 0x7FFFFFFF) > 0x7F800000
    r = np.where(nan, ((b >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32)


def dequantized(g: np.ndarray, storage: int) -> np.ndarray:
    """The Gaussian records as a scene of the given storage holds them: quat, scale and albedo rounded to half or
    bfloat16 (nearest even) and back; mu, kappa, part_surface and flags unchanged."""
    if storage == STORAGE_F32:
        return g
    out = g.copy()
    for f in ("quat", "scale", "albedo"):
        v = np.ascontiguousarray(g[f], np.float32)
        out[f] = v.astype(np.float16).astype(np.float32) if storage == STORAGE_F16 else _bf16(v)
    return out
