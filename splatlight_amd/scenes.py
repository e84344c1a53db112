"""Seeded synthetic scenes for the splat frame path (numpy, PCG64).

The reference ships no scene generator; these scenes restate the workloads the path is measured on. Every
convention a generator has to pick is fixed here once, and the oracle and the HIP path both read the records
this module writes:

* World frame: right-handed, z up, metres (src/math.js:15-16). The Cornell box is specified y-up; a y-up point
  (x, y, z) maps to (x, -z, y), so the y-up camera (0, 1, 3.5) -> (0, -3.5, 1) looking at (0, 0, 1). The hill is
  specified z-up already (z = 0.3 sin x cos y) and is used as is.
* Opacity: records store peak extinction kappa (shaders/common.wgsl:8). kappa = -ln(1 - opacity) / (sqrt(2 pi)
  s_min) gives the stated opacity through the centre along the short axis.
* Scales: log-scales drawn per axis; the "wall-normal" axis (axis 0 before sorting) is shifted by -1.5, then
  the three are sorted ascending because the record's x is the short axis (shaders/common.wgsl:10).
* Rotations: uniform unit quaternions (x, y, z, w) from normalised 4D Gaussians.
* Roughness U[0.2, 1] is quantised to 8 materials with flake (1, r^2, r^2) and 8 phase samples, the way
  setRoughness writes them (src/lighting.js:1215-1227). All are matte, so roughness does not change the colour
  (runtime/wgsl/emissive.wgsl:898).
* Lights: emissive Gaussians in the `emitters` buffer (world space, albedo = Le), drawn in the splat slots after
  `count + 1` (shaders/project.wgsl:123-127) and lighting the scene through the emitter tree
  (src/lighting.js:99-169) with EP.scale 1; the Cornell lamp emits one-sided toward its short axis (down), the
  hill's lanterns two-sided. The hill's sun: 35 degrees up, azimuth 30 degrees from +x, angular std 0.27 degrees.
* Parts: the Cornell room (walls), short box and tall box; the hill's ground, trunk and canopy. Shadows between
  parts come from the opacity maps (a part's own occlusion is its visibility grid's).
* Leaf nodes are derived from the Gaussian: mean mu, radius 3 max(s), cinv = R diag(1/s^2) R^T, kappa, smax =
  max(s); the baked SGGX S_hat is taken as the identity and the visibility grid as 255 (V = 1) because their
  bakers are not part of the reference.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import records as R


@dataclasses.dataclass
class Scene:
    name: str
    gaussians: np.ndarray            # GAUSSIAN_DTYPE [G]
    parts: np.ndarray                # PART_DTYPE [P]
    materials: np.ndarray            # MATERIAL_DTYPE [M]
    emitters: np.ndarray             # GAUSSIAN_DTYPE [E], world space, albedo = Le
    eye: tuple
    target: tuple
    fov_deg: float
    width: int
    height: int
    far: float
    sun: dict | None = None          # {"dir": unit vector toward the sun, "E": rgb, "std": rad, "res": int}
    sky: np.ndarray | None = None    # 16 x rgb real SH of the sky radiance (Env.sh), seen through the probes
    one_sided: bool = False          # emitters emit only toward +short axis (EmitParams.one_sided)
    emit_set: dict | None = None     # an asset's emissive part as the light set (lighting.asset_emitters + "part")
    room: dict | None = None         # an asset's room box {"x", "y", "z": [lo, hi]} (the probe lattice's)

    @property
    def count(self) -> int:
        return int(self.gaussians.shape[0])

    def view(self) -> np.ndarray:
        return R.look_at(self.eye, self.target, (0.0, 0.0, 1.0))

    def frame_bytes(self, width=None, height=None, eye=None, target=None, view_mode=R.VIEW_LIT,
                    bounce_gain=0.0, shutter=0.0, blur_coverage=True, exposure=1.0, tone=R.TONE_AGX) -> bytes:
        w = width or self.width
        h = height or self.height
        v = R.look_at(eye or self.eye, target or self.target, (0.0, 0.0, 1.0))
        return R.render_frame(view=v, width=w, height=h, fovy=math.radians(self.fov_deg), near=0.02,
                              far=self.far, count=self.count, view_mode=view_mode,
                              emitter_count=self.emitters.shape[0], bounce_gain=bounce_gain, shutter=shutter,
                              blur_coverage=blur_coverage, exposure=exposure, tone=tone)


def _quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def _scales(rng, n, mu, sd):
    ls = rng.normal(mu, sd, (n, 3))
    ls[:, 0] -= 1.5
    return np.sort(np.exp(ls), axis=1)


def _opacity_kappa(rng, n, smin):
    op = 1.0 / (1.0 + np.exp(-rng.normal(1.5, 1.0, n)))
    op = np.minimum(op, 0.999)
    return -np.log1p(-op) / (math.sqrt(2 * math.pi) * smin)


def _materials(rough_levels=8):
    m = np.zeros(rough_levels, R.MATERIAL_DTYPE)
    for i in range(rough_levels):
        r = 0.2 + (i + 0.5) * 0.8 / rough_levels
        a = max(r * r, 1e-4)
        m["flake"][i] = (1.0, a, a)
        m["phase_samples"][i] = 8
        m["vfloor"][i] = 0.02
    return m


def _pack(rng, mu, scale_mu, scale_sd, albedo):
    n = mu.shape[0]
    g = np.zeros(n, R.GAUSSIAN_DTYPE)
    s = _scales(rng, n, scale_mu, scale_sd)
    g["mu"] = mu
    g["quat"] = _quats(rng, n)
    g["scale"] = s
    g["kappa"] = _opacity_kappa(rng, n, s[:, 0])
    g["albedo"] = albedo
    rough = rng.uniform(0.2, 1.0, n)
    g["flags"] = np.minimum(((rough - 0.2) / 0.8 * 8).astype(np.int64), 7)
    g["part_surface"] = 0
    return g


def _emitter_records(mu, le, scale, quat=None):
    n = mu.shape[0]
    e = np.zeros(n, R.GAUSSIAN_DTYPE)
    e["mu"] = mu
    e["scale"] = np.sort(np.asarray(scale, np.float64), axis=-1)
    e["quat"] = quat if quat is not None else np.array([0.0, 0.0, 0.0, 1.0])
    e["kappa"] = -math.log(0.05) / (math.sqrt(2 * math.pi) * float(np.min(scale)))
    e["albedo"] = le
    e["part_surface"] = 0
    e["flags"] = 0
    return e


def _parts(g, part_of):
    """Parts table: identity poses, rest bounds (centre, radius) and Gaussian ranges; part_of must be sorted."""
    g["part_surface"] = part_of.astype(np.uint32)
    out = []
    for p in range(int(part_of.max()) + 1):
        idx = np.nonzero(part_of == p)[0]
        mu = g["mu"][idx].astype(np.float64)
        lo, hi = mu.min(0), mu.max(0)
        c = 0.5 * (lo + hi)
        r = float(np.linalg.norm(mu - c, axis=1).max() + 3 * g["scale"][idx].max())
        out.append(R.identity_part(int(idx[0]), int(idx[-1]) + 1, bound=(*c, r), flags=1))   # static
    return np.concatenate(out)


def _far(g):
    mu = g["mu"].astype(np.float64)
    return max(50.0, 1.5 * float(np.linalg.norm(mu.max(0) - mu.min(0))))


def cornell(n: int = 200_000, seed: int = 0, width: int = 640, height: int = 480, scale_mu: float = -4.5) -> Scene:
    """Cornell-like box: 5 walls of a 2 m cube plus a 0.6 m and a 1.2 m box, 16 ceiling emitters of radiance 12."""
    rng = np.random.default_rng(seed)
    # surfaces (z-up): (centre, u axis, v axis, half extents, albedo or None)
    red, green = (0.8, 0.1, 0.1), (0.1, 0.8, 0.1)
    surf = [
        ((0, 0, 0.0), (1, 0, 0), (0, 1, 0), (1, 1), None),      # floor
        ((0, 0, 2.0), (1, 0, 0), (0, 1, 0), (1, 1), None),      # ceiling
        ((0, 1.0, 1.0), (1, 0, 0), (0, 0, 1), (1, 1), None),    # back wall
        ((-1.0, 0, 1.0), (0, 1, 0), (0, 0, 1), (1, 1), red),    # left wall
        ((1.0, 0, 1.0), (0, 1, 0), (0, 0, 1), (1, 1), green),   # right wall
    ]
    for (cx, cy), hgt in (((0.40, -0.35), 0.6), ((-0.35, 0.30), 1.2)):
        hw = 0.3
        surf.append(((cx, cy, hgt), (1, 0, 0), (0, 1, 0), (hw, hw), None))                  # top
        surf.append(((cx - hw, cy, hgt / 2), (0, 1, 0), (0, 0, 1), (hw, hgt / 2), None))    # four sides
        surf.append(((cx + hw, cy, hgt / 2), (0, 1, 0), (0, 0, 1), (hw, hgt / 2), None))
        surf.append(((cx, cy - hw, hgt / 2), (1, 0, 0), (0, 0, 1), (hw, hgt / 2), None))
        surf.append(((cx, cy + hw, hgt / 2), (1, 0, 0), (0, 0, 1), (hw, hgt / 2), None))
    area = np.array([4 * a * b for (_, _, _, (a, b), _) in surf])
    which = rng.choice(len(surf), size=n, p=area / area.sum())
    which.sort()
    mu = np.zeros((n, 3))
    alb = rng.uniform(0.0, 1.0, (n, 3))
    uv = rng.uniform(-1.0, 1.0, (n, 2))
    for k, (c, u, v, (a, b), col) in enumerate(surf):
        m = which == k
        mu[m] = np.asarray(c) + uv[m, :1] * a * np.asarray(u) + uv[m, 1:] * b * np.asarray(v)
        if col is not None:
            alb[m] = col
    g = _pack(rng, mu, scale_mu, 0.4, alb)
    parts = _parts(g, np.select([which < 5, which < 10], [0, 1], 2))      # room, short box, tall box
    # 16 emitters: a 4 x 4 grid under the ceiling, short axis down (x -> -z)
    gx, gy = np.meshgrid(np.linspace(-0.3, 0.3, 4), np.linspace(-0.3, 0.3, 4))
    emu = np.stack([gx.ravel(), gy.ravel(), np.full(16, 1.97)], 1)
    s45 = math.sqrt(0.5)
    em = _emitter_records(emu, np.full((16, 3), 12.0), (0.005, 0.04, 0.04), quat=(0.0, s45, 0.0, s45))
    return Scene("cornell", g, parts, _materials(), em, (0.0, -3.5, 1.0), (0.0, 0.0, 1.0), 40.0,
                 width, height, _far(g), one_sided=True)


def hill(n: int = 1_000_000, seed: int = 1, width: int = 1920, height: int = 1080, n_emit: int = 64,
         shadow_res: int = 2048) -> Scene:
    """Outdoor hill: z = 0.3 sin x cos y over a 20 m square, a 5 m trunk and a 4 m canopy; 64 emitters of
    radiance U[1, 20] and a sun at 35 degrees elevation."""
    rng = np.random.default_rng(seed)
    n_trunk = n // 20
    n_canopy = (3 * n) // 20
    n_ground = n - n_trunk - n_canopy
    xy = rng.uniform(-10.0, 10.0, (n_ground, 2))
    ground = np.column_stack([xy, 0.3 * np.sin(xy[:, 0]) * np.cos(xy[:, 1])])
    th = rng.uniform(0, 2 * math.pi, n_trunk)
    trunk = np.column_stack([0.3 * np.cos(th), 0.3 * np.sin(th), rng.uniform(0.0, 5.0, n_trunk)])
    d = rng.standard_normal((n_canopy, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    canopy = d * np.array([2.0, 2.0, 1.5]) + np.array([0.0, 0.0, 5.5])
    mu = np.concatenate([ground, trunk, canopy])
    alb = rng.uniform(0.0, 1.0, (n, 3))
    g = _pack(rng, mu, -3.5, 0.5, alb)
    parts = _parts(g, np.repeat(np.arange(3), [n_ground, n_trunk, n_canopy]))   # ground, trunk, canopy
    exy = rng.uniform(-8.0, 8.0, (n_emit, 2))
    ez = 0.3 * np.sin(exy[:, 0]) * np.cos(exy[:, 1]) + rng.uniform(0.3, 2.0, n_emit)
    le = rng.uniform(1.0, 20.0, n_emit)
    em = _emitter_records(np.column_stack([exy, ez]), np.repeat(le[:, None], 3, 1), (0.03, 0.03, 0.03))
    el, az = math.radians(35.0), math.radians(30.0)
    sun = {"dir": (math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)),
           "E": (3.0, 2.9, 2.7), "std": math.radians(0.27), "res": shadow_res}
    # sky: radiance A + B z, brighter toward the zenith, as real SH bands 0-1
    sky = np.zeros((16, 3))
    sky[0] = np.array([0.35, 0.45, 0.60]) / 0.282095
    sky[2] = np.array([0.15, 0.20, 0.30]) / 0.488603
    return Scene("hill", g, parts, _materials(), em, (0.0, -15.0, 4.0), (0.0, 0.0, 1.5), 40.0,
                 width, height, _far(g), sun, sky=sky)


def leaf_nodes(g: np.ndarray) -> np.ndarray:
    """Leaf Node records derived from the Gaussians (runtime/wgsl/common.wgsl:12-19)."""
    n = g.shape[0]
    q = g["quat"].astype(np.float64)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    Rm = np.empty((n, 3, 3))
    Rm[:, 0, 0] = 1 - 2 * (y * y + z * z)
    Rm[:, 1, 0] = 2 * (x * y + w * z)
    Rm[:, 2, 0] = 2 * (x * z - w * y)
    Rm[:, 0, 1] = 2 * (x * y - w * z)
    Rm[:, 1, 1] = 1 - 2 * (x * x + z * z)
    Rm[:, 2, 1] = 2 * (y * z + w * x)
    Rm[:, 0, 2] = 2 * (x * z + w * y)
    Rm[:, 1, 2] = 2 * (y * z - w * x)
    Rm[:, 2, 2] = 1 - 2 * (x * x + y * y)
    s = g["scale"].astype(np.float64)
    ci = np.einsum("nij,nj,nkj->nik", Rm, 1.0 / (s * s), Rm)
    nd = np.zeros(n, R.NODE_DTYPE)
    nd["mean"] = g["mu"]
    nd["radius"] = 3 * s.max(1)
    nd["cinv_diag"] = np.stack([ci[:, 0, 0], ci[:, 1, 1], ci[:, 2, 2]], 1)
    nd["cinv_off"] = np.stack([ci[:, 0, 1], ci[:, 0, 2], ci[:, 1, 2]], 1)
    nd["kappa"] = g["kappa"]
    nd["smax"] = s.max(1)
    nd["sn_diag"] = 1.0
    nd["part"] = g["part_surface"] & 0xFFFF
    return nd


def vgrid_open(n: int) -> np.ndarray:
    return np.full((n, R.VGRID_BYTES), 255, np.uint8)


def tiny(n: int = 64, seed: int = 7, width: int = 64, height: int = 48, scale_mu: float = -2.0) -> Scene:
    """A small Cornell-like scene (larger Gaussians) for golden fixtures and quick parity checks."""
    s = cornell(n=n, seed=seed, width=width, height=height, scale_mu=scale_mu)
    s.name = "tiny"
    return s


def orbit_eye(sc: Scene, radius: float, azimuth: float, elevation: float) -> tuple:
    """A camera position on an orbit of the given radius around the scene's target (z-up)."""
    t = np.asarray(sc.target, np.float64)
    d = np.array([math.cos(elevation) * math.cos(azimuth), math.cos(elevation) * math.sin(azimuth),
                  math.sin(elevation)])
    return tuple(float(x) for x in t + radius * d)


BATCH_SIZES = ((256, 256), (640, 480), (1280, 720), (1920, 1080))


def batch_workload(n_scenes: int = 8, n_views: int = 64, lo: int = 50_000, hi: int = 2_000_000, seed: int = 0,
                   sizes=BATCH_SIZES):
    """Scenes of seeds 1..n_scenes with Gaussian counts log-uniform in [lo, hi], alternately Cornell-like (odd
    seeds) and hill (even seeds), and n_views views: a scene, a size from `sizes` and a camera on a random orbit
    of radius U[2, 6] m around the scene's target. Returns (scenes, [(scene index, width, height, eye)])."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(lo), math.log(hi), n_scenes)).astype(int)
    scenes = []
    for k in range(n_scenes):
        sd = k + 1
        scenes.append(cornell(n=int(counts[k]), seed=sd) if sd % 2 else hill(n=int(counts[k]), seed=sd))
    views = []
    for _ in range(n_views):
        si = int(rng.integers(n_scenes))
        w, h = sizes[int(rng.integers(len(sizes)))]
        eye = orbit_eye(scenes[si], rng.uniform(2.0, 6.0), rng.uniform(0.0, 2 * math.pi),
                        rng.uniform(math.radians(10.0), math.radians(40.0)))
        views.append((si, int(w), int(h), eye))
    return scenes, views


def serving_workload(n_scenes: int = 32, lo: int = 100_000, hi: int = 3_000_000, first_seed: int = 10,
                     seed: int = 0):
    """Resident scenes of seeds first_seed.. with Gaussian counts log-uniform in [lo, hi], Cornell-like (even
    seeds) and hill (odd seeds) alternately."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(lo), math.log(hi), n_scenes)).astype(int)
    out = []
    for k in range(n_scenes):
        sd = first_seed + k
        out.append(cornell(n=int(counts[k]), seed=sd) if sd % 2 == 0 else hill(n=int(counts[k]), seed=sd))
    return out


def request_stream(scs, rate: float, count: int, seed: int = 1, sizes=BATCH_SIZES):
    """Poisson arrivals at `rate` per second: [(arrival s, scene index, width, height, eye)]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for k in range(count):
        si = int(rng.integers(len(scs)))
        w, h = sizes[int(rng.integers(len(sizes)))]
        eye = orbit_eye(scs[si], rng.uniform(2.0, 6.0), rng.uniform(0.0, 2 * math.pi),
                        rng.uniform(math.radians(10.0), math.radians(40.0)))
        out.append((float(t[k]), si, int(w), int(h), eye))
    return out
