"""CPU pins of the display oracle (shaders/display.wgsl:22-189) and of the Env / SkyBg packing
(src/environment.js:111-153): what the mathematics fixes, independent of the restatement's own expressions.

* the background sky's SH recurrence equals sky_radiance's 16 explicit basis functions (bands 0-3) and scipy's
  complex spherical harmonics made real without the Condon-Shortley phase (bands 0-8, every index l*l + l + m);
* the glow and the lights' discs peak at irr / (2 pi s^2) and integrate over the sphere to irr (1 - e^{-2/s^2});
  a visible-disc std replaces the light's; a horizon lobe is cos(d) exp(-sin(d)^2 / (2 sigma^2)) along the horizon;
  below the 1-degree blend the ground colour, above it the sky;
* the view direction through a pixel from the camera's own basis (eye, target, up, focal);
* the heat ramp's named colours; sRGB's 8-bit table points; each tone curve's fixed points, limits and invariants
  (ACES reaching 1 at the root of 0.08 x^2 - 0.56 x - 0.14, Reinhard's x / (1 + x) and hue, PBR Neutral's identity
  band and toe, Hable's zero and white point, AgX's black, saturation and grey axis);
* the whole pass: covered pixels are the tone-mapped colour, bare ones add sky and discs at their view direction,
  bytes are floor(255 v + 0.5);
* the uniforms: SkyBg's words at the struct's offsets for the hill asset's own environment.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import special

import oracle
from splatlight_amd import environment as E
from splatlight_amd import records as R

HERE = os.path.dirname(os.path.abspath(__file__))
ENVS = json.load(open(os.path.join(HERE, "golden", "environments.json")))


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def sphere_dirs(n, seed=0):
    return unit(np.random.default_rng(seed).standard_normal((n, 3))).astype(np.float32)


def bg_words(coeffs=None, glow=None, ring=None, ground=None, ring_par=(0.0, 0.0), disc_std=(0, 0, 0, 0)):
    u = np.zeros(380, np.uint32)
    f = u.view(np.float32)
    if coeffs is not None:
        for k, c in enumerate(coeffs):
            f[4 * k: 4 * k + 3] = c
        u[324] = len(coeffs)
    f[328:332] = disc_std
    if ring is not None:
        for k, c in enumerate(ring):
            f[4 * (83 + k): 4 * (83 + k) + 3] = c
        u[325] = len(ring)
        f[364:368] = ground
        f[368:370] = ring_par
    if glow is not None:
        irr, std, d = glow
        f[372:376] = [*irr, std]
        f[376:379] = d
    return u.tobytes()


def env_words(sh=None, lights=()):
    u = np.zeros(104, np.uint32)
    f = u.view(np.float32)
    if sh is not None:
        for k, c in enumerate(sh):
            f[4 * k: 4 * k + 3] = c
        u[97] = 1
    for k, (d, std, irr) in enumerate(lights):
        f[64 + 8 * k: 68 + 8 * k] = [*d, std]
        f[68 + 8 * k: 72 + 8 * k] = [*irr, 1.0]
    u[96] = len(lights)
    return u.tobytes()


def real_sh(l, m, w):
    """Real SH, index l*l + l + m, m < 0 sin and m > 0 cos, no Condon-Shortley phase, from scipy's complex ones."""
    polar = np.arccos(np.clip(w[:, 2], -1, 1))
    az = np.arctan2(w[:, 1], w[:, 0])
    y = special.sph_harm_y(l, abs(m), polar, az)
    if m == 0:
        return y.real
    s = math.sqrt(2) * (-1) ** abs(m)
    return s * (y.real if m > 0 else y.imag)


# ---- sky -----------------------------------------------------------------------------------------------------------

def test_sky_recurrence_matches_explicit_bands_0_3():
    rng = np.random.default_rng(1)
    sh = rng.uniform(-0.3, 0.3, (16, 3)).astype(np.float32)
    sh[0] = [3.0, 3.0, 3.0]                 # positive everywhere, so the max(., 0) does not hide a term
    w = sphere_dirs(4000)
    a = oracle.sky(env_words(sh=sh), None, w, 0)
    b = oracle.sky(None, bg_words(coeffs=sh), w, 1)
    np.testing.assert_allclose(b, a, rtol=0, atol=2e-5)


def test_sky_recurrence_matches_scipy_bands_0_8():
    w = sphere_dirs(3000, seed=2)
    base = np.zeros((81, 3), np.float32)
    base[0] = [10.0, 0.0, 0.0]
    ref0 = oracle.sky(None, bg_words(coeffs=base), w, 1)[:, 0]
    np.testing.assert_allclose(ref0, 10 * 0.28209479177387814, rtol=1e-6)
    for l in range(9):
        for m in range(-l, l + 1):
            c = base.copy()
            c[l * l + l + m, 0] += 1.0
            got = oracle.sky(None, bg_words(coeffs=c), w, 1)[:, 0] - ref0
            np.testing.assert_allclose(got, real_sh(l, m, w.astype(np.float64)), rtol=0, atol=3e-5,
                                       err_msg=f"l={l} m={m}")


def test_sky_falls_back_to_lighting_sky_and_clamps_at_zero():
    rng = np.random.default_rng(3)
    sh = rng.uniform(-1, 1, (16, 3)).astype(np.float32)
    w = sphere_dirs(500, seed=3)
    a = oracle.sky(env_words(sh=sh), None, w, 0)
    b = oracle.sky(env_words(sh=sh), bg_words(), w, 1)     # SkyBg without coefficients: the lighting sky
    np.testing.assert_array_equal(a, b)
    assert a.min() == 0.0 and (a == 0).any()


@pytest.mark.parametrize("std_deg", [3.0, 12.0, 30.0])
def test_glow_peak_and_integral(std_deg):
    s = math.radians(std_deg)
    d = unit([0.3, -0.5, 0.8])
    irr = (0.5, 1.0, 2.0)
    bg = bg_words(coeffs=np.zeros((1, 3), np.float32), glow=(irr, s, d))
    peak = oracle.sky(None, bg, d[None].astype(np.float32), 1)[0]
    np.testing.assert_allclose(peak, np.array(irr) / (2 * math.pi * s * s), rtol=2e-6)
    # the integral over the sphere, in coordinates about d: theta by midpoints, phi uniform
    nt, npf = 2000, 64
    th = (np.arange(nt) + 0.5) * math.pi / nt
    ph = np.arange(npf) * 2 * math.pi / npf
    t1 = unit(np.cross(d, [0.0, 0.0, 1.0]))
    t2 = np.cross(d, t1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = (np.cos(T)[..., None] * d + (np.sin(T) * np.cos(P))[..., None] * t1 + (np.sin(T) * np.sin(P))[..., None] * t2)
    v = oracle.sky(None, bg, w.reshape(-1, 3).astype(np.float32), 1).reshape(nt, npf, 3).astype(np.float64)
    integral = (v * (np.sin(T) * (math.pi / nt) * (2 * math.pi / npf))[..., None]).sum((0, 1))
    np.testing.assert_allclose(integral, np.array(irr) * (1 - math.exp(-2 / (s * s))), rtol=2e-3)


def test_light_discs_peak_and_visible_std():
    d = unit([-0.433013, 0.75, 0.5])
    irr = (4.0, 3.0, 2.0)
    s_light, s_disc = math.radians(1.1), math.radians(0.2)
    env = env_words(lights=[(d, s_light, irr)])
    # at the light's own direction the float32 dot product is 1 within an ulp, which 1 / s^2 magnifies: the peak
    # is irr / (2 pi s^2) to within that (2 ulp / s^2 relative)
    d32 = d.astype(np.float32)
    for bg, s in ((None, s_light), (bg_words(disc_std=(s_disc, 0, 0, 0)), s_disc)):
        p = oracle.sky(env, bg, d32[None], 2)[0]
        np.testing.assert_allclose(p, np.array(irr) / (2 * math.pi * s * s), rtol=2 * 2 ** -24 / (s * s) + 2e-6)
    # one disc std away: e^{(cos s - 1) / s^2} ~ e^{-1/2}
    off = (math.cos(s_light) * d + math.sin(s_light) * unit(np.cross(d, [0, 0, 1])))[None].astype(np.float32)
    q = oracle.sky(env, None, off, 2)[0]
    np.testing.assert_allclose(q / (np.array(irr) / (2 * math.pi * s_light ** 2)), math.exp(-0.5), rtol=1e-3)


def test_horizon_lobe_profile_and_ground():
    n = 8
    along, across = math.radians(25), math.radians(8)
    par = (1 / (2 * along ** 2), 1 / (2 * across ** 2))
    ground = (0.1, 0.2, 0.3, math.sin(math.radians(1)))
    for k in range(n):
        ring = np.zeros((n, 3), np.float32)
        ring[k] = [1.0, 0.5, 0.25]
        bg = bg_words(coeffs=np.zeros((1, 3), np.float32), ring=ring, ground=ground, ring_par=par)
        a = 2 * math.pi * k / n
        dl = np.linspace(-1.2, 1.2, 25)                     # azimuth offsets from the lobe
        e_hi = 2 * math.radians(1.0)                         # above the 1-degree blend: t = 1 (sky only)
        w = np.stack([np.cos(a + dl) * math.cos(e_hi), np.sin(a + dl) * math.cos(e_hi),
                      np.full_like(dl, math.sin(e_hi))], 1)
        got = oracle.sky(None, bg, w.astype(np.float32), 1)[:, 0]
        c = math.cos(e_hi)
        want = c * np.maximum(np.cos(dl), 0) * np.exp(-par[0] * (c * np.sin(dl)) ** 2 - par[1] * math.sin(e_hi) ** 2)
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-7, err_msg=f"lobe {k}")
    # below the blend: the ground; at the horizon: the mean of ground and sky
    ring = np.zeros((n, 3), np.float32)
    bg = bg_words(coeffs=np.zeros((1, 3), np.float32), ring=ring, ground=ground, ring_par=par)
    w = np.array([[math.cos(-0.1), 0, math.sin(-0.1)], [0, -0.8, -0.6], [0, 0, -1]], np.float32)
    np.testing.assert_allclose(oracle.sky(None, bg, w, 1), np.tile(ground[:3], (3, 1)), rtol=1e-6)
    w = np.array([[0.6, 0.8, 0.0]], np.float32)
    np.testing.assert_allclose(oracle.sky(None, bg, w, 1)[0], np.array(ground[:3]) * 0.5, rtol=1e-6)


# ---- camera, heat, sRGB, tone curves ----------------------------------------------------------------------------------

def test_view_dir_from_camera_basis():
    eye, target, up = np.array([3.0, -4.0, 2.0]), np.array([0.5, 1.0, 0.3]), np.array([0.0, 0.0, 1.0])
    W, H, fov = 320, 200, math.radians(50)
    fb = R.render_frame(view=R.look_at(eye, target, up), width=W, height=H, fovy=fov, count=1)
    f = H / (2 * math.tan(fov / 2))
    fwd = unit(target - eye)
    right = unit(np.cross(fwd, up))
    upv = np.cross(right, fwd)
    rng = np.random.default_rng(4)
    pxy = np.concatenate([[[W / 2, H / 2], [0.5, 0.5], [W - 0.5, H - 0.5]], rng.uniform(0, [W, H], (200, 2))])
    got = oracle.view_dir(fb, pxy.astype(np.float32))
    want = unit(fwd + ((pxy[:, 0] - W / 2) / f)[:, None] * right - ((pxy[:, 1] - H / 2) / f)[:, None] * upv)
    np.testing.assert_allclose(got, want, atol=2e-6)
    np.testing.assert_allclose(got[0], fwd, atol=1e-6)          # the centre looks at the target
    assert got[1] @ upv > 0 and got[2] @ upv < 0                  # row 0 is the top


def test_heat_ramp_named_colours():
    got = oracle.heat([0.0, 0.25, 0.5, 0.75, 1.0, -1.0, 2.0])
    want = [[0, 0, 0], [0, 0, 1], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 0], [1, 0, 0]]
    np.testing.assert_array_equal(got, np.array(want, np.float32))


def test_srgb_table_points():
    # the sRGB transfer's published anchors: 0, 1, the knee (0.0031308 -> 0.04045), and 8-bit codes 128 and 188
    # (linear 0.2158605 and 0.5028865)
    x = np.array([0.0, 1.0, 0.0031308, 0.2158605, 0.5028865, 0.5], np.float32)
    y = oracle.srgb_encode(x)
    np.testing.assert_allclose(y[:3], [0.0, 1.0, 0.04045], atol=2e-6)
    np.testing.assert_allclose(y[3:5] * 255, [128.0, 188.0], atol=2e-3)
    np.testing.assert_allclose(y[5], 0.735357, atol=2e-6)
    lo, hi = oracle.srgb_encode(np.array([0.0031308, np.nextafter(np.float32(0.0031308), np.float32(1))], np.float32))
    assert abs(lo - hi) < 1e-6                                    # the two branches meet at the knee


def grey(v):
    v = np.asarray(v, np.float32)
    return np.stack([v, v, v], 1)


def test_tone_clamp_and_aces():
    x = grey([-1.0, 0.0, 0.3, 1.0, 7.0])
    np.testing.assert_array_equal(oracle.tone_map(x, R.TONE_CLAMP)[:, 0], np.float32([0, 0, 0.3, 1.0, 1.0]))
    root = (0.56 + math.sqrt(0.56 ** 2 + 4 * 0.08 * 0.14)) / (2 * 0.08)     # where the fit reaches 1
    v = np.concatenate([np.linspace(0, root * 0.999, 400), [root * 1.001, 20.0, 1e4]]).astype(np.float32)
    y = oracle.tone_map(grey(v), R.TONE_ACES)[:, 0]
    assert y[0] == 0.0 and np.all(np.diff(y[:400]) > 0) and y[399] < 1.0 and np.all(y[400:] == 1.0)


def test_tone_reinhard_luminance():
    v = np.linspace(0, 50, 300).astype(np.float32)
    np.testing.assert_allclose(oracle.tone_map(grey(v), R.TONE_REINHARD)[:, 0], v / (1 + v), rtol=2e-6)
    c = np.array([[0.3, 0.6, 0.1], [2.0, 1.0, 0.5]], np.float32)           # hue: channel ratios kept below the clamp
    y = oracle.tone_map(c, R.TONE_REINHARD)
    np.testing.assert_allclose(y / y[:, :1], c / c[:, :1], rtol=2e-6)


def test_tone_pbr_neutral():
    v = np.linspace(0.08, 0.79, 200).astype(np.float32)                   # identity band: x - 0.04
    np.testing.assert_allclose(oracle.tone_map(grey(v), R.TONE_NEUTRAL)[:, 0], v - 0.04, atol=2e-7)
    t = np.linspace(0, 0.079, 100).astype(np.float32)                      # toe: 6.25 x^2
    np.testing.assert_allclose(oracle.tone_map(grey(t), R.TONE_NEUTRAL)[:, 0], 6.25 * t.astype(np.float64) ** 2,
                               atol=2e-7)
    s = np.float32(0.8)                                                    # continuous where compression starts
    y = oracle.tone_map(grey([np.nextafter(s, np.float32(0)), s, np.nextafter(s, np.float32(1))]), R.TONE_NEUTRAL)
    assert np.ptp(y) < 1e-6
    big = oracle.tone_map(np.array([[1e5, 5e4, 10.0]], np.float32), R.TONE_NEUTRAL)[0]
    np.testing.assert_allclose(big, [1.0, 1.0, 1.0], atol=1e-3)          # very bright: white
    v = np.linspace(0.8, 40, 400).astype(np.float32)
    y = oracle.tone_map(grey(v), R.TONE_NEUTRAL)[:, 0]
    assert np.all(np.diff(y) > 0) and y[-1] < 1.0


def test_tone_hable():
    y = oracle.tone_map(grey([0.0, 5.6, 6.0, 100.0]), R.TONE_FILMIC)[:, 0]
    assert abs(y[0]) < 1e-6 and abs(y[1] - 1.0) < 1e-6 and y[2] == 1.0 and y[3] == 1.0
    v = np.linspace(0.01, 5.5, 300).astype(np.float32)
    assert np.all(np.diff(oracle.tone_map(grey(v), R.TONE_FILMIC)[:, 0]) > 0)


def test_tone_agx():
    v = np.geomspace(1e-3, 1e3, 400).astype(np.float32)
    y = oracle.tone_map(grey(v), R.TONE_AGX)
    assert np.all(np.ptp(y, axis=1) < 2e-3 * np.maximum(y[:, 0], 1e-3))     # grey stays (nearly) grey
    enc = v < 15.0                                      # the log2 encode saturates at 2^4.026069 ~ 16.3 (after inset)
    assert np.all(np.diff(y[enc, 0]) > 0) and np.all(y[v > 17.0] == y[-1])
    black = oracle.tone_map(grey([0.0, 1e-5]), R.TONE_AGX)
    assert np.all(black < 1e-6)
    top = oracle.tone_map(grey([1e4]), R.TONE_AGX)[0]
    np.testing.assert_allclose(top, 0.99846 ** 2.2, atol=2e-3)            # the sigmoid's end value, to the 2.2 power
    hot = oracle.tone_map(np.array([[8.0, 2.0, 0.5]], np.float32), R.TONE_AGX)[0]
    assert hot[0] > hot[1] > hot[2] and hot[0] < 1.0                        # bright orange keeps its hue


# ---- the pass ------------------------------------------------------------------------------------------------------

def test_display_pass_composition():
    W, H = 96, 64
    env_p = E.parse_environment(ENVS["hill"])
    env = E.Environment(env_p)
    fb = R.render_frame(view=R.look_at((0, -6, 1.5), (0, 0, 1), (0, 0, 1)), width=W, height=H,
                        fovy=math.radians(60), count=1, exposure=1.5, tone=R.TONE_AGX)
    rng = np.random.default_rng(5)
    hdr = rng.uniform(0, 3, (H, W, 4)).astype(np.float32)
    hdr[..., 3] = rng.choice([0.0, 0.3, 1.0 - 1.0 / 2048, 1.0, 1.2], (H, W))
    out, by = oracle.display(fb, hdr, W, H, env.env, env.bg)
    np.testing.assert_array_equal(by[..., :3], np.floor(np.clip(out[..., :3], 0, 1) * 255 + 0.5).astype(np.uint8))
    assert np.all(by[..., 3] == 255)
    covered = hdr[..., 3] >= 1.0 - 1.0 / 1024
    want = oracle.srgb_encode(oracle.tone_map(np.maximum(hdr[covered][:, :3] * np.float32(1.5), 0), R.TONE_AGX))
    np.testing.assert_array_equal(out[covered][:, :3].ravel(), want)
    py, px = np.nonzero(~covered)
    w = oracle.view_dir(fb, np.stack([px + 0.5, py + 0.5], 1).astype(np.float32))
    sky = oracle.sky(env.env, env.bg, w, 1) + oracle.sky(env.env, env.bg, w, 2)
    bare = (1 - np.clip(hdr[py, px, 3], 0, 1))[:, None]
    lit = (hdr[py, px, :3] + bare * sky).astype(np.float32)
    want = oracle.srgb_encode(oracle.tone_map(np.maximum(lit * np.float32(1.5), 0), R.TONE_AGX)).reshape(-1, 3)
    np.testing.assert_allclose(out[py, px, :3], want, atol=2e-6)
    # without an environment nothing is added; view 2 is the heat of the count
    out0, _ = oracle.display(fb, hdr, W, H, None, None)
    want = oracle.srgb_encode(oracle.tone_map(np.maximum(hdr[..., :3].reshape(-1, 3) * np.float32(1.5), 0),
                                              R.TONE_AGX))
    np.testing.assert_array_equal(out0[..., :3].ravel(), want)
    fb2 = R.render_frame(view=R.look_at((0, -6, 1.5), (0, 0, 1), (0, 0, 1)), width=W, height=H,
                         fovy=math.radians(60), count=1, view_mode=2)
    out2, _ = oracle.display(fb2, hdr, W, H, env.env, env.bg)
    want = oracle.heat(np.log2(1 + hdr[..., 0].astype(np.float64)).ravel() / 7)
    np.testing.assert_allclose(out2[..., :3].reshape(-1, 3), want, atol=2e-6)


def test_sky_bg_packing_hill():
    p = E.parse_environment(ENVS["hill"])
    env = E.Environment(p)
    u = np.frombuffer(env.bg, np.uint32)
    f = u.view(np.float32)
    sky = ENVS["hill"]["sky"]
    assert len(env.bg) == 1520 and len(env.env) == 416
    assert u[4 * 81] == 81 and u[4 * 81 + 1] == 8
    np.testing.assert_allclose(f[0:3], sky["background_sh_rgb"][0], rtol=1e-7)
    np.testing.assert_allclose(f[4 * 80: 4 * 80 + 3], sky["background_sh_rgb"][80], rtol=1e-7)
    np.testing.assert_allclose(f[4 * 82], math.radians(ENVS["hill"]["lights"][0]["display_std_deg"]), rtol=1e-7)
    for k in range(8):
        np.testing.assert_allclose(f[4 * (83 + k): 4 * (83 + k) + 3], sky["horizon"]["rgb"][k], rtol=1e-7)
    np.testing.assert_allclose(f[4 * 91: 4 * 92], [*sky["ground_rgb"], math.sin(math.radians(1))], rtol=1e-7)
    np.testing.assert_allclose(f[4 * 92: 4 * 92 + 2], [1 / (2 * math.radians(25) ** 2), 1 / (2 * math.radians(8) ** 2)],
                               rtol=1e-6)
    np.testing.assert_allclose(f[4 * 93: 4 * 94], [*sky["glow"]["irradiance_rgb"], math.radians(12)], rtol=1e-7)
    np.testing.assert_allclose(f[4 * 94: 4 * 94 + 3], unit(ENVS["hill"]["lights"][0]["dir"]), atol=1e-6)
    ue = np.frombuffer(env.env, np.uint32)
    assert ue[96] == 1 and ue[97] == 1
    # off: no sky words at all, the disc std stays (the reference writes it whatever the switch)
    assert env.set(on=False) and not env.set(on=False)
    u = np.frombuffer(env.bg, np.uint32)
    assert u[4 * 81] == 0 and u[4 * 81 + 1] == 0 and np.frombuffer(env.env, np.uint32)[96] == 0
    assert u.view(np.float32)[4 * 82] > 0


def test_sun_angles_round_trip():
    for d in ([-0.433013, 0.75, 0.5], [0.1, -0.2, 0.97], [1.0, 0.0, 0.0]):
        np.testing.assert_allclose(E.sun_dir(*E.sun_angles(unit(d))), unit(d), atol=1e-12)
    np.testing.assert_allclose(E.sun_dir(0, 90), [0, 0, 1], atol=1e-12)
    np.testing.assert_allclose(E.sun_dir(90, 0), [-1, 0, 0], atol=1e-12)   # azimuth from +y toward -x
