"""The environment's host side: src/environment.js:14-153 for the two uniforms the frame path reads.

parse_environment follows parseEnvironment (src/environment.js:23-38), sun_dir / sun_angles follow sunDir /
sunAngles (:54-61), and Environment.write packs Env (416 B, the lighting and display uniform) and SkyBg
(1,520 B: 81 + 2 + 8 + 2 + 2 vec4, the display pass's background sky) exactly as Environment.write does
(:111-153), so the display pass (sl_display_encode) reads the same bytes the reference's display.wgsl binds.
"""
from __future__ import annotations

import math

import numpy as np

from .lighting import ENV_BYTES

SKY_BG_BYTES = (81 + 2 + 8 + 2 + 2) * 16
DEG = math.pi / 180


def _normalize(v):
    l = math.hypot(*v) or 1.0
    return [x / l for x in v]


def parse_environment(env: dict | None) -> dict | None:
    """src/environment.js:23-38: an asset.json `environment` block (or an env/<name>.json) -> the parsed form."""
    if not env:
        return None
    sky = env.get("sky") or {}
    sh = sky.get("sh_rgb") or []
    bg = sky.get("background_sh_rgb")
    bg = bg[:81] if bg is not None else None
    hz = sky.get("horizon")
    horizon = ({"rgb": hz["rgb"][:8], "along": hz.get("along_deg", 25) * DEG, "across": hz.get("across_deg", 8) * DEG}
               if hz else None)
    ground = sky.get("ground_rgb")
    g = sky.get("glow")
    glow = ({"irradiance": g["irradiance_rgb"], "std": g.get("angular_std_deg", 12) * DEG,
             "followSun": g.get("follows_sun", True), "dir": _normalize(g["dir"]) if g.get("dir") else None}
            if g else None)
    lights = [{"dir": _normalize(L["dir"]), "irradiance": L.get("irradiance_rgb", [1, 1, 1]),
               "std": L.get("angular_std_deg", 0.27) * DEG, "displayStd": L.get("display_std_deg", 0) * DEG,
               "shadows": L.get("casts_shadows", True), "followSun": L.get("follows_sun", False)}
              for L in (env.get("lights") or [])[:4]]
    return {"sh": sh, "bg": bg, "horizon": horizon, "ground": ground, "glow": glow, "lights": lights}


def sun_dir(az: float, el: float):
    """src/environment.js:54-57: the direction toward the light, azimuth in degrees from +y toward -x."""
    a, e = az * DEG, el * DEG
    return [-math.sin(a) * math.cos(e), math.cos(a) * math.cos(e), math.sin(e)]


def sun_angles(d):
    """src/environment.js:58-61."""
    return [math.atan2(-d[0], d[1]) / DEG, math.asin(max(-1.0, min(1.0, d[2]))) / DEG]


class Environment:
    """The uniform-writing part of src/environment.js:63-160 (the sun maps live in lighting.sun_plan)."""

    def __init__(self, env: dict):
        self.base = env
        self.skyGain = 1.0
        self.lightGain = 1.0
        self.sun = sun_angles(env["lights"][0]["dir"]) if env["lights"] else [0.0, 45.0]
        self.on = True
        self.env = self.bg = None
        self.key = None
        self.write()

    def set(self, on=None, sunAz=None, sunEl=None, skyGain=None, lightGain=None):
        """src/environment.js:104-109: True when anything changed (the uniforms were rewritten)."""
        self.on = self.on if on is None else on
        self.sun = [self.sun[0] if sunAz is None else sunAz, self.sun[1] if sunEl is None else sunEl]
        self.skyGain = self.skyGain if skyGain is None else skyGain
        self.lightGain = self.lightGain if lightGain is None else lightGain
        return self.write()

    def write(self):
        """src/environment.js:111-153: self.env = Env bytes, self.bg = SkyBg bytes; False when nothing changed."""
        key = (self.on, *self.sun, self.skyGain, self.lightGain)
        if key == self.key:
            return False
        self.key = key
        u = np.zeros(ENV_BYTES // 4, np.uint32)
        f = u.view(np.float32)
        b = self.base
        if self.on:
            for k, c in enumerate(b["sh"]):
                f[4 * k: 4 * k + 3] = [v * self.skyGain for v in c]
            for k, L in enumerate(b["lights"]):
                d = sun_dir(*self.sun) if k == 0 or L["followSun"] else L["dir"]
                f[64 + 8 * k: 68 + 8 * k] = [*d, L["std"]]
                f[68 + 8 * k: 72 + 8 * k] = [*[v * self.lightGain for v in L["irradiance"]], 1.0 if L["shadows"] else 0.0]
            u[96] = len(b["lights"])
            u[97] = 1 if len(b["sh"]) else 0
        bu = np.zeros(SKY_BG_BYTES // 4, np.uint32)
        bf = bu.view(np.float32)
        bg = b["bg"]
        if self.on and bg:
            for k, c in enumerate(bg):
                bf[4 * k: 4 * k + 3] = [v * self.skyGain for v in c]
            bu[4 * 81] = len(bg)
        for k, L in enumerate(b["lights"]):
            bf[4 * 82 + k] = L["displayStd"]
        hz, ground = b["horizon"], b["ground"]
        if self.on and bg and hz and ground:
            for k, c in enumerate(hz["rgb"]):
                bf[4 * (83 + k): 4 * (83 + k) + 3] = [v * self.skyGain for v in c]
            bu[4 * 81 + 1] = len(hz["rgb"])
            bf[4 * 91: 4 * 92] = [*[v * self.skyGain for v in ground], math.sin(1 * DEG)]
            bf[4 * 92: 4 * 92 + 2] = [1 / (2 * hz["along"] ** 2), 1 / (2 * hz["across"] ** 2)]
        glow = b["glow"]
        if self.on and bg and glow:
            bf[4 * 93: 4 * 94] = [*[v * self.skyGain for v in glow["irradiance"]], glow["std"]]
            bf[4 * 94: 4 * 94 + 3] = sun_dir(*self.sun) if glow["followSun"] or not glow["dir"] else glow["dir"]
        self.env, self.bg = u.tobytes(), bu.tobytes()
        return True

    def lights(self):
        """src/environment.js:156-160."""
        if not self.on:
            return []
        return [{"dir": sun_dir(*self.sun) if k == 0 or L["followSun"] else L["dir"], "std": L["std"],
                 "irradiance": [v * self.lightGain for v in L["irradiance"]]} for k, L in enumerate(self.base["lights"])]
