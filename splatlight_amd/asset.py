"""The asset loader of the "gs-runtime" format (src/asset.js:1-88) and its display-fill side table.

An asset is a directory with asset.json and raw little-endian .bin buffers that go to the device unchanged.
check_header / check_buffer refuse what the renderer's record layouts cannot read, with the reference's messages
(src/asset.js:18-38); load_asset reads the seven buffers, the optional blockers, the lamps as plain records and the
fingerprint baked sidecars are keyed by (src/asset.js:51-88). load_fill reads the display fill table of a
true-density asset (onepath.json / onepath.bin `fill`, src/lighting.js:961-985).
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import records as R

MAJOR = "0"
# The fields the record layouts assume (src/asset.js:8-16).
EXPECT = {
    "gaussians": {"stride": 64, "fields": {"mu": 0, "kappa": 12, "quat": 16, "scale": 32, "part_surface": 44,
                                           "albedo": 48}},
    "parts": {"stride": 128, "fields": {"xf": 0, "xf_prev": 48, "bound": 96, "root": 112, "g_lo": 116, "g_hi": 120,
                                        "flags": 124}},
    "lamps": {"stride": 96, "fields": {"center": 0, "kind": 12, "e1": 16, "radius": 28, "e2": 32, "normal": 48,
                                       "radiance": 64, "center_prev": 80}},
    "materials": {"stride": 32, "fields": {"flake": 0, "phase_samples": 12, "vfloor": 16, "flags": 20}},
    "nodes": {"stride": 96, "fields": {}},
    "kids": {"stride": 4, "fields": {}},
    "vgrid": {"stride": 128, "fields": {}},
}
BUFFERS = ["gaussians", "parts", "lamps", "materials", "nodes", "kids", "vgrid"]


class AssetError(ValueError):
    pass


def check_header(header: dict) -> None:
    """src/asset.js:18-31."""
    if header.get("format") != "gs-runtime":
        raise AssetError(f"Unknown asset format {header.get('format')}")
    if str(header.get("version")).split(".")[0] != MAJOR:
        raise AssetError(f"Unsupported asset version {header.get('version')}")
    for name, want in EXPECT.items():
        b = (header.get("buffers") or {}).get(name)
        if not b:
            raise AssetError(f"Asset has no {name} buffer")
        if b.get("stride") != want["stride"]:
            raise AssetError(f"{name} stride {b.get('stride')}, renderer expects {want['stride']}")
        for f, off in want["fields"].items():
            got = ((b.get("fields") or {}).get(f) or {}).get("offset")
            if got != off:
                raise AssetError(f"{name}.{f} at {got}, renderer expects {off}")
        c = b.get("count")
        if not isinstance(c, int) or isinstance(c, bool) or c <= 0 or c > (1 << 24):
            raise AssetError(f"Bad {name} count {c}")


def check_buffer(header: dict, name: str, data: bytes) -> bytes:
    """src/asset.js:33-39."""
    b = header["buffers"][name]
    if len(data) != b["count"] * b["stride"]:
        raise AssetError(f"{b['file']} is {len(data)} B, expected {b['count']} x {b['stride']} B")
    return data


def read_lamps(header: dict, data: bytes) -> list:
    """src/asset.js:41-49: the lamp records as plain dicts."""
    a = np.frombuffer(data, np.uint8).reshape(-1, 96)
    out = []
    for i in range(header["buffers"]["lamps"]["count"]):
        f = a[i].view(np.float32)
        u = a[i].view(np.uint32)
        out.append({"center": f[0:3].tolist(), "kind": int(u[3]), "e1": f[4:7].tolist(), "radius": float(f[7]),
                    "e2": f[8:11].tolist(), "normal": f[12:15].tolist(), "radiance": f[16:19].tolist()})
    return out


def fingerprint(header: dict, gaussians: bytes) -> dict:
    """src/asset.js:80-88: count, part ranges, gaussians.bin bytes and sum_i w_i (i * 2654435761 + 1) mod 2^32 over
    its u32 words."""
    w = np.frombuffer(gaussians, "<u4").astype(np.uint64)
    i = np.arange(w.size, dtype=np.uint64)
    mult = (i * np.uint64(2654435761) + np.uint64(1)) & np.uint64(0xFFFFFFFF)
    h = int(np.sum((w * mult) & np.uint64(0xFFFFFFFF), dtype=np.uint64) & np.uint64(0xFFFFFFFF))
    return {"count": header["buffers"]["gaussians"]["count"], "parts": [p["gaussians"] for p in header["parts"]],
            "bytes": len(gaussians), "hash": h}


def load_asset(base: str) -> dict:
    """src/asset.js:51-78: header, raw buffers (bytes), the records as structured arrays, lamps, counts and the
    fingerprint. Raises AssetError on a layout the renderer cannot read."""
    with open(os.path.join(base, "asset.json")) as f:
        header = json.load(f)
    check_header(header)
    raw = {}
    for name in BUFFERS:
        with open(os.path.join(base, header["buffers"][name]["file"]), "rb") as f:
            raw[name] = check_buffer(header, name, f.read())
    raw["blockers"] = None                   # optional own-part blockers, 64 B per Gaussian
    bk = header.get("blockers")
    if bk and bk.get("file") and bk.get("stride") == 64 and bk.get("count") == header["buffers"]["gaussians"]["count"]:
        p = os.path.join(base, bk["file"])
        if os.path.exists(p) and os.path.getsize(p) == bk["count"] * 64:
            with open(p, "rb") as f:
                raw["blockers"] = f.read()
    return {
        "header": header, "raw": raw,
        "gaussians": np.frombuffer(raw["gaussians"], R.GAUSSIAN_DTYPE),
        "parts": np.frombuffer(raw["parts"], R.PART_DTYPE),
        "materials": np.frombuffer(raw["materials"], R.MATERIAL_DTYPE),
        "nodes": np.frombuffer(raw["nodes"], R.NODE_DTYPE),
        "kids": np.frombuffer(raw["kids"], "<u4"),
        "vgrid": np.frombuffer(raw["vgrid"], np.uint8).reshape(-1, 128),
        "lamps": read_lamps(header, raw["lamps"]),
        "count": header["buffers"]["gaussians"]["count"], "partCount": header["buffers"]["parts"]["count"],
        "fingerprint": fingerprint(header, raw["gaussians"]),
    }


def emit_part(header: dict) -> int:
    """The emitters' part (src/lod.js:108-109): header.emitters.part, else the part named "lamp", else none."""
    if header.get("emitters"):
        return int(header["emitters"]["part"])
    for k, p in enumerate(header.get("parts", [])):
        if p.get("name") == "lamp":
            return k
    return 0xFFFFFFFF


def load_fill(base: str, count: int):
    """The display fill table (src/lighting.js:961-970): u32 offsets [G + 1], then (source index, f32 weight bits)
    pairs; None when the asset has none or it is for another Gaussian count."""
    p = os.path.join(base, "onepath.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        meta = json.load(f)
    if not meta.get("fill") or meta.get("G") != count:
        return None
    with open(os.path.join(base, "onepath.bin"), "rb") as f:
        f.seek(meta["fill"]["offset"])
        data = f.read(meta["fill"]["bytes"])
    return np.frombuffer(data, "<u4").copy()


def check_fill(fill: np.ndarray, count: int) -> None:
    """What the fill kernel relies on: G + 1 ascending offsets, the pairs they span, every source index < G, and
    targets and sources disjoint (fill.wgsl:1-6)."""
    off = fill[:count + 1].astype(np.int64)
    if off[0] != 0 or np.any(np.diff(off) < 0):
        raise AssetError("fill offsets are not ascending from 0")
    n = int(off[-1])
    if fill.size != count + 1 + 2 * n:
        raise AssetError(f"fill table has {fill.size} words, offsets span {n} pairs")
    src = fill[count + 1::2][:n]
    if n and int(src.max()) >= count:
        raise AssetError("fill source index out of range")
    targets = np.nonzero(np.diff(off) > 0)[0]
    if np.intersect1d(targets, src).size:
        raise AssetError("fill targets and sources overlap")
