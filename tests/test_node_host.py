"""The JavaScript host's parameter logic (splatlight_amd/node/host.js and asset.js), on the CPU: on asset directories
cut from the shipped hill and Cornell assets (tests/jsasset.py), the bytes the shim would hand to the C ABI for a
first frame equal the Python host's (splatlight_amd/lighting.py, environment.py, bounce.py, lod.py, records.py),
which the GPU tests hold against the oracle.

Integer words (counts, indices, flags, offsets) must be equal. Float words are computed in float64 in both hosts and
stored as float32; the two languages' libm (exp, pow, hypot, atan2, cos, sin) may round differently in the last
double bit, so a float32 word may differ by at most 2 ulp (measured: see the printed counts)."""
import json
import math
import os
import shutil
import subprocess
import sys
import types

import numpy as np
import pytest

import jsasset
from splatlight_amd import bounce as BN
from splatlight_amd import environment as E
from splatlight_amd import lighting as L
from splatlight_amd import lod as LOD
from splatlight_amd import records as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None, reason="no node")


def ulps(a, b):
    """Per-word float32 distance in ulp (0 where both are equal, including matching NaN/inf)."""
    ia = np.frombuffer(np.ascontiguousarray(a, np.float32).tobytes(), np.int32).astype(np.int64)
    ib = np.frombuffer(np.ascontiguousarray(b, np.float32).tobytes(), np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def same_words(js, py, float_words, what, max_ulp=2):
    """js, py: byte strings of u32 words; float_words: a boolean mask (or index list) of the float32 words."""
    a = np.frombuffer(bytes(js), np.uint32)
    b = np.frombuffer(bytes(py), np.uint32)
    assert a.shape == b.shape, f"{what}: {a.size} words against {b.size}"
    fm = np.zeros(a.size, bool)
    fm[float_words] = True
    np.testing.assert_array_equal(a[~fm], b[~fm], err_msg=f"{what}: integer words")
    d = ulps(a[fm].view(np.float32), b[fm].view(np.float32))
    assert d.max(initial=0) <= max_ulp, f"{what}: float words up to {d.max()} ulp apart (word {np.nonzero(fm)[0][d.argmax()]})"
    return int((d > 0).sum())


def node_cov(rec):
    """R diag(s^2) R^T in float64 of Gaussian records (float32 [n, 16]: quat 4-7, scale 8-10)."""
    x, y, z, w = [rec[:, 4 + k].astype(np.float64) for k in range(4)]
    Rm = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                   np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                   np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    s2 = rec[:, 8:11].astype(np.float64) ** 2
    return np.einsum("nij,nj,nkj->nik", Rm, s2, Rm)


def run_host(tmp_path, name, eye, fwd):
    adir, out = tmp_path / name, tmp_path / (name + "_out")
    out.mkdir()
    header, bufs = jsasset.write(name, str(adir))
    spec = tmp_path / (name + "_spec.json")
    spec.write_text(json.dumps({"eye": list(eye), "fwd": list(fwd)}))
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node_host.js"), str(adir), str(out), str(spec)], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rd = lambda n: (out / n).read_bytes() if (out / n).exists() else None
    return header, bufs, rd, json.loads((out / "meta.json").read_text())


def camera(name):
    cam = jsasset.lodcut.raw(name)["camera"]
    eye = np.asarray(cam[:3], np.float64)
    d = np.asarray(cam[3:6], np.float64) - eye
    return tuple(float(v) for v in eye), tuple(float(v) for v in d / np.linalg.norm(d))


@pytest.mark.parametrize("name", ["cornell", "hill"])
def test_js_host_writers_equal_the_python_host(tmp_path, name):
    eye, fwd = camera(name)
    h, b, rd, meta = run_host(tmp_path, name, eye, fwd)
    g = b["gaussians"].view(R.GAUSSIAN_DTYPE).reshape(-1)
    parts = b["parts"].view(R.PART_DTYPE).reshape(-1)
    nodes = b["nodes"].view(R.NODE_DTYPE).reshape(-1)
    P, G, N = parts.shape[0], g.shape[0], nodes.shape[0]
    diffs = {}
    # emitter set, tree, EmitParams, LSParams
    if h.get("emitters"):
        em = h["emitters"]
        g0, g1 = h["parts"][em["part"]]["gaussians"]
        one_sided = bool(b["materials"][em["material"], 5] & 4)
        es = L.asset_emitters(g, g0, g1, em["emission"])
        np.testing.assert_array_equal(np.frombuffer(rd("emitter_records.bin"), np.float32).reshape(-1, 16),
                                      es["records"].view(np.float32).reshape(-1, 16))
        tn, root, cov = L.build_emitter_tree(es["records"])
        ep = R.emit_params(count=g1 - g0, root=root, scale=es["scale"], theta=4.0, mode=1, node_count=tn.shape[0],
                           skip_part=em["part"], one_sided=int(one_sided), normal=es["normal"], n5=8, edge_soft=0.2)
        ls = L.ls_params(tn, root, cov, es["normal"], one_sided, P, (1 << P) - 1, skip_part=em["part"])
        lw = np.ones(84, bool)
        lw[[3, 7, 11, 15, 17, 18, 19, 21, 22, 27, 31, 32, 35] + list(range(36, 52))] = False
        diffs["LSParams"] = same_words(rd("LS.bin"), ls, lw, "LSParams")
    else:
        dark = np.zeros((1, 16), np.float32)
        dark[0, 3], dark[0, 7], dark[0, 8:11] = 1, 1, 0.01
        tn, root, cov = L.build_emitter_tree(dark)
        ep = R.emit_params(count=0, root=root, scale=0.0, theta=4.0, mode=1, node_count=1, n5=8, edge_soft=0.2)
        assert rd("LS.bin") is None                       # no lamp: the emitter map never runs
    nw = np.ones(24, bool)
    nw[[11, 15]] = False
    diffs["emitter nodes"] = same_words(rd("emit_nodes.bin"), L.pack_emit_nodes(tn, cov).tobytes(),
                                        np.tile(nw, tn.shape[0]), "emitter nodes")
    ew = np.zeros(16, bool)
    ew[[2, 3, 8, 9, 10, 12]] = True
    diffs["EmitParams"] = same_words(rd("EP.bin"), ep, ew, "EmitParams")
    # FrameUniforms
    boxes = sum(p["name"] in ("short_box", "tall_box", "floor") for p in h["parts"])
    U = R.frame_uniforms(frame=0, gaussian_count=G, part_count=P, lamp_count=1, theta=8.0, alpha_min=1 / 16,
                         alpha_part=0.5, eye=eye, material_count=b["materials"].shape[0] + boxes, iso=7,
                         stratum_step=7, light_samples_axis=(h.get("shading") or {}).get("light_samples_axis", 4))
    same_words(rd("U.bin"), U, [4, 5, 6, 8, 9, 10], "FrameUniforms", max_ulp=0)
    # environment: Env, SkyBg, the sun plan
    if h.get("environment"):
        env = E.Environment(E.parse_environment(h["environment"]))
        diffs["Env"] = same_words(rd("ENV.bin"), env.env, np.arange(104) < 96, "Env")
        bw = np.ones(380, bool)
        bw[[4 * 81, 4 * 81 + 1]] = False
        diffs["SkyBg"] = same_words(rd("BG.bin"), env.bg, bw, "SkyBg")
        lights = [(Lt["dir"], Lt["std"], True) for Lt in env.lights()]
        SM, desc, _, _ = L.sun_plan(parts, lights, eye, fwd, res=512, cap_bytes=120e6)
        sw = np.zeros(108, bool)
        sw[8:11] = True
        sw[12:28] = True
        diffs["SunParams"] = same_words(rd("SM.bin"), SM, sw, "SunParams")
        dw = np.zeros(8, bool)
        dw[4:7] = True
        diffs["sun layers"] = same_words(rd("desc.bin"), desc, np.tile(dw, len(desc) // 32), "sun layers")
    # probes: layout, ProbeParams of the first frame and of a later round
    sc = types.SimpleNamespace(gaussians=g, parts=parts, name=h.get("name", name))
    lay = BN.probe_layout(sc, h)
    base = BN.probe_params(lay, N, P)
    assert meta["probes"] == BN.probe_count(lay, P) and meta["terrain"] == lay["terrain"]
    pw = np.zeros(base.size, bool)
    pw[[0, 1, 2, 4, 5, 6, 8, 9, 10, 11, 12, 14, 15]] = True
    pw[20:76] = True
    pw[76:202] = True
    pw[212:228] = True
    pw[228:1252] = True
    for k in (3, 7, 13, 16, 17, 18, 19):
        pw[k] = False
    diffs["ProbeParams"] = same_words(rd("PP0.bin"), BN.frame_params(base, 0, True, True), pw, "ProbeParams round 0")
    same_words(rd("PP9.bin"), BN.frame_params(base, 9, False, False), pw, "ProbeParams round 9")
    # LOD node records: every word but the rotation to 4 ulp; the rotation through the covariance it gives (an
    # eigenvector's sign is free where two Jacobi runs round differently, so R may differ by a half turn)
    ng = LOD.node_gaussians(nodes, b["kids"], g)["records"].view(np.uint32).reshape(-1, 16)
    jg = np.frombuffer(rd("node_gaussians.bin"), np.uint32).reshape(-1, 16)
    gw = np.ones(16, bool)
    gw[[11, 15]] = False
    gw[4:8] = False
    keep = np.r_[0:4, 8:16]
    diffs["LOD node records"] = same_words(jg[:, keep].tobytes(), ng[:, keep].tobytes(),
                                           np.tile(gw[keep], N - G), "LOD node records", max_ulp=4)
    cj, cp = node_cov(jg.view(np.float32)), node_cov(ng.view(np.float32))
    scale = np.abs(cp).reshape(-1, 9).max(1)[:, None, None]
    err = float(np.max(np.abs(cj - cp) / scale))
    assert err < 1e-5, f"LOD node covariances differ by {err:.2e} of their largest entry"
    diffs["LOD rotations"] = int(np.any(jg[:, 4:8] != ng[:, 4:8], axis=1).sum())
    print(f"{name}: float words that differ by 1-2 ulp: {diffs}", file=sys.stderr, flush=True)


def test_js_loader_refuses_what_the_reference_refuses(tmp_path):
    h, _ = jsasset.write("cornell", str(tmp_path / "a"))
    bad = dict(h, version="1.0")
    (tmp_path / "a" / "asset.json").write_text(json.dumps(bad))
    r = subprocess.run([NODE, "-e", "require('./splatlight_amd/node/asset.js').loadAsset(process.argv[1])"
                        ".then(()=>process.exit(0), (e)=>{console.log(e.message); process.exit(3);})",
                        str(tmp_path / "a")], cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "Unsupported asset version 1.0" in r.stdout
    h2 = json.loads(json.dumps(h))
    h2["buffers"]["gaussians"]["count"] += 1
    (tmp_path / "a" / "asset.json").write_text(json.dumps(h2))
    r = subprocess.run([NODE, "-e", "require('./splatlight_amd/node/asset.js').loadAsset(process.argv[1])"
                        ".then(()=>process.exit(0), (e)=>{console.log(e.message); process.exit(3);})",
                        str(tmp_path / "a")], cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "gaussians.bin is" in r.stdout and "expected" in r.stdout
