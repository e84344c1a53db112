"""A lit frame through the N-API addon and its JS shim (SplatRenderer / Lighting as the reference's host calls them)
on the GPU: the bytes of its rgba16f target equal those of the same frame rendered through ctypes, and its
counts are the same."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from splatlight_amd import build as B
from splatlight_amd import lighting as L
from splatlight_amd import native as N
from splatlight_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(NODE is None or not os.path.exists(B.NODE_ADDON), reason="no node / addon")]


def test_js_host_frame_equals_python_frame(ctx, tmp_path):
    sc = scenes.cornell(n=8_000, seed=21, width=200, height=150)
    nd = scenes.leaf_nodes(sc.gaussians)
    inp = L.lighting_inputs(sc)
    fb = sc.frame_bytes()
    files = {"gaussians.bin": sc.gaussians, "parts.bin": sc.parts, "materials.bin": sc.materials, "nodes.bin": nd,
             "emitters.bin": sc.emitters, "frame.bin": fb, "U.bin": inp["U"], "EP.bin": inp["EP"],
             "emit_nodes.bin": np.ascontiguousarray(inp["emit_nodes"], np.float32), "LS.bin": inp["LS"]}
    for k in ("ENV", "SM", "desc"):
        if inp[k] is not None:
            files[f"{k}.bin"] = inp[k]
    for name, v in files.items():
        (tmp_path / name).write_bytes(v if isinstance(v, (bytes, bytearray)) else np.ascontiguousarray(v).tobytes())
    (tmp_path / "meta.json").write_text(json.dumps({"width": sc.width, "height": sc.height}))
    print(f"node frame: {sc.count} Gaussians {sc.width}x{sc.height}", file=sys.stderr, flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node_frame.js"), str(tmp_path)], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    js_img = (tmp_path / "hdr.bin").read_bytes()
    js = json.loads((tmp_path / "stats.json").read_text())
    assert set(js["timers"]) >= {"project", "radix", "pairs", "tilesort", "ranges", "splat"}
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials, nodes=nd)
    out = ctx.alloc(sc.width * sc.height * 8)
    try:
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        N.frame_encode(ctx, scene, fb, sc.width, sc.height, out, N.SL_OUT_RGBA16F)
        st = N.frame_finish(ctx)
        py_img = ctx.read(out, sc.width * sc.height * 8, np.uint8).tobytes()
    finally:
        ctx.free(out)
        scene.destroy()
    assert (js["stats"]["visible"], js["stats"]["pairs"]) == (st.visible, st.pairs)
    assert js_img == py_img
    print(f"  JS host: visible={st.visible} pairs={st.pairs}, target bytes identical to the ctypes path",
          file=sys.stderr, flush=True)


def test_js_host_bounce_frame_equals_python_frame(ctx, tmp_path):
    """The same through initBounce / encodeBounce: the sky and bounce through probes on the hill, frame at gain 1."""
    from splatlight_amd import bounce as BN
    from splatlight_amd import tree
    sc = scenes.hill(n=6_000, seed=23, width=160, height=120, shadow_res=256)
    t = tree.scene_tree(sc)
    lay = BN.probe_layout(sc)
    base = BN.probe_params(lay, t.count, sc.parts.shape[0])
    pp0 = BN.frame_params(base, 0, True, True)
    inp = L.lighting_inputs(sc)
    fb = sc.frame_bytes(bounce_gain=1.0)
    files = {"gaussians.bin": sc.gaussians, "parts.bin": sc.parts, "materials.bin": sc.materials,
             "nodes.bin": t.nodes[:sc.count], "tree_nodes.bin": t.nodes, "kids.bin": t.kids, "PP0.bin": pp0,
             "PP.bin": pp0, "emitters.bin": sc.emitters, "frame.bin": fb, "U.bin": inp["U"], "EP.bin": inp["EP"],
             "emit_nodes.bin": np.ascontiguousarray(inp["emit_nodes"], np.float32), "LS.bin": inp["LS"]}
    for k in ("ENV", "SM", "desc"):
        if inp[k] is not None:
            files[f"{k}.bin"] = inp[k]
    for name, v in files.items():
        (tmp_path / name).write_bytes(v if isinstance(v, (bytes, bytearray)) else np.ascontiguousarray(v).tobytes())
    (tmp_path / "meta.json").write_text(json.dumps({"width": sc.width, "height": sc.height, "maxDepth": t.max_depth}))
    print(f"node bounce frame: {sc.count} Gaussians {sc.width}x{sc.height}", file=sys.stderr, flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node_frame.js"), str(tmp_path)], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    js_img = (tmp_path / "hdr.bin").read_bytes()
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials, nodes=t.nodes[:sc.count])
    b = BN.Bounce(scene, sc, t)
    out = ctx.alloc(sc.width * sc.height * 8)
    try:
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        b.encode()
        N.frame_encode(ctx, scene, fb, sc.width, sc.height, out, N.SL_OUT_RGBA16F)
        N.frame_finish(ctx)
        py_img = ctx.read(out, sc.width * sc.height * 8, np.uint8).tobytes()
    finally:
        ctx.free(out)
        scene.destroy()
    assert js_img == py_img
    print("  JS host with bounce: target bytes identical to the ctypes path", file=sys.stderr, flush=True)


def test_js_host_encodes_next_frame_before_awaiting(ctx, tmp_path):
    """Lighting and frame encoded for four cameras back to back before any frame's counts are awaited (as the
    reference's capture loop does): each frame's promise resolves to that frame's own counts, the same as when every
    frame is awaited before the next, and the last target's bytes are the same."""
    sc = scenes.cornell(n=6_000, seed=22, width=160, height=120)
    nd = scenes.leaf_nodes(sc.gaussians)
    inp = L.lighting_inputs(sc)
    files = {"gaussians.bin": sc.gaussians, "parts.bin": sc.parts, "materials.bin": sc.materials, "nodes.bin": nd,
             "emitters.bin": sc.emitters, "U.bin": inp["U"], "EP.bin": inp["EP"],
             "emit_nodes.bin": np.ascontiguousarray(inp["emit_nodes"], np.float32), "LS.bin": inp["LS"]}
    for k, r in enumerate((2.5, 3.5, 5.0, 3.0)):
        files[f"frame{k}.bin"] = sc.frame_bytes(eye=scenes.orbit_eye(sc, r, -1.2 + 0.4 * k, 0.3))
    for name, v in files.items():
        (tmp_path / name).write_bytes(v if isinstance(v, (bytes, bytearray)) else np.ascontiguousarray(v).tobytes())
    (tmp_path / "meta.json").write_text(json.dumps({"width": sc.width, "height": sc.height, "frames": 4}))
    print("node pipelined frames: 4 cameras", file=sys.stderr, flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node_pipeline.js"), str(tmp_path)], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads((tmp_path / "piped.json").read_text())
    piped = [(f["visible"], f["pairs"]) for f in out["piped"]]
    serial = [(f["visible"], f["pairs"]) for f in out["serial"]]
    print(f"  pipelined {piped}, awaited {serial}", file=sys.stderr, flush=True)
    assert piped == serial
    assert len(set(serial)) > 1                 # the cameras differ, so a stale count would show
    assert out["hdrEqual"]
