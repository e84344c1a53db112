"""The JavaScript host on the GPU: the N-API addon and the shim (splatlight_amd/node/renderer.js) driven as
src/main.js drives the reference's classes, on asset directories cut from the shipped Cornell and hill assets
(tests/jsasset.py; materials matte, as the probe bounce toward specular receivers is not on this path): loadAsset,
nodeGaussians, SplatRenderer, Lighting (its box materials, iso_init and probe bounce), loadFill, Lod, bindColour,
Environment, bindEnv, setEmitters, then per frame lighting.encode, encodeBounce, encodeFill, lod.encode and
renderer.encode into a display target.

Bars: the HDR target and the display bytes of every frame equal, bit for bit, those of the same calls made through
ctypes with the bytes the shim handed over (the uniform writers themselves are held to the Python host's in
tests/test_node_host.py, and the Python path to the oracle in the other GPU tests); the counts are the same. A
second test encodes frames back to back before awaiting any (the reference's capture loop) and gets each frame's
own counts."""
import json
import math
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import jsasset
import lodcut
from splatlight_amd import build as B
from splatlight_amd import native as N
from splatlight_amd import records as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(NODE is None or not os.path.exists(B.NODE_ADDON), reason="no node / addon")]
W, H = 320, 200


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def look(eye, target):
    d = np.asarray(target, np.float64) - np.asarray(eye, np.float64)
    return [float(v) for v in d / np.linalg.norm(d)]


def frames_for(name, tmp_path, sc, n_nodes, lod_px):
    """Three frames: the asset's camera (everything dirty), the same camera again (nothing moved), and a moved
    camera with the lighting still; RenderFrames written beside the asset directory."""
    eyes = [sc.eye, sc.eye, tuple(float(e + 0.15 * (t - e)) for e, t in zip(sc.eye, sc.target))]
    out = []
    for k, (eye, changed, still) in enumerate(zip(eyes, (True, False, False), (False, False, True))):
        fb = R.render_frame(view=R.look_at(eye, sc.target, (0.0, 0.0, 1.0)), width=W, height=H,
                            fovy=math.radians(sc.fov_deg), far=sc.far, count=sc.count, bounce_gain=1.0,
                            lod=(lod_px, sc.count + 1, n_nodes) if lod_px > 0 else None)
        (tmp_path / f"frame{k}.bin").write_bytes(fb)
        v = np.frombuffer(fb, np.float32)[:16]
        out.append({"eye": list(eye), "fwd": look(eye, sc.target), "view": [float(x) for x in v],
                    "frame": f"frame{k}.bin", "changed": changed, "lightStill": still, "bytes": fb})
    return out


@pytest.mark.parametrize("name", ["cornell", "hill"])
def test_js_host_drives_the_reference_surface(ctx, tmp_path, name):
    adir, out = tmp_path / "asset", tmp_path / "out"
    out.mkdir()
    header, bufs = jsasset.write(name, str(adir), matte=True)   # the hill's shiny grass and leaf: matte
    sc, nodes, kids, _, emit = lodcut.load(name, W, H)
    G, Nn = sc.count, nodes.shape[0]
    frames = frames_for(name, tmp_path, sc, Nn - G, 2.0)
    spec = {"width": W, "height": H, "lodPx": 2.0, "fovy": math.radians(sc.fov_deg),
            "frames": [{k: v for k, v in f.items() if k != "bytes"} for f in frames]}
    (tmp_path / "spec.json").write_text(json.dumps(spec))
    log(f"node drop-in {name}: {G} Gaussians, {Nn - G} LOD nodes, {W}x{H}, 3 frames")
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node_dropin.js"), str(adir), str(out),
                        str(tmp_path / "spec.json")], cwd=ROOT, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    res = json.loads((out / "result.json").read_text())
    rd = lambda n: (out / n).read_bytes() if (out / n).exists() else None
    assert res["drawn"] == 0 and res["nodeBase"] == G + 1 and res["lodNodeCount"] == Nn - G
    boxes = sum(p["name"] in ("short_box", "tall_box", "floor") for p in header["parts"])
    assert res["materialCount"] == bufs["materials"].shape[0] + boxes
    assert res["fill"] == (bufs["fill"] is not None)
    assert [s["xfer"] for s in res["stats"]] == [True, False, False]
    assert [s["lightStill"] for s in res["stats"]] == [False, False, True]
    assert {"project", "radix", "pairs", "tilesort", "ranges", "splat"} <= set(res["timers"])
    # the same calls through ctypes, with the bytes the shim handed over
    gauss = np.frombuffer(rd("gaussians.bin"), R.GAUSSIAN_DTYPE).copy()
    mats = np.frombuffer(rd("materials.bin"), R.MATERIAL_DTYPE).copy()
    if boxes:
        moved = np.nonzero(gauss["flags"] != sc.gaussians["flags"])[0]
        assert moved.size and np.all((gauss["flags"][moved] & 0xFFFF) >= bufs["materials"].shape[0])
    scene = N.Scene(ctx, gauss, sc.parts, materials=mats, nodes=nodes[:G], vgrid=bufs["vgrid"])
    scene.set_lod(nodes, kids, np.frombuffer(rd("node_gaussians.bin"), R.GAUSSIAN_DTYPE), emit_part=emit)
    if bufs["fill"] is not None:
        scene.set_fill(bufs["fill"])
    scene.bounce_init(nodes, kids, int(header["walk"]["max_depth"]), rd("PP_init.bin"))
    hdr, disp = ctx.alloc(W * H * 8), ctx.alloc(W * H * 4)
    try:
        for k, f in enumerate(frames):
            st = res["stats"][k]
            scene.light(rd(f"U{k}.bin"), rd(f"EP{k}.bin"), np.frombuffer(rd(f"emit_nodes{k}.bin"), np.float32),
                        rd(f"LS{k}.bin"), rd(f"ENV{k}.bin"), rd(f"SM{k}.bin"), rd(f"desc{k}.bin"))
            scene.bounce_encode(rd(f"PP{k}.bin"), (N.SL_BOUNCE_XFER if st["xfer"] else 0) |
                                (N.SL_BOUNCE_LIGHT_STILL if st["lightStill"] else 0))
            if bufs["fill"] is not None:
                scene.fill()
            N.frame_encode(ctx, scene, f["bytes"], W, H, hdr, N.SL_OUT_RGBA16F)
            ps = N.frame_finish(ctx)
            N.display_encode(ctx, f["bytes"], W, H, hdr, disp, N.SL_OUT_RGBA16F, N.SL_DISPLAY_RGBA8,
                             rd(f"ENV{k}.bin"), rd("BG.bin"))
            py_hdr = ctx.read(hdr, W * H * 8, np.uint8).tobytes()
            py_disp = ctx.read(disp, W * H * 4, np.uint8).tobytes()
            assert (st["st"]["visible"], st["st"]["pairs"]) == (ps.visible, ps.pairs), f"frame {k}"
            assert rd(f"hdr{k}.bin") == py_hdr, f"frame {k}: HDR target"
            assert rd(f"display{k}.bin") == py_disp, f"frame {k}: display bytes"
            cov = np.frombuffer(py_disp, np.uint8).reshape(H, W, 4)[..., :3].mean()
            log(f"  frame {k}: visible={ps.visible} pairs={ps.pairs}, HDR and display bytes identical to the ctypes "
                f"path (mean display {cov:.1f})")
    finally:
        ctx.free(hdr)
        ctx.free(disp)
        scene.destroy()


def test_js_host_encodes_next_frame_before_awaiting(ctx, tmp_path):
    """Lighting and frame encoded for four cameras back to back before any frame's counts are awaited (as the
    reference's capture loop does): each frame's promise resolves to that frame's own counts, the same as when every
    frame is awaited before the next, and the last target's bytes are the same."""
    adir = tmp_path / "asset"
    jsasset.write("cornell", str(adir))
    sc, nodes, kids, _, _ = lodcut.load("cornell", W, H)
    t = np.asarray(sc.target, np.float64)
    files = []
    for k, (rad, az) in enumerate(((2.5, -1.2), (3.5, -0.8), (5.0, -0.4), (3.0, 0.0))):
        eye = tuple(float(v) for v in t + rad * np.array([math.cos(0.3) * math.cos(az - math.pi / 2),
                                                          math.cos(0.3) * math.sin(az - math.pi / 2), math.sin(0.3)]))
        fb = R.render_frame(view=R.look_at(eye, sc.target, (0.0, 0.0, 1.0)), width=W, height=H,
                            fovy=math.radians(sc.fov_deg), far=sc.far, count=sc.count)
        (tmp_path / f"frame{k}.bin").write_bytes(fb)
        files.append({"eye": list(eye), "frame": f"frame{k}.bin"})
    (tmp_path / "spec.json").write_text(json.dumps({"width": W, "height": H, "frames": files}))
    log("node pipelined frames: 4 cameras over the Cornell cut")
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node_pipeline.js"), str(adir), str(tmp_path / "spec.json"),
                        str(tmp_path / "piped.json")], cwd=ROOT, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    res = json.loads((tmp_path / "piped.json").read_text())
    piped = [(f["visible"], f["pairs"]) for f in res["piped"]]
    serial = [(f["visible"], f["pairs"]) for f in res["serial"]]
    log(f"  pipelined {piped}, awaited {serial}")
    assert piped == serial
    assert len(set(serial)) > 1                 # the cameras differ, so a stale count would show
    assert res["hdrEqual"] and res["displayEqual"]
