"""The N-API addon and its JS shim, on the CPU: it loads, exports the binding, and the session lock behaves as
acquireGpuSession (src/gpu-session.js:3-14): the second holder is refused at once with the reference's message."""
import os
import shutil
import subprocess

import pytest

from splatlight_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(B.NODE_ADDON), reason="no node / addon not built")


def run_js(code):
    return subprocess.run([NODE, "-e", code], cwd=ROOT, capture_output=True, text=True, timeout=60)


def test_addon_exports_the_binding():
    r = run_js("const a=require('./splatlight_amd/node/splatlight.node'); console.log(Object.keys(a).sort().join(','))")
    assert r.returncode == 0, r.stderr
    names = set(r.stdout.strip().split(","))
    assert {"sessionAcquire", "sessionRelease", "ctxCreate", "bufAlloc", "bufRead", "sceneCreate", "sceneWriteParts",
            "sceneWriteColour", "lightingIsoInit", "lightingEncode", "frameEncode", "frameFinish",
            "timersRead", "bounceInit", "bounceEncode", "sceneWriteBounce", "sceneWriteGaussians",
            "sceneWriteMaterials", "sceneSetLod", "sceneSetFill", "fillEncode", "displayEncode", "sync"} <= names


def test_shim_session_lock_matches_the_reference():
    r = run_js("""
const {acquireGpuSession} = require('./splatlight_amd/node/renderer.js');
(async () => {
  const release = await acquireGpuSession('splatlight-node-cpu-test');
  let msg = '';
  try { await acquireGpuSession('splatlight-node-cpu-test'); } catch (e) { msg = e.message; }
  release();
  const again = await acquireGpuSession('splatlight-node-cpu-test');
  again();
  console.log(msg);
})().catch((e) => { console.error(e); process.exit(1); });
""")
    assert r.returncode == 0, r.stderr
    assert "Another tab on this origin owns the GPU session. Close that experiment tab first." in r.stdout
