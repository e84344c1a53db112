"""C-ABI checks that need no GPU: the library loads, exports every declared symbol, and its host-side error
paths behave as the reference's (session lock, containment of a missing device)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from splatlight_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "splatlight.h")).read()
    return sorted(set(re.findall(r"\b(sl_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    decl = declared_symbols()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(N.EXPORTS) == decl
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (sl_[a-z_0-9]+)$", out, re.M))
    assert set(decl) <= exported


def test_library_is_gfx950_code():
    # the fat binary of our own build names its one code object
    blob = open(N.LIB_PATH, "rb").read()
    assert b"hipv4-amdgcn-amd-amdhsa--gfx950" in blob


def test_session_lock_is_exclusive_and_fails_fast():
    L = N.lib()
    name = b"sl-test-session-%d" % os.getpid()
    t1, t2 = C.c_int64(), C.c_int64()
    assert L.sl_session_acquire(name, C.byref(t1)) == N.SL_OK
    # a second holder is refused at once with the reference's message (src/gpu-session.js:9)
    assert L.sl_session_acquire(name, C.byref(t2)) == 6
    assert L.sl_session_error().decode() == (
        "Another tab on this origin owns the GPU session. Close that experiment tab first.")
    assert L.sl_session_release(t1.value) == N.SL_OK
    assert L.sl_session_acquire(name, C.byref(t2)) == N.SL_OK
    assert L.sl_session_release(t2.value) == N.SL_OK


def test_null_context_is_rejected():
    L = N.lib()
    assert L.sl_sync(None) == 1
    assert L.sl_ctx_status(None) == 1
    assert L.sl_last_error(None) == b"no context"


def test_context_without_device_fails_loudly():
    if os.path.exists("/dev/kfd") and os.environ.get("HIP_VISIBLE_DEVICES", "") != "-1":
        pytest.skip("a device may be present")
    with pytest.raises(N.SplatlightError):
        N.Context(0)
