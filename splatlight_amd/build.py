"""Build the HIP library (gfx950) in-tree: splatlight_amd/libsplatlight_hip.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsplatlight_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SL_OFFLOAD_ARCH", "gfx950")


def hip_flags() -> list[str]:
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include")]


def build_lib(verbose: bool = False) -> str:
    src = os.path.join(HERE, "csrc", "splatlight.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *hip_flags(), "-shared", "-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libsplatlight_hip.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


NODE_ADDON = os.path.join(HERE, "node", "splatlight.node")
NODE_INCLUDE = os.environ.get("SL_NODE_INCLUDE", "/usr/include/node")


def build_node_addon() -> str | None:
    """The N-API addon the JavaScript host binds (splatlight_amd/node/splatlight.node), linked against the library;
    None when no Node headers are installed."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    src = os.path.join(HERE, "node", "splatlight_node.cc")
    tmp = NODE_ADDON + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"),
           "-I" + NODE_INCLUDE, src, "-L" + HERE, "-lsplatlight_hip", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for the N-API addon")
    os.replace(tmp, NODE_ADDON)
    return NODE_ADDON


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_node_addon())
