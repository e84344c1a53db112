"""Benchmark of the splat frame path on one MI355X.

A step is one fully relit frame: the emitter-light opacity maps and the sun maps re-splatted for every part
(the light moved), their lookups, shade_emit, env_direct, then project -> depth order -> pair emission -> tile
sort -> tile ranges -> composite. The scene is resident in HBM before timing starts. Workload: the outdoor hill,
1M Gaussians, 64 emissive Gaussians and a sun with a 2048^2 map, 1920x1080 (16x16 tiles: 8,160). The rate with
the lighting cached (nothing moved, the reference's still-skip case) is reported beside it.

Prints one JSON line (value = Mpix/s over the timed frames) with the composite kernel's HBM roofline and the
CPU restatement's rate on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--half", type=int, default=1, help="1: rgba16f target (the reference's), 0: rgba32f")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB, 5 times (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size: what FETCH_SIZE / WRITE_SIZE read for it scales the other kernels'."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"copy_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4, "ms": round(dt * 1e3, 4),
                      "GB_s": round(2 * n * 4 / dt / 1e9, 1)}))


def cpu_baseline(sc, fb, k=8):
    """The CPU restatement (oracle, one core) on a bounded sample of one relit frame, scaled to the frame.

    Per-Gaussian relighting (emitter map splat + lookup, shade_emit, sun map splat + lookup, env_direct) runs on
    every k-th Gaussian and is scaled by k; the per-texel map resolves are timed alone and not scaled; project,
    depth order and tile lists run on the whole frame; the fp16-target blend runs on every k-th pixel row and is
    scaled by k. iso_norm (once per asset) is not timed, as on the device.
    """
    import dataclasses

    import oracle
    from splatlight_amd import lighting as L
    from splatlight_amd import scenes
    W, H = sc.width, sc.height
    sub = dataclasses.replace(sc, gaussians=np.ascontiguousarray(sc.gaussians[::k]))
    inp = L.lighting_inputs(sub, dirty_all=True)
    nd = scenes.leaf_nodes(sub.gaussians)
    vg = scenes.vgrid_open(sub.count)
    iso = oracle.iso_init(sub.gaussians, nd, vg)
    t0 = time.perf_counter()
    acc, _, T = oracle.lsm(inp["U"], inp["LS"], sub.gaussians, nd, sub.parts)
    col = oracle.shade_emit(inp["U"], inp["EP"], inp["emit_nodes"], sub.gaussians, nd, vg, sub.parts, T,
                            sub.materials, iso)
    sacc, _, envT = oracle.sun(inp["U"], inp["SM"], inp["desc"], sub.gaussians, nd, sub.parts)
    oracle.env_direct(inp["U"], inp["ENV"], sub.gaussians, nd, vg, sub.parts, sub.materials, iso, envT, col)
    t_light = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.lsm_resolve(inp["LS"], acc)
    oracle.sun_resolve(inp["SM"], inp["desc"], sacc)
    t_resolve = time.perf_counter() - t0
    colour = np.concatenate([sc.gaussians["albedo"], np.ones((sc.count, 1))], 1).astype(np.float32)
    t0 = time.perf_counter()
    sp, keys, vis = oracle.project(fb, sc, colour=colour)
    _, _, order, _ = oracle.sort_keys(keys)
    oracle.tile_lists(sp, order, vis, W, H)
    t_frame = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.composite(fb, sp, order, vis, W, H, mode=2, row_step=k)
    t_blend = time.perf_counter() - t0
    est = k * (t_light - t_resolve) + t_resolve + t_frame + k * t_blend
    return {"value": round(W * H / est / 1e6, 4), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"one relit frame ({sc.count} Gaussians, {W}x{H}, emitter map + {2048}^2 sun maps) on one "
                      f"core: per-Gaussian relighting of every {k}th Gaussian ({t_light:.1f} s, x{k}), map resolves "
                      f"({t_resolve:.1f} s), project/order/tile lists of the whole frame ({t_frame:.1f} s), the "
                      f"fp16-target blend of every {k}th pixel row ({t_blend:.1f} s, x{k}); estimated frame "
                      f"{est:.1f} s"}


def main():
    a = parse()
    if a.copy_mib:
        copy_calibration(a.copy_mib)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (this build has no multi-GPU path)",
              file=sys.stderr)
        sys.exit(2)
    from splatlight_amd import lighting as L
    from splatlight_amd import native as N
    from splatlight_amd import scenes

    sc = scenes.hill(n=a.n, seed=1, width=a.width, height=a.height)
    fb = sc.frame_bytes()
    W, H = sc.width, sc.height
    fmt = N.SL_OUT_RGBA16F if a.half else N.SL_OUT_RGBA32F
    bpx = 8 if a.half else 16
    inp = L.lighting_inputs(sc, dirty_all=True)

    ctx = N.Context(0)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials,
                    nodes=scenes.leaf_nodes(sc.gaussians), vgrid=scenes.vgrid_open(sc.count))
    scene.iso_init()
    out = ctx.alloc(W * H * bpx)

    def lit_frame():
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"],
                    timers=True)
        N.frame_encode(ctx, scene, fb, W, H, out, fmt, timers=True)
        return N.frame_finish(ctx)

    for _ in range(a.warmup):
        st = lit_frame()
    ctx.sync()
    frame_ms, comp_ms, passes = [], [], {}
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t = time.perf_counter()
        st = lit_frame()
        frame_ms.append((time.perf_counter() - t) * 1e3)
        tm = N.timers(ctx)
        comp_ms.append(tm["splat"])
        for k, v in tm.items():
            passes.setdefault(k, []).append(v)
    ctx.sync()
    total = time.perf_counter() - t0
    # the same frame with the lighting cached (nothing moved): the splat path alone
    t1 = time.perf_counter()
    for _ in range(a.steps):
        N.frame_encode(ctx, scene, fb, W, H, out, fmt)
        N.frame_finish(ctx)
    view_only = (time.perf_counter() - t1) / a.steps
    P, V, T = st.pairs, st.visible, st.tiles
    mpix = W * H * a.steps / total / 1e6
    comp = float(np.mean(comp_ms))
    comp_bytes = 48 * P + 4 * P + 8 * T + W * H * bpx        # splat fetch + pair value per pair, ranges, target
    achieved = comp_bytes / (comp * 1e-3) / 1e9
    res = {
        "metric": "Mpix/s + Gaussian-tile pairs/s, 1080p/1M Gaussians; p50/p99 latency at fixed req/s",
        "value": round(mpix, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(total / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded hill scene)",
        "config": {"workload": f"hill {sc.count} Gaussians, 64 emitters + sun (2048^2 map), {W}x{H} fully relit "
                               f"frame, 16x16 tiles",
                   "gaussians": sc.count, "width": W, "height": H, "target": "rgba16f" if a.half else "rgba32f"},
        "view_only_mpix_s": round(W * H / view_only / 1e6, 3),
        "view_only_ms": round(view_only * 1e3, 4),
        "pairs_per_s": round(P * a.steps / total, 1),
        "pairs": P,
        "visible": V,
        "frame_ms_p50": round(float(np.percentile(frame_ms, 50)), 4),
        "frame_ms_p99": round(float(np.percentile(frame_ms, 99)), 4),
        "pass_ms": {k: round(float(np.mean(v)), 4) for k, v in passes.items()},
        "roofline": {"bound": "hbm", "kernel": "sl_composite", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "algorithmic_bytes": comp_bytes, "avg_ms": round(comp, 5)},
    }
    ctx.free(out)
    scene.destroy()
    ctx.close()
    if a.cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(sc, fb)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
