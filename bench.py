"""Benchmark of the splat frame path on one MI355X.

A step is one fully relit frame: the emitter-light opacity maps and the sun maps re-splatted for every part
(the light moved), their lookups, shade_emit, env_direct, then project -> depth order -> pair emission -> tile
sort -> tile ranges -> composite. The scene is resident in HBM before timing starts. Workload: the outdoor hill,
1M Gaussians, 64 emissive Gaussians and a sun with a 2048^2 map, 1920x1080 (16x16 tiles: 8,160). The rate with
the lighting cached (nothing moved, the reference's still-skip case) is reported beside it.

Prints one JSON line (value = Mpix/s over the timed frames) with the roofline of the kernel launch that takes the
most time per frame and the CPU restatement's rate on a bounded sample of the same workload.
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
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_hill1m_1080p.txt"),
                    help="rocprof traffic summary written by --pmc-summary; its figure fills roofline.traffic")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="summarise rocprofv3 --pmc output directories into --traffic and exit")
    ap.add_argument("--pmc-command", default="", help="the profiling command, recorded in the summary")
    ap.add_argument("--pmc-valu", default="",
                    help="with --pmc-summary: a rocprofv3 --pmc run of SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS "
                         "SQ_WAVES GRBM_GUI_ACTIVE, summarised beside the traffic")
    ap.add_argument("--stats-summary", nargs=2, metavar=("KERNEL_STATS_CSV", "OUT"),
                    help="write a markdown summary of a rocprofv3 --kernel-trace --stats run and exit")
    ap.add_argument("--batched", type=int, default=1, help="also measure the batched form (64 views, 8 scenes)")
    ap.add_argument("--serving", type=int, default=1, help="also measure the served form (Poisson arrivals)")
    ap.add_argument("--rate", type=float, default=1000.0, help="serving arrival rate, requests per second")
    ap.add_argument("--storage", type=int, default=1, help="also measure f16 / bf16 storage at 4K over 4M Gaussians")
    ap.add_argument("--bounce", type=int, default=1, help="also measure the relit frame with sky and bounce through "
                                                          "probes (lighting, bounce, frame at bounce gain 1)")
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


def cpu_frame_seconds(sc, fb, oracle, L, scenes):
    """One fully relit frame of `sc` through the CPU restatement (oracle), timed end to end: the emitter map (splat,
    resolve, mips, lookup), shade_emit, the sun maps (when the scene has a sun) and env_direct, then project, the
    stable depth order and the back-to-front blend into an fp16 target rounded after every blend (the reference's
    rgba16float). iso_norm (once per asset, as on the device) is set up outside the timed frame."""
    inp = L.lighting_inputs(sc, dirty_all=True)
    nd = scenes.leaf_nodes(sc.gaussians)
    vg = scenes.vgrid_open(sc.count)
    iso = oracle.iso_init(sc.gaussians, nd, vg)
    t0 = time.perf_counter()
    _, _, T = oracle.lsm(inp["U"], inp["LS"], sc.gaussians, nd, sc.parts)
    col = oracle.shade_emit(inp["U"], inp["EP"], inp["emit_nodes"], sc.gaussians, nd, vg, sc.parts, T,
                            sc.materials, iso)
    if inp["SM"] is not None:
        _, _, envT = oracle.sun(inp["U"], inp["SM"], inp["desc"], sc.gaussians, nd, sc.parts)
        col = oracle.env_direct(inp["U"], inp["ENV"], sc.gaussians, nd, vg, sc.parts, sc.materials, iso, envT, col)
    sp, keys, vis = oracle.project(fb, sc, colour=col)
    _, _, order, _ = oracle.sort_keys(keys)
    oracle.composite(fb, sp, order, vis, sc.width, sc.height, mode=2)
    return time.perf_counter() - t0


def cpu_baseline(sc, fb):
    """The CPU restatement (oracle, C with OpenMP over every host thread it is given) rendering whole relit frames
    end to end, no sampling: the measured workload (hill 1M, 64 emitters, 2048^2 sun map, 1080p), best of two
    frames, and the Cornell box workload (200k Gaussians, 16 emitters, 640x480) beside it. The reference's WGSL on
    a software WebGPU adapter is not available (no Node WebGPU module, no Vulkan ICD)."""
    import oracle
    from splatlight_amd import lighting as L
    from splatlight_amd import scenes
    threads = oracle.threads()
    t1 = min(cpu_frame_seconds(sc, fb, oracle, L, scenes) for _ in range(2))
    c0 = scenes.cornell(n=200_000, seed=0)
    t0 = min(cpu_frame_seconds(c0, c0.frame_bytes(), oracle, L, scenes) for _ in range(2))
    return {"value": round(sc.width * sc.height / t1 / 1e6, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"whole relit frames, end to end, on {threads} host threads (OpenMP): hill {sc.count} "
                      f"Gaussians {sc.width}x{sc.height} with emitter map and 2048^2 sun maps {t1:.2f} s a frame "
                      f"(best of 2); Cornell box 200k Gaussians 640x480 with 16 emitters {t0:.2f} s a frame "
                      f"({640 * 480 / t0 / 1e6:.3f} Mpix/s)",
            "config0_mpix_s": round(640 * 480 / t0 / 1e6, 4), "frame_s": round(t1, 3), "config0_frame_s": round(t0, 3)}


def batched_leg(N, L, scenes, rounds=5):
    """The batched form on 64 views over 8 resident scenes (Gaussian counts log-uniform in [50k, 2M], Cornell-like
    and hill alternately, seeds 1-8; sizes from 256^2 to 1920x1080; cameras on orbits of radius U[2, 6] m), each
    scene relit once (the lighting is view-independent), against the plain per-frame path over the same views."""
    scs, views = scenes.batch_workload(seed=0)
    ctx = N.Context(0)
    handles = []
    for sc in scs:
        h = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials,
                    nodes=scenes.leaf_nodes(sc.gaussians), vgrid=scenes.vgrid_open(sc.count))
        inp = L.lighting_inputs(sc, dirty_all=True)
        h.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        handles.append(h)
    frames = [scs[si].frame_bytes(width=w, height=h, eye=eye) for si, w, h, eye in views]
    outs = [ctx.alloc(w * h * 8) for _, w, h, _ in views]
    batch = [(handles[si], fb, w, h, o, N.SL_OUT_RGBA16F) for (si, w, h, _), fb, o in zip(views, frames, outs)]
    mpix = sum(w * h for _, w, h, _ in views) / 1e6

    def plain():
        for (sc, fb, w, h, o, fmt) in batch:
            N.frame_encode(ctx, sc, fb, w, h, o, fmt)
            N.frame_finish(ctx)

    def batched():
        N.batch_encode(ctx, batch, graph=True)
        return N.batch_finish(ctx, len(batch))

    plain()
    from splatlight_amd import autotune
    tuning, table = autotune.tune(ctx, batch, repeats=3, cache=None)
    batched()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(rounds):
        plain()
    t_plain = (time.perf_counter() - t0) / rounds
    lat = []
    for _ in range(rounds * 4):
        t = time.perf_counter()
        st = batched()
        lat.append(time.perf_counter() - t)
    t_batch = float(np.mean(lat))
    pairs = sum(x.pairs for x in st)
    for o in outs:
        ctx.free(o)
    for h in handles:
        h.destroy()
    ctx.close()
    return {"workload": f"{len(views)} views over {len(scs)} scenes "
                        f"({min(sc.count for sc in scs)}-{max(sc.count for sc in scs)} Gaussians), sizes 256^2 to "
                        f"1920x1080, rgba16f, hipGraph replay",
            "mpix_per_batch": round(mpix, 3), "pairs_per_batch": int(pairs),
            "batched_mpix_s": round(mpix / t_batch, 1), "plain_mpix_s": round(mpix / t_plain, 1),
            "speedup_vs_plain": round(t_plain / t_batch, 3),
            "batch_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "batch_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "plain_ms_per_batch": round(t_plain * 1e3, 3), "tuning": tuning,
            "tuning_ms": {f"{c['tile_px']}px/{c['radix_bits']}b/{c['composite_batch']}/{c['radix_items']}/"
                          f"{c['grid_mul']}": round(ms, 3) for c, ms in table}}


def serving_leg(N, scenes, rate=1000.0, seconds=2.0, budget=16 << 30, lanes=3, max_batch=64):
    """The served form: Poisson arrivals at `rate` over 32 resident scenes (100k-3M Gaussians, seeds 10-41, colour
    precomputed), 4 sizes from 256^2 to 1920x1080, random orbit poses; admission by a fixed HBM budget, continuous
    batching over `lanes` lanes. Against the plain per-frame path fed the same arrivals one frame at a time. Also
    the saturated rate of both (every request queued at once)."""
    scs = scenes.serving_workload()
    ctx = N.Context(0)
    handles = []
    for sc in scs:
        h = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
        rng = np.random.default_rng(sc.count)
        h.write_colour(np.concatenate([sc.gaussians["albedo"] * rng.uniform(0.5, 3.0, (sc.count, 1)),
                                       np.ones((sc.count, 1))], 1).astype(np.float32))
        handles.append(h)
    reqs = scenes.request_stream(scs, rate, int(rate * seconds))
    frames = [scs[si].frame_bytes(width=w, height=h, eye=eye) for _, si, w, h, eye in reqs]
    ring = lanes * max_batch                      # targets per size: enough for everything in flight
    targets = {(w, h): [ctx.alloc(w * h * 8) for _ in range(ring)] for w, h in scenes.BATCH_SIZES}
    mpix = sum(w * h for _, _, w, h, _ in reqs) / 1e6
    # the lanes' tuning: the autotuner on a batch of the first max_batch requests (the lanes copy the owner
    # context's tuning when the server starts); the plain path keeps the default tuning
    from splatlight_amd import autotune
    default_tuning = N.get_tuning(ctx)
    sample = [(handles[si], fb, w, h, targets[(w, h)][k % ring], N.SL_OUT_RGBA16F)
              for k, ((_, si, w, h, _), fb) in enumerate(zip(reqs[:max_batch], frames[:max_batch]))]
    lane_tuning, _ = autotune.tune(ctx, sample, repeats=3, cache=None,
                                   candidates=[c for c in autotune.CANDIDATES if c["radix_bits"] == 8])
    N.set_tuning(ctx, **default_tuning)

    def served(arrivals):
        N.set_tuning(ctx, **lane_tuning)
        srv = N.Server(ctx, hbm_budget=budget, lanes=lanes, max_batch_views=max_batch)
        N.set_tuning(ctx, **default_tuning)
        done, sizes = [], []
        t0 = time.perf_counter()
        for k, ((t_arr, si, w, h, _), fb) in enumerate(zip(reqs, frames)):
            if arrivals:
                while time.perf_counter() - t0 < t_arr:
                    done += srv.poll()
            srv.submit(handles[si], fb, w, h, targets[(w, h)][k % ring], N.SL_OUT_RGBA16F)
        while len(done) < len(reqs):
            done += srv.poll(wait_us=2000)
        srv.close()
        lat = np.array([r.done_s - r.submit_s for r in done])
        span = max(r.done_s for r in done) - min(r.submit_s for r in done)
        return lat, span, np.mean([r.batch_views for r in done])

    def plain(arrivals):
        lat = []
        t0 = time.perf_counter()
        for k, ((t_arr, si, w, h, _), fb) in enumerate(zip(reqs, frames)):
            if arrivals:
                while time.perf_counter() - t0 < t_arr:
                    pass
            ta = t0 + t_arr if arrivals else t0
            N.frame_encode(ctx, handles[si], fb, w, h, targets[(w, h)][k % ring], N.SL_OUT_RGBA16F)
            N.frame_finish(ctx)
            lat.append(time.perf_counter() - ta)
        return np.array(lat), time.perf_counter() - t0

    served(False)                                  # warm-up: workspaces of every lane grown
    lat_s, _, views_s = served(True)
    lat_p, _ = plain(True)
    _, span_sat, views_sat = served(False)
    _, span_plain = plain(False)
    for lst in targets.values():
        for o in lst:
            ctx.free(o)
    for h in handles:
        h.destroy()
    ctx.close()
    pc = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 3)
    return {"workload": f"Poisson {rate:.0f} req/s for {seconds:.0f} s over {len(scs)} resident scenes "
                        f"({min(sc.count for sc in scs)}-{max(sc.count for sc in scs)} Gaussians), sizes 256^2 to "
                        f"1920x1080, rgba16f; HBM budget {budget >> 20} MiB, {lanes} lanes, <= {max_batch} views/batch",
            "requests": len(reqs), "mpix": round(mpix, 2),
            "served_p50_ms": pc(lat_s, 50), "served_p99_ms": pc(lat_s, 99), "served_mean_batch": round(views_s, 2),
            "plain_p50_ms": pc(lat_p, 50), "plain_p99_ms": pc(lat_p, 99),
            "saturated_served_mpix_s": round(mpix / span_sat, 1), "saturated_mean_batch": round(views_sat, 2),
            "saturated_plain_mpix_s": round(mpix / span_plain, 1),
            "saturated_speedup": round(span_plain / span_sat, 3), "lane_tuning": lane_tuning}


def storage_leg(N, L, scenes, steps=5):
    """Reduced-precision storage at 4K: 4M Gaussians of the hill (seed 9), 3840x2160 rgba16f, quat / scale /
    albedo stored f32, f16 or bf16 (means and opacity f32; all arithmetic f32): relit and cached-lighting frame
    rates and device bytes per storage."""
    sc = scenes.hill(n=4_000_000, seed=9, width=3840, height=2160)
    fb = sc.frame_bytes()
    W, H = sc.width, sc.height
    inp = L.lighting_inputs(sc, dirty_all=True)
    nodes = scenes.leaf_nodes(sc.gaussians)
    vgrid = scenes.vgrid_open(sc.count)
    out = {"workload": f"hill {sc.count} Gaussians, {W}x{H} rgba16f, 64 emitters + sun (2048^2 map)"}
    ctx = N.Context(0)
    target = ctx.alloc(W * H * 8)
    for storage, name in ((0, "f32"), (1, "f16"), (2, "bf16")):
        scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials, nodes=nodes,
                        vgrid=vgrid, storage=storage)
        scene.iso_init()

        def lit():
            scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
            N.frame_encode(ctx, scene, fb, W, H, target, N.SL_OUT_RGBA16F)
            return N.frame_finish(ctx)

        st = lit()
        t0 = time.perf_counter()
        for _ in range(steps):
            lit()
        t_lit = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        for _ in range(2 * steps):
            N.frame_encode(ctx, scene, fb, W, H, target, N.SL_OUT_RGBA16F)
            N.frame_finish(ctx)
        t_view = (time.perf_counter() - t0) / (2 * steps)
        out[name] = {"relit_mpix_s": round(W * H / t_lit / 1e6, 1), "relit_ms": round(t_lit * 1e3, 3),
                     "view_only_mpix_s": round(W * H / t_view / 1e6, 1), "view_only_ms": round(t_view * 1e3, 3),
                     "pairs": int(st.pairs), "attribute_bytes_per_gaussian": 64 if storage == 0 else 48,
                     "scene_device_mib": round(scene.device_bytes / 2**20, 1)}
        scene.destroy()
    ctx.free(target)
    ctx.close()
    return out


def bounce_leg(N, L, scenes, sc, steps=16, warmup=3, half=True):
    """The hill frame with the sky and bounce through probes each frame (Lighting.encodeBounce after the lighting,
    the frame adding gain x bounce): the bounce's GPU time, its probe gather, and the whole relit frame's rate with
    the frames issued as a render loop issues them."""
    from splatlight_amd import bounce as BN
    from splatlight_amd import tree
    t0 = time.perf_counter()
    tr = tree.scene_tree(sc)
    build_s = time.perf_counter() - t0
    W, H = sc.width, sc.height
    fmt = N.SL_OUT_RGBA16F if half else N.SL_OUT_RGBA32F
    ctx = N.Context(0)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials,
                    nodes=tr.nodes[:sc.count], vgrid=scenes.vgrid_open(sc.count))
    b = BN.Bounce(scene, sc, tr)
    fb = sc.frame_bytes(bounce_gain=1.0)
    out = ctx.alloc(W * H * (8 if half else 16))
    inp = L.lighting_inputs(sc, dirty_all=True)
    bt, gt, ft = [], [], []

    def frame(timed):
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
        b.encode(timers=timed)
        N.frame_encode(ctx, scene, fb, W, H, out, fmt)
        N.frame_finish(ctx)
        if timed:
            tm = N.timers(ctx)
            bt.append(tm["bounce"])
            gt.append(tm["probe_gather"])

    def frames(n, record=None):
        # n frames as a render loop issues them: frame i+1's lighting and bounce are encoded before frame i is
        # finished (the bounce joins frame i's composite on the context's stream; the lighting runs beside it)
        def lit():
            scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
            b.encode()
        if n <= 0:
            return
        lit()
        N.frame_encode(ctx, scene, fb, W, H, out, fmt)
        t = time.perf_counter()
        for i in range(n):
            if i + 1 < n:
                lit()
            N.frame_finish(ctx)
            if record is not None:
                record.append((time.perf_counter() - t) * 1e3)
                t = time.perf_counter()
            if i + 1 < n:
                N.frame_encode(ctx, scene, fb, W, H, out, fmt)

    frames(warmup)
    ctx.sync()
    t1 = time.perf_counter()
    frames(steps, ft)
    ctx.sync()
    total = time.perf_counter() - t1
    for _ in range(3):                     # the bounce's pass timers, from frames issued one after another
        frame(True)
    rays = b.n_probes * 128
    res = {"workload": f"hill {sc.count} Gaussians, tree of {tr.count} nodes, {b.n_probes} probes x 128 rays, "
                       f"sky + sun + 64 emitters, {W}x{H}",
           "relit_frame_ms": round(total / steps * 1e3, 4), "mpix_s": round(W * H * steps / total / 1e6, 3),
           "frame_ms_p50": round(float(np.percentile(ft, 50)), 4),
           "bounce_ms": round(float(np.mean(bt)), 4), "probe_gather_ms": round(float(np.mean(gt)), 4),
           "probe_rays_per_s": round(rays / (float(np.mean(gt)) * 1e-3), 1), "probes": b.n_probes,
           "nodes": tr.count, "tree_build_s": round(build_s, 2)}
    ctx.free(out)
    scene.destroy()
    ctx.close()
    return res


def kernel_bytes(passes, P, T, target_bytes, ls):
    """Per-launch algorithmic HBM bytes and HIP-event milliseconds of the heaviest kernel launches of a relit frame.

    sl_composite<256u, 4u, 4u> (one launch; 256 splats staged per round over 16x16-pixel tiles, the default tuning): a
    48 B splat fetch and a 4 B pair value per pair, 8 B per tile range, the target.
    sl_map_accumulate<0u> (emitter map) and <1u> (sun maps), one launch each: a 64 B record and a 4 B list entry
    per (record, map tile) pair, 4 B per tile key, 12 B per texel of every map tile that has records.
    """
    out = [{"kernel": "sl_composite<256u, 4u, 4u>", "ms": float(np.mean(passes["splat"])),
            "bytes": 48 * P + 4 * P + 8 * T + target_bytes}]
    for kind, m, key in ((0, "lsm", "lsm_sum"), (1, "sun", "sun_sum")):
        if key in passes:
            out.append({"kernel": f"sl_map_accumulate<{kind}u>", "ms": float(np.mean(passes[key])),
                        "bytes": 68 * ls[f"{m}_pairs"] + 4 * (ls[f"{m}_tiles"] + 1)
                        + 256 * 12 * ls[f"{m}_tiles_used"]})
    return out


def roofline(passes, P, T, target_bytes, ls, traffic_file):
    """The roofline object of the kernel launch that takes the most time per frame (HIP events on the stream the
    kernel is launched on)."""
    k = max(kernel_bytes(passes, P, T, target_bytes, ls), key=lambda e: e["ms"])
    achieved = k["bytes"] / (k["ms"] * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": k["kernel"], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
           "traffic": summary_entry(traffic_file, "TRAFFIC", k["kernel"], "bytes_per_launch"),
           "algorithmic_bytes": int(k["bytes"]), "avg_ms": round(k["ms"], 5), "launches_per_frame": 1}
    valu = summary_entry(traffic_file, "VALU", k["kernel"], None)
    if valu:
        out["valu"] = valu
    return out


def short_name(kernel):
    """rocprof's kernel name without 'void ' and the argument list: sl_map_accumulate<1u>."""
    n = kernel[5:] if kernel.startswith("void ") else kernel
    return n.replace("(anonymous namespace)::", "").split("(")[0]


def summary_entry(path, tag, kernel, field):
    """A per-kernel figure from the rocprof summary written by --pmc-summary (its `tag` lines), or None."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(tag + " "):
                d = json.loads(line[len(tag) + 1:])
                if d.get("kernel") == kernel:
                    return d if field is None else d[field]
    return None


def stats_summary(csv_path, out, command):
    """Markdown table of a rocprofv3 --kernel-trace --stats kernel_stats.csv, with the command that made it."""
    import csv
    rows = list(csv.DictReader(open(csv_path)))
    lines = ["# Kernel time, relit frame: hill 1M Gaussians, 64 emitters + sun 2048^2, 1920x1080 rgba16f", "",
             "Command (on one MI355X):", "", "    " + command, "",
             "Per relit frame: one launch each of sl_map_accumulate<0u> (emitter map) and <1u> (sun maps), "
             "sl_shade_emit, sl_env_direct and the frame path; the cached-lighting frames add frame-path launches "
             "only. The emitter-map chain and the sun-map chain run on two streams at once, so kernel times include "
             "sharing the GPU. The bounce leg's frames add the sky and bounce through probes (tree of 2M nodes, "
             "1,339 probes x 128 rays): sl_bounce_lo_dir, sl_leaf_sh, sl_node_up_sh (one launch per deep tree "
             "depth), sl_node_up_tail (the shallow depths), sl_probe_gather, sl_probe_finish, sl_bounce_sh each "
             "frame, sl_iso_xfer and sl_leaf_weights once.", "",
             "| kernel | calls | avg us | min us | max us | share |", "|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f}% |")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, out, command, valu_dir=""):
    """Per-kernel HBM bytes from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch), each scaled by
    what the counter reads for a plain copy kernel of known size, and, from a valu_dir run, instruction counts per
    dispatch; writes a short text summary."""
    import csv
    import glob
    from collections import defaultdict

    def per_kernel(d, counter):
        vals = defaultdict(list)
        for p in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(p)):
                if r["Counter_Name"] == counter:
                    vals[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return vals

    def copy_scale(d, counter, nbytes):
        v = per_kernel(d, counter)
        k = max(v, key=lambda name: np.mean(v[name]))
        return nbytes / (np.mean(v[k]) * 1024.0), k, float(np.mean(v[k]))

    keep = ("sl_composite<256u, 4u, 4u>", "sl_map_accumulate<0u>", "sl_map_accumulate<1u>", "sl_shade_emit_wave",
            "sl_shade_emit<48u>")
    copy_bytes = 1024 << 20
    fs, fk, fv = copy_scale(copy_fetch_dir, "FETCH_SIZE", copy_bytes)
    ws, wk, wv = copy_scale(copy_write_dir, "WRITE_SIZE", copy_bytes)
    fetch, write = per_kernel(fetch_dir, "FETCH_SIZE"), per_kernel(write_dir, "WRITE_SIZE")
    lines = ["# HBM traffic per launch, rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), hill 1M Gaussians 1920x1080", "",
             "Command (each counter in its own run):", "", "    " + command, "",
             f"Calibration: a 1 GiB device copy ({fk}) reads FETCH_SIZE {fv:.0f} KiB and WRITE_SIZE {wv:.0f} KiB per "
             f"dispatch, so FETCH_SIZE is scaled by {fs:.3f} and WRITE_SIZE by {ws:.3f}.", "",
             "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes per launch (scaled) |",
             "|---|---|---|---|---|"]
    rows = []
    for name in sorted(set(fetch) | set(write), key=lambda n: -np.mean(fetch.get(n, [0]))):
        f_kib, w_kib = float(np.mean(fetch.get(name, [0]))), float(np.mean(write.get(name, [0])))
        b = f_kib * 1024 * fs + w_kib * 1024 * ws
        rows.append((name, f_kib, w_kib, b))
        lines.append(f"| {name[:48]} | {len(fetch.get(name, []))} | {f_kib:.0f} | {w_kib:.0f} | {b:.4g} |")
    lines.append("")
    for name, f_kib, w_kib, b in rows:
        if name in keep:
            lines.append("TRAFFIC " + json.dumps({"kernel": name, "bytes_per_launch": int(b),
                                                  "fetch_kib": round(f_kib, 1), "write_kib": round(w_kib, 1),
                                                  "fetch_scale": round(fs, 4), "write_scale": round(ws, 4)}))
    if valu_dir:
        cnt = {c: per_kernel(valu_dir, c) for c in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_WAVES",
                                                     "GRBM_GUI_ACTIVE")}
        lines += ["", "Instructions per dispatch (wave instructions, summed over the chip; rocprofv3 --pmc "
                  "SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVES GRBM_GUI_ACTIVE, one run). GRBM_GUI_ACTIVE "
                  "sums the 8 XCDs' busy cycles, so CU-cycles = 32 x GRBM_GUI_ACTIVE (256 CUs) and VALU issue per "
                  "CU-cycle = SQ_INSTS_VALU / CU-cycles.", "",
                  "| kernel | dispatches | VALU | SALU | LDS | waves | GRBM_GUI_ACTIVE | VALU per CU-cycle |",
                  "|---|---|---|---|---|---|---|---|"]
        valu_rows = []
        for name in sorted(cnt["SQ_INSTS_VALU"], key=lambda n: -np.mean(cnt["SQ_INSTS_VALU"][n])):
            m = {c: float(np.mean(cnt[c].get(name, [0]))) for c in cnt}
            rate = m["SQ_INSTS_VALU"] / (32.0 * m["GRBM_GUI_ACTIVE"]) if m["GRBM_GUI_ACTIVE"] else 0.0
            valu_rows.append((name, m, rate))
            lines.append(f"| {name[:48]} | {len(cnt['SQ_INSTS_VALU'][name])} | {m['SQ_INSTS_VALU']:.4g} | "
                         f"{m['SQ_INSTS_SALU']:.4g} | {m['SQ_INSTS_LDS']:.4g} | {m['SQ_WAVES']:.0f} | "
                         f"{m['GRBM_GUI_ACTIVE']:.4g} | {rate:.3f} |")
        lines.append("")
        for name, m, rate in valu_rows:
            if name in keep:
                lines.append("VALU " + json.dumps({"kernel": name, "insts_per_launch": int(m["SQ_INSTS_VALU"]),
                                                   "issue_per_cu_cycle": round(rate, 3)}))
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    a = parse()
    if a.pmc_summary:
        pmc_summary(*a.pmc_summary, a.traffic, a.pmc_command, a.pmc_valu)
        return
    if a.stats_summary:
        stats_summary(*a.stats_summary, a.pmc_command)
        return
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

    def lit_frame(frame_timers):
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"],
                    timers=True)
        N.frame_encode(ctx, scene, fb, W, H, out, fmt, timers=frame_timers)
        return N.frame_finish(ctx)

    def light():
        scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"],
                    timers=True)

    def frames(n, record=None):
        # n relit frames as a render loop issues them: frame i+1's lighting is encoded before frame i is finished,
        # so it runs beside frame i's composite (the composite reads nothing the lighting writes)
        if n <= 0:
            return None
        light()
        N.frame_encode(ctx, scene, fb, W, H, out, fmt)
        st, t = None, time.perf_counter()
        for i in range(n):
            if i + 1 < n:
                light()
            st = N.frame_finish(ctx)
            if record is not None:
                record(time.perf_counter() - t)
                t = time.perf_counter()
            if i + 1 < n:
                N.frame_encode(ctx, scene, fb, W, H, out, fmt)
        return st

    frames(a.warmup)
    ctx.sync()
    # timed: the lighting's pass timers (the map sums' HIP events for the roofline) are recorded live in every
    # frame and read after each frame's finish (the last completed lighting's); the frame path's pass timers come
    # from a few untimed frames after, so the timed frames carry only the events they use
    frame_ms, passes = [], {}

    def record(dt):
        frame_ms.append(dt * 1e3)
        for k, v in N.timers(ctx).items():
            passes.setdefault(k, []).append(v)
    t0 = time.perf_counter()
    st = frames(a.steps, record)
    ctx.sync()
    total = time.perf_counter() - t0
    frame_passes = {}
    for _ in range(min(a.steps, 10)):
        lit_frame(True)
        for k, v in N.timers(ctx).items():
            if k not in passes:
                frame_passes.setdefault(k, []).append(v)
    passes.update(frame_passes)
    # the same frame with the lighting cached (nothing moved): the splat path alone
    t1 = time.perf_counter()
    for _ in range(a.steps):
        N.frame_encode(ctx, scene, fb, W, H, out, fmt)
        N.frame_finish(ctx)
    view_only = (time.perf_counter() - t1) / a.steps
    P, V, T = st.pairs, st.visible, st.tiles
    ls = scene.lighting_stats()
    mpix = W * H * a.steps / total / 1e6
    roof = roofline(passes, P, T, W * H * bpx, ls, a.traffic)
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
        "roofline": roof,
        "map_splat": {"emitter_pairs": ls["lsm_pairs"], "sun_pairs": ls["sun_pairs"],
                      "sum_ms": {"emitter": round(float(np.mean(passes.get("lsm_sum", [0]))), 4),
                                 "sun": round(float(np.mean(passes.get("sun_sum", [0]))), 4)}},
    }
    ctx.free(out)
    scene.destroy()
    ctx.close()
    if a.bounce:
        res["bounce"] = bounce_leg(N, L, scenes, sc)
    # batched and served throughput against the plain per-frame path at this workload (one 1080p view over 1M
    # Gaussians, view only: the batches and requests are unlit), beside the same-views ratios each leg measures
    plain_1080p = W * H / view_only / 1e6
    if a.batched:
        res["batched"] = batched_leg(N, L, scenes)
        res["batched"]["plain_1080p_1m_mpix_s"] = round(plain_1080p, 1)
        res["batched"]["speedup_vs_plain_1080p_1m"] = round(res["batched"]["batched_mpix_s"] / plain_1080p, 3)
    if a.serving:
        res["serving"] = serving_leg(N, scenes, rate=a.rate)
        res["serving"]["plain_1080p_1m_mpix_s"] = round(plain_1080p, 1)
        res["serving"]["saturated_speedup_vs_plain_1080p_1m"] = round(
            res["serving"]["saturated_served_mpix_s"] / plain_1080p, 3)
    if a.storage:
        res["storage_4k"] = storage_leg(N, L, scenes)
    if a.cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(sc, fb)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
