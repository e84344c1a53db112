"""The four measured workloads at their own sizes on the GPU, against the oracle (multithreaded, so the full sizes
finish in seconds) and against themselves:

* hill, 1M Gaussians, 64 emitters, a sun with a 2048^2 map, 1920x1080, fully relit: iso_norm, the emitter map's
  and the sun maps' fixed-point sums (bit-exact), their mips, the transmittances, the lit colours, then the frame:
  keys, order, splat words and pair lists bit-exact, the image within 1/1024 and >= 55 dB;
* 64 views over 8 scenes of 50k-2M Gaussians, 256^2 to 1080p, each scene relit, in one batch replayed from its
  hipGraph: every view bit-identical to the same view rendered alone, and a 1080p view against the oracle;
* hill, 4M Gaussians at 3840x2160 with quaternion, scale and albedo stored as f16 and as bf16: keys, order, splat
  words and pair lists bit-exact against the oracle on the stored (dequantized) records, the image within 1/1024;
* serving over 32 resident scenes (100k-3M Gaussians, 31M in all), Poisson arrivals at 1,000 requests/s in 4 sizes
  on random orbits: every served result bit-identical to the view rendered alone.
"""
import math
import sys
import time

import numpy as np
import pytest

import oracle
from splatlight_amd import lighting as L
from splatlight_amd import native as N
from splatlight_amd import records as R
from splatlight_amd import scenes
from test_frame_gpu import check_integer_artefacts, psnr
from test_lighting_gpu import check_lighting

pytestmark = pytest.mark.gpu

TOL_ABS = 1.0 / 1024.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def frame_artefacts(ctx, scene, fb, W, H, fmt=N.SL_OUT_RGBA32F):
    px = 16 if fmt == N.SL_OUT_RGBA32F else 8
    out = ctx.alloc(W * H * px)
    try:
        N.frame_encode(ctx, scene, fb, W, H, out, fmt)
        st = N.frame_finish(ctx)
        img = ctx.read(out, W * H * px, np.float32 if px == 16 else np.float16).astype(np.float32).reshape(H, W, 4)
        art = {"keys": N.frame_read(ctx, N.ART_KEYS), "order": N.frame_read(ctx, N.ART_ORDER),
               "splats": N.frame_read(ctx, N.ART_SPLATS).reshape(-1, 12), "tiles": N.frame_read(ctx, N.ART_PAIR_TILES),
               "slots": N.frame_read(ctx, N.ART_PAIR_SLOTS), "ranges": N.frame_read(ctx, N.ART_RANGES).reshape(-1, 2)}
    finally:
        ctx.free(out)
    return st, img, art


def check_frame(sc, fb, colour, st, img, art, gaussians=None):
    """Integer artefacts bit-exact and the image within 1/1024 and >= 55 dB of the oracle's float32 blend."""
    gs = sc.gaussians if gaussians is None else gaussians
    view = scenes.Scene(sc.name, gs, sc.parts, sc.materials, sc.emitters, sc.eye, sc.target, sc.fov_deg, sc.width,
                        sc.height, sc.far)
    t = time.perf_counter()
    o_splats, o_keys, o_vis = oracle.project(fb, view, colour=colour)
    _, _, o_order, _ = oracle.sort_keys(o_keys)
    check_integer_artefacts(view, st, art, o_keys, o_order, o_vis, o_splats)
    ref = oracle.composite(fb, o_splats, o_order, o_vis, sc.width, sc.height)
    err = float(np.abs(img - ref).max())
    p = psnr(img, ref)
    log(f"  frame: visible {st.visible}, pairs {st.pairs}; keys, order, splats, pairs bit-exact; image max|err| "
        f"{err:.3g}, psnr {p:.1f} dB (oracle {time.perf_counter() - t:.1f} s)")
    assert err <= TOL_ABS and p >= 55.0


def test_config1_relit_hill_1m_1080p(ctx):
    log("config 1: hill 1M Gaussians, 64 emitters, sun 2048^2, 1920x1080, fully relit")
    t = time.perf_counter()
    sc = scenes.hill(n=1_000_000, seed=1)
    assert sc.emitters.shape[0] == 64 and sc.sun["res"] == 2048 and (sc.width, sc.height) == (1920, 1080)
    scene, inp, o_col = check_lighting(ctx, sc)
    try:
        log(f"  lighting checked ({time.perf_counter() - t:.1f} s)")
        col = scene.read(N.SCENE_COLOUR).reshape(-1, 4)    # the device's lit colours drive the frame check
        fb = sc.frame_bytes()
        st, img, art = frame_artefacts(ctx, scene, fb, sc.width, sc.height)
        check_frame(sc, fb, col, st, img, art)
    finally:
        scene.destroy()


def test_config2_batch_64_views_8_scenes(ctx):
    log("config 2: 64 views over 8 relit scenes (50k-2M), one batch replayed from its hipGraph")
    scs, views = scenes.batch_workload(seed=0)
    assert len(scs) == 8 and len(views) == 64
    handles = []
    try:
        for sc in scs:
            h = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials,
                        nodes=scenes.leaf_nodes(sc.gaussians), vgrid=scenes.vgrid_open(sc.count))
            inp = L.lighting_inputs(sc, dirty_all=True)
            h.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"])
            handles.append(h)
        frames = [scs[si].frame_bytes(width=w, height=hh, eye=eye) for si, w, hh, eye in views]
        big = next(k for k, (_, w, hh, _) in enumerate(views) if (w, hh) == (1920, 1080))
        fmts = [N.SL_OUT_RGBA32F if k == big else N.SL_OUT_RGBA16F for k in range(64)]
        px = [16 if f == N.SL_OUT_RGBA32F else 8 for f in fmts]
        outs = [ctx.alloc(w * hh * p) for (_, w, hh, _), p in zip(views, px)]
        try:
            batch = [(handles[si], fb, w, hh, o, f) for (si, w, hh, _), fb, o, f in zip(views, frames, outs, fmts)]
            for rep in range(2):                    # capture, then replay
                N.batch_encode(ctx, batch, graph=True)
                bst = N.batch_finish(ctx, 64)
            got = [ctx.read(o, w * hh * p, np.uint8) for o, (_, w, hh, _), p in zip(outs, views, px)]
            for k, ((si, w, hh, _), fb, o, f, p) in enumerate(zip(views, frames, outs, fmts, px)):
                N.frame_encode(ctx, handles[si], fb, w, hh, o, f)
                st = N.frame_finish(ctx)
                assert np.array_equal(ctx.read(o, w * hh * p, np.uint8), got[k]), f"view {k}: batch != alone"
                assert (bst[k].visible, bst[k].pairs) == (st.visible, st.pairs), k
            log(f"  64 views: batched images and counts identical to each view alone "
                f"({sum(w * hh for _, w, hh, _ in views) / 1e6:.1f} Mpix, {sum(s.pairs for s in bst)} pairs)")
            si, w, hh, _ = views[big]
            col = handles[si].read(N.SCENE_COLOUR).reshape(-1, 4)
            st, img, art = frame_artefacts(ctx, handles[si], frames[big], w, hh)
            sc = scs[si]
            one = scenes.Scene(sc.name, sc.gaussians, sc.parts, sc.materials, sc.emitters, sc.eye, sc.target,
                               sc.fov_deg, w, hh, sc.far)
            log(f"  view {big}: scene {si} ({sc.count} Gaussians) at 1920x1080 against the oracle")
            check_frame(one, frames[big], col, st, img, art)
        finally:
            for o in outs:
                ctx.free(o)
    finally:
        for h in handles:
            h.destroy()


@pytest.mark.parametrize("storage,name", [(R.STORAGE_F16, "f16"), (R.STORAGE_BF16, "bf16")])
def test_config3_reduced_storage_4m_4k(ctx, storage, name):
    log(f"config 3: hill 4M Gaussians at 3840x2160, {name} storage")
    sc = scenes.hill(n=4_000_000, seed=9, width=3840, height=2160)
    col = np.concatenate([sc.gaussians["albedo"], np.ones((sc.count, 1))], 1).astype(np.float32)
    scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, storage=storage)
    try:
        scene.write_colour(col)
        fb = sc.frame_bytes()
        st, img, art = frame_artefacts(ctx, scene, fb, sc.width, sc.height)
        check_frame(sc, fb, col, st, img, art, gaussians=R.dequantized(sc.gaussians, storage))
    finally:
        scene.destroy()


def test_config4_serving_32_scenes(ctx):
    log("config 4: serving over 32 resident scenes (100k-3M), Poisson 1000 req/s, 4 sizes")
    t = time.perf_counter()
    scs = scenes.serving_workload()
    assert len(scs) == 32
    log(f"  scenes generated ({sum(s.count for s in scs)} Gaussians, {time.perf_counter() - t:.1f} s)")
    handles = []
    try:
        for sc in scs:
            h = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters)
            h.write_colour(np.concatenate([sc.gaussians["albedo"], np.ones((sc.count, 1))], 1).astype(np.float32))
            handles.append(h)
        reqs = scenes.request_stream(scs, rate=1000.0, count=96, seed=3)
        frames = [scs[si].frame_bytes(width=w, height=hh, eye=eye) for _, si, w, hh, eye in reqs]
        outs = [ctx.alloc(w * hh * 8) for _, _, w, hh, _ in reqs]
        handles.append(None)        # the targets are freed with the scenes
        srv = N.Server(ctx, hbm_budget=4 << 30, lanes=2, max_batch_views=32)
        try:
            ids = {}
            t0 = time.perf_counter()
            for k, ((at, si, w, hh, _), fb, o) in enumerate(zip(reqs, frames, outs)):
                while time.perf_counter() - t0 < at:
                    time.sleep(0.0002)
                ids[srv.submit(handles[si], fb, w, hh, o, N.SL_OUT_RGBA16F)] = k
            done = []
            while len(done) < len(reqs) and time.perf_counter() - t0 < 120:
                done += srv.poll(wait_us=20_000)
            assert len(done) == len(reqs) and all(r.status == 0 for r in done)
            lat = sorted(r.done_s - r.submit_s for r in done)
            log(f"  {len(done)} served; batch sizes up to {max(r.batch_views for r in done)}; latency p50 "
                f"{lat[len(lat) // 2] * 1e3:.2f} ms, p99 {lat[int(0.99 * (len(lat) - 1))] * 1e3:.2f} ms")
            served = [ctx.read(o, w * hh * 8, np.uint8) for o, (_, _, w, hh, _) in zip(outs, reqs)]
        finally:
            srv.close()
        for r in done:
            k = ids[r.id]
            _, si, w, hh, _ = reqs[k]
            N.frame_encode(ctx, handles[si], frames[k], w, hh, outs[k], N.SL_OUT_RGBA16F)
            st = N.frame_finish(ctx)
            assert np.array_equal(ctx.read(outs[k], w * hh * 8, np.uint8), served[k]), f"request {k}"
            assert (r.stats.visible, r.stats.pairs) == (st.visible, st.pairs), k
        log("  every served result identical to its view rendered alone")
    finally:
        if handles and handles[-1] is None:
            handles.pop()
            for o in outs:
                ctx.free(o)
        for h in handles:
            h.destroy()
