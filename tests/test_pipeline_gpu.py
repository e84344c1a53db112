"""Relit frames issued the way a render loop issues them: frame i+1's lighting is encoded before frame i is finished,
so it runs beside frame i's composite (which reads nothing the lighting writes). Every frame must equal the same
lighting and frame rendered on their own, bit for bit; lighting timers read after a frame's finish come from a
completed lighting; and a frame whose pairs overflow after the next lighting was encoded is reported, not redrawn
with the next frame's colours.
"""
import math
import sys

import numpy as np
import pytest

from splatlight_amd import lighting as L
from splatlight_amd import native as N
from splatlight_amd import scenes

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make(ctx, sc):
    return N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials,
                   nodes=scenes.leaf_nodes(sc.gaussians), vgrid=scenes.vgrid_open(sc.count))


def orbit(sc, k):
    """A camera on a circle around the scene's target, one step of 0.3 rad per k."""
    t, e = np.asarray(sc.target, np.float64), np.asarray(sc.eye, np.float64)
    d = e - t
    a = 0.3 * k
    x = d[0] * math.cos(a) - d[1] * math.sin(a)
    y = d[0] * math.sin(a) + d[1] * math.cos(a)
    return (float(t[0] + x), float(t[1] + y), float(e[2]))


def light(scene, inp, timers=False):
    scene.light(inp["U"], inp["EP"], inp["emit_nodes"], inp["LS"], inp["ENV"], inp["SM"], inp["desc"],
                timers=timers)


def test_pipelined_frames_equal_frames_alone(ctx):
    sc = scenes.hill(n=30_000, seed=3, width=320, height=180, shadow_res=256)
    W, H = sc.width, sc.height
    eyes = [orbit(sc, k) for k in range(4)]
    inps = [L.lighting_inputs(sc, eye=e, frame=k) for k, e in enumerate(eyes)]
    fbs = [sc.frame_bytes(eye=e) for e in eyes]
    scene = make(ctx, sc)
    outs = [ctx.alloc(W * H * 16) for _ in eyes]
    try:
        alone = []
        for inp, fb, o in zip(inps, fbs, outs):
            light(scene, inp)
            N.frame_encode(ctx, scene, fb, W, H, o)
            N.frame_finish(ctx)
            alone.append(ctx.read(o, W * H * 16, np.float32).copy())
        # the render loop: lighting k + 1 encoded before frame k's finish
        light(scene, inps[0], timers=True)
        N.frame_encode(ctx, scene, fbs[0], W, H, outs[0])
        piped = []
        for k in range(len(eyes)):
            if k + 1 < len(eyes):
                light(scene, inps[k + 1], timers=True)
            st = N.frame_finish(ctx)
            tm = N.timers(ctx)
            assert tm.get("sun_sum", 0.0) > 0.0 and tm.get("lsm_sum", 0.0) > 0.0 and tm.get("lighting", 0.0) > 0.0
            piped.append(ctx.read(outs[k], W * H * 16, np.float32).copy())
            log(f"  frame {k}: visible {st.visible}, pairs {st.pairs}")
            if k + 1 < len(eyes):
                N.frame_encode(ctx, scene, fbs[k + 1], W, H, outs[k + 1])
        for k, (a, b) in enumerate(zip(alone, piped)):
            assert np.array_equal(a, b), f"frame {k} differs from the frame alone"
        assert not np.array_equal(alone[0], alone[1])      # the frames do differ
    finally:
        for o in outs:
            ctx.free(o)
        scene.destroy()


def test_pipelined_bounce_frames_equal_frames_alone(ctx):
    # the next frame's lighting and probe bounce encoded before this frame's finish: the bounce joins the
    # composite on the context's stream, so every frame equals the same lighting, bounce and frame issued one
    # after another (each run on a fresh scene and bounce, whose probe rounds then match)
    from splatlight_amd import bounce as BN
    from splatlight_amd import tree
    sc = scenes.hill(n=8_000, seed=4, width=320, height=240, shadow_res=256)
    W, H = sc.width, sc.height
    t = tree.scene_tree(sc)
    eyes = [orbit(sc, k) for k in range(3)]
    inps = [L.lighting_inputs(sc, eye=e, frame=k) for k, e in enumerate(eyes)]
    fbs = [sc.frame_bytes(eye=e, bounce_gain=1.0) for e in eyes]

    def run(piped):
        scene = N.Scene(ctx, sc.gaussians, sc.parts, emitters=sc.emitters, materials=sc.materials,
                        nodes=t.nodes[:sc.count], vgrid=scenes.vgrid_open(sc.count))
        b = BN.Bounce(scene, sc, t)
        outs = [ctx.alloc(W * H * 16) for _ in eyes]
        imgs = []
        try:
            if not piped:
                for inp, fb, o in zip(inps, fbs, outs):
                    light(scene, inp)
                    b.encode()
                    N.frame_encode(ctx, scene, fb, W, H, o)
                    N.frame_finish(ctx)
                    imgs.append(ctx.read(o, W * H * 16, np.float32).copy())
                return imgs, b.n_probes
            light(scene, inps[0])
            b.encode()
            N.frame_encode(ctx, scene, fbs[0], W, H, outs[0])
            for k in range(len(eyes)):
                if k + 1 < len(eyes):
                    light(scene, inps[k + 1])
                    b.encode()
                N.frame_finish(ctx)
                imgs.append(ctx.read(outs[k], W * H * 16, np.float32).copy())
                if k + 1 < len(eyes):
                    N.frame_encode(ctx, scene, fbs[k + 1], W, H, outs[k + 1])
            return imgs, b.n_probes
        finally:
            for o in outs:
                ctx.free(o)
            scene.destroy()

    alone, n_probes = run(False)
    piped, _ = run(True)
    for k, (a, p) in enumerate(zip(alone, piped)):
        assert np.array_equal(a, p), f"bounce frame {k} differs from the frame issued alone"
    assert not np.array_equal(alone[0], alone[1])
    log(f"  {len(eyes)} bounce frames, {sc.count} Gaussians, {n_probes} probes")


def test_pipelined_overflow_is_reported_then_fits():
    # big splats: the first frame's pairs overflow a fresh context's pair buffers
    sc = scenes.cornell(n=2_000, seed=5, width=640, height=480, scale_mu=-1.0)
    W, H = sc.width, sc.height
    inp = L.lighting_inputs(sc)
    fb = sc.frame_bytes()
    ctx = N.Context(0)
    scene = make(ctx, sc)
    out = ctx.alloc(W * H * 16)
    try:
        light(scene, inp)
        N.frame_encode(ctx, scene, fb, W, H, out)
        light(scene, inp)                                 # the next frame's lighting, before the finish
        with pytest.raises(N.SplatlightError) as ei:
            N.frame_finish(ctx)
        assert "not re-run" in str(ei.value)
        # the buffers were grown: the next frame fits, and equals a frame rendered on its own
        N.frame_encode(ctx, scene, fb, W, H, out)
        st = N.frame_finish(ctx)
        assert st.overflow == 0 and st.pairs > 8 * sc.count + 4096
        piped = ctx.read(out, W * H * 16, np.float32).copy()
        light(scene, inp)
        N.frame_encode(ctx, scene, fb, W, H, out)
        N.frame_finish(ctx)
        assert np.array_equal(piped, ctx.read(out, W * H * 16, np.float32))
        log(f"  pairs {st.pairs} for {sc.count} Gaussians")
    finally:
        ctx.free(out)
        scene.destroy()
        ctx.close()
