"""SplatRenderer over the HIP library: the method surface of src/renderer.js:34-525 for the splat frame path.

Same method names and argument meaning as the JavaScript class, so a host written against the reference
calls this the same way; the GPU work behind each call is the HIP pipeline of libsplatlight_hip.so.
"""
from __future__ import annotations

import numpy as np

from . import native as N
from . import records as R


class SplatRenderer:
    """src/renderer.js:34 SplatRenderer(device, canvasFormat, shaders, asset, lodNodes, particleCount).

    `device` is a native.Context; `asset` a dict with "gaussians" (64 B records), "parts" (128 B records) and
    optionally "emitters", "materials", "nodes", "vgrid", "kids" and "emitPart". lodNodes: lod.node_gaussians()'s
    result; with it the scene holds the tree (asset "nodes", all of them, and "kids") and frames whose
    RenderFrame.lod is set draw its cut, the internal nodes in the slots from nodeBase on (src/renderer.js:42-46).
    Particles are not on this path, so a non-zero particleCount raises.
    """

    PASSES = ["clear", "project", "radix", "pairs", "tilesort", "ranges", "splat"]

    def __init__(self, device: N.Context, canvasFormat=None, shaders=None, asset=None, lodNodes=None,
                 particleCount: int = 0, storage: int = N.SL_STORAGE_F32):
        if particleCount:
            raise N.SplatlightError(5, "particles are not on the splat frame path")
        self.device = device
        g = asset["gaussians"]
        if g.dtype != R.GAUSSIAN_DTYPE:
            g = np.frombuffer(np.ascontiguousarray(g).tobytes(), R.GAUSSIAN_DTYPE)
        p = asset["parts"]
        if p.dtype != R.PART_DTYPE:
            p = np.frombuffer(np.ascontiguousarray(p).tobytes(), R.PART_DTYPE)
        self.count = g.shape[0]
        self.partCount = p.shape[0]
        nodes = asset.get("nodes")
        leaf_nodes = None if nodes is None else np.asarray(nodes).view(R.NODE_DTYPE).reshape(-1)[:self.count]
        self.scene = N.Scene(device, g, p, emitters=asset.get("emitters"), materials=asset.get("materials"),
                             nodes=leaf_nodes, vgrid=asset.get("vgrid"), storage=storage)
        self.nodeBase = self.count + 1 + self.scene.n_emit
        self.lodNodeCount = 0
        if lodNodes is not None:
            self.scene.set_lod(nodes, asset["kids"], lodNodes["records"], emit_part=asset.get("emitPart", 0xFFFFFFFF))
            self.lodNodeCount = lodNodes["N"] - lodNodes["G"]
        self.splatCount = self.nodeBase + self.lodNodeCount
        self.hdr = None
        self.width = self.height = 0
        self.hdrFormat = N.SL_OUT_RGBA16F if canvasFormat in (None, "rgba16float") else N.SL_OUT_RGBA32F
        self.displayFormat = N.SL_DISPLAY_BGRA8 if canvasFormat == "bgra8unorm" else N.SL_DISPLAY_RGBA8
        self.envBuf = None
        self.skyBgBuf = None
        self.stats = None

    def bindColour(self, colour, bounce=None, emitters=None):
        """src/renderer.js:264: the lit colour (vec4 per Gaussian) and the bounce (3 vec4 per Gaussian)."""
        if colour is not None:
            self.scene.write_colour(np.asarray(colour, np.float32))
        self.scene.write_bounce(None if bounce is None else np.asarray(bounce, np.float32))

    def bindEnv(self, envBuf, skyBgBuf=None):
        """src/renderer.js:400-404: the environment uniform (Env, 416 B) and the background sky (SkyBg, 1520 B) the
        display pass draws behind the splats; skyBgBuf None keeps the one bound before."""
        self.envBuf = None if envBuf is None else bytes(envBuf)
        if skyBgBuf is not None:
            self.skyBgBuf = bytes(skyBgBuf)

    def resize(self, width: int, height: int):
        """src/renderer.js:281: (re)create the linear HDR target."""
        if self.hdr is not None:
            self.device.free(self.hdr)
        self.width, self.height = int(width), int(height)
        px = 16 if self.hdrFormat == N.SL_OUT_RGBA32F else 8
        self.hdr = self.device.alloc(self.width * self.height * px)

    def writeParts(self, partBytes):
        """src/renderer.js:407: the per-frame part table."""
        self.scene.write_parts(np.frombuffer(bytes(partBytes), R.PART_DTYPE))

    def encode(self, frameBytes: bytes, viewMode: int | None = None, timers: bool = False, target: int | None = None):
        """src/renderer.js:411: one frame into the HDR target, then, with a target (a device buffer of width *
        height * 4 bytes), the display pass into it (exposure, tone curve, sRGB; the bound environment's sky
        behind the splats). viewMode, when given, must match the frame's."""
        if self.hdr is None:
            raise N.SplatlightError(1, "resize() before encode()")
        if viewMode is not None and int(np.frombuffer(frameBytes, np.uint32)[79]) != viewMode:
            raise N.SplatlightError(1, "viewMode differs from RenderFrame.view_mode")
        N.frame_encode(self.device, self.scene, frameBytes, self.width, self.height, self.hdr, self.hdrFormat,
                       timers)
        self.stats = N.frame_finish(self.device)
        if target is not None:
            N.display_encode(self.device, frameBytes, self.width, self.height, self.hdr, target, self.hdrFormat,
                             self.displayFormat, self.envBuf, self.skyBgBuf)
            self.device.sync()
        return self.stats

    def captureLinear(self) -> bytes:
        """The HDR target as captureLinear posts it (src/main.js:909-919): width * height rgba float16, rows
        top-down, 8 bytes per pixel, no header (the .f16 capture file). An rgba32float target is rounded to
        float16 (nearest even) first."""
        if self.hdr is None:
            raise N.SplatlightError(1, "resize() before captureLinear()")
        if self.hdrFormat == N.SL_OUT_RGBA16F:
            return self.device.read(self.hdr, self.width * self.height * 8, np.uint8).tobytes()
        a = self.device.read(self.hdr, self.width * self.height * 16, np.float32)
        return a.astype("<f2").tobytes()

    def readHdr(self) -> np.ndarray:
        """The HDR target as float32 [H, W, 4], rows top-down (captureLinear's layout, src/main.js:871-920)."""
        if self.hdrFormat == N.SL_OUT_RGBA32F:
            a = self.device.read(self.hdr, self.width * self.height * 16, np.float32)
        else:
            a = self.device.read(self.hdr, self.width * self.height * 8, np.float16).astype(np.float32)
        return a.reshape(self.height, self.width, 4)

    def readDisplay(self, target: int) -> np.ndarray:
        """A display target as uint8 [H, W, 4] in its byte order (rgba8unorm or bgra8unorm)."""
        return self.device.read(target, self.width * self.height * 4, np.uint8).reshape(self.height, self.width, 4)

    def timers(self) -> dict:
        return N.timers(self.device)

    def destroy(self):
        if self.hdr is not None:
            self.device.free(self.hdr)
            self.hdr = None
        self.scene.destroy()


def read_f16(data, width: int, height: int) -> np.ndarray:
    """A .f16 capture (captureLinear's bytes or a file path) as float32 [height, width, 4], rows top-down."""
    if isinstance(data, str):
        with open(data, "rb") as f:
            data = f.read()
    if len(data) != width * height * 8:
        raise ValueError(f".f16 capture of {len(data)} bytes is not {width}x{height} rgba float16")
    return np.frombuffer(data, "<f2").astype(np.float32).reshape(height, width, 4)
