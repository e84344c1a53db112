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
    optionally "emitters", "materials", "nodes", "vgrid". LOD nodes and particles are not on this path, so a
    non-empty lodNodes or particleCount raises.
    """

    PASSES = ["clear", "project", "radix", "pairs", "tilesort", "ranges", "splat"]

    def __init__(self, device: N.Context, canvasFormat=None, shaders=None, asset=None, lodNodes=None,
                 particleCount: int = 0, storage: int = N.SL_STORAGE_F32):
        if lodNodes is not None or particleCount:
            raise N.SplatlightError(5, "LOD nodes and particles are not on the splat frame path")
        self.device = device
        g = asset["gaussians"]
        if g.dtype != R.GAUSSIAN_DTYPE:
            g = np.frombuffer(np.ascontiguousarray(g).tobytes(), R.GAUSSIAN_DTYPE)
        p = asset["parts"]
        if p.dtype != R.PART_DTYPE:
            p = np.frombuffer(np.ascontiguousarray(p).tobytes(), R.PART_DTYPE)
        self.count = g.shape[0]
        self.partCount = p.shape[0]
        self.scene = N.Scene(device, g, p, emitters=asset.get("emitters"), materials=asset.get("materials"),
                             nodes=asset.get("nodes"), vgrid=asset.get("vgrid"), storage=storage)
        self.splatCount = self.count + 1 + self.scene.n_emit
        self.hdr = None
        self.width = self.height = 0
        self.hdrFormat = N.SL_OUT_RGBA16F if canvasFormat in (None, "rgba16float") else N.SL_OUT_RGBA32F
        self.stats = None

    def bindColour(self, colour, bounce=None, emitters=None):
        """src/renderer.js:264: the lit colour (vec4 per Gaussian) and the bounce (3 vec4 per Gaussian)."""
        if colour is not None:
            self.scene.write_colour(np.asarray(colour, np.float32))
        self.scene.write_bounce(None if bounce is None else np.asarray(bounce, np.float32))

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

    def encode(self, frameBytes: bytes, viewMode: int | None = None, timers: bool = False):
        """src/renderer.js:411: one frame into the HDR target. viewMode, when given, must match the frame's."""
        if self.hdr is None:
            raise N.SplatlightError(1, "resize() before encode()")
        if viewMode is not None and int(np.frombuffer(frameBytes, np.uint32)[79]) != viewMode:
            raise N.SplatlightError(1, "viewMode differs from RenderFrame.view_mode")
        N.frame_encode(self.device, self.scene, frameBytes, self.width, self.height, self.hdr, self.hdrFormat,
                       timers)
        self.stats = N.frame_finish(self.device)
        return self.stats

    def readHdr(self) -> np.ndarray:
        """The HDR target as float32 [H, W, 4], rows top-down (captureLinear's layout, src/main.js:871-920)."""
        if self.hdrFormat == N.SL_OUT_RGBA32F:
            a = self.device.read(self.hdr, self.width * self.height * 16, np.float32)
        else:
            a = self.device.read(self.hdr, self.width * self.height * 8, np.float16).astype(np.float32)
        return a.reshape(self.height, self.width, 4)

    def timers(self) -> dict:
        return N.timers(self.device)

    def destroy(self):
        if self.hdr is not None:
            self.device.free(self.hdr)
            self.hdr = None
        self.scene.destroy()
