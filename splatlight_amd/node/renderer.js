'use strict';
// The reference host's renderer / lighting / session classes (src/renderer.js, src/lighting.js,
// src/gpu-session.js) over the MI355X addon: the same constructors and methods, the same byte layouts in
// (RenderFrame 1200 B from writeFrame, the lighting uniforms from writeUniform / writeEmitParams /
// LightShadow.writeParams / Environment.write / the sun layer plan, 64 B Gaussian records, 128 B parts); the GPU
// work goes through the C ABI instead of WebGPU.
const addon = require('./splatlight.node');

const PASSES = ['clear', 'project', 'radix', 'pairs', 'tilesort', 'ranges', 'splat'];   // src/renderer.js:7-8
const OUT_RGBA32F = 0, OUT_RGBA16F = 1;

// acquireGpuSession (src/gpu-session.js:3-14): resolves to release(), or rejects at once with the reference's
// message when another holder has the session.
async function acquireGpuSession(name) {
  const token = addon.sessionAcquire(name);
  return () => addon.sessionRelease(token);
}

class Device {
  constructor(ordinal = 0) { this.ctx = addon.ctxCreate(ordinal); }
}

class SplatRenderer {
  // asset: {gaussians, parts, materials?, nodes?, vgrid?, emitters?, storage?} as ArrayBuffers / typed arrays
  constructor(device, canvasFormat, shaders, asset) {
    this.device = device;
    this.format = canvasFormat === 'rgba32float' ? OUT_RGBA32F : OUT_RGBA16F;   // the target is rgba16float
    this.scene = addon.sceneCreate(device.ctx, asset);
    this.width = 0;
    this.height = 0;
    this.target = null;
  }
  // colour: Float32Array of rgba per Gaussian; with Lighting on the same renderer the colour is already bound
  bindColour(colour) { if (colour) addon.sceneWriteColour(this.scene, colour); }
  resize(width, height) {
    this.width = width;
    this.height = height;
    this.target = addon.bufAlloc(this.device.ctx, width * height * (this.format === OUT_RGBA32F ? 16 : 8));
  }
  writeParts(partBytes) { addon.sceneWriteParts(this.scene, partBytes); }
  // frameBytes: the 1200 B RenderFrame (its view_mode word selects the view); resolves to the frame's counts
  encode(encoder, target, frameBytes, viewMode, timers) {
    addon.frameEncode(this.scene, frameBytes, this.width, this.height, this.target, this.format, !!timers);
    return addon.frameFinish(this.device.ctx);
  }
  readHdr() { return addon.bufRead(this.target); }
  timers() { return addon.timersRead(this.device.ctx); }
}

class Lighting {
  constructor(device, renderer) {
    this.renderer = renderer;
    addon.lightingIsoInit(renderer.scene);
  }
  // uniforms: {frameUniforms, emitParams, emitNodes, lsParams?, env?, sunParams?, sunLayers?} (bytes)
  encode(encoder, uniforms, timers) { addon.lightingEncode(this.renderer.scene, uniforms, !!timers); }
  // initBounce (src/lighting.js:723-840): the asset's tree {nodes, kids, maxDepth} and the ProbeParams bytes the
  // host's layout code writes (probeLayout, D14, pinvD14 unchanged)
  initBounce(tree, probeParams) {
    addon.bounceInit(this.renderer.scene, {nodes: tree.nodes, kids: tree.kids, maxDepth: tree.maxDepth, probeParams});
  }
  // encodeBounce (src/lighting.js:1020-1103) after this frame's encode: probeParams with this frame's words
  // (alpha, force_near, subset, phase); xfer: the transfer is stale (forceNear or a part moved)
  encodeBounce(encoder, timers, lightStill, probeParams, xfer) {
    addon.bounceEncode(this.renderer.scene, probeParams, {timers: !!timers, lightStill: !!lightStill, xfer: !!xfer});
  }
}

module.exports = {acquireGpuSession, Device, SplatRenderer, Lighting, PASSES, OUT_RGBA16F, OUT_RGBA32F};
