'use strict';
// The reference host's renderer, lighting, environment, LOD and session classes (src/renderer.js, src/lighting.js,
// src/light_shadow.js, src/sun_map.js, src/environment.js, src/lod.js, src/gpu-session.js) over the MI355X addon,
// with their constructors and methods as src/main.js calls them:
//
//   const asset = await loadAsset(dir);
//   const renderer = new SplatRenderer(device, format, shaders, asset, lodNodes);
//   const lighting = new Lighting(device, shaders, asset, renderer.gaussBuf, renderer.partBuf);
//   await lighting.loadFill(dir, shaders);
//   const lod = lodNodes ? new Lod(device, shaders, asset, lighting, renderer, lodNodes) : null;
//   renderer.bindColour(lighting.buffers.colour, lighting.buffers.lo_b, lighting.buffers.emitters);
//   const environment = envDesc ? new Environment(device, shaders, lighting, envDesc, asset) : null;
//   renderer.bindEnv(lighting.buffers.env, environment?.bgBuf);
//   lighting.setEmitters(center);
//   ... per frame (src/main.js:827-848): lighting.camEye / camFwd; lighting.encode(encoder, eye, timers, changed,
//   lightStill); lighting.encodeBounce(encoder, timers, lightStill); lighting.encodeFill(encoder); lod?.encode(...);
//   renderer.encode(encoder, target, frameBytes, viewMode, timers)
//
// The uniform writers run here in JavaScript as they do in the reference (host.js); the bytes they write go to the C
// ABI instead of WebGPU buffers. The lighting's device buffers (colour, bounce) live in the scene the renderer made,
// so `buffers` hands out handles to them; a uniform "buffer" holds the bytes last written to it. The encoder and
// shader-source arguments are accepted and unused (work is enqueued on the context's stream as each call is made).
const fs = require('fs');
const path = require('path');
const addon = require('./splatlight.node');
const H = require('./host.js');
const { loadAsset } = require('./asset.js');

const PASSES = ['clear', 'project', 'radix', 'pairs', 'tilesort', 'ranges', 'splat'];   // src/renderer.js:7-8
const OUT_RGBA32F = 0, OUT_RGBA16F = 1, DISPLAY_RGBA8 = 0, DISPLAY_BGRA8 = 1;
const MAX_MATERIALS = 32;                                // src/lighting.js:37
const BOX_PARTS = ['short_box', 'tall_box', 'floor'];    // parts that get their own material (src/lighting.js:380-392)
const EMIT_THETA = 4;                                    // src/lighting.js:53

// acquireGpuSession (src/gpu-session.js:3-14): resolves to release(), or rejects at once with the reference's
// message when another holder has the session.
async function acquireGpuSession(name) {
  const token = addon.sessionAcquire(name);
  return () => addon.sessionRelease(token);
}

class Device {
  constructor(ordinal = 0) {
    this.ctx = addon.ctxCreate(ordinal);
    this.queue = { onSubmittedWorkDone: async () => addon.sync(this.ctx) };
  }
  // a display target (the canvas texture's stand-in): width * height * 4 bytes in rgba8unorm or bgra8unorm order
  createTarget(width, height, format = 'rgba8unorm') {
    return { buf: addon.bufAlloc(this.ctx, width * height * 4), width, height,
      format: format === 'bgra8unorm' ? DISPLAY_BGRA8 : DISPLAY_RGBA8, read: function () { return addon.bufRead(this.buf); } };
  }
}

// the emitters' part for the LOD's lit-node rule (src/lod.js:108-109): header.emitters.part, else a part named lamp
function emitPartOf(header) {
  if (header.emitters) return header.emitters.part >>> 0;
  const k = header.parts.findIndex((p) => p.name === 'lamp');
  return k >= 0 ? k : 0xffffffff;
}

class SplatRenderer {
  // src/renderer.js:34-60: the scene (records, part table, and for the lighting the asset's materials, leaf nodes and
  // V grids) goes to the device once; lodNodes (nodeGaussians' result) adds the tree for the screen-space LOD, its
  // internal nodes drawn from slot nodeBase on. Particles are not on this path.
  constructor(device, canvasFormat, shaders, asset, lodNodes = null, particleCount = 0, storage = 0) {
    if (particleCount) throw new Error('particles are not on the splat frame path');
    this.device = device;
    this.count = asset.count;
    this.partCount = asset.partCount;
    this.scene = addon.sceneCreate(device.ctx, {
      gaussians: asset.gaussians, parts: asset.raw.parts, materials: asset.raw.materials, nodes: asset.raw.nodes,
      vgrid: asset.raw.vgrid, storage });
    this.gaussBuf = { scene: this.scene, label: 'gaussians' };
    this.partBuf = { scene: this.scene, label: 'parts' };
    this.nodeBase = this.count + 1;          // no emitter records are drawn on this path (setEmitters returns 0)
    this.lodNodeCount = 0;
    if (lodNodes) {
      addon.sceneSetLod(this.scene, { nodes: asset.raw.nodes, kids: asset.raw.kids, nodeGaussians: lodNodes.records,
        emitPart: emitPartOf(asset.header), sizeRule: 1 });
      this.lodNodeCount = lodNodes.N - lodNodes.G;
    }
    this.hdrFormat = canvasFormat === 'rgba32float' ? OUT_RGBA32F : OUT_RGBA16F;   // the HDR target is rgba16float
    this.width = 0;
    this.height = 0;
    this.hdr = null;
    this.envBuf = null;
    this.bgBuf = null;
  }
  // src/renderer.js:264-268: the lit colour, the bounce and the drawn emitters. The lighting's buffers are this scene's
  // own (nothing to bind); host arrays are uploaded (colour: vec4 per Gaussian, bounce: 3 vec4 per Gaussian).
  bindColour(colourBuf, bounceBuf = null, emitterBuf = null) {
    const own = (b) => b && b.scene === this.scene;
    if (colourBuf && !own(colourBuf)) addon.sceneWriteColour(this.scene, colourBuf);
    if (bounceBuf && !own(bounceBuf)) addon.sceneWriteBounce(this.scene, bounceBuf);
    if (emitterBuf && !own(emitterBuf)) throw new Error('bindColour: emitter records are drawn only from the lighting');
  }
  // src/renderer.js:400-404: the environment uniform and the background sky the display pass draws behind the splats
  bindEnv(envBuf, bgBuf) {
    this.envBuf = envBuf || null;
    if (bgBuf !== undefined) this.bgBuf = bgBuf || null;
  }
  resize(width, height) {
    this.width = width;
    this.height = height;
    this.hdr = addon.bufAlloc(this.device.ctx, width * height * (this.hdrFormat === OUT_RGBA32F ? 16 : 8));
  }
  writeParts(partBytes) { addon.sceneWriteParts(this.scene, partBytes); }
  // src/renderer.js:411-524: the frame into the HDR target, then with a target the display pass into it. Resolves to
  // the frame's counts ({slots, visible, pairs, overflow, tiles}); frames may be encoded before earlier ones resolve.
  encode(encoder, target, frameBytes, viewMode, timers) {
    if (!this.hdr) throw new Error('resize() before encode()');
    if (viewMode !== undefined && viewMode !== null) {
      const dv = ArrayBuffer.isView(frameBytes) ? new DataView(frameBytes.buffer, frameBytes.byteOffset, frameBytes.byteLength)
        : new DataView(frameBytes);
      if (dv.getUint32(79 * 4, true) !== viewMode) throw new Error('viewMode differs from RenderFrame.view_mode');
    }
    addon.frameEncode(this.scene, frameBytes, this.width, this.height, this.hdr, this.hdrFormat, !!timers);
    if (!target) return addon.frameFinish(this.device.ctx);
    return addon.frameFinish(this.device.ctx, {
      frameBytes, env: this.envBuf ? this.envBuf.bytes : null, skyBg: this.bgBuf ? this.bgBuf.bytes : null,
      width: this.width, height: this.height, hdrFormat: this.hdrFormat, outFormat: target.format, hdr: this.hdr,
      out: target.buf });
  }
  // captureLinear's bytes (src/main.js:909-919): the HDR target as rgba float16, rows top-down
  readHdr() { return addon.bufRead(this.hdr); }
  timers() { return addon.timersRead(this.device.ctx); }
}

// The emitter light's opacity map (src/light_shadow.js:60-160): which parts re-splat and the LSParams bytes.
class LightShadow {
  constructor(lighting) {
    this.lighting = lighting;
    this.dirtyAll = true;
  }
  setLight() { this.dirtyAll = true; }
  // LSParams for this frame: every part after a light change, else the parts that moved; null without a light yet
  params(movedMask, still) {
    const L = this.lighting;
    if (!L.emitTree || !L.emitLight) return null;
    const dirty = still ? 0 : this.dirtyAll ? 2 ** L.partCount - 1 : movedMask;
    if (!still) this.dirtyAll = false;
    return H.lsParams(L.emitTree, L.emitLight, L.partCount, dirty >>> 0);
  }
}

class Lighting {
  // src/lighting.js:319-527: the asset's counts and materials (plus one per box part, its Gaussians pointed at it and
  // re-uploaded), the emitter set of asset.json `emitters` (none: no lamp), iso_init, and the probe bounce
  constructor(device, sources, asset, gaussBuf, partBuf) {
    this.device = device;
    this.scene = gaussBuf.scene;
    if (!this.scene || partBuf.scene !== this.scene) throw new Error('Lighting needs the renderer\'s gaussBuf and partBuf');
    const h = asset.header, r = asset.raw;
    this.count = asset.count;
    this.partCount = asset.partCount;
    this.lampCount = h.buffers.lamps.count;
    this.materialCount = h.buffers.materials.count + h.parts.filter((p) => BOX_PARTS.includes(p.name)).length;
    this.lightSamplesAxis = h.shading && h.shading.light_samples_axis !== undefined ? h.shading.light_samples_axis : 4;
    this.frame = 0;
    this.gaussianBytes = asset.gaussians;
    this.partRanges = h.parts.map((p) => p.gaussians);
    if (h.emitters) {
      const mf = new Uint32Array(r.materials)[h.emitters.material * 8 + 5];
      this.assetEmitters = { part: h.emitters.part, emission: h.emitters.emission, oneSided: (mf & 4) !== 0 };
    }
    this.emitOneSided = this.assetEmitters ? this.assetEmitters.oneSided : false;
    this.lampOff = !this.assetEmitters;
    this.emitTheta = EMIT_THETA;
    this.emitMode = 1;
    this.emitEdgeSoft = 0.2;
    this.asgSpec = 1;
    this.emitKey = null;
    this.transmit = false;
    this.specSplit = false;
    this.sheen = false;
    // materials: the asset's, then one per box part starting from material 0 (setBoxMaterial)
    this.assetMaterialCount = h.buffers.materials.count;
    this.materialBytes = new ArrayBuffer(Math.max(MAX_MATERIALS, this.assetMaterialCount + 3) * 32);
    new Uint8Array(this.materialBytes).set(new Uint8Array(r.materials));
    this.boxMaterial = {};
    const gu = new Uint32Array(asset.gaussians);
    for (const name of BOX_PARTS) {
      const p = h.parts.findIndex((q) => q.name === name);
      if (p < 0) continue;
      const m = this.assetMaterialCount + Object.keys(this.boxMaterial).length;
      new Uint8Array(this.materialBytes, m * 32, 32).set(new Uint8Array(r.materials, 0, 32));
      this.boxMaterial[name] = m;
      const [g0, g1] = h.parts[p].gaussians;
      for (let g = g0; g < g1; g++) gu[g * 16 + 15] = ((gu[g * 16 + 15] & 0xffff0000) | m) >>> 0;
    }
    addon.sceneWriteMaterials(this.scene, this.materialBytes);
    if (Object.keys(this.boxMaterial).length) addon.sceneWriteGaussians(this.scene, asset.gaussians);
    this.buffers = {
      colour: { scene: this.scene, label: 'colour' }, lo_b: { scene: this.scene, label: 'bounce' },
      emitters: { scene: this.scene, label: 'emitters' }, env: { bytes: null, label: 'environment' },
      uniform: { bytes: new ArrayBuffer(64), label: 'frame uniforms' },
    };
    this.uniform = this.buffers.uniform.bytes;
    this.lightShadow = new LightShadow(this);
    this.movedMask = 0;
    this.partXf = null;
    this.camEye = null;
    this.camFwd = null;
    this.runIsoInit();
    this.initBounce(device, sources, asset);
  }

  runIsoInit() { addon.lightingIsoInit(this.scene); }

  // initBounce (src/lighting.js:723-797): the probe layout, ProbeParams, and the tree (nodes.bin, kids.bin) with
  // node_up's depth lists up to asset.json walk.max_depth
  initBounce(device, sources, asset) {
    const h = asset.header;
    this.probeLayout = H.probeLayout(asset);
    this.nodeCount = asset.raw.nodes.byteLength / 96;
    this.probeCount = H.probeCount(this.probeLayout, this.partCount);
    this.maxDepth = h.walk && h.walk.max_depth !== undefined ? h.walk.max_depth : 6;
    this.probeBase = H.probeParams(this.probeLayout, this.nodeCount, this.partCount);
    this.round = 0;
    this.forceNear = true;
    this.firstGather = true;
    this.nearField = false;
    addon.bounceInit(this.scene, { nodes: asset.raw.nodes, kids: asset.raw.kids, maxDepth: this.maxDepth,
      probeParams: H.probeFrame(this.probeBase, 0, true) });
  }

  // loadFill (src/lighting.js:961-985): the display fill table of onepath.json / onepath.bin in the asset's directory;
  // none (or one for another Gaussian count) leaves encodeFill a no-op
  async loadFill(base, sources) {
    const mp = path.join(base, 'onepath.json');
    if (!fs.existsSync(mp)) return;
    const meta = JSON.parse(fs.readFileSync(mp, 'utf8'));
    if (!meta.fill || meta.G !== this.count) return;
    const bin = fs.readFileSync(path.join(base, 'onepath.bin'));
    const table = new Uint32Array(bin.buffer.slice(bin.byteOffset + meta.fill.offset, bin.byteOffset + meta.fill.offset + meta.fill.bytes));
    addon.sceneSetFill(this.scene, table);
    this.fill = { words: table.length };
  }
  encodeFill(encoder) { if (this.fill) addon.fillEncode(this.scene); }

  // setEmitters (src/lighting.js:612-620): the asset's emissive part moved to `center`, else one dark emitter.
  // Returns the count of emitter records drawn (0: asset emitters are drawn with the scene).
  setEmitters(center) {
    const key = [...center, this.emitTheta, this.emitMode, this.emitOneSided, this.emitEdgeSoft, this.asgSpec].join();
    if (key === this.emitKey) return this.emitterDrawn;
    this.emitKey = key;
    return this.assetEmitters ? this.setAssetEmitters(center) : this.setDarkEmitter(center);
  }

  // writeEmitParams (src/lighting.js:657-673)
  writeEmitParams(n, root, scale, nodeCount, skipPart, oneSided, normal) {
    this.emitParams = H.emitParams({ n, root, scale, theta: this.emitTheta, mode: this.emitMode, nodeCount, skipPart,
      oneSided, normal, asgSpec: this.asgSpec, edgeSoft: this.emitEdgeSoft });
    this.emitterScale = scale;
    this.emitterCount = n;
    this.emitLight = { skipPart, oneSided, normal };
    this.lightShadow.setLight();
  }

  // setDarkEmitter (src/lighting.js:675-688): one black leaf at `center`; count 0 (the shading skips its walk)
  setDarkEmitter(center) {
    this.emitTree = H.buildEmitterTree(H.darkEmitter(center), 1);
    this.emitNodes = H.packEmitNodes(this.emitTree.nodes, this.emitTree.cov);
    this.writeEmitParams(0, this.emitTree.root, 0, this.emitTree.nodes.length / 16, 0xffffffff, 0, [0, 0, 1]);
    this.emitterDrawn = 0;
    return 0;
  }

  // setAssetEmitters (src/lighting.js:690-721): the emissive part's records moved to `center` are the light set;
  // they are drawn with the scene
  setAssetEmitters(center) {
    const { part, emission } = this.assetEmitters;
    const [g0, g1] = this.partRanges[part];
    const set = H.assetEmitters(this.gaussianBytes, g0, g1, emission, center);
    this.emitTree = H.buildEmitterTree(set.records, set.n);
    this.emitNodes = H.packEmitNodes(this.emitTree.nodes, this.emitTree.cov);
    this.writeEmitParams(set.n, this.emitTree.root, set.scale, this.emitTree.nodes.length / 16, part,
      this.emitOneSided ? 1 : 0, set.normal);
    this.emitterDrawn = 0;
    return 0;
  }

  // writeUniform (src/lighting.js:1229-1238)
  writeUniform(eye) {
    this.uniform = H.frameUniforms({ frame: this.frame++, count: this.count, partCount: this.partCount,
      lampCount: this.lampCount, eye, lightSamplesAxis: this.lightSamplesAxis, materialCount: this.materialCount,
      iso: 1 | 2 | (this.asgSpec ? 4 : 0) | (this.transmit ? 8 : 0) | (this.specSplit ? 16 : 0) | (this.sheen ? 32 : 0) });
    this.buffers.uniform.bytes = this.uniform;
  }

  // encode (src/lighting.js:1258-1274): the emitter map of what moved (none while the light is still or without a
  // lamp), shade_emit, then the environment's sun maps and direct light. changed: a part or the light moved.
  encode(encoder, eye, timers, changed = true, lightStill = false) {
    if (!this.emitParams) throw new Error('setEmitters() before encode()');
    this.writeUniform(eye);
    const moved = this.movedMask >>> 0;
    const still = lightStill && !changed && moved === 0;
    const d = { frameUniforms: this.uniform, emitParams: this.emitParams, emitNodes: this.emitNodes,
      lsParams: this.lampOff ? null : this.lightShadow.params(moved, still) };
    if (this.env) Object.assign(d, this.env.directInputs(moved, still));
    this.last = d;
    addon.lightingEncode(this.scene, d, !!timers);
  }

  // isoMoved (src/lighting.js:1006-1016): a moved part turned (its rotation rows differ from last frame's)
  isoMoved(mask) {
    if (!mask) return false;
    const pf = this.partXf;
    if (!pf || this.nearField) return true;
    for (let p = 0; p < Math.min(this.partCount, 32); p++) {
      if (((mask >>> p) & 1) === 0) continue;
      const o = p * 32;
      for (let r = 0; r < 3; r++) for (let j = 0; j < 3; j++) if (pf[o + 4 * r + j] !== pf[o + 12 + 4 * r + j]) return true;
    }
    return false;
  }

  // encodeBounce (src/lighting.js:1020-1103): this frame's ProbeParams words; the transfer rebuilt on the first frame
  // or when a moved part turned; lightStill keeps the probes and only reads the bounce toward the eye
  encodeBounce(encoder, timers, lightStill = false) {
    lightStill = lightStill && !this.forceNear && !this.firstGather;
    const pp = H.probeFrame(this.probeBase, this.round, this.forceNear);
    this.round++;
    const xfer = this.forceNear || this.isoMoved(this.movedMask >>> 0);
    this.forceNear = false;
    this.firstGather = false;
    this.lastProbe = { params: pp, xfer, lightStill };
    addon.bounceEncode(this.scene, pp, { timers: !!timers, xfer, lightStill });
  }
}

// The environment (src/environment.js:63-160) with its sun maps (src/sun_map.js): Env and SkyBg bytes, and each
// frame's sun layer plan. Registers itself as lighting.env.
class Environment {
  constructor(device, sources, lighting, env, asset = null) {
    this.lighting = lighting;
    this.base = env;
    this.skyGain = 1;
    this.lightGain = 1;
    this.sun = env.lights[0] ? H.sunAngles(env.lights[0].dir) : [0, 45];
    this.on = true;
    this.key = null;
    this.bgBuf = { bytes: null, label: 'sky background' };
    this.plan = asset && env.lights.length ? H.sunLayout(asset.header.parts, asset.raw.parts, env.lights) : null;
    this.lightKey = null;
    this.write();
    lighting.env = this;
  }
  // the UI's sun angles and gains; true when anything changed
  set({ on = this.on, sunAz = this.sun[0], sunEl = this.sun[1], skyGain = this.skyGain, lightGain = this.lightGain }) {
    this.on = on; this.sun = [sunAz, sunEl]; this.skyGain = skyGain; this.lightGain = lightGain;
    return this.write();
  }
  write() {
    const key = [this.on, ...this.sun, this.skyGain, this.lightGain].join();
    if (key === this.key) return false;
    this.key = key;
    const { env, bg } = H.envBytes(this.base, this);
    this.lighting.buffers.env.bytes = env;
    this.bgBuf.bytes = bg;
    return true;
  }
  // src/environment.js:156-160: the lights as they shine now
  lights() {
    if (!this.on) return [];
    return this.base.lights.map((L, k) => ({ dir: k === 0 || L.followSun ? H.sunDir(this.sun[0], this.sun[1]) : L.dir,
      std: L.std, irradiance: L.irradiance.map((v) => v * this.lightGain) }));
  }
  // encodeDirect's inputs (src/environment.js:181-189, src/sun_map.js:104-159): Env, and with distant lights the
  // layer plan (still: the maps hold, no layer dirty)
  directInputs(movedMask, still) {
    const out = { env: this.lighting.buffers.env.bytes };
    if (!this.plan) return out;
    const lights = this.lights();
    const key = lights.map((L) => [...L.dir, L.std].join()).join('|') + (this.on ? '' : 'off');
    const changed = key !== this.lightKey;
    if (!still) this.lightKey = key;
    const L = this.lighting;
    const frame = H.sunFrame(this.plan, lights, this.base.lights.map((x) => x.shadows), this.on, L.camEye, L.camFwd,
      movedMask, changed, still);
    out.sunParams = frame.params;
    out.sunLayers = frame.layers;
    return out;
  }
}

// The screen-space LOD (src/lod.js:103-196): the renderer holds the tree (SplatRenderer's lodNodes); lod_select and
// lod_colour run inside the frame whose RenderFrame.lod words are set (threshold, renderer.nodeBase,
// renderer.lodNodeCount), so encode / encodeSelect / encodeColour only record the camera's cut parameters.
class Lod {
  constructor(device, sources, asset, lighting, renderer, info) {
    if (!renderer.lodNodeCount && info.N > info.G) throw new Error('Lod: the renderer was made without lodNodes');
    this.renderer = renderer;
    this.info = info;
    this.colourAll = false;
  }
  encodeSelect(encoder, view, threshold, focal, near) { this.cut = { view, threshold, focal, near }; }
  encodeColour(encoder) {}
  encode(encoder, view, threshold, focal, near) { this.encodeSelect(encoder, view, threshold, focal, near); }
}

module.exports = {
  acquireGpuSession, Device, SplatRenderer, Lighting, Environment, Lod, loadAsset, nodeGaussians: H.nodeGaussians,
  parseEnvironment: H.parseEnvironment, PASSES, OUT_RGBA16F, OUT_RGBA32F, DISPLAY_RGBA8, DISPLAY_BGRA8,
};
