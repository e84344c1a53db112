'use strict';
// A JavaScript host driving the shim the way src/main.js does (construction at :117-152, each frame at :827-848), on
// an asset directory: loadAsset, nodeGaussians, SplatRenderer, Lighting, loadFill, Lod, bindColour, Environment,
// bindEnv, setEmitters; then per frame the camera for the sun cascades, lighting.encode, encodeBounce, encodeFill,
// lod.encode and renderer.encode into a display target. Writes what each call handed to the C ABI and what came back
// (HDR and display bytes, counts) for the Python side to replay. Usage: node node_dropin.js <asset dir> <out dir>
// <spec.json>; spec: {width, height, lodPx, frames: [{eye, fwd, view, frame (RenderFrame file), changed,
// lightStill}]}.
const fs = require('fs');
const path = require('path');
const { Device, SplatRenderer, Lighting, Environment, Lod, loadAsset, nodeGaussians, parseEnvironment } =
  require('../splatlight_amd/node/renderer.js');
const H = require('../splatlight_amd/node/host.js');

async function main() {
  const [dir, out, specFile] = process.argv.slice(2);
  const spec = JSON.parse(fs.readFileSync(specFile, 'utf8'));
  const put = (name, b) => {
    if (!b) return;
    const u8 = ArrayBuffer.isView(b) ? new Uint8Array(b.buffer, b.byteOffset, b.byteLength) : new Uint8Array(b);
    fs.writeFileSync(path.join(out, name), Buffer.from(u8));
  };
  const asset = await loadAsset(dir);
  const header = asset.header;
  const lodNodes = spec.lodPx > 0 ? nodeGaussians(asset) : null;
  const device = new Device(0);
  const renderer = new SplatRenderer(device, 'rgba16float', null, asset, lodNodes);
  const lighting = new Lighting(device, null, asset, renderer.gaussBuf, renderer.partBuf);
  await lighting.loadFill(dir, null);
  const lod = lodNodes ? new Lod(device, null, asset, lighting, renderer, lodNodes) : null;
  renderer.bindColour(lighting.buffers.colour, lighting.buffers.lo_b, lighting.buffers.emitters);
  const envDesc = parseEnvironment(header.environment);
  const environment = envDesc ? new Environment(device, null, lighting, envDesc, asset) : null;
  renderer.bindEnv(lighting.buffers.env, environment ? environment.bgBuf : undefined);
  renderer.resize(spec.width, spec.height);
  const target = device.createTarget(spec.width, spec.height);
  // the lamp at its rest centroid (the emitter part unmoved), else the origin
  let center = [0, 0, 0];
  if (header.emitters) {
    const [g0, g1] = header.parts[header.emitters.part].gaussians;
    center = H.assetEmitters(asset.gaussians, g0, g1, header.emitters.emission, [0, 0, 0]).centroid;
  }
  const drawn = lighting.setEmitters(center);
  put('gaussians.bin', asset.gaussians);               // after the box parts were pointed at their own materials
  put('materials.bin', lighting.materialBytes);
  if (lodNodes) put('node_gaussians.bin', lodNodes.records);
  put('PP_init.bin', H.probeFrame(lighting.probeBase, 0, true));
  const stats = [];
  for (let k = 0; k < spec.frames.length; k++) {
    const f = spec.frames[k];
    const frameBytes = fs.readFileSync(path.join(dir, '..', f.frame));
    lighting.camEye = f.eye;
    lighting.camFwd = f.fwd;
    lighting.encode(null, f.eye, true, f.changed, f.lightStill);
    const d = lighting.last;
    put(`U${k}.bin`, d.frameUniforms); put(`EP${k}.bin`, d.emitParams); put(`emit_nodes${k}.bin`, d.emitNodes);
    put(`LS${k}.bin`, d.lsParams); put(`ENV${k}.bin`, d.env); put(`SM${k}.bin`, d.sunParams);
    put(`desc${k}.bin`, d.sunLayers);
    lighting.encodeBounce(null, true, f.lightStill);
    put(`PP${k}.bin`, lighting.lastProbe.params);
    lighting.encodeFill(null);
    if (lod) lod.encode(null, f.view, spec.lodPx, spec.height / (2 * Math.tan(spec.fovy / 2)), 0.02);
    const st = await renderer.encode(null, target, frameBytes, 3, true);
    stats.push({ st, xfer: lighting.lastProbe.xfer, lightStill: lighting.lastProbe.lightStill });
    put(`hdr${k}.bin`, renderer.readHdr());
    put(`display${k}.bin`, target.read());
  }
  if (environment) put('BG.bin', environment.bgBuf.bytes);
  fs.writeFileSync(path.join(out, 'result.json'), JSON.stringify({ stats, drawn, timers: renderer.timers(),
    materialCount: lighting.materialCount, boxes: lighting.boxMaterial, fill: !!lighting.fill,
    nodeBase: renderer.nodeBase, lodNodeCount: renderer.lodNodeCount }));
}
main().catch((e) => { console.error(e && e.stack ? e.stack : String(e)); process.exit(1); });
