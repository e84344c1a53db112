'use strict';
// Frames encoded back to back without awaiting the last one's counts (the reference's capture loop, src/main.js:
// 886-906, encodes the lighting and the next frame before the previous frame's readback resolves), then the same
// frames one at a time, each awaited; writes both sets of counts and whether the last frame's HDR and display bytes
// agree. Usage: node node_pipeline.js <asset dir> <spec.json> <out.json>; spec: {width, height, frames: [{eye, frame}]}.
const fs = require('fs');
const path = require('path');
const { Device, SplatRenderer, Lighting, loadAsset } = require('../splatlight_amd/node/renderer.js');

async function main() {
  const [dir, specFile, outFile] = process.argv.slice(2);
  const spec = JSON.parse(fs.readFileSync(specFile, 'utf8'));
  const asset = await loadAsset(dir);
  const device = new Device(0);
  const renderer = new SplatRenderer(device, 'rgba16float', null, asset);
  const lighting = new Lighting(device, null, asset, renderer.gaussBuf, renderer.partBuf);
  renderer.bindColour(lighting.buffers.colour, lighting.buffers.lo_b, lighting.buffers.emitters);
  renderer.resize(spec.width, spec.height);
  const target = device.createTarget(spec.width, spec.height);
  lighting.setEmitters([0, 0, 1]);
  const frames = spec.frames.map((f) => ({ eye: f.eye, bytes: fs.readFileSync(path.join(specFile, '..', f.frame)) }));
  const pending = [];
  for (const f of frames) {
    lighting.encode(null, f.eye, false, true, false);
    pending.push(renderer.encode(null, target, f.bytes, 3, false));
  }
  const piped = await Promise.all(pending);
  const pipedHdr = Buffer.from(renderer.readHdr()), pipedDisp = Buffer.from(target.read());
  const serial = [];
  for (const f of frames) {
    lighting.encode(null, f.eye, false, true, false);
    serial.push(await renderer.encode(null, target, f.bytes, 3, false));
  }
  const serialHdr = Buffer.from(renderer.readHdr()), serialDisp = Buffer.from(target.read());
  fs.writeFileSync(outFile, JSON.stringify({ piped, serial, hdrEqual: pipedHdr.equals(serialHdr),
    displayEqual: pipedDisp.equals(serialDisp) }));
}
main().catch((e) => { console.error(e && e.stack ? e.stack : String(e)); process.exit(1); });
