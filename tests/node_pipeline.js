'use strict';
// Frames encoded back to back without awaiting the last one's counts (the reference's capture loop, src/main.js:
// 886-906, encodes the lighting and the next frame before the previous frame's readback resolves), then the same
// frames one at a time, each awaited; writes both sets of counts and the last target's bytes of each run.
const fs = require('fs');
const path = require('path');
const {Device, SplatRenderer, Lighting} = require('../splatlight_amd/node/renderer.js');

async function main() {
  const dir = process.argv[2];
  const rd = (n) => { const p = path.join(dir, n); return fs.existsSync(p) ? new Uint8Array(fs.readFileSync(p)) : null; };
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json')));
  const device = new Device(0);
  const renderer = new SplatRenderer(device, 'rgba16float', null, {
    gaussians: rd('gaussians.bin'), parts: rd('parts.bin'), materials: rd('materials.bin'),
    nodes: rd('nodes.bin'), emitters: rd('emitters.bin')});
  renderer.resize(meta.width, meta.height);
  const lighting = new Lighting(device, renderer);
  const uniforms = {frameUniforms: rd('U.bin'), emitParams: rd('EP.bin'), emitNodes: rd('emit_nodes.bin'),
                    lsParams: rd('LS.bin'), env: rd('ENV.bin'), sunParams: rd('SM.bin'), sunLayers: rd('desc.bin')};
  const frames = [];
  for (let k = 0; k < meta.frames; k++) frames.push(rd(`frame${k}.bin`));
  // pipelined: nothing awaited until every frame is encoded
  const pending = [];
  for (const fb of frames) {
    lighting.encode(null, uniforms, false);
    pending.push(renderer.encode(null, null, fb, 3, false));
  }
  const piped = await Promise.all(pending);
  const pipedHdr = Buffer.from(renderer.readHdr());
  // one at a time
  const serial = [];
  for (const fb of frames) {
    lighting.encode(null, uniforms, false);
    serial.push(await renderer.encode(null, null, fb, 3, false));
  }
  const serialHdr = Buffer.from(renderer.readHdr());
  fs.writeFileSync(path.join(dir, 'piped.json'), JSON.stringify({piped, serial,
                                                                 hdrEqual: pipedHdr.equals(serialHdr)}));
}
main().catch((e) => { console.error(String(e)); process.exit(1); });
