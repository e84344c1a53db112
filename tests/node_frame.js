'use strict';
// Renders one lit frame (with the sky and bounce through probes when the tree is given) through the JS shim from
// byte files written by tests/test_node_gpu.py; writes the HDR target and the frame's counts next to them.
const fs = require('fs');
const path = require('path');
const {acquireGpuSession, Device, SplatRenderer, Lighting} = require('../splatlight_amd/node/renderer.js');

async function main() {
  const dir = process.argv[2];
  const rd = (n) => { const p = path.join(dir, n); return fs.existsSync(p) ? new Uint8Array(fs.readFileSync(p)) : null; };
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json')));
  const release = await acquireGpuSession('splatlight-node-test');
  try {
    const device = new Device(0);
    const renderer = new SplatRenderer(device, 'rgba16float', null, {
      gaussians: rd('gaussians.bin'), parts: rd('parts.bin'), materials: rd('materials.bin'),
      nodes: rd('nodes.bin'), emitters: rd('emitters.bin')});
    renderer.resize(meta.width, meta.height);
    const lighting = new Lighting(device, renderer);
    // with a tree (tree_nodes.bin, kids.bin) the bounce runs after the lighting, as encodeBounce follows encode
    const bounce = rd('tree_nodes.bin') !== null;
    if (bounce) {
      renderer.writeParts(rd('parts.bin'));
      lighting.initBounce({nodes: rd('tree_nodes.bin'), kids: rd('kids.bin'), maxDepth: meta.maxDepth}, rd('PP0.bin'));
    }
    lighting.encode(null, {frameUniforms: rd('U.bin'), emitParams: rd('EP.bin'), emitNodes: rd('emit_nodes.bin'),
                           lsParams: rd('LS.bin'), env: rd('ENV.bin'), sunParams: rd('SM.bin'),
                           sunLayers: rd('desc.bin')}, false);
    if (bounce) lighting.encodeBounce(null, true, false, rd('PP.bin'), true);
    const stats = await renderer.encode(null, null, rd('frame.bin'), 3, true);
    fs.writeFileSync(path.join(dir, 'hdr.bin'), Buffer.from(renderer.readHdr()));
    fs.writeFileSync(path.join(dir, 'stats.json'), JSON.stringify({stats, timers: renderer.timers()}));
  } finally {
    release();
  }
}
main().catch((e) => { console.error(String(e)); process.exit(1); });
