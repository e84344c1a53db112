'use strict';
// The JavaScript host's parameter logic (splatlight_amd/node/host.js, asset.js) on an asset directory, without the
// addon: loads the asset as loadAsset does and writes the bytes the shim's Lighting / Environment / Lod would hand to
// the C ABI for the first frame at spec.eye looking along spec.fwd. Usage: node node_host.js <asset dir> <out dir>
// <spec.json>.
const fs = require('fs');
const path = require('path');
const H = require('../splatlight_amd/node/host.js');
const { loadAsset } = require('../splatlight_amd/node/asset.js');

async function main() {
  const [dir, out, specFile] = process.argv.slice(2);
  const spec = JSON.parse(fs.readFileSync(specFile, 'utf8'));
  const asset = await loadAsset(dir);
  const h = asset.header, P = asset.partCount;
  const put = (name, bytes) => fs.writeFileSync(path.join(out, name), Buffer.from(bytes.buffer || bytes, bytes.byteOffset || 0, bytes.byteLength));
  // the LOD's node records, from the records as loaded (before the lighting points box parts at their own materials)
  put('node_gaussians.bin', new Uint8Array(H.nodeGaussians(asset).records));
  let tree, ep, ls = null;
  if (h.emitters) {
    const em = h.emitters, [g0, g1] = h.parts[em.part].gaussians;
    const oneSided = (new Uint32Array(asset.raw.materials)[em.material * 8 + 5] & 4) !== 0;
    const c0 = H.assetEmitters(asset.gaussians, g0, g1, em.emission, [0, 0, 0]).centroid;
    const set = H.assetEmitters(asset.gaussians, g0, g1, em.emission, c0);
    tree = H.buildEmitterTree(set.records, set.n);
    ep = H.emitParams({ n: set.n, root: tree.root, scale: set.scale, theta: 4, mode: 1, nodeCount: tree.nodes.length / 16,
      skipPart: em.part, oneSided: oneSided ? 1 : 0, normal: set.normal, asgSpec: 1, edgeSoft: 0.2 });
    ls = H.lsParams(tree, { normal: set.normal, skipPart: em.part, oneSided }, P, 2 ** P - 1);
    put('emitter_records.bin', set.records);
  } else {
    tree = H.buildEmitterTree(H.darkEmitter([0, 0, 0]), 1);
    ep = H.emitParams({ n: 0, root: tree.root, scale: 0, theta: 4, mode: 1, nodeCount: 1, skipPart: 0xffffffff,
      oneSided: 0, normal: [0, 0, 1], asgSpec: 1, edgeSoft: 0.2 });
  }
  put('emit_nodes.bin', H.packEmitNodes(tree.nodes, tree.cov));
  put('EP.bin', new Uint8Array(ep));
  if (ls) put('LS.bin', new Uint8Array(ls));
  const boxes = h.parts.filter((p) => ['short_box', 'tall_box', 'floor'].includes(p.name)).length;
  const samples = h.shading && h.shading.light_samples_axis !== undefined ? h.shading.light_samples_axis : 4;
  put('U.bin', new Uint8Array(H.frameUniforms({ frame: 0, count: asset.count, partCount: P, lampCount: h.buffers.lamps.count,
    eye: spec.eye, lightSamplesAxis: samples, materialCount: h.buffers.materials.count + boxes, iso: 7 })));
  const envDesc = H.parseEnvironment(h.environment);
  if (envDesc) {
    const state = { on: true, sun: envDesc.lights[0] ? H.sunAngles(envDesc.lights[0].dir) : [0, 45], skyGain: 1, lightGain: 1 };
    const { env, bg } = H.envBytes(envDesc, state);
    put('ENV.bin', new Uint8Array(env));
    put('BG.bin', new Uint8Array(bg));
    if (envDesc.lights.length) {
      const plan = H.sunLayout(h.parts, asset.raw.parts, envDesc.lights);
      const lights = envDesc.lights.map((L, k) => ({ dir: k === 0 || L.followSun ? H.sunDir(state.sun[0], state.sun[1]) : L.dir, std: L.std }));
      const fr = H.sunFrame(plan, lights, envDesc.lights.map((L) => L.shadows), true, spec.eye, spec.fwd, 0, true);
      put('SM.bin', new Uint8Array(fr.params));
      put('desc.bin', new Uint8Array(fr.layers));
    }
  }
  const lay = H.probeLayout(asset);
  const base = H.probeParams(lay, asset.raw.nodes.byteLength / 96, P);
  put('PP0.bin', new Uint8Array(H.probeFrame(base, 0, true)));
  put('PP9.bin', new Uint8Array(H.probeFrame(base, 9, false)));
  fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify({ probes: H.probeCount(lay, P), terrain: lay.terrain,
    vols: lay.vols, fingerprint: asset.fingerprint }));
}
main().catch((e) => { console.error(e && e.stack ? e.stack : String(e)); process.exit(1); });
