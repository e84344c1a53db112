'use strict';
// The gs-runtime asset loader (src/asset.js:1-88) for a Node host: asset.json plus raw little-endian .bin buffers,
// read from a directory instead of fetched. Same checks and messages (checkHeader / checkBuffer), same returned
// shape ({header, raw, gaussians, parts, lamps, count, partCount, fingerprint}), plus `base` (the directory) for
// Lighting.loadFill.
const fs = require('fs');
const path = require('path');

const MAJOR = '0';
// the fields the record layouts assume (src/asset.js:8-16)
const EXPECT = {
  gaussians: { stride: 64, fields: { mu: 0, kappa: 12, quat: 16, scale: 32, part_surface: 44, albedo: 48 } },
  parts: { stride: 128, fields: { xf: 0, xf_prev: 48, bound: 96, root: 112, g_lo: 116, g_hi: 120, flags: 124 } },
  lamps: { stride: 96, fields: { center: 0, kind: 12, e1: 16, radius: 28, e2: 32, normal: 48, radiance: 64, center_prev: 80 } },
  materials: { stride: 32, fields: { flake: 0, phase_samples: 12, vfloor: 16, flags: 20 } },
  nodes: { stride: 96, fields: {} },
  kids: { stride: 4, fields: {} },
  vgrid: { stride: 128, fields: {} },
};
const BUFFERS = ['gaussians', 'parts', 'lamps', 'materials', 'nodes', 'kids', 'vgrid'];

function checkHeader(header) {
  if (header.format !== 'gs-runtime') throw new Error(`Unknown asset format ${header.format}`);
  if (String(header.version).split('.')[0] !== MAJOR) throw new Error(`Unsupported asset version ${header.version}`);
  for (const name of Object.keys(EXPECT)) {
    const want = EXPECT[name], b = header.buffers && header.buffers[name];
    if (!b) throw new Error(`Asset has no ${name} buffer`);
    if (b.stride !== want.stride) throw new Error(`${name} stride ${b.stride}, renderer expects ${want.stride}`);
    for (const f of Object.keys(want.fields)) {
      const got = b.fields && b.fields[f] ? b.fields[f].offset : undefined;
      if (got !== want.fields[f]) throw new Error(`${name}.${f} at ${got}, renderer expects ${want.fields[f]}`);
    }
    if (!Number.isInteger(b.count) || b.count <= 0 || b.count > (1 << 24)) throw new Error(`Bad ${name} count ${b.count}`);
  }
}

function checkBuffer(header, name, bytes) {
  const b = header.buffers[name];
  if (bytes.byteLength !== b.count * b.stride) throw new Error(`${b.file} is ${bytes.byteLength} B, expected ${b.count} x ${b.stride} B`);
  return bytes;
}

// the lamp records as plain objects (src/asset.js:41-49)
function readLamps(header, bytes) {
  const v = new DataView(bytes), out = [];
  const vec = (o) => [0, 1, 2].map((k) => v.getFloat32(o + 4 * k, true));
  for (let i = 0; i < header.buffers.lamps.count; i++) {
    const o = i * 96;
    out.push({ center: vec(o), kind: v.getUint32(o + 12, true), e1: vec(o + 16), radius: v.getFloat32(o + 28, true),
      e2: vec(o + 32), normal: vec(o + 48), radiance: vec(o + 64) });
  }
  return out;
}

// count, part ranges, byte size and sum_i w_i (i * 2654435761 + 1) mod 2^32 over the u32 words (src/asset.js:80-88)
function fingerprint(header, gaussians) {
  const w = new Uint32Array(gaussians);
  let h = 0;
  for (let i = 0; i < w.length; i++) h = (h + Math.imul(w[i], Math.imul(i, 2654435761) + 1)) >>> 0;
  return { count: header.buffers.gaussians.count, parts: header.parts.map((p) => p.gaussians), bytes: gaussians.byteLength, hash: h };
}

const readBytes = (file) => { const b = fs.readFileSync(file); return b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength); };

async function loadAsset(base) {
  const hp = path.join(base, 'asset.json');
  if (!fs.existsSync(hp)) throw new Error(`Could not load ${base}/asset.json (404)`);
  const header = JSON.parse(fs.readFileSync(hp, 'utf8'));
  checkHeader(header);
  const raw = {};
  for (const name of BUFFERS) {
    const file = path.join(base, header.buffers[name].file);
    if (!fs.existsSync(file)) throw new Error(`Could not load ${base}/${header.buffers[name].file} (404)`);
    raw[name] = checkBuffer(header, name, readBytes(file));
  }
  raw.blockers = null;        // optional own-part blockers, 64 B per Gaussian
  const bk = header.blockers;
  if (bk && bk.file && bk.stride === 64 && bk.count === header.buffers.gaussians.count) {
    const file = path.join(base, bk.file);
    if (fs.existsSync(file) && fs.statSync(file).size === bk.count * 64) raw.blockers = readBytes(file);
  }
  return {
    header, raw, gaussians: raw.gaussians, parts: raw.parts, lamps: readLamps(header, raw.lamps),
    count: header.buffers.gaussians.count, partCount: header.buffers.parts.count,
    fingerprint: fingerprint(header, raw.gaussians), base,
  };
}

module.exports = { checkHeader, checkBuffer, readLamps, fingerprint, loadAsset };
