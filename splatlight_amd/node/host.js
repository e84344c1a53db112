'use strict';
// The reference host's parameter logic in JavaScript, for the shim in renderer.js: the emitter tree
// (src/lighting.js:55-169), the light shadow's light frame (src/light_shadow.js:113-140, 169-180), the environment's
// parse and uniforms (src/environment.js:23-61, 111-153), the sun layer plan (src/sun_map.js:24-159), the probe
// layout and ProbeParams (src/lighting.js:170-316, 723-797, 1020-1034) and the LOD's node records (src/lod.js:9-110).
// Everything is computed in float64, as the reference's JavaScript does, and stored as float32 where it stores it,
// so the byte layouts the addon hands to the C ABI are the reference's.

// ---- emitter tree ------------------------------------------------------------------------------------------------

// Ein(3) = int_0^3 (1 - e^-x) / x dx by the reference's 3000-step midpoint sum (src/lighting.js:55-56)
const EIN3 = (() => {
  let acc = 0;
  for (let i = 0; i < 3000; i++) {
    const x = (i + 0.5) * 0.001;
    acc += (1 - Math.exp(-x)) / x * 0.001;
  }
  return acc;
})();
const ASG_WIDEN = EIN3 / (1 - Math.exp(-3));          // src/lighting.js:67
const PAIRS = [[0, 0], [1, 1], [2, 2], [0, 1], [0, 2], [1, 2]];   // symmetric 3x3 as [xx yy zz xy xz yz]

function symDet(m) {
  const a = m[0], d = m[1], f = m[2], b = m[3], c = m[4], e = m[5];
  return a * (d * f - e * e) - b * (b * f - e * c) + c * (b * e - d * c);
}
function symInv(m) {
  const a = m[0], d = m[1], f = m[2], b = m[3], c = m[4], e = m[5], det = symDet(m);
  return [(d * f - e * e) / det, (a * f - c * c) / det, (a * d - b * b) / det,
    (c * e - b * f) / det, (b * e - c * d) / det, (b * c - a * e) / det];
}
function symQuad(m, w) {
  const x = w[0], y = w[1], z = w[2];
  return m[0] * x * x + m[1] * y * y + m[2] * z * z + 2 * (m[3] * x * y + m[4] * x * z + m[5] * y * z);
}
// principal axes by 12 cyclic Jacobi sweeps (src/lighting.js:79-93); returns the eigenvectors
function symAxes(m) {
  const A = [[m[0], m[3], m[4]], [m[3], m[1], m[5]], [m[4], m[5], m[2]]];
  const V = [[1, 0, 0], [0, 1, 0], [0, 0, 1]];
  const rot = (M, p, q, c, s, rows) => {
    for (let k = 0; k < 3; k++) {
      if (rows) { const x = M[p][k], y = M[q][k]; M[p][k] = c * x - s * y; M[q][k] = s * x + c * y; }
      else { const x = M[k][p], y = M[k][q]; M[k][p] = c * x - s * y; M[k][q] = s * x + c * y; }
    }
  };
  for (let sweep = 0; sweep < 12; sweep++) {
    for (const [p, q] of [[0, 1], [0, 2], [1, 2]]) {
      if (Math.abs(A[p][q]) < 1e-30) continue;
      const th = 0.5 * Math.atan2(2 * A[p][q], A[q][q] - A[p][p]);
      const c = Math.cos(th), s = Math.sin(th);
      rot(A, p, q, c, s, false);
      rot(A, p, q, c, s, true);
      rot(V, p, q, c, s, false);
    }
  }
  return [0, 1, 2].map((k) => [V[0][k], V[1][k], V[2][k]]);
}
const lum = (c) => 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];

// buildEmitterTree (src/lighting.js:99-169): n emitter records (16 floats each, Float32Array) -> EmitNode records
// (16 floats, u32 words 11 and 15), the root and each node's ASG covariance (8 floats). Median split along the
// longest axis; a leaf's M = (2 pi Ein(3))^2 det C C^-1, a parent's matches its children's intensity along its own
// principal axes.
function buildEmitterTree(rec, n) {
  if (!(n >= 1)) throw new Error(`emitter tree needs at least one emitter (got ${n})`);
  const nodes = new Float32Array((2 * n - 1) * 16), words = new Uint32Array(nodes.buffer);
  const cov = new Float32Array((2 * n - 1) * 8);
  const made = [];
  let next = 0;
  const at = (i, k) => rec[i * 16 + k];
  const leaf = (i) => {
    const x = at(i, 4), y = at(i, 5), z = at(i, 6), w = at(i, 7);
    const R = [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]];
    const s2 = [at(i, 8) ** 2, at(i, 9) ** 2, at(i, 10) ** 2];
    const C = PAIRS.map(([r, k]) => R[r][0] * R[k][0] * s2[0] + R[r][1] * R[k][1] * s2[1] + R[r][2] * R[k][2] * s2[2]);
    const g = (2 * Math.PI * EIN3) ** 2 * symDet(C);
    return {
      mu: [at(i, 0), at(i, 1), at(i, 2)], C, M: symInv(C).map((v) => v * g),
      mass: at(i, 3) * (2 * Math.PI) ** 1.5 * Math.sqrt(symDet(C)), Cs: C.map((v) => v * ASG_WIDEN),
      Le: [at(i, 12), at(i, 13), at(i, 14)], radius: 3 * Math.max(at(i, 8), at(i, 9), at(i, 10)), lo: 0xffffffff, hi: i,
    };
  };
  const emit = (nd) => {
    const id = next++, o = id * 16;
    nodes.set([nd.mu[0], nd.mu[1], nd.mu[2], nd.radius, nd.M[0], nd.M[1], nd.M[2], 0, nd.M[3], nd.M[4], nd.M[5]], o);
    words[o + 11] = nd.lo;
    nodes.set(nd.Le, o + 12);
    words[o + 15] = nd.hi;
    cov.set([nd.Cs[0], nd.Cs[1], nd.Cs[2], 0, nd.Cs[3], nd.Cs[4], nd.Cs[5], 0], id * 8);
    made[id] = nd;
    return id;
  };
  // children's mass-weighted sum of a per-child term (0 + first + second, the reference's reduce order)
  const msum = (kids, term) => { let s = 0; for (const ch of kids) s += term(ch); return s; };
  const build = (idx) => {
    if (idx.length === 1) return emit(leaf(idx[0]));
    const lo = [Infinity, Infinity, Infinity], hi = [-Infinity, -Infinity, -Infinity];
    for (const i of idx) for (let k = 0; k < 3; k++) { lo[k] = Math.min(lo[k], at(i, k)); hi[k] = Math.max(hi[k], at(i, k)); }
    let ax = 0;
    for (let k = 1; k < 3; k++) if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
    const order = idx.slice().sort((p, q) => at(p, ax) - at(q, ax));   // stable: ties keep their order
    const half = order.length >> 1;
    const ia = build(order.slice(0, half)), ib = build(order.slice(half));
    const kids = [made[ia], made[ib]];
    const m = kids[0].mass + kids[1].mass;
    const mu = [0, 1, 2].map((k) => msum(kids, (ch) => ch.mass * ch.mu[k]) / m);
    const moment = (key) => PAIRS.map(([r, k], j) =>
      msum(kids, (ch) => ch.mass * (ch[key][j] + (ch.mu[r] - mu[r]) * (ch.mu[k] - mu[k]))) / m);
    const C = moment('C'), Cs = moment('Cs');
    const axes = symAxes(C);
    const reach = (ch, w) => Math.sqrt(Math.max(symQuad(ch.M, w), 0));
    const tot = axes.map((w) => msum(kids, (ch) => lum(ch.Le) * reach(ch, w)));
    const wsum = kids.map((ch) => { let s = 0; for (const w of axes) s += reach(ch, w); return s; });
    const Le = [0, 1, 2].map((c) => (0 + kids[0].Le[c] * wsum[0] + kids[1].Le[c] * wsum[1]) / (0 + wsum[0] + wsum[1]));
    const L = Math.max(lum(Le), 1e-12);
    const M = [0, 0, 0, 0, 0, 0];
    axes.forEach((w, k) => {
      const lam = (tot[k] / L) ** 2;
      PAIRS.forEach(([r, c], j) => { M[j] += lam * w[r] * w[c]; });
    });
    let radius = -Infinity;
    for (const ch of kids) radius = Math.max(radius, Math.hypot(ch.mu[0] - mu[0], ch.mu[1] - mu[1], ch.mu[2] - mu[2]) + ch.radius);
    return emit({ mu, C, Cs, M, mass: m, Le, radius, lo: ia, hi: ib });
  };
  const root = build(Array.from({ length: n }, (_, i) => i));
  return { nodes, root, cov };
}

// packEmitNodes (src/lighting.js:61-65): 24 floats (96 B) per node, the record then its ASG covariance
function packEmitNodes(nodes, cov) {
  const n = nodes.length / 16, out = new Float32Array(n * 24);
  for (let i = 0; i < n; i++) {
    out.set(nodes.subarray(i * 16, i * 16 + 16), i * 24);
    out.set(cov.subarray(i * 8, i * 8 + 8), i * 24 + 16);
  }
  return out;
}

// setDarkEmitter's record (src/lighting.js:675-688): one black emitter at `center` so the light cut has a root
function darkEmitter(center) {
  const rec = new Float32Array(16);
  rec.set(center, 0); rec[3] = 1; rec.set([0, 0, 0, 1], 4); rec.set([0.01, 0.01, 0.01], 8);
  return rec;
}

// setAssetEmitters' light set (src/lighting.js:690-721): the emissive part's records [g0, g1) moved so their rest
// centroid sits at `center`, with Le = emission; the first record's short axis as the emitting normal; EP.scale the
// panel's area (the grid of emitter centres along the first record's y and z axes, each extent stretched by
// cols / (cols - 1) over its distinct columns at 1/1000 of the span) over the emitters' summed footprint
function assetEmitters(gaussianBytes, g0, g1, emission, center) {
  const n = g1 - g0;
  const src = new Float32Array(gaussianBytes, g0 * 64, n * 16);
  const rec = new Float32Array(src);
  const c0 = [0, 1, 2].map((k) => { let a = 0; for (let i = 0; i < n; i++) a += src[i * 16 + k]; return a / n; });
  for (let i = 0; i < n; i++) {
    for (let k = 0; k < 3; k++) rec[i * 16 + k] += center[k] - c0[k];
    rec.set(emission, i * 16 + 12);
  }
  const x = rec[4], y = rec[5], z = rec[6], w = rec[7];
  const normal = [1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)];
  const axes = [[2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
    [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]];
  const extent = axes.map((a) => {
    const vals = [];
    for (let i = 0; i < n; i++) vals.push(a[0] * rec[i * 16] + a[1] * rec[i * 16 + 1] + a[2] * rec[i * 16 + 2]);
    let lo = Infinity, hi = -Infinity;
    for (const v of vals) { lo = Math.min(lo, v); hi = Math.max(hi, v); }
    const cols = new Set(vals.map((v) => Math.round((v - lo) / (hi - lo + 1e-9) * 1000))).size;
    return (hi - lo) * cols / Math.max(cols - 1, 1);
  });
  let foot = 0;
  for (let i = 0; i < n; i++) foot += 2 * Math.PI * rec[i * 16 + 9] * rec[i * 16 + 10] * EIN3;
  return { records: rec, n, normal, scale: extent[0] * extent[1] / foot, centroid: c0 };
}

// EmitParams (64 B) as writeEmitParams packs it (src/lighting.js:657-673)
function emitParams(o) {
  const ep = new ArrayBuffer(64), u = new Uint32Array(ep), f = new Float32Array(ep);
  u[0] = o.n; u[1] = o.root; f[2] = o.scale; f[3] = o.theta; u[4] = o.mode; u[5] = o.nodeCount;
  u[6] = o.skipPart; u[7] = o.oneSided;
  f.set(o.normal, 8); u[11] = o.asgSpec << 3;
  f[12] = o.edgeSoft;
  return ep;
}

// FrameUniforms (64 B) as writeUniform packs it (src/lighting.js:1229-1238); THETA 8, ALPHA_MIN 1/16, ALPHA_PART 0.5,
// STRATUM_STEP 7 (src/lighting.js:212)
function frameUniforms(o) {
  const buf = new ArrayBuffer(64), f = new Float32Array(buf), u = new Uint32Array(buf);
  u[0] = o.frame; u[1] = o.count; u[2] = o.partCount; u[3] = o.lampCount;
  f[4] = 8; f[5] = 1 / 16; f[6] = 0.5; u[7] = 1;
  f.set(o.eye, 8); u[11] = 7;
  u[12] = o.lightSamplesAxis; u[13] = o.materialCount;
  u[15] = o.iso;
  return buf;
}

// ---- light shadow (emitter map) ------------------------------------------------------------------------------------

const LS_RES = 512, LS_PARAMS_BYTES = 336;
// method B / C fields LightShadow.writeParams also writes (src/light_shadow.js:18): kept as the reference's
const LS_OPT = { tiles: 32, tileCap: 1024, cutTheta: 0.25, refine: 8, budget: 256, maxLobes: 1 << 17, slices: 10,
  d0: 0.25, d1: 3.5, combine: 2 };

// ascending eigenvalues of a symmetric 3x3 in closed form (src/light_shadow.js:169-180)
function eig3(m) {
  const a = m[0], d = m[1], f = m[2], b = m[3], c = m[4], e = m[5];
  const p1 = b * b + c * c + e * e, q = (a + d + f) / 3;
  const p2 = (a - q) ** 2 + (d - q) ** 2 + (f - q) ** 2 + 2 * p1, p = Math.sqrt(p2 / 6);
  if (p < 1e-30) return [q, q, q];
  const B = [(a - q) / p, (d - q) / p, (f - q) / p, b / p, c / p, e / p];
  const r = (B[0] * (B[1] * B[2] - B[5] * B[5]) - B[3] * (B[3] * B[2] - B[5] * B[4]) + B[4] * (B[3] * B[5] - B[1] * B[4])) / 2;
  const phi = Math.acos(Math.min(1, Math.max(-1, r))) / 3;
  const e1 = q + 2 * p * Math.cos(phi), e3 = q + 2 * p * Math.cos(phi + 2 * Math.PI / 3);
  return [e3, 3 * q - e1 - e3, e1];
}

// mip chain of a res^2 map: [(offset, side)] and the texels of one layer
function mipLevels(res) {
  const levels = [];
  let off = 0;
  for (let k = res; k >= 1; k >>= 1) { levels.push([off, k]); off += k * k; }
  return { levels, texels: off };
}

// the emitter map's resolution: halved while the parts' tiles overflow a dispatch (src/light_shadow.js:89-92)
function lsRes(partCount) {
  let r = LS_RES;
  while (r > 64 && Math.ceil(partCount * r * r / 64) > 65535) r >>= 1;
  return r;
}

// LSParams (336 B) as LightShadow.writeParams packs it (src/light_shadow.js:113-140): the light frame from the tree
// root's mean and ASG covariance and the emitting normal. Word 52 on (method B's per-part surface counts) is 1.
function lsParams(tree, light, partCount, dirty) {
  const buf = new ArrayBuffer(LS_PARAMS_BYTES), f = new Float32Array(buf), u = new Uint32Array(buf);
  const res = lsRes(partCount), { levels, texels } = mipLevels(res);
  const c = tree.cov.subarray(tree.root * 8, tree.root * 8 + 8);
  const center = [tree.nodes[tree.root * 16], tree.nodes[tree.root * 16 + 1], tree.nodes[tree.root * 16 + 2]];
  const nl = Math.hypot(light.normal[0], light.normal[1], light.normal[2]);
  const n = light.normal.map((v) => v / nl);
  const a = Math.abs(n[0]) < 0.9 ? [1, 0, 0] : [0, 1, 0];
  let t1 = [n[1] * a[2] - n[2] * a[1], n[2] * a[0] - n[0] * a[2], n[0] * a[1] - n[1] * a[0]];
  const tl = Math.hypot(t1[0], t1[1], t1[2]);
  t1 = t1.map((v) => v / tl);
  const t2 = [n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]];
  const ev = eig3([c[0], c[1], c[2], c[4], c[5], c[6]]);
  f.set(center, 0); u[3] = res;
  f.set(n, 4); u[7] = levels.length;
  f.set(t1, 8); u[11] = texels;
  f.set(t2, 12); u[15] = dirty >>> 0;
  f[16] = Math.sqrt(0.5 * (ev[1] + ev[2])); u[17] = light.skipPart; u[18] = light.oneSided ? 1 : 0; u[19] = partCount;
  f[20] = LS_OPT.cutTheta; u[21] = LS_OPT.tiles; u[22] = LS_OPT.tileCap; f[23] = LS_OPT.refine;
  f.set([c[0], c[1], c[2]], 24); u[27] = LS_OPT.maxLobes;
  f.set([c[4], c[5], c[6]], 28); u[31] = LS_OPT.budget;
  u[32] = LS_OPT.slices; f[33] = LS_OPT.d0; f[34] = Math.log(LS_OPT.d1 / LS_OPT.d0) / LS_OPT.slices; u[35] = LS_OPT.combine;
  levels.forEach(([o], k) => { u[36 + k] = o; });
  f.fill(1, 52, 84);
  return buf;
}

// ---- environment ---------------------------------------------------------------------------------------------------

// a ?? b of newer JavaScript (Node 12 has no nullish operator)
const or = (a, b) => (a === undefined || a === null ? b : a);
const DEG = Math.PI / 180, ENV_BYTES = 416, SKY_BG_BYTES = (81 + 2 + 8 + 2 + 2) * 16;
const unit = (v) => { const l = Math.hypot(v[0], v[1], v[2]) || 1; return [v[0] / l, v[1] / l, v[2] / l]; };

// parseEnvironment (src/environment.js:23-38): asset.json's `environment` block -> the parsed form
function parseEnvironment(env) {
  if (!env) return null;
  const sky = or(env.sky, {}), hz = sky.horizon, g = sky.glow;
  return {
    sh: or(sky.sh_rgb, []),
    bg: sky.background_sh_rgb ? sky.background_sh_rgb.slice(0, 81) : null,
    horizon: hz ? { rgb: hz.rgb.slice(0, 8), along: or(hz.along_deg, 25) * DEG, across: or(hz.across_deg, 8) * DEG } : null,
    ground: or(sky.ground_rgb, null),
    glow: g ? { irradiance: g.irradiance_rgb, std: or(g.angular_std_deg, 12) * DEG, followSun: or(g.follows_sun, true),
      dir: g.dir ? unit(g.dir) : null } : null,
    lights: or(env.lights, []).slice(0, 4).map((L) => ({
      dir: unit(L.dir), irradiance: or(L.irradiance_rgb, [1, 1, 1]), std: or(L.angular_std_deg, 0.27) * DEG,
      displayStd: or(L.display_std_deg, 0) * DEG, shadows: or(L.casts_shadows, true), followSun: or(L.follows_sun, false) })),
  };
}
// sunDir / sunAngles (src/environment.js:54-61): azimuth from +y toward -x, degrees
function sunDir(az, el) {
  const a = az * DEG, e = el * DEG;
  return [-Math.sin(a) * Math.cos(e), Math.cos(a) * Math.cos(e), Math.sin(e)];
}
function sunAngles(d) {
  return [Math.atan2(-d[0], d[1]) / DEG, Math.asin(Math.max(-1, Math.min(1, d[2]))) / DEG];
}

// Env (416 B) and SkyBg (1,520 B) as Environment.write packs them (src/environment.js:111-153)
function envBytes(base, state) {
  const env = new ArrayBuffer(ENV_BYTES), f = new Float32Array(env), u = new Uint32Array(env);
  const lightDir = (L, k) => (k === 0 || L.followSun ? sunDir(state.sun[0], state.sun[1]) : L.dir);
  if (state.on) {
    base.sh.forEach((c, k) => f.set(c.map((v) => v * state.skyGain), 4 * k));
    base.lights.forEach((L, k) => {
      f.set([...lightDir(L, k), L.std], 64 + 8 * k);
      f.set([...L.irradiance.map((v) => v * state.lightGain), L.shadows ? 1 : 0], 68 + 8 * k);
    });
    u[96] = base.lights.length;
    u[97] = base.sh.length ? 1 : 0;
  }
  const bg = new ArrayBuffer(SKY_BG_BYTES), bf = new Float32Array(bg), bu = new Uint32Array(bg);
  const sky = (c) => c.map((v) => v * state.skyGain);
  if (state.on && base.bg) {
    base.bg.forEach((c, k) => bf.set(sky(c), 4 * k));
    bu[4 * 81] = base.bg.length;
  }
  base.lights.forEach((L, k) => { bf[4 * 82 + k] = L.displayStd; });
  const hz = base.horizon, ground = base.ground;
  if (state.on && base.bg && hz && ground) {
    hz.rgb.forEach((c, k) => bf.set(sky(c), 4 * (83 + k)));
    bu[4 * 81 + 1] = hz.rgb.length;
    bf.set([...sky(ground), Math.sin(1 * DEG)], 4 * 91);
    bf.set([1 / (2 * hz.along ** 2), 1 / (2 * hz.across ** 2)], 4 * 92);
  }
  const glow = base.glow;
  if (state.on && base.bg && glow) {
    bf.set([...glow.irradiance.map((v) => v * state.skyGain), glow.std], 4 * 93);
    bf.set(glow.followSun || !glow.dir ? sunDir(state.sun[0], state.sun[1]) : glow.dir, 4 * 94);
  }
  return { env, bg };
}

// ---- sun maps ------------------------------------------------------------------------------------------------------

const SUN_PARAMS_BYTES = 432, SUN_CASCADES = [8, 32], SUN_CAP = 120e6;
const cross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
const dot = (a, b) => a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
// common.wgsl basis_k: a unit vector perpendicular to w
function basisK(w) {
  const c = cross(Math.abs(w[2]) >= 0.999 ? [1, 0, 0] : [0, 0, 1], w), l = Math.hypot(c[0], c[1], c[2]);
  return c.map((v) => v / l);
}

// The layer plan of SunMap's constructor (src/sun_map.js:24-60): per part (its rest bound from parts.bin) and light,
// the cascades of a large static part and a map fitted to the part; the resolution halved under the 120 MB cap.
function sunLayout(headerParts, partBytes, lights, res = 512) {
  const K = Math.min(lights.length, 4), P = Math.min(headerParts.length, 32);
  const bound = [];
  for (let p = 0; p < P; p++) bound.push(Array.from(new Float32Array(partBytes, p * 128 + 96, 4)));
  const ext = Math.max(...bound.map((b) => Math.hypot(b[0], b[1], b[2]) + b[3]));
  const sigMax = Math.max(...lights.slice(0, K).map((L) => L.std));
  const desc = [], partRange = [];
  for (let p = 0; p < P; p++) {
    const first = desc.length, big = bound[p][3] > SUN_CASCADES[0] && headerParts[p].mobility === 'static';
    const g = headerParts[p].gaussians;
    if (g && g[1] > g[0]) {
      for (let k = 0; k < K; k++) {
        if (big) SUN_CASCADES.forEach((h) => { if (h < bound[p][3] * 1.5) desc.push({ p, k, kind: 1, half: h, center: null }); });
        desc.push({ p, k, kind: 0 });
      }
    }
    partRange.push([first, desc.length - first]);
  }
  while (res > 64 && (desc.length * res * res * 12 > SUN_CAP || desc.length * res * res * 16 * 4 / 3 > SUN_CAP)) res >>= 1;
  const { levels, texels } = mipLevels(res);
  return { K, P, bound, margin: 3 * sigMax * 2 * ext, desc, partRange, res, levels, texels };
}

// One frame of SunMap.encode's host side (src/sun_map.js:104-159): each layer's dirty flag (the lights changed, its
// part moved, or a cascade drifted a quarter of its size from its centre ahead of the camera), the LayerDesc rows
// and SunParams. hold: the lighting is still, so nothing is re-evaluated and no layer is dirty (the maps hold; the
// reference skips the pass, and a lookup over unchanged maps gives the same T).
function sunFrame(plan, lights, casts, on, eye, fwd, movedMask, lightsChanged, hold = false) {
  let any = false;
  const dirty = plan.desc.map((d) => {
    if (hold) return false;
    let dd = lightsChanged || (d.p < 32 && ((movedMask >>> d.p) & 1) === 1);
    if (d.kind === 1) {
      const L = lights[d.k] ? lights[d.k].dir : [0, 0, 1];
      const t1 = basisK(L), t2 = cross(L, t1);
      const at = eye && fwd ? eye.map((v, i) => v + fwd[i] * d.half * 0.75) : plan.bound[d.p].slice(0, 3);
      const c = [dot(at, t1), dot(at, t2)];
      if (!d.center || lightsChanged || Math.hypot(c[0] - d.center[0], c[1] - d.center[1]) > d.half / 4) {
        d.center = c;
        dd = true;
      }
    }
    any = any || dd;
    return dd;
  });
  const layers = new ArrayBuffer(plan.desc.length * 32), du = new Uint32Array(layers), df = new Float32Array(layers);
  plan.desc.forEach((d, l) => {
    du.set([d.p, d.k, d.kind, dirty[l] ? 1 : 0], l * 8);
    if (d.kind === 1 && d.center) df.set([d.center[0], d.center[1], d.half], l * 8 + 4);
  });
  const params = new ArrayBuffer(SUN_PARAMS_BYTES), u = new Uint32Array(params), f = new Float32Array(params);
  // words 6-7 and 11 (the reference's splat range list and node count for its dispatch) stay 0: nothing reads them
  u.set([plan.desc.length, plan.res, plan.levels.length, plan.texels, plan.K, plan.P, 0, 0], 0);
  f.set([plan.margin, 0, 0], 8);
  for (let k = 0; k < 4; k++) {
    const L = lights[k];
    f.set(L ? [...L.dir, casts[k] && on ? L.std : -1] : [0, 0, 1, -1], 12 + 4 * k);
  }
  plan.partRange.forEach(([a, n], p) => { u[28 + 2 * p] = a; u[29 + 2 * p] = n; });
  plan.levels.forEach(([o], k) => { u[92 + k] = o; });
  return { params, layers, any };
}

// ---- probes --------------------------------------------------------------------------------------------------------

const PPK = 8, PROBES_AXIS = 8, PROBE_THETA = 4, TMAX = 6, NEAR_R = 0.3, FULL_ROUNDS = 8;
const PROBE_PARAMS_BYTES = 832 + 16 + 64 + 4096 + 528 + 4096;
// D14: the 6 axes (vertical ones 1e-4 off the pole) and the 8 diagonals (src/lighting.js:186-187)
const D14 = [[1, 0, 0], [0, 1, 0], [1e-4, 0, 1], [-1, 0, 0], [0, -1, 0], [1e-4, 0, -1]];
for (const x of [1, -1]) for (const y of [1, -1]) for (const z of [1, -1]) D14.push([x / Math.sqrt(3), y / Math.sqrt(3), z / Math.sqrt(3)]);
const sh2 = (d) => {
  const x = d[0], y = d[1], z = d[2];
  return [0.282095, 0.488603 * y, 0.488603 * z, 0.488603 * x, 1.092548 * x * y, 1.092548 * y * z,
    0.315392 * (3 * z * z - 1), 1.092548 * x * z, 0.546274 * (x * x - y * y)];
};
// the 9 x 14 least-squares SH2 fit on D14 (src/lighting.js:190-206): Gauss-Jordan with partial pivoting on
// [A^T A | I], then (A^T A)^-1 A^T
function pinvD14() {
  const A = D14.map(sh2);
  const M = [];
  for (let i = 0; i < 9; i++) {
    const row = new Array(18).fill(0);
    for (let j = 0; j < 9; j++) { let s = 0; for (const r of A) s += r[i] * r[j]; row[j] = s; }
    row[9 + i] = 1;
    M.push(row);
  }
  for (let c = 0; c < 9; c++) {
    let piv = c;
    for (let r = c + 1; r < 9; r++) if (Math.abs(M[r][c]) > Math.abs(M[piv][c])) piv = r;
    const t = M[c]; M[c] = M[piv]; M[piv] = t;
    const d = M[c][c];
    for (let j = 0; j < 18; j++) M[c][j] /= d;
    for (let r = 0; r < 9; r++) {
      if (r === c) continue;
      const f = M[r][c];
      for (let j = 0; j < 18; j++) M[r][j] -= f * M[c][j];
    }
  }
  const out = [];
  for (let m = 0; m < 9; m++) {
    for (const row of A) { let s = 0; for (let j = 0; j < 9; j++) s += M[m][9 + j] * row[j]; out.push(s); }
  }
  return out;   // 9 x 14, row-major
}

const isStaticPart = (p, pu, q) => (p.mobility ? p.mobility !== 'movable' : (pu[q * 32 + 31] & 1) !== 0);

// probeLayout (src/lighting.js:220-290): the room lattice inset 0.125 m, or terrain columns whose probes stand at
// fixed heights over the highest terrain Gaussian of each column (empty columns filled from their neighbours, ring
// by ring), plus the probe volumes on static objects
function probeLayout(asset) {
  const h = asset.header, pr = or(h.probes, {});
  const terrain = pr.layout === 'terrain' || (!h.room && !!h.bounds && pr.layout !== 'room');
  const box = h.room || (h.bounds ? { x: [h.bounds.lo[0], h.bounds.hi[0]], y: [h.bounds.lo[1], h.bounds.hi[1]],
    z: [h.bounds.lo[2], h.bounds.hi[2]] } : null);
  if (!box) throw new Error('asset.json has neither room nor bounds (probe lattice)');
  const ax = [box.x, box.y, box.z];
  const tmax = h.environment ? Math.max(TMAX, Math.hypot(...ax.map(([a, b]) => b - a))) : TMAX;
  if (!terrain) {
    const n = PROBES_AXIS;
    const half = ax.map(([a, b]) => (b - a) / 2 - 0.125);
    const lo = ax.map(([a, b], k) => (a + b) / 2 - half[k]), step = half.map((v) => 2 * v / (n - 1));
    return { terrain: false, dims: [n, n, n], lo, step, levels: Array.from({ length: n }, (_, k) => lo[2] + k * step[2]),
      ground: new Float32Array(n * n), tmax, vols: [] };
  }
  const dims = or(pr.dims, [16, 16, 5]).slice();
  if (dims[0] * dims[1] > 1024 || dims[2] > 16) throw new Error('probe terrain layout: at most 1024 columns and 16 levels');
  const levels = or(pr.heights_m, [0.3, 1, 3, 8, 20]);
  dims[2] = levels.length;
  const step = [(ax[0][1] - ax[0][0]) / dims[0], (ax[1][1] - ax[1][0]) / dims[1]];
  const lo = [ax[0][0] + step[0] / 2, ax[1][0] + step[1] / 2, 0];
  const parts = h.parts, gf = new Float32Array(asset.gaussians);
  let tp = or(pr.terrain_parts, []).map((nm) => parts.findIndex((p) => p.name === nm)).filter((i) => i >= 0);
  if (!tp.length) {           // the static part with the widest footprint (a sample of at most ~4k centres)
    let best = -1, bestA = -1;
    parts.forEach((p, i) => {
      if (p.mobility === 'movable') return;
      const [g0, g1] = p.gaussians;
      let x0 = Infinity, x1 = -Infinity, y0 = Infinity, y1 = -Infinity;
      for (let g = g0; g < g1; g += Math.max(1, (g1 - g0) >> 12)) {
        const x = gf[g * 16], y = gf[g * 16 + 1];
        x0 = Math.min(x0, x); x1 = Math.max(x1, x); y0 = Math.min(y0, y); y1 = Math.max(y1, y);
      }
      const A = (x1 - x0) * (y1 - y0);
      if (A > bestA) { bestA = A; best = i; }
    });
    tp = best >= 0 ? [best] : [];
  }
  const W = dims[1], ground = new Float32Array(dims[0] * W).fill(-Infinity);
  for (const i of tp) {
    const [g0, g1] = parts[i].gaussians;
    for (let g = g0; g < g1; g++) {
      const ci = Math.floor((gf[g * 16] - ax[0][0]) / step[0]), cj = Math.floor((gf[g * 16 + 1] - ax[1][0]) / step[1]);
      if (ci < 0 || cj < 0 || ci >= dims[0] || cj >= W) continue;
      ground[ci * W + cj] = Math.max(ground[ci * W + cj], gf[g * 16 + 2]);
    }
  }
  for (let ring = 0; ring < dims[0] + W && ground.some((v) => v === -Infinity); ring++) {
    const prev = ground.slice();
    for (let i = 0; i < dims[0]; i++) {
      for (let j = 0; j < W; j++) {
        if (prev[i * W + j] !== -Infinity) continue;
        let s = 0, c = 0;
        for (const [di, dj] of [[1, 0], [-1, 0], [0, 1], [0, -1]]) {
          const a = i + di, b = j + dj;
          if (a < 0 || b < 0 || a >= dims[0] || b >= W || prev[a * W + b] === -Infinity) continue;
          s += prev[a * W + b];
          c++;
        }
        if (c) ground[i * W + j] = s / c;
      }
    }
  }
  for (let q = 0; q < ground.length; q++) if (ground[q] === -Infinity) ground[q] = ax[2][0];
  return { terrain: true, dims, lo, step: [step[0], step[1], 0], levels, ground, tmax, vols: probeVolumes(asset, tp, ax) };
}

// probeVolumes (src/lighting.js:295-316): asset.json probes.volumes by part name, else every static non-terrain part
// whose bounding radius is 1 m to a quarter of the scene's half-diagonal: ~1.2 m spacing, 2-6 per axis
function probeVolumes(asset, terrainParts, ax) {
  const h = asset.header, pr = or(h.probes, {});
  const pf = new Float32Array(asset.parts), pu = new Uint32Array(asset.parts);
  const out = [];
  if (pr.volumes) {
    for (const v of pr.volumes) {
      const q = h.parts.findIndex((p) => p.name === v.part);
      if (q >= 0 && q < 32) out.push({ part: q, dims: v.dims });
    }
    return out;
  }
  const halfDiag = 0.5 * Math.hypot(...ax.map(([a, b]) => b - a));
  h.parts.forEach((p, q) => {
    if (q >= 32 || terrainParts.includes(q) || !isStaticPart(p, pu, q)) return;
    const R = pf[q * 32 + 27];
    if (R < 1 || R > 0.25 * halfDiag) return;
    const n = Math.max(2, Math.min(6, Math.round(R / 1.2)));
    out.push({ part: q, dims: [n, n, n] });
  });
  return out;
}

const probeCount = (lay, partCount) =>
  lay.dims[0] * lay.dims[1] * lay.dims[2] + PPK * partCount + lay.vols.reduce((a, v) => a + v.dims[0] * v.dims[1] * v.dims[2], 0);

// ProbeParams (9,632 B) as initBounce writes it (src/lighting.js:776-797)
function probeParams(lay, nodeCount, partCount) {
  const buf = new ArrayBuffer(PROBE_PARAMS_BYTES), f = new Float32Array(buf), u = new Uint32Array(buf);
  f.set(lay.lo, 0); u[3] = lay.dims[0];
  f.set(lay.step, 4); u[7] = nodeCount;
  f.set([0.5 * Math.hypot(lay.step[0], lay.step[1], lay.step[2]), PROBE_THETA, lay.tmax, NEAR_R], 8);
  u.set([lay.dims[0], lay.dims[1], lay.dims[2], lay.terrain ? 1 : 0], 208);
  f.set(lay.levels, 212);
  f.set(lay.ground, 228);
  let off = PPK * partCount, vol = 0;
  for (const v of lay.vols) {
    u.set([off, ...v.dims], 1252 + 4 * v.part);
    const n = v.dims[0] * v.dims[1] * v.dims[2];
    off += n;
    vol += n;
  }
  u[1380] = PPK * partCount;
  u[1381] = 0;
  f.set([1, 0, 0.02, 0.95], 12);
  u[16] = 1;
  u[18] = PPK * partCount + vol;
  u[19] = 1;
  D14.forEach((d, k) => f.set([d[0], d[1], d[2], 0], 20 + 4 * k));
  f.set(pinvD14(), 76);
  return buf;
}

// the words encodeBounce rewrites each frame (src/lighting.js:1020-1034): alpha 1 (each gather taken as is),
// force_near, the subset and phase (every ray while round < FULL_ROUNDS), the near-field reflection words off
function probeFrame(base, round, forceNear) {
  const buf = base.slice(0), f = new Float32Array(buf), u = new Uint32Array(buf);
  f[12] = 1;
  u[13] = forceNear ? 1 : 0;
  u[205] = 0;
  const sub = 1;     // ?probesubset: 1, every ray each frame
  u[16] = round < FULL_ROUNDS ? 1 : sub;
  u[17] = round % u[16];
  u[204] = 0;
  u[206] = 8;
  u[207] = 0;
  return buf;
}

// ---- LOD: the tree's internal nodes as Gaussian records ------------------------------------------------------------

// symmetric 3x3 eigen-decomposition by Jacobi sweeps until the off-diagonal vanishes (src/lod.js:9-26): values
// ascending with their vectors
function jacobiEig(A0) {
  const a = A0.map((r) => r.slice()), V = [[1, 0, 0], [0, 1, 0], [0, 0, 1]];
  for (let it = 0; it < 20; it++) {
    let off = 0;
    for (let p = 0; p < 3; p++) for (let q = p + 1; q < 3; q++) off += a[p][q] * a[p][q];
    if (off < 1e-30) break;
    for (let p = 0; p < 3; p++) {
      for (let q = p + 1; q < 3; q++) {
        if (Math.abs(a[p][q]) < 1e-30) continue;
        const th = 0.5 * Math.atan2(2 * a[p][q], a[q][q] - a[p][p]), c = Math.cos(th), s = Math.sin(th);
        for (let k = 0; k < 3; k++) { const x = a[k][p], y = a[k][q]; a[k][p] = c * x - s * y; a[k][q] = s * x + c * y; }
        for (let k = 0; k < 3; k++) { const x = a[p][k], y = a[q][k]; a[p][k] = c * x - s * y; a[q][k] = s * x + c * y; }
        for (let k = 0; k < 3; k++) { const x = V[k][p], y = V[k][q]; V[k][p] = c * x - s * y; V[k][q] = s * x + c * y; }
      }
    }
  }
  const idx = [0, 1, 2].sort((i, j) => a[i][i] - a[j][j]);
  return { values: idx.map((i) => a[i][i]), vectors: idx.map((i) => [V[0][i], V[1][i], V[2][i]]) };
}
// rotation with columns c[0..2] -> quaternion (x, y, z, w) (src/lod.js:29-41)
function quatFromColumns(c) {
  const m00 = c[0][0], m11 = c[1][1], m22 = c[2][2];
  const m10 = c[0][1], m20 = c[0][2], m01 = c[1][0], m21 = c[1][2], m02 = c[2][0], m12 = c[2][1];
  const tr = m00 + m11 + m22;
  let x, y, z, w, s;
  if (tr > 0) { s = 2 * Math.sqrt(tr + 1); w = s / 4; x = (m21 - m12) / s; y = (m02 - m20) / s; z = (m10 - m01) / s; }
  else if (m00 > m11 && m00 > m22) { s = 2 * Math.sqrt(1 + m00 - m11 - m22); x = s / 4; w = (m21 - m12) / s; y = (m01 + m10) / s; z = (m02 + m20) / s; }
  else if (m11 > m22) { s = 2 * Math.sqrt(1 + m11 - m00 - m22); y = s / 4; w = (m02 - m20) / s; x = (m01 + m10) / s; z = (m12 + m21) / s; }
  else { s = 2 * Math.sqrt(1 + m22 - m00 - m11); z = s / 4; w = (m10 - m01) / s; x = (m02 + m20) / s; y = (m12 + m21) / s; }
  const l = Math.hypot(x, y, z, w) || 1;
  return [x / l, y / l, z / l, w / l];
}

// nodeGaussians (src/lod.js:46-110) without node V grids: each internal node of the tree (nodes.bin: N 96 B records,
// leaves [0, G) first; kids.bin) as a 64 B Gaussian record: the node's mean and kappa, its covariance (the inverse
// of the baked C^-1) as scales and a rotation, its part, the mass-weighted albedo of its children and the first
// child's material. Parents after children (deepest depth first).
function nodeGaussians(asset) {
  const G = asset.count, raw = asset.raw, N = raw.nodes.byteLength / 96;
  const nf = new Float32Array(raw.nodes), nu = new Uint32Array(raw.nodes), kids = new Uint32Array(raw.kids);
  const gf = new Float32Array(asset.gaussians), gu = new Uint32Array(asset.gaussians);
  const mass = new Float64Array(N), alb = new Float64Array(N * 3), mat = new Uint32Array(N);
  const cinv = (n) => {
    const o = n * 24;
    return [[nf[o + 4], nf[o + 8], nf[o + 9]], [nf[o + 8], nf[o + 5], nf[o + 10]], [nf[o + 9], nf[o + 10], nf[o + 6]]];
  };
  const det3 = (M) => M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
    + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  for (let n = 0; n < G; n++) {
    const o = n * 24, cd = [nf[o + 4], nf[o + 5], nf[o + 6]], co = [nf[o + 8], nf[o + 9], nf[o + 10]];
    const det = cd[0] * (cd[1] * cd[2] - co[2] * co[2]) - co[0] * (co[0] * cd[2] - co[2] * co[1]) + co[1] * (co[0] * co[2] - cd[1] * co[1]);
    mass[n] = nf[o + 7] / Math.sqrt(Math.max(det, 1e-30));
    for (let c = 0; c < 3; c++) alb[n * 3 + c] = gf[n * 16 + 12 + c];
    mat[n] = gu[n * 16 + 15] & 0xffff;
  }
  let maxDepth = 0;
  const inner = [];
  for (let n = G; n < N; n++) { inner.push(n); maxDepth = Math.max(maxDepth, nu[n * 24 + 21]); }
  for (let n = 0; n < G; n++) maxDepth = Math.max(maxDepth, nu[n * 24 + 21]);
  inner.sort((p, q) => nu[q * 24 + 21] - nu[p * 24 + 21]);    // deepest first; ties in index order (stable)
  const out = new ArrayBuffer(Math.max(1, N - G) * 64), of = new Float32Array(out), ou = new Uint32Array(out);
  for (const n of inner) {
    const lo = nu[n * 24 + 19], hi = nu[n * 24 + 20];
    let m = 0;
    for (let k = lo; k < hi; k++) m += mass[kids[k]];
    mass[n] = m;
    for (let c = 0; c < 3; c++) {
      let s = 0;
      for (let k = lo; k < hi; k++) s += mass[kids[k]] * alb[kids[k] * 3 + c];
      alb[n * 3 + c] = s / Math.max(m, 1e-30);
    }
    mat[n] = hi > lo ? mat[kids[lo]] : 0;
    const M = cinv(n), det = det3(M);
    const inv = [[M[1][1] * M[2][2] - M[1][2] * M[2][1], M[0][2] * M[2][1] - M[0][1] * M[2][2], M[0][1] * M[1][2] - M[0][2] * M[1][1]],
      [M[1][2] * M[2][0] - M[1][0] * M[2][2], M[0][0] * M[2][2] - M[0][2] * M[2][0], M[0][2] * M[1][0] - M[0][0] * M[1][2]],
      [M[1][0] * M[2][1] - M[1][1] * M[2][0], M[0][1] * M[2][0] - M[0][0] * M[2][1], M[0][0] * M[1][1] - M[0][1] * M[1][0]]]
      .map((r) => r.map((v) => v / det));
    const { values, vectors } = jacobiEig(inv);
    const cols = vectors.map((v) => v.slice());
    const d = cols[0][0] * (cols[1][1] * cols[2][2] - cols[1][2] * cols[2][1]) - cols[1][0] * (cols[0][1] * cols[2][2] - cols[0][2] * cols[2][1])
      + cols[2][0] * (cols[0][1] * cols[1][2] - cols[0][2] * cols[1][1]);
    if (d < 0) cols[2] = cols[2].map((v) => -v);
    const r = (n - G) * 16;
    of.set([nf[n * 24], nf[n * 24 + 1], nf[n * 24 + 2], nf[n * 24 + 7]], r);
    of.set(quatFromColumns(cols), r + 4);
    of.set(values.map((v) => Math.sqrt(Math.max(v, 1e-12))), r + 8);
    ou[r + 11] = nu[n * 24 + 15] & 0xffff;
    of.set([alb[n * 3], alb[n * 3 + 1], alb[n * 3 + 2]], r + 12);
    ou[r + 15] = mat[n];
  }
  return { records: out.slice(0, (N - G) * 64), N, G, maxDepth };
}

module.exports = {
  EIN3, ASG_WIDEN, buildEmitterTree, packEmitNodes, darkEmitter, assetEmitters, emitParams, frameUniforms, eig3, mipLevels, lsParams, LS_PARAMS_BYTES,
  ENV_BYTES, SKY_BG_BYTES, parseEnvironment, sunDir, sunAngles, envBytes,
  SUN_PARAMS_BYTES, basisK, sunLayout, sunFrame,
  PPK, FULL_ROUNDS, PROBE_PARAMS_BYTES, D14, pinvD14, probeLayout, probeVolumes, probeCount, probeParams, probeFrame,
  isStaticPart, nodeGaussians,
};
