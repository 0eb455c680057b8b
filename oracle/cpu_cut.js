'use strict';
/**
 * ORACLE — test and measurement infrastructure only: the CPU baseline of the cut, timed beside the GPU.
 *
 * A Node restatement of the reference's CPU cut (src/core/runtime/cut.ts:26-215: fp64, Math.hypot, the
 * view·model matrix precomposed, instanceMeshletRange / instanceGuaranteedRange, lodFadeRange from
 * lodBlend.ts:52-61) extended with the GPU pass's cell and instance culls (GeometryPool.ts:742-757,
 * 877-893: frustum, minPixelRadius, maxDrawDistance) and the meshlets' frustum test (GeometryPool.ts:940-943),
 * so that a world-scale cut finishes. Camera cut only, no occlusion (cut.ts has neither shadow cut nor HZB).
 * Work is counted in the GPU's unit: a cluster test is one meshlet of an instance's candidate range
 * [start, end) (cut.ts:146-166), or one fast-path meshlet.
 *
 * Usage: node cpu_cut.js <dir> [--seconds S] [--threads N]
 *   <dir> holds the pool in the reference layout (mesh_table.f32, instances.f32, cells.f32, bounds.f32,
 *   info.u32), views.bin (vg_view records, 292 bytes each) and scales.f32 (each view's camera LOD scale,
 *   the capacity controller's state). Prints one JSON line: the CPU model, and for one thread (the views in
 *   order until S seconds, repeated for at least a second, after a warm-up cut) and for N worker threads sharded by cells (the same views,
 *   median of 3 after a warm-up) the cluster tests, drawn clusters, seconds and tests per second.
 */
const fs = require('fs');
const os = require('os');
const path = require('path');
const { Worker, isMainThread, parentPort, workerData } = require('worker_threads');

const ERROR_INFINITY = 1e30;
const FAST_PATH_MAX_MESHLETS = 4;
const FAST_PATH_MARGIN = 1e-3;
const LOD_FADE_STEPS = 255;
const MESH_STRIDE = 6 * 4;
const VIEW_BYTES = 292;

/** The vg_view fields the cut reads (include/vgcut.h), as doubles. */
function readView(buf, k, scale) {
  const dv = new DataView(buf.buffer, buf.byteOffset + k * VIEW_BYTES, VIEW_BYTES);
  const f = (at) => dv.getFloat32(at, true);
  const arr = (at, n) => Array.from({ length: n }, (_, i) => f(at + i * 4));
  const errorThreshold = f(244);
  // loadBands (GeometryPool.ts:652-665): the camera band at the controller's scale, widened above 1.25 px.
  const low = errorThreshold * scale;
  const t = Math.min(Math.max((low - 1.25) / (2.5 - 1.25), 0), 1);
  const width = 1 + (f(252) - 1) * (t * t * (3 - 2 * t));
  return { view: arr(0, 16), planes: arr(128, 24), projScale: f(236), near: f(240), low, blend: width > 1.001 ? width : 1,
    minPixelRadius: f(260), frustum: dv.getUint32(264, true) };
}

const hypot3 = (x, y, z) => Math.hypot(x, y, z);

/** Column-major a * b for 4x4 (a: the view matrix, b: the instance matrix at `o` of `m`). */
function mul(a, m, o) {
  const r = new Float64Array(16);
  for (let j = 0; j < 4; j++)
    for (let i = 0; i < 4; i++)
      r[j * 4 + i] = a[i] * m[o + j * 4] + a[4 + i] * m[o + j * 4 + 1] + a[8 + i] * m[o + j * 4 + 2] + a[12 + i] * m[o + j * 4 + 3];
  return r;
}

function frustumDistance(planes, x, y, z) {
  let best = Infinity;
  for (let p = 0; p < 6; p++) best = Math.min(best, planes[p * 4] * x + planes[p * 4 + 1] * y + planes[p * 4 + 2] * z + planes[p * 4 + 3]);
  return best;
}

/** lodFadeRange (lodBlend.ts:52-61): whether the cluster is drawn for some threshold of the band. */
function inBand(own, parent, threshold, blend) {
  const high = blend > 1 ? threshold * blend : threshold;
  if (!(own <= high && parent > threshold)) return false;
  if (!(blend > 1)) return true;
  const scale = LOD_FADE_STEPS / Math.log(blend);
  const step = (e) => Math.min(LOD_FADE_STEPS, Math.max(0, Math.round(Math.log(Math.max(e, 1e-30) / threshold) * scale)));
  const hi = Number.isFinite(parent) ? step(parent) : LOD_FADE_STEPS;
  return hi > step(own);
}

/** The cut of cells [c0, c1) for one view: {tests, drawn}. */
function cutCells(P, v, c0, c1) {
  const { meshTable: T, instances: I, cells: C, bounds: B } = P;
  const V = v.view;
  const high = v.blend > 1 ? v.low * v.blend : v.low;
  let tests = 0;
  let drawn = 0;
  for (let cell = c0; cell < c1; cell++) {
    const cs = cell * 8;
    const slot = C[cs + 5];
    const mt = slot * MESH_STRIDE;
    if (T[mt + 20] === 0) continue; // disabled mesh
    const R = C[cs + 3];
    const [sx, sy, sz] = [C[cs], C[cs + 1], C[cs + 2]];
    if (v.frustum && frustumDistance(v.planes, sx, sy, sz) < -R) continue;
    const maxDraw = T[mt + 12];
    const minPixels = Math.max(v.minPixelRadius, T[mt + 13]);
    const maxR = C[cs + 4];
    const gap = hypot3(V[0] * sx + V[4] * sy + V[8] * sz + V[12], V[1] * sx + V[5] * sy + V[9] * sz + V[13],
      V[2] * sx + V[6] * sy + V[10] * sz + V[14]) - R;
    if (gap > maxR && (maxR * v.projScale) / Math.max(gap, 1e-6) < minPixels) continue;
    if (gap > maxDraw) continue;
    const levels = T[mt + 14];
    const levelBase = T[mt + 15] * 4;
    const [ox, oy, oz, or] = [T[mt], T[mt + 1], T[mt + 2], T[mt + 3]];
    const [lx, ly, lz, lr] = [T[mt + 4], T[mt + 5], T[mt + 6], T[mt + 7]];
    const [kx, ky, kz, kr] = [T[mt + 8], T[mt + 9], T[mt + 10], T[mt + 11]];
    const first = C[cs + 6];
    const count = C[cs + 7];
    for (let i = first; i < first + count; i++) {
      const m = i * 20;
      const s = Math.max(hypot3(I[m], I[m + 1], I[m + 2]), hypot3(I[m + 4], I[m + 5], I[m + 6]), hypot3(I[m + 8], I[m + 9], I[m + 10]));
      const cx = I[m] * ox + I[m + 4] * oy + I[m + 8] * oz + I[m + 12];
      const cy = I[m + 1] * ox + I[m + 5] * oy + I[m + 9] * oz + I[m + 13];
      const cz = I[m + 2] * ox + I[m + 6] * oy + I[m + 10] * oz + I[m + 14];
      const radius = or * s;
      if (v.frustum && frustumDistance(v.planes, cx, cy, cz) < -radius) continue;
      const vd = hypot3(V[0] * cx + V[4] * cy + V[8] * cz + V[12], V[1] * cx + V[5] * cy + V[9] * cz + V[13],
        V[2] * cx + V[6] * cy + V[10] * cz + V[14]);
      if (vd > radius && (radius * v.projScale) / vd < minPixels) continue;
      if (vd - radius > maxDraw) continue;
      // The instance's LOD test in cut.ts's form: view * model precomposed, errors scaled by s.
      const VM = mul(V, I, m);
      const dl = hypot3(VM[0] * lx + VM[4] * ly + VM[8] * lz + VM[12], VM[1] * lx + VM[5] * ly + VM[9] * lz + VM[13],
        VM[2] * lx + VM[6] * ly + VM[10] * lz + VM[14]);
      const dMin = Math.max(dl - lr * s, v.near);
      const dMax = Math.max(dl + lr * s, v.near);
      const k = s * v.projScale;
      let lo = -1;
      let hi = -1;
      let possible = 0;
      let sure = 0;
      for (let l = 0; l < levels; l++) {
        const e = levelBase + l * 8;
        const minOwn = T[e + 2];
        const maxParent = T[e + 3];
        if (!((minOwn * k) / dMax <= high && (maxParent >= ERROR_INFINITY * 0.5 || (maxParent * k) / dMin > v.low))) continue;
        if (lo < 0) lo = l;
        hi = l;
        possible += T[e + 1];
        const maxOwn = T[e + 4];
        const minParent = T[e + 5];
        if ((maxOwn * k) / dMin <= v.low * (1 - FAST_PATH_MARGIN) &&
            (minParent >= ERROR_INFINITY * 0.5 || (minParent * k) / dMax > high * (1 + FAST_PATH_MARGIN))) sure += T[e + 1];
      }
      if (lo < 0) continue;
      const start = T[levelBase + lo * 8];
      const end = T[levelBase + hi * 8] + T[levelBase + hi * 8 + 1];
      const ccx = I[m] * kx + I[m + 4] * ky + I[m + 8] * kz + I[m + 12];
      const ccy = I[m + 1] * kx + I[m + 5] * ky + I[m + 9] * kz + I[m + 13];
      const ccz = I[m + 2] * kx + I[m + 6] * ky + I[m + 10] * kz + I[m + 14];
      if (possible === end - start && sure === possible && possible <= FAST_PATH_MAX_MESHLETS &&
          (!v.frustum || frustumDistance(v.planes, ccx, ccy, ccz) >= kr * s)) {
        tests += possible;
        drawn += possible;
        continue;
      }
      tests += end - start;
      for (let j = start; j < end; j++) {
        const b = j * 16;
        if (v.frustum) {
          const ux = I[m] * B[b + 12] + I[m + 4] * B[b + 13] + I[m + 8] * B[b + 14] + I[m + 12];
          const uy = I[m + 1] * B[b + 12] + I[m + 5] * B[b + 13] + I[m + 9] * B[b + 14] + I[m + 13];
          const uz = I[m + 2] * B[b + 12] + I[m + 6] * B[b + 13] + I[m + 10] * B[b + 14] + I[m + 14];
          if (frustumDistance(v.planes, ux, uy, uz) < -(B[b + 15] * s)) continue;
        }
        const own = projected(VM, B, b, B[b + 8], s, v);
        const parent = projected(VM, B, b + 4, B[b + 9], s, v);
        if (inBand(own, parent, v.low, v.blend)) drawn++;
      }
    }
  }
  return { tests, drawn };
}

/** cut.ts:32-39: the projected error of a sphere at `o` of B with error `e`, through view * model. */
function projected(VM, B, o, e, s, v) {
  if (e >= ERROR_INFINITY * 0.5) return Infinity;
  const cx = B[o];
  const cy = B[o + 1];
  const cz = B[o + 2];
  const x = VM[0] * cx + VM[4] * cy + VM[8] * cz + VM[12];
  const y = VM[1] * cx + VM[5] * cy + VM[9] * cz + VM[13];
  const z = VM[2] * cx + VM[6] * cy + VM[10] * cz + VM[14];
  return (e * s * v.projScale) / Math.max(Math.hypot(x, y, z) - B[o + 3] * s, v.near);
}

function shared(file, Type) {
  const b = fs.readFileSync(file);
  const sab = new SharedArrayBuffer(b.length);
  new Uint8Array(sab).set(b);
  return new Type(sab);
}

const poolOf = (bufs) => ({ meshTable: new Float32Array(bufs.meshTable), instances: new Float32Array(bufs.instances),
  cells: new Float32Array(bufs.cells), bounds: new Float32Array(bufs.bounds) });

if (!isMainThread) {
  const P = poolOf(workerData.bufs);
  const views = workerData.views.map((x) => x);
  parentPort.on('message', (msg) => {
    if (msg === 'stop') {
      process.exit(0);
    }
    let tests = 0;
    let drawn = 0;
    for (const v of views) {
      const r = cutCells(P, v, workerData.c0, workerData.c1);
      tests += r.tests;
      drawn += r.drawn;
    }
    parentPort.postMessage({ tests, drawn });
  });
} else {
  const args = process.argv.slice(2);
  const dir = args[0];
  const opt = (name, dflt) => {
    const i = args.indexOf(name);
    return i >= 0 ? Number(args[i + 1]) : dflt;
  };
  const seconds = opt('--seconds', 12);
  const threads = opt('--threads', os.cpus().length);
  const f32 = (name) => shared(path.join(dir, name), Float32Array);
  const bufs = { meshTable: f32('mesh_table.f32').buffer, instances: f32('instances.f32').buffer, cells: f32('cells.f32').buffer,
    bounds: f32('bounds.f32').buffer };
  const P = poolOf(bufs);
  const viewBytes = fs.readFileSync(path.join(dir, 'views.bin'));
  const scaleBytes = fs.readFileSync(path.join(dir, 'scales.f32'));
  const scales = new Float32Array(scaleBytes.buffer.slice(scaleBytes.byteOffset, scaleBytes.byteOffset + scaleBytes.length));
  const views = Array.from({ length: viewBytes.length / VIEW_BYTES }, (_, k) => readView(viewBytes, k, scales[k]));
  const cells = P.cells.length / 8;
  const now = () => Number(process.hrtime.bigint()) / 1e9;

  // One thread: a warm-up cut of the first view, then the views in order until `seconds` have passed; a
  // view list that takes less than a second (or `seconds`, if smaller) is cut again until it has.
  cutCells(P, views[0], 0, cells);
  let t0 = now();
  const single = { views: 0, tests: 0, drawn: 0, seconds: 0 };
  let done = false;
  while (!done) {
    for (const v of views) {
      const r = cutCells(P, v, 0, cells);
      single.views++;
      single.tests += r.tests;
      single.drawn += r.drawn;
      if (now() - t0 > seconds) {
        done = true;
        break;
      }
    }
    if (now() - t0 >= Math.min(1, seconds)) done = true;
  }
  single.seconds = now() - t0;
  single.tests_per_s = single.tests / single.seconds;

  // N worker threads over the same views, each a contiguous range of cells of every view (the pool shared,
  // not copied).
  const sample = views.slice(0, Math.min(single.views, views.length));
  const workers = [];
  for (let w = 0; w < threads; w++) {
    const c0 = Math.floor((w * cells) / threads);
    const c1 = Math.floor(((w + 1) * cells) / threads);
    workers.push(new Worker(__filename, { workerData: { bufs, views: sample, c0, c1 } }));
  }
  const round = () => Promise.all(workers.map((wk) => new Promise((resolve) => {
    wk.once('message', resolve);
    wk.postMessage('go');
  })));
  const runs = [];
  round().then(async () => {
    for (let k = 0; k < 3; k++) {
      t0 = now();
      const parts = await round();
      const secs = now() - t0;
      const tests = parts.reduce((a, p) => a + p.tests, 0);
      const drawn = parts.reduce((a, p) => a + p.drawn, 0);
      runs.push({ views: sample.length, tests, drawn, seconds: secs, tests_per_s: tests / secs });
    }
    runs.sort((a, b) => a.tests_per_s - b.tests_per_s);
    for (const wk of workers) wk.postMessage('stop');
    const cpus = os.cpus();
    console.log(JSON.stringify({ cpu_model: cpus.length ? cpus[0].model : 'unknown', logical_cpus: cpus.length, node: process.version,
      single_thread: single, threads: Object.assign({ n: threads }, runs[1]) }));
  });
}
