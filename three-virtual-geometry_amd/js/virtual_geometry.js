'use strict';
/**
 * The Node host of the cut: the reference's VirtualGeometry / VirtualMesh surface for this path
 * (src/core/runtime/VirtualGeometry.ts:62-633, VirtualMesh.ts:39-331) over the N-API addon of libvgcut.so
 * (../napi/vgcut.node), written as plain CommonJS for Node 12 (the image has no TypeScript compiler; the
 * JSDoc types are what the .ts signatures say).
 *
 * Kept from the reference: the levers (errorThreshold, lodBlend, maxErrorThreshold, triangleBudget,
 * thresholdRange, minPixelRadius, frustumCulling, shadowErrorScale, shadowSweep, statsInterval), createMesh
 * with per-mesh maxDrawDistance / minPixelRadius, update(camera) with the skip of unchanged frames
 * (cutInputsChanged), readStats, and the budget glide (adaptThreshold / glideTo / stepThreshold).
 * update(camera, viewportHeight) takes the viewport's height in the pixels errors are measured in (CSS
 * pixels in the reference) instead of a renderer. Raster, materials and the scene-graph hooks are not
 * part of this path.
 *
 * The view math restates three.js's formulas in their operation order (the same doubles as the Python
 * host's vgcut/camera.py): PerspectiveCamera / OrthographicCamera projection (WebGPU depth range),
 * Matrix4.lookAt, multiplyMatrices, Frustum.setFromProjectionMatrix. A camera from three.js itself works
 * too: update() reads matrixWorldInverse.elements, projectionMatrix.elements, near and
 * isPerspectiveCamera.
 */
const path = require('path');

const THRESHOLD_RATE_UP = 1.03; // VirtualGeometry.ts:21-22
const THRESHOLD_RATE_DOWN = 1.01;
const NO_DRAW_DISTANCE = 1e30; // GeometryPool.ts:111
const STATE_FIXED = 10000; // GeometryPool.ts:131
const S_SCALE = 0;
const S_FILL = 2;
const FILL_HIGH = 0.9; // GeometryPool.ts:136-137
const FILL_LOW = 0.75;
const CUT_COUNTERS = 5;
const PER_MESH_COUNTERS = 12;
const C_SELECTED = 0;
const C_DRAWN = 1;
const C_SURVIVED = 2;
const C_REQUESTED = 4;
/** Bytes of one vg_view record (include/vgcut.h). */
const VIEW_BYTES = 292;
const DEG2RAD = Math.PI / 180;

/** VirtualGeometry.ts:60 */
const refineThreshold = (threshold, use) => threshold * Math.pow(use / 0.82, 0.25);

// ------------------------------------------------------------------ three.js view math (fp64, column-major)

function normalize3(x, y, z) {
  const n = Math.sqrt(x * x + y * y + z * z);
  const inv = 1 / (n || 1);
  return [x * inv, y * inv, z * inv];
}

function cross(ax, ay, az, bx, by, bz) {
  return [ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx];
}

/** World matrix of a camera at `eye` looking at `target`: Matrix4.lookAt's basis and the translation. */
function lookAtWorld(eye, target, up) {
  up = up || [0, 1, 0];
  let [zx, zy, zz] = [eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]];
  if (zx * zx + zy * zy + zz * zz === 0) zz = 1;
  [zx, zy, zz] = normalize3(zx, zy, zz);
  let [xx, xy, xz] = cross(up[0], up[1], up[2], zx, zy, zz);
  if (xx * xx + xy * xy + xz * xz === 0) {
    if (Math.abs(up[2]) === 1) zx += 0.0001;
    else zz += 0.0001;
    [zx, zy, zz] = normalize3(zx, zy, zz);
    [xx, xy, xz] = cross(up[0], up[1], up[2], zx, zy, zz);
  }
  [xx, xy, xz] = normalize3(xx, xy, xz);
  const [yx, yy, yz] = cross(zx, zy, zz, xx, xy, xz);
  return Float64Array.from([xx, xy, xz, 0, yx, yy, yz, 0, zx, zy, zz, 0, eye[0], eye[1], eye[2], 1]);
}

/** Inverse of a rotation + translation (column-major). */
function invertRigid(m) {
  const out = new Float64Array(16);
  for (let i = 0; i < 3; i++) {
    out[i] = m[i * 4];
    out[4 + i] = m[i * 4 + 1];
    out[8 + i] = m[i * 4 + 2];
    out[12 + i] = -(m[i * 4] * m[12] + m[i * 4 + 1] * m[13] + m[i * 4 + 2] * m[14]);
  }
  out[15] = 1;
  return out;
}

/** PerspectiveCamera.updateProjectionMatrix (zoom 1) + Matrix4.makePerspective, WebGPU depth range. */
function perspective(fov, aspect, near, far) {
  const top = near * Math.tan(DEG2RAD * 0.5 * fov);
  const height = 2 * top;
  const width = aspect * height;
  const left = -0.5 * width;
  const right = left + width;
  const bottom = top - height;
  const te = new Float64Array(16);
  te[0] = 2 * near / (right - left);
  te[5] = 2 * near / (top - bottom);
  te[8] = (right + left) / (right - left);
  te[9] = (top + bottom) / (top - bottom);
  te[10] = -far / (far - near);
  te[11] = -1;
  te[14] = -far * near / (far - near);
  return te;
}

/** OrthographicCamera.updateProjectionMatrix (zoom 1) + Matrix4.makeOrthographic, WebGPU depth range. */
function orthographic(left, right, top, bottom, near, far) {
  const dx = (right - left) / 2;
  const dy = (top - bottom) / 2;
  const cx = (right + left) / 2;
  const cy = (top + bottom) / 2;
  [left, right, top, bottom] = [cx - dx, cx + dx, cy + dy, cy - dy];
  const w = 1 / (right - left);
  const h = 1 / (top - bottom);
  const p = 1 / (far - near);
  const te = new Float64Array(16);
  te[0] = 2 * w;
  te[5] = 2 * h;
  te[10] = -1 * p;
  te[12] = -((right + left) * w);
  te[13] = -((top + bottom) * h);
  te[14] = -(near * p);
  te[15] = 1;
  return te;
}

/** Column-major a * b (Matrix4.multiplyMatrices). */
function multiply(a, b) {
  const out = new Float64Array(16);
  for (let j = 0; j < 4; j++)
    for (let i = 0; i < 4; i++)
      out[j * 4 + i] = a[i] * b[j * 4] + a[4 + i] * b[j * 4 + 1] + a[8 + i] * b[j * 4 + 2] + a[12 + i] * b[j * 4 + 3];
  return out;
}

/** Frustum.setFromProjectionMatrix(m, WebGPUCoordinateSystem): 6 x (nx, ny, nz, constant), normalised. */
function frustumPlanes(me) {
  const rows = [
    [me[3] - me[0], me[7] - me[4], me[11] - me[8], me[15] - me[12]],
    [me[3] + me[0], me[7] + me[4], me[11] + me[8], me[15] + me[12]],
    [me[3] + me[1], me[7] + me[5], me[11] + me[9], me[15] + me[13]],
    [me[3] - me[1], me[7] - me[5], me[11] - me[9], me[15] - me[13]],
    [me[3] - me[2], me[7] - me[6], me[11] - me[10], me[15] - me[14]],
    [me[2], me[6], me[10], me[14]],
  ];
  const out = new Float64Array(24);
  rows.forEach(([x, y, z, w], i) => {
    const inv = 1 / Math.sqrt(x * x + y * y + z * z);
    out[i * 4] = x * inv;
    out[i * 4 + 1] = y * inv;
    out[i * 4 + 2] = z * inv;
    out[i * 4 + 3] = w * inv;
  });
  return out;
}

/** A camera with three.js's fields (projectionMatrix, matrixWorld, matrixWorldInverse, near, far). */
class Camera {
  constructor(projection, near, far, isPerspective) {
    this.projectionMatrix = { elements: Float64Array.from(projection) };
    this.matrixWorld = { elements: Float64Array.from([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]) };
    this.matrixWorldInverse = { elements: Float64Array.from(this.matrixWorld.elements) };
    this.near = near;
    this.far = far;
    this.isPerspectiveCamera = isPerspective;
    this.isOrthographicCamera = !isPerspective;
  }

  /** Places the camera at `eye` looking at `target`. */
  lookAt(eye, target, up) {
    this.matrixWorld.elements = lookAtWorld(eye, target, up);
    this.matrixWorldInverse.elements = invertRigid(this.matrixWorld.elements);
    return this;
  }
}

const perspectiveCamera = (fov, aspect, near, far) => new Camera(perspective(fov, aspect, near, far), near, far, true);
const orthographicCamera = (left, right, top, bottom, near, far) =>
  new Camera(orthographic(left, right, top, bottom, near, far), near, far, false);

// ------------------------------------------------------------------ vg_view records

/**
 * The per-view uniforms VirtualGeometry.update() sets (VirtualGeometry.ts:400-437) for `camera` at a viewport
 * `viewportHeight` pixels high, with the context's levers in `o`.
 * @returns {object} the vg_view fields
 */
function viewRecord(camera, viewportHeight, o) {
  o = o || {};
  const view = camera.matrixWorldInverse.elements;
  const proj = camera.projectionMatrix.elements;
  const errorThreshold = o.errorThreshold === undefined ? 1 : o.errorThreshold;
  const maxErrorThreshold = o.maxErrorThreshold === undefined ? 64 : o.maxErrorThreshold;
  const occlusion = o.occlusion && camera.isPerspectiveCamera ? (o.occlusion === 2 ? 2 : 1) : 0;
  const range = o.thresholdRange || [0.5, 3];
  return {
    viewMatrix: view,
    projectionMatrix: proj,
    frustumPlanes: frustumPlanes(multiply(proj, view)),
    shadowSweep: o.shadowSweep || [0, 0, 0],
    projScale: proj[5] * viewportHeight * 0.5,
    near: camera.near,
    errorThreshold,
    shadowErrorScale: o.shadowErrorScale === undefined ? 3 : o.shadowErrorScale,
    lodBlendMax: Math.max(1, o.lodBlend === undefined ? 2 : o.lodBlend),
    lodScaleLimit: Math.max(1, maxErrorThreshold / errorThreshold),
    minPixelRadius: o.minPixelRadius === undefined ? 0.7 : o.minPixelRadius,
    frustumCulling: o.frustumCulling === false ? 0 : 1,
    occlusion,
    slot: o.slot || 0,
    triangleBudget: o.triangleBudget || 0,
    thresholdRange: range,
    maxErrorThreshold,
  };
}

/** Writes a vg_view record (include/vgcut.h layout, little-endian) at `offset` of a DataView. */
function packView(dv, offset, r) {
  const f = (at, x) => dv.setFloat32(offset + at, x, true);
  const u = (at, x) => dv.setUint32(offset + at, x >>> 0, true);
  for (let i = 0; i < 16; i++) f(i * 4, r.viewMatrix[i]);
  for (let i = 0; i < 16; i++) f(64 + i * 4, r.projectionMatrix[i]);
  for (let i = 0; i < 24; i++) f(128 + i * 4, r.frustumPlanes[i]);
  for (let i = 0; i < 3; i++) f(224 + i * 4, r.shadowSweep[i]);
  f(236, r.projScale);
  f(240, r.near);
  f(244, r.errorThreshold);
  f(248, r.shadowErrorScale);
  f(252, r.lodBlendMax);
  f(256, r.lodScaleLimit);
  f(260, r.minPixelRadius);
  u(264, r.frustumCulling);
  u(268, r.occlusion);
  u(272, r.slot);
  u(276, r.triangleBudget);
  f(280, r.thresholdRange[0]);
  f(284, r.thresholdRange[1]);
  f(288, r.maxErrorThreshold);
}

/** vg_view records of several views in one ArrayBuffer (what the addon's cut takes). */
function packViews(records) {
  const buf = new ArrayBuffer(VIEW_BYTES * records.length);
  const dv = new DataView(buf);
  records.forEach((r, i) => packView(dv, i * VIEW_BYTES, r));
  return buf;
}

// ------------------------------------------------------------------ stats (readCounters + readStats)

/**
 * GeometryPool.readCounters (GeometryPool.ts:1135-1151) over a view slot's counter words: per cut
 * (selected, drawn, survived, requestedTriangles), occluded instances, controller scale and fill.
 */
function countersOf(words, slots) {
  const stateBase = PER_MESH_COUNTERS + 2 * slots * 4;
  const state = (i) => words[stateBase + i] / STATE_FIXED;
  const cut = (c) => ({
    selected: words[2 + c * CUT_COUNTERS + C_SELECTED],
    drawn: words[2 + c * CUT_COUNTERS + C_DRAWN],
    survived: words[2 + c * CUT_COUNTERS + C_SURVIVED],
    requestedTriangles: words[2 + c * CUT_COUNTERS + C_REQUESTED] / 3,
  });
  return { camera: cut(0), shadow: cut(1), occluded: words[1], scale: [state(S_SCALE), state(S_SCALE + 1)],
    fill: [state(S_FILL), state(S_FILL + 1)] };
}

/** readCounters' "still steering" test (GeometryPool.ts:1141-1142). */
const controllerActive = (c) => c.fill.some((f, k) => f > FILL_HIGH || (f < FILL_LOW && c.scale[k] > 1.0001));

/**
 * VirtualGeometry.readStats (VirtualGeometry.ts:591-632) for one pool: the visible meshes' totals and the
 * counters of its last cut against its capacity {triangles, shadowTriangles, meshlets}.
 */
function statsFromCounters(c, capacity, meshes) {
  const result = {
    drawnMeshlets: 0, drawnTriangles: 0, fullDetailTriangles: 0, instances: 0, visibleInstances: 0,
    requestedTriangles: 0, overflow: false, capacityUse: 0, occludedInstances: 0, shadowTriangles: 0, thresholdScale: 1,
  };
  for (const mesh of meshes) {
    if (!mesh.visible) continue;
    result.fullDetailTriangles += mesh.fullDetailTriangles;
    result.instances += mesh.instanceCount;
  }
  const cap = capacity;
  result.drawnMeshlets += c.camera.drawn;
  result.drawnTriangles += Math.min(c.camera.requestedTriangles, cap.triangles);
  result.requestedTriangles += c.camera.requestedTriangles;
  result.visibleInstances += c.camera.survived;
  result.occludedInstances += c.occluded;
  result.shadowTriangles += Math.min(c.shadow.requestedTriangles, cap.shadowTriangles);
  result.thresholdScale = Math.max(result.thresholdScale, c.scale[0]);
  if (c.camera.selected > cap.meshlets || c.camera.requestedTriangles > cap.triangles) result.overflow = true;
  if (c.shadow.selected > cap.meshlets || c.shadow.requestedTriangles > cap.shadowTriangles) result.overflow = true;
  result.capacityUse = Math.max(result.capacityUse, c.camera.requestedTriangles / cap.triangles, c.camera.selected / cap.meshlets,
    c.shadow.requestedTriangles / cap.shadowTriangles, c.shadow.selected / cap.meshlets);
  return result;
}

// ------------------------------------------------------------------ meshes and the context

const leafTrianglesOf = (data) =>
  (data.stats && data.stats.leafTriangles) || data.leafTriangles || (data.indices ? data.indices.length / 3 : 0);

/** Instances of one DAG (VirtualMesh.ts:91-331, the cut's side). */
class VirtualMesh {
  /**
   * @param {VirtualGeometry} context
   * @param {object} data VirtualMeshData (buildVirtualMesh's output, typed arrays)
   * @param {{matrices: Float32Array, colors?: Float32Array}} instances
   * @param {{maxDrawDistance?: number, minPixelRadius?: number}} [options]
   */
  constructor(context, data, instances, options) {
    options = options || {};
    this.context = context;
    this.data = data;
    this.matrices = Float32Array.from(instances.matrices);
    this.colors = instances.colors ? Float32Array.from(instances.colors) : undefined;
    this.instanceCount = this.matrices.length / 16;
    this.fullDetailTriangles = leafTrianglesOf(data) * this.instanceCount;
    this.slot = context.addon.addMesh(context.scene, data);
    this._maxDrawDistance = NO_DRAW_DISTANCE;
    this._minPixelRadius = 0;
    this.visible = true;
    this.castShadow = false;
    this.occlusionCulling = true;
    this._flags = '';
    this.commitInstances();
    if (options.maxDrawDistance !== undefined) this.maxDrawDistance = options.maxDrawDistance;
    if (options.minPixelRadius !== undefined) this.minPixelRadius = options.minPixelRadius;
    this.syncFlags();
  }

  /** Cull distance in world units (Infinity: never), VirtualMesh.ts:137-144. */
  get maxDrawDistance() {
    return this._maxDrawDistance >= NO_DRAW_DISTANCE ? Infinity : this._maxDrawDistance;
  }

  set maxDrawDistance(distance) {
    this._maxDrawDistance = Number.isFinite(distance) ? Math.max(0, distance) : NO_DRAW_DISTANCE;
    this.syncFlags();
  }

  /** Per-mesh minimum on-screen radius in pixels, VirtualMesh.ts:147-152. */
  get minPixelRadius() {
    return this._minPixelRadius;
  }

  set minPixelRadius(pixels) {
    this._minPixelRadius = Math.max(0, pixels);
    this.syncFlags();
  }

  /** Updates one instance's column-major matrix (16 numbers); commitInstances() uploads a batch of them. */
  setMatrixAt(index, elements) {
    this.matrices.set(elements, index * 16);
  }

  /** Uploads the instances; the library re-sorts them in Morton order (gpuToInput / gpuIndexOf follow). */
  commitInstances() {
    this.gpuToInput = this.context.addon.setInstances(this.context.scene, this.slot, this.matrices, this.colors);
    this.gpuIndexOf = new Uint32Array(this.instanceCount);
    for (let g = 0; g < this.instanceCount; g++) this.gpuIndexOf[this.gpuToInput[g]] = g;
    this.context.cutDirty = true;
  }

  /** The mesh table flags (GeometryPool.writeFlags): written only when a value changes. */
  syncFlags() {
    const flags = {
      maxDrawDistance: this._maxDrawDistance, minPixelRadius: this._minPixelRadius, enabled: !!this.visible,
      shadow: !!this.castShadow, occlusion: !!(this.occlusionCulling && this.context.occlusion),
    };
    const key = JSON.stringify(flags);
    if (key === this._flags) return;
    this.context.addon.setFlags(this.context.scene, this.slot, flags);
    this._flags = key;
    this.context.cutDirty = true;
  }
}

/** Context: capacities, levers, the per-frame update, stats and the budget glide (VirtualGeometry.ts:77-633). */
class VirtualGeometry {
  /**
   * @param {{maxTriangles?: number, maxShadowTriangles?: number, maxMeshlets?: number, maxViews?: number,
   *   meshSlots?: number, device?: number, addon?: object, addonPath?: string}} [options]
   */
  constructor(options) {
    options = options || {};
    this.addon = options.addon || require(options.addonPath || path.join(__dirname, '..', 'napi', 'vgcut.node'));
    this.capacity = {
      triangles: options.maxTriangles || 10000000,
      shadowTriangles: options.maxShadowTriangles || 4000000,
      meshlets: Math.min(options.maxMeshlets || 1 << 20, 1 << 24),
    };
    this.meshSlots = options.meshSlots || 64;
    this.ctx = this.addon.createContext({
      maxTriangles: this.capacity.triangles, maxShadowTriangles: this.capacity.shadowTriangles, maxMeshlets: this.capacity.meshlets,
      maxViews: options.maxViews || 1, meshSlots: this.meshSlots, device: options.device === undefined ? 0 : options.device,
    });
    this.scene = this.addon.createScene(this.ctx);
    this.meshes = [];
    // levers (VirtualGeometry.ts:78-183)
    this.errorThreshold = 1;
    this.lodBlend = 2;
    this.maxErrorThreshold = 64;
    this.triangleBudget = 0;
    this.thresholdRange = [0.5, 3];
    this.minPixelRadius = 0.7;
    this.frustumCulling = true;
    this.shadowErrorScale = 3;
    this.shadowSweep = [0, 0, 0];
    this.occlusion = false;
    this.statsInterval = 10;
    /**
     * How triangleBudget sets the threshold: 'glide' as the reference (a stats readback every statsInterval
     * frames feeds adaptThreshold), or 'histogram': each cut picks its threshold on the GPU from a histogram of
     * projected errors (the smallest of the budget table's thresholds whose cut fits the budget).
     */
    this.budgetMode = 'glide';
    this.lastStats = null;
    this.frameCount = 0;
    this.skippedFrames = 0;
    this.cutDirty = true;
    this._controllerActive = false;
    this._outOfBand = 0;
    this._thresholdTarget = null;
    this._thresholdWritten = NaN;
    this._statsThreshold = 1;
    this._cutInputs = null;
  }

  /** VirtualGeometry.createMesh (VirtualGeometry.ts:231-233); the material is the raster's and not used here. */
  createMesh(data, material, instances, options) {
    const mesh = new VirtualMesh(this, data, instances, options);
    this.meshes.push(mesh);
    return mesh;
  }

  /** VirtualGeometry.unregister (VirtualGeometry.ts:370-376): the mesh's pool slot is freed for later meshes. */
  unregister(mesh) {
    const i = this.meshes.indexOf(mesh);
    if (i < 0) return;
    this.addon.removeMesh(this.scene, mesh.slot);
    this.meshes.splice(i, 1);
    this.cutDirty = true;
  }

  /** The depth of last frame's occluders for a view slot (OcclusionCulling.beginFrame + the HZB build). */
  setHzb(depth, width, height, slot) {
    this.addon.setHzb(this.scene, slot || 0, depth, width, height);
    this.cutDirty = true;
  }

  /** The vg_view record this context's levers give `camera` (for batched cuts of several views). */
  viewFor(camera, viewportHeight, slot) {
    return viewRecord(camera, viewportHeight, {
      errorThreshold: this.errorThreshold, shadowErrorScale: this.shadowErrorScale, lodBlend: this.lodBlend,
      maxErrorThreshold: this.maxErrorThreshold, minPixelRadius: this.minPixelRadius, frustumCulling: this.frustumCulling,
      occlusion: this.occlusion, shadowSweep: this.shadowSweep, slot: slot || 0,
      triangleBudget: this.budgetMode === 'histogram' ? this.triangleBudget : 0, thresholdRange: this.thresholdRange,
    });
  }

  /**
   * VirtualGeometry.update (VirtualGeometry.ts:392-507) for one camera: the glide step, the view constants, the
   * cut on the GPU (skipped when nothing it reads changed and the capacity controller has settled), and every
   * statsInterval frames a stats readback feeding the budget glide. Resolves when the cut is done.
   */
  update(camera, viewportHeight, slot) {
    slot = slot || 0;
    this.stepThreshold();
    for (const mesh of this.meshes) mesh.syncFlags();
    const record = this.viewFor(camera, viewportHeight, slot);
    const changed = this.cutInputsChanged(record);
    let work;
    if (!changed && !this.cutDirty && !this._controllerActive) {
      this.skippedFrames++;
      work = Promise.resolve();
    } else {
      this.cutDirty = false;
      work = this.addon.cut(this.scene, packViews([record]), 1);
    }
    return work.then(() => {
      this.frameCount++;
      const interval = this._thresholdTarget !== null ? Math.min(2, this.statsInterval) : this.statsInterval;
      const counters = this.readCounters(slot);
      this._controllerActive = controllerActive(counters);
      if (this.meshes.length > 0 && this.frameCount % interval === 0) {
        this._statsThreshold = this.errorThreshold;
        this.lastStats = statsFromCounters(counters, this.capacity, this.meshes);
        if (this.triangleBudget > 0 && this.budgetMode === 'glide') this.adaptThreshold(this.lastStats);
      }
    });
  }

  /** The counter words of a view slot's last cut, as readCounters reports them. */
  readCounters(slot) {
    const words = new Uint32Array(this.addon.counterWords(this.ctx));
    const args = new Uint32Array(this.addon.drawArgsWords(this.ctx));
    this.addon.read(this.scene, slot || 0, words, args, new Uint32Array(0));
    return countersOf(words, this.meshSlots);
  }

  /** VirtualGeometry.readStats (VirtualGeometry.ts:591-632): stats of a view slot's last cut. */
  readStats(slot) {
    return statsFromCounters(this.readCounters(slot), this.capacity, this.meshes);
  }

  /** cutInputsChanged (VirtualGeometry.ts:515-542): exact comparison of everything the cut reads from the context. */
  cutInputsChanged(record) {
    const cur = [].concat(Array.from(record.viewMatrix), Array.from(record.projectionMatrix), [record.near, record.projScale,
      record.errorThreshold, this.lodBlend, record.minPixelRadius, record.frustumCulling, record.occlusion], record.shadowSweep,
    [record.triangleBudget, record.lodScaleLimit]);
    const prev = this._cutInputs;
    this._cutInputs = cur;
    if (!prev || prev.length !== cur.length) return true;
    for (let k = 0; k < cur.length; k++) if (cur[k] !== prev[k]) return true;
    return false;
  }

  /** adaptThreshold (VirtualGeometry.ts:549-568): the glide's target from a stats readback. */
  adaptThreshold(stats) {
    const t = this._statsThreshold;
    const now = this.errorThreshold;
    const use = stats.capacityUse;
    if (this._thresholdTarget !== null && this._thresholdTarget < now && use >= 0.75) this._thresholdTarget = null;
    if (stats.drawnTriangles <= 0) return;
    const ratio = stats.drawnTriangles / this.triangleBudget;
    if (Math.abs(ratio - 1) < 0.15) {
      this._outOfBand = 0;
      return;
    }
    if (++this._outOfBand < 2) return;
    const [lo, hi] = this.thresholdRange;
    let next = Math.min(hi, Math.max(lo, t * Math.sqrt(ratio)));
    if (next < t) next = Math.max(next, refineThreshold(t, use));
    this.glideTo(next);
  }

  /** glideTo (VirtualGeometry.ts:570-572). */
  glideTo(target) {
    this._thresholdTarget = target === this.errorThreshold ? null : target;
  }

  /** stepThreshold (VirtualGeometry.ts:575-588): every frame, a few percent toward the target. */
  stepThreshold() {
    const t = this.errorThreshold;
    if (t !== this._thresholdWritten && !Number.isNaN(this._thresholdWritten)) this._thresholdTarget = null;
    const target = this._thresholdTarget;
    if (target !== null) {
      const next = target > t ? Math.min(target, t * THRESHOLD_RATE_UP) : Math.max(target, t / THRESHOLD_RATE_DOWN);
      this.errorThreshold = next;
      if (next === target) this._thresholdTarget = null;
    }
    this._thresholdWritten = this.errorThreshold;
  }
}

module.exports = {
  VirtualGeometry, VirtualMesh, Camera, perspectiveCamera, orthographicCamera, lookAtWorld, invertRigid, perspective, orthographic,
  multiply, frustumPlanes, viewRecord, packView, packViews, countersOf, statsFromCounters, controllerActive, refineThreshold,
  VIEW_BYTES, NO_DRAW_DISTANCE,
};
