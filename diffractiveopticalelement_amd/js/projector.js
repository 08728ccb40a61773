// Drop-in for the hot path of src/projector.js.  Targets and ROIs are built
// on the GPU in FP64 (returned as the reference's arrays); design() runs the
// whole phase-retrieval loop on the GPU; simulate() propagates on the GPU;
// screenMetrics() judges the device-resident screen on the GPU in FP64 (the
// screen is never downloaded); defocusCurve() / focusStack() propagate every
// distance in one batch; designMultiPlane() builds all plane targets in one
// batched call and runs the multi-plane loop on the GPU.
//
// DEVICE_PROJECTOR holds those methods; installDeviceProjector(Cls) puts them
// on a class with the reference's ProjectorSystem state (cfg, nx, ny, dx, dy,
// lambda, z, target, result, illumination()), which is how the reference's
// own class takes them (its other methods stay as they are).  The
// ProjectorSystem exported here is the minimal such class: the reference's
// constructor, setTarget and illumination, plus the device methods.
import { ComplexField } from './field.js';
import { native, dtypeCode } from './native.js';
import { propagateAngularSpectrumInPlace, focalPlaneSpectrum } from './propagate.js';
import { designDOE, illuminationProfile, gerchbergSaxtonMultiPlane, errorDiffusionQuantize } from './doe.js';

function pick(v, d) {
  return v === undefined || v === null ? d : v;
}

// 5x7 bitmap font (data: the rows the reference's target raster is made of,
// ref: projector.js:48-94).
export const GLYPHS = {
  A: '01110/10001/10001/11111/10001/10001/10001',
  B: '11110/10001/10001/11110/10001/10001/11110',
  C: '01110/10001/10000/10000/10000/10001/01110',
  D: '11110/10001/10001/10001/10001/10001/11110',
  E: '11111/10000/10000/11110/10000/10000/11111',
  F: '11111/10000/10000/11110/10000/10000/10000',
  G: '01110/10001/10000/10111/10001/10001/01111',
  H: '10001/10001/10001/11111/10001/10001/10001',
  I: '01110/00100/00100/00100/00100/00100/01110',
  J: '00111/00010/00010/00010/00010/10010/01100',
  K: '10001/10010/10100/11000/10100/10010/10001',
  L: '10000/10000/10000/10000/10000/10000/11111',
  M: '10001/11011/10101/10101/10001/10001/10001',
  N: '10001/11001/10101/10011/10001/10001/10001',
  O: '01110/10001/10001/10001/10001/10001/01110',
  P: '11110/10001/10001/11110/10000/10000/10000',
  Q: '01110/10001/10001/10001/10101/10010/01101',
  R: '11110/10001/10001/11110/10100/10010/10001',
  S: '01111/10000/10000/01110/00001/00001/11110',
  T: '11111/00100/00100/00100/00100/00100/00100',
  U: '10001/10001/10001/10001/10001/10001/01110',
  V: '10001/10001/10001/10001/10001/01010/00100',
  W: '10001/10001/10001/10101/10101/11011/10001',
  X: '10001/10001/01010/00100/01010/10001/10001',
  Y: '10001/10001/01010/00100/00100/00100/00100',
  Z: '11111/00001/00010/00100/01000/10000/11111',
  0: '01110/10001/10011/10101/11001/10001/01110',
  1: '00100/01100/00100/00100/00100/00100/01110',
  2: '01110/10001/00001/00010/00100/01000/11111',
  3: '11111/00010/00100/00010/00001/10001/01110',
  4: '00010/00110/01010/10010/11111/00010/00010',
  5: '11111/10000/11110/00001/00001/10001/01110',
  6: '00110/01000/10000/11110/10001/10001/01110',
  7: '11111/00001/00010/00100/01000/01000/01000',
  8: '01110/10001/10001/01110/10001/10001/01110',
  9: '01110/10001/10001/01111/00001/00010/01100',
  ' ': '00000/00000/00000/00000/00000/00000/00000',
  '-': '00000/00000/00000/11111/00000/00000/00000',
  '.': '00000/00000/00000/00000/00000/01100/01100',
  ':': '00000/01100/01100/00000/01100/01100/00000',
  '!': '00100/00100/00100/00100/00100/00000/00100',
  '+': '00000/00100/00100/11111/00100/00100/00000',
  '/': '00001/00010/00010/00100/01000/01000/10000',
  '%': '11001/11010/00010/00100/01000/01011/10011',
  '*': '00000/10101/01110/11111/01110/10101/00000',
};

// ref: projector.js:97-123
export function renderText(text, { scale = 8, spacing = 1 } = {}) {
  const chars = text.toUpperCase().split('');
  const width = chars.length * (5 + spacing) - spacing;
  const height = 7;
  const bits = new Float64Array(width * height);
  chars.forEach((ch, ci) => {
    const rows = (GLYPHS[ch] === undefined ? GLYPHS[' '] : GLYPHS[ch]).split('/');
    for (let y = 0; y < 7; y++) {
      for (let x = 0; x < 5; x++) if (rows[y][x] === '1') bits[x + ci * (5 + spacing) + y * width] = 1;
    }
  });
  const w2 = width * scale;
  const h2 = height * scale;
  const data = new Float64Array(w2 * h2);
  for (let y = 0; y < h2; y++) {
    const row = Math.floor(y / scale) * width;
    for (let x = 0; x < w2; x++) data[x + y * w2] = bits[Math.floor(x / scale) + row];
  }
  return { data, width: w2, height: h2 };
}

// Reference's arrays are any array-like; the device takes FP64 rasters of at
// least width * height samples (a short raster reads NaN, as the reference's
// out-of-range reads do).
function f64Raster(src, w, h) {
  const n = Math.max(0, Math.floor(w)) * Math.max(0, Math.floor(h));
  if (src instanceof Float64Array && src.length >= n) return src;
  const out = new Float64Array(Math.max(n, src.length)).fill(NaN);
  for (let k = 0; k < src.length; k++) out[k] = src[k];
  return out;
}

function f64Plane(a) {
  return a instanceof Float64Array ? a : Float64Array.from(a);
}

// ref: projector.js:126-154.  Bilinear resample on the GPU (doe_target_make,
// raster kind): 'contain' or 'cover' placement, background outside the
// raster, gain.
export function resampleRaster(src, sw, sh, dw, dh, { fit = 'contain', background = 0, gain = 1 } = {}) {
  const dev = native.targetMake([{ kind: 'raster', raster: f64Raster(src, sw, sh), width: sw, height: sh,
    cover: fit !== 'contain', background, gain }], dw, dh);
  const out = new Float64Array(dw * dh);
  native.fieldDownload(dev, out, new Float64Array(dw * dh));
  return out;
}

// Element-wise equality (NaN equal to NaN) of a stored FP64 copy and an array.
function sameValues(copy, a) {
  if (copy.length !== a.length) return false;
  for (let k = 0; k < a.length; k++) {
    const x = copy[k];
    const y = a[k];
    if (x !== y && !(x !== x && y !== y)) return false;
  }
  return true;
}

// The makeTarget kinds evaluated per pixel on the device.
const ANALYTIC = new Set(['disc', 'ring', 'dots', 'grid', 'checker', 'lines', 'zone']);

// Dark cells of a spot array as (col, row) int pairs; entries that cannot
// equal a cell index (non-integers) never match, as in the reference.
function darkCells(missing) {
  const pairs = [];
  for (const m of missing || []) {
    if (m && Number.isInteger(m[0]) && Number.isInteger(m[1]) && Math.abs(m[0]) < 2 ** 31 && Math.abs(m[1]) < 2 ** 31) {
      pairs.push(m[0], m[1]);
    }
  }
  return Int32Array.from(pairs);
}

// ref: projector.js:169-255.  The irradiance and, unless the caller passes
// one, the ROI are computed on the GPU in FP64 (doe_target_make,
// doe_target_auto_roi) and returned as host arrays.
// Device spec and display name of a makeTarget kind.
function targetSpec(kind, opts) {
  if (kind === 'image') {
    const { raster, width, height, gamma = 1 } = opts;
    if (!raster) throw new Error('makeTarget("image"): pass {raster, width, height}');
    return { spec: { kind: 'raster', raster: f64Raster(raster, width, height), width, height, gamma },
      name: opts.name === undefined ? 'image' : opts.name };
  }
  if (kind === 'text') {
    const { text = 'DOE', scale = 6, spacing = 1 } = opts;
    const g = renderText(text, { scale, spacing });
    return { spec: { kind: 'raster', raster: g.data, width: g.width, height: g.height }, name: `text:${text}` };
  }
  if (ANALYTIC.has(kind)) {
    return { spec: Object.assign({}, opts, { kind, missing: darkCells(opts.missing), raster: undefined }), name: kind };
  }
  throw new Error(`makeTarget: unknown kind "${kind}"`);
}

// autoRoi options of a target spec (the reference's dilate / softEdge names)
function roiOptions(opts) {
  return {
    dilate: Number(opts.dilate === undefined ? 2 : opts.dilate),
    soft: Number(opts.softEdge === undefined ? 2 : opts.softEdge),
    threshold: 0.05,
  };
}

export function makeTarget(kind, nx, ny, opts = {}) {
  const { spec, name } = targetSpec(kind, opts);
  const dev = opts.roi ? native.targetMake([spec], nx, ny) : native.targetMake([spec], nx, ny, roiOptions(opts));
  const irr = new Float64Array(nx * ny);
  const devRoi = new Float64Array(nx * ny);
  native.fieldDownload(dev, irr, devRoi);
  const roi = opts.roi ? opts.roi : devRoi;
  const grid = kind === 'dots' || kind === 'grid' ? { cols: opts.cols === undefined ? 8 : opts.cols, rows: opts.rows === undefined ? 8 : opts.rows } : null;
  return { irradiance: irr, roi, name, grid };
}

// Irradiance and ROI of many planes (designMultiPlane's specs: a kind with
// its options, or {irradiance, roi?}) with as few device calls as possible:
// the kind planes sharing ROI options are one doe_target_make_roi launch;
// the given irradiances without an ROI are one batched upload and one
// doe_target_auto_roi launch.
function planeTargets(specs, nx, ny) {
  const n = nx * ny;
  const out = new Array(specs.length);
  const groups = new Map(); // ROI options -> [plane index, device spec]
  const custom = [];
  specs.forEach((p, q) => {
    if (p.irradiance) {
      if (p.roi !== undefined && p.roi !== null) out[q] = { irradiance: p.irradiance, roi: p.roi };
      else custom.push(q);
      return;
    }
    const key = JSON.stringify(roiOptions(p));
    if (!groups.has(key)) groups.set(key, []);
    groups.get(key).push([q, targetSpec(p.kind, p).spec]);
  });
  for (const [key, members] of groups) {
    const dev = native.targetMake(members.map((m) => m[1]), nx, ny, JSON.parse(key));
    const irr = new Float64Array(n * members.length);
    const roi = new Float64Array(n * members.length);
    native.fieldDownload(dev, irr, roi);
    members.forEach(([q], b) => {
      out[q] = { irradiance: irr.subarray(b * n, (b + 1) * n), roi: roi.subarray(b * n, (b + 1) * n) };
    });
  }
  if (custom.length > 0) {
    const dev = native.fieldCreate(nx, ny, custom.length, 1);
    const irr = new Float64Array(n * custom.length);
    custom.forEach((q, b) => irr.set(f64Plane(specs[q].irradiance).subarray(0, n), b * n));
    native.fieldUpload(dev, irr, irr);
    native.targetAutoRoi(dev, 2, 2, 0.05);
    const roi = new Float64Array(n * custom.length);
    native.fieldDownload(dev, irr, roi);
    custom.forEach((q, b) => {
      out[q] = { irradiance: specs[q].irradiance, roi: roi.subarray(b * n, (b + 1) * n) };
    });
  }
  return out;
}

// Threshold, separable dilation, box-blurred edge (ref: projector.js:262-315),
// on the GPU: the ROI equals the reference's bit for bit.
export function autoRoi(irr, nx, ny, { dilate = 2, soft = 2, threshold = 0.05 } = {}) {
  const dev = native.fieldCreate(nx, ny, 1, 1);
  const plane = f64Plane(irr);
  native.fieldUpload(dev, plane, plane);
  native.targetAutoRoi(dev, Number(dilate), Number(soft), Number(threshold));
  const out = new Float64Array(nx * ny);
  native.fieldDownload(dev, new Float64Array(nx * ny), out);
  return out;
}

export const DEFAULT_PROJECTOR = {
  n: 256,
  pitch: 8e-6,
  lambda: 532e-9,
  distance: 0.03,
  aperture: 0.95,
  apodization: 2,
  source: 'tophat',
  waist: 0.6,
  inner: 0.6,
  outer: 0.98,
  mode: 'image-plane',
  algorithm: 'auto',
  iterations: 60,
  levels: 0,
  seed: 12345,
  quantize: false,
  target: { kind: 'text', text: 'DOE' },
};

// The hot-path methods of ProjectorSystem (ref: projector.js:438-498,
// 554-653, 675-712) on the device.
class DeviceProjector {
  // ref: projector.js:438-467
  design(opts = {}) {
    const pick = (v, d) => (v === undefined || v === null ? d : v);
    const result = designDOE({
      nx: this.nx,
      ny: this.ny,
      dx: this.dx,
      dy: this.dy,
      lambda: this.lambda,
      z: this.z,
      mode: this.cfg.mode,
      algorithm: this.cfg.algorithm,
      iterations: pick(opts.iterations, this.cfg.iterations),
      levels: pick(opts.levels, this.cfg.levels),
      seed: pick(opts.seed, this.cfg.seed),
      illumination: this.illumination(),
      target: this.target.irradiance,
      roi: this.target.roi,
      onProgress: opts.onProgress,
      progressEvery: pick(opts.progressEvery, 10),
      beta: pick(opts.beta, 0.9),
      escape: opts.escape,
      initPhase: opts.initPhase,
    });
    result.iterations = pick(opts.iterations, this.cfg.iterations);
    result.target = this.target;
    this.result = result;
    return result;
  }

  // ref: projector.js:475-498
  simulate({ z = null, phase = null, bandLimit = false, pad = 1, mode = null } = {}) {
    if (!this.result) throw new Error('ProjectorSystem.simulate: call design() first');
    const dist = z === null ? this.z : z;
    const useMode = mode === null ? this.cfg.mode : mode;
    const r = this.result;
    if (useMode === 'far-field') {
      if (!(dist > 0)) throw new Error('ProjectorSystem.simulate: far-field mode needs a positive throw');
      const field = new ComplexField(this.nx, this.ny, this.dx, this.dy);
      field.setAmplitudePhase(r.illumination, phase || r.phase);
      return focalPlaneSpectrum(field, this.lambda, dist);
    }
    let field = new ComplexField(this.nx, this.ny, this.dx, this.dy);
    field.setAmplitudePhase(r.illumination, phase || r.phase);
    if (pad > 1) field = field.resample(this.nx * pad, this.ny * pad);
    propagateAngularSpectrumInPlace(field, this.lambda, dist, { bandLimit });
    if (pad > 1) field = field.resample(this.nx, this.ny);
    return field;
  }

  // Device copy of a target's irradiance (re) and ROI (im), FP64, cached per
  // (irradiance, roi) pair and re-uploaded only when the arrays or their
  // contents changed since the last call (callers may edit a target in place
  // between metric calls).
  _targetDevice(irradiance, mask) {
    const c = this._tgtDev;
    const m = mask || new Float64Array(irradiance.length).fill(1);
    if (c && c.irr === irradiance && c.mask === mask && sameValues(c.irrCopy, irradiance) &&
        sameValues(c.maskCopy, m)) {
      return c.dev;
    }
    const dev = native.fieldCreate(this.nx, this.ny, 1, 1);
    const irr64 = Float64Array.from(irradiance);
    const m64 = Float64Array.from(m);
    native.fieldUpload(dev, irr64, m64);
    this._tgtDev = { irr: irradiance, mask, dev, irrCopy: irr64, maskCopy: m64 };
    return dev;
  }

  // ref: projector.js:554-614.  Every figure is reduced on the GPU in FP64
  // (doe_screen_metrics); the incident power is the element's sum |E| dx dy,
  // i.e. that of the design's illumination.
  screenMetrics(field, { target = null, roi = null, phase = null } = {}) {
    const a = this._metricsArgs(target, roi, phase);
    return native.screenMetrics(field._device(), a.tgt, a.el, field.dx, field.dy, this.dx, this.dy,
      a.cols, a.rows, a.minPeak, true);
  }

  // target device copy, FP64 element (its sums keep the host illumination's
  // precision) and spot-grid options shared by the metric entry points
  _metricsArgs(target, roi, phase) {
    const tgt = target || this.target;
    const mask = roi || tgt.roi;
    const element = new ComplexField(this.nx, this.ny, this.dx, this.dy);
    element.setAmplitudePhase(this.result.illumination, phase || this.result.phase);
    const grid = tgt.grid || null;
    return {
      tgt: this._targetDevice(tgt.irradiance, mask),
      el: element._device(1),
      cols: grid ? (grid.cols === undefined ? 8 : grid.cols) : 0,
      rows: grid ? (grid.rows === undefined ? 8 : grid.rows) : 0,
      minPeak: grid && grid.minPeak !== undefined ? grid.minPeak : 0.05,
    };
  }

  // Image-plane screens of simulate({ z, pad }) for every z, as one device
  // batch of nx x ny fields: the element is padded and broadcast once and one
  // angular-spectrum call propagates every slot by its own distance
  // (ref: projector.js:492-497, once per z there).
  _screensBatch(zs, pad) {
    if (!this.result) throw new Error('ProjectorSystem.simulate: call design() first');
    const r = this.result;
    const dt = dtypeCode();
    const el = new ComplexField(this.nx, this.ny, this.dx, this.dy);
    el.setAmplitudePhase(r.illumination, r.phase);
    const P = zs.length;
    const big = pad > 1
      ? native.fieldCreate(this.nx * pad, this.ny * pad, P, dt)
      : native.fieldCreate(this.nx, this.ny, P, dt);
    native.fieldResample(el._device(dt), big);
    native.asm(big, big, this.dx, this.dy, this.lambda, Float64Array.from(zs), false, true, 0.02);
    if (!(pad > 1)) return big;
    const small = native.fieldCreate(this.nx, this.ny, P, dt);
    native.fieldResample(big, small);
    return small;
  }

  // ref: projector.js:622-641 (image plane: all distances in one batch,
  // metrics reduced on the device; far field: one lens transform per z)
  defocusCurve({ points = 21, span = 0.6, pad = 2, mode = null } = {}) {
    const z0 = this.z * (1 - span);
    const z1 = this.z * (1 + span);
    const zs = [];
    for (let i = 0; i < points; i++) zs.push(z0 + ((z1 - z0) * i) / (points - 1));
    const entry = (z, m) => ({ z, zMm: z * 1e3, fidelity: m.fidelity, correlation: m.correlation, ssim: m.ssim, rmse: m.rmse, efficiency: m.efficiency, uniformity: m.uniformity });
    const useMode = mode === null ? this.cfg.mode : mode;
    if (useMode === 'far-field' || zs.length === 0) {
      return zs.map((z) => entry(z, this.screenMetrics(this.simulate({ z, pad, mode }))));
    }
    const screens = this._screensBatch(zs, pad);
    const a = this._metricsArgs(null, null, null);
    const ms = native.screenMetricsBatch(screens, a.tgt, a.el, this.dx, this.dy, this.dx, this.dy,
      a.cols, a.rows, a.minPeak, true);
    return zs.map((z, i) => entry(z, ms[i]));
  }

  // ref: projector.js:644-653 (image plane: one batch)
  focusStack({ points = 9, span = 0.5, pad = 2, mode = null } = {}) {
    const zs = [];
    for (let i = 0; i < points; i++) zs.push(this.z * (1 - span) + (this.z * 2 * span * i) / (points - 1));
    const useMode = mode === null ? this.cfg.mode : mode;
    if (useMode === 'far-field' || zs.length === 0) {
      return zs.map((z) => ({ z, intensity: this.simulate({ z, pad, mode }).intensity() }));
    }
    const screens = this._screensBatch(zs, pad);
    const n = this.nx * this.ny;
    const re = new Float64Array(n * zs.length);
    const im = new Float64Array(n * zs.length);
    native.fieldDownload(screens, re, im);
    return zs.map((z, i) => {
      const I = new Float64Array(n);
      for (let k = 0, q = i * n; k < n; k++, q++) I[k] = re[q] * re[q] + im[q] * im[q];
      return { z, intensity: I };
    });
  }

  // ref: projector.js:675-712.  Every plane's target and ROI come from as few
  // batched device calls as the planes allow (planeTargets); the loop runs on
  // the device (doe_gs_multiplane_run).  The result carries the reference's
  // fields (the loop's phase and histories, the planes, the illumination and
  // the grid), and becomes this.result like a design()'s.
  designMultiPlane(planeSpecs, opts = {}) {
    const n = this.nx * this.ny;
    const illumination = this.illumination();
    const made = planeTargets(planeSpecs, this.nx, this.ny);
    const planes = planeSpecs.map((p, q) => {
      const amp = new Float64Array(n);
      const irr = made[q].irradiance;
      for (let k = 0; k < n; k++) amp[k] = irr[k] > 0 ? Math.sqrt(irr[k]) : 0;
      return { z: p.z, targetAmp: amp, roi: made[q].roi, weight: pick(p.weight, 1) };
    });
    const loop = gerchbergSaxtonMultiPlane(illumination, planes, {
      nx: this.nx, ny: this.ny, dx: this.dx, dy: this.dy, lambda: this.lambda,
      iterations: pick(opts.iterations, this.cfg.iterations),
      seed: pick(opts.seed, this.cfg.seed),
      initPhase: opts.initPhase,
      onProgress: opts.onProgress,
    });
    const levels = pick(opts.levels, this.cfg.levels);
    const result = Object.assign({}, loop, {
      illumination, planes, mode: 'multi-plane', target: this.target, roi: this.target.roi,
      lambda: this.lambda, z: this.z, nx: this.nx, ny: this.ny, dx: this.dx, dy: this.dy,
    });
    if (levels > 1) result.phase = errorDiffusionQuantize(result.phase, levels);
    this.result = result;
    return result;
  }
}

// Put the device methods on a ProjectorSystem-like class (the reference's
// own, or the one below); returns the class.
export function installDeviceProjector(Cls) {
  for (const k of Object.getOwnPropertyNames(DeviceProjector.prototype)) {
    if (k !== 'constructor') {
      Object.defineProperty(Cls.prototype, k, Object.getOwnPropertyDescriptor(DeviceProjector.prototype, k));
    }
  }
  return Cls;
}

export class ProjectorSystem {
  // ref: projector.js:347-366
  constructor(cfg = {}) {
    this.cfg = Object.assign({}, DEFAULT_PROJECTOR, cfg, {
      target: Object.assign({}, DEFAULT_PROJECTOR.target, cfg.target || {}),
    });
    this.nx = this.cfg.n;
    this.ny = this.cfg.n;
    this.dx = this.cfg.pitch;
    this.dy = this.cfg.pitch;
    this.lambda = this.cfg.lambda;
    if (cfg.distance === undefined && this.cfg.mode === 'far-field') {
      this.cfg.distance = (this.cfg.n * this.cfg.pitch * this.cfg.pitch) / this.cfg.lambda;
    }
    this.z = this.cfg.distance;
    this.target = null;
    this.result = null;
    this.setTarget(this.cfg.target);
  }

  setTarget(spec) {
    if (spec && spec.irradiance) {
      this.target = {
        irradiance: spec.irradiance,
        roi: spec.roi ? spec.roi : autoRoi(spec.irradiance, this.nx, this.ny),
        name: spec.name === undefined ? 'custom' : spec.name,
      };
    } else {
      this.target = makeTarget(spec.kind, this.nx, this.ny, spec);
    }
    return this.target;
  }

  matchDistance() {
    return (this.nx * this.dx * this.dx) / this.lambda;
  }

  // ref: projector.js:387-397
  illumination() {
    const radius = (Math.min(this.nx * this.dx, this.ny * this.dy) / 2) * this.cfg.aperture;
    return illuminationProfile(this.cfg.source, this.nx, this.ny, this.dx, this.dy, {
      radius,
      soft: this.cfg.apodization * this.dx,
      waist: radius * this.cfg.waist,
      inner: this.cfg.inner === undefined ? 0.6 : this.cfg.inner,
      outer: this.cfg.outer === undefined ? 0.98 : this.cfg.outer,
    });
  }
}
installDeviceProjector(ProjectorSystem);
