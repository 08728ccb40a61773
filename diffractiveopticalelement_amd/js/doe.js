// Drop-in for the phase-retrieval part of src/doe.js.  The GS / HIO /
// weighted-GS / Fourier-CGH loops run entirely on the GPU (one call of the
// addon per design: every iteration stays in HBM); the orchestration of
// designDOE (target power matching, initial phase, algorithm choice) is the
// host JavaScript of the reference's contract.
import { ComplexField, discMask, rectMask } from './field.js';
import { propagateAngularSpectrumInPlace, lensPhase, wrapPhase } from './propagate.js';
import { native, loopDtype } from './native.js';

export const DEFAULT_DESIGN = {
  lambda: 532e-9,
  z: 0.03,
  iterations: 60,
  seed: 12345,
  feedback: 0,
  beta: 0.9,
  historyEvery: 1,
  randomRestarts: 1,
  onProgress: null,
};

const MODE_IMAGE_PLANE = 0;
const MODE_FAR_FIELD = 1;
const ALG_GS = 0;
const ALG_WGS = 1;

function def(v, d) {
  return v === undefined || v === null ? d : v;
}

// xorshift32, seed>>>0 || 1 (ref: doe.js:65-75)
export function makeRng(seed = 1) {
  let s = seed >>> 0 || 1;
  return function next() {
    s ^= s << 13;
    s >>>= 0;
    s ^= s >>> 17;
    s ^= s << 5;
    s >>>= 0;
    return s / 4294967296;
  };
}

// (u*2 - 1)*pi from the xorshift stream (ref: doe.js:81-86), generated on
// the GPU (doe_random_phase: each thread jumps ahead to its chunk of the
// stream), equal to the host stream bit for bit.
export function randomPhase(nx, ny, seed = 1) {
  const n = nx * ny;
  const p = new Float64Array(n);
  if (n === 0) return p;
  const f = native.fieldCreate(n, 1, 1, 1);
  native.randomPhase(f, seed >>> 0);
  native.fieldDownload(f, p, new Float64Array(n));
  return p;
}

export function flatPhase(nx, ny) {
  return new Float64Array(nx * ny);
}

export function kinoformLens(nx, ny, dx, dy, lambda, f) {
  return lensPhase(new ComplexField(nx, ny, dx, dy), lambda, f, { wrap: true });
}

// ref: doe.js:167-174
export function quantizePhase(phi, levels = 4) {
  const step = (2 * Math.PI) / levels;
  const out = new Float64Array(phi.length);
  for (let k = 0; k < phi.length; k++) out[k] = Math.round(phi[k] / step) * step;
  return wrapPhase(out);
}

// Floyd–Steinberg on the wrapped phase (ref: doe.js:185-226), on the GPU
// (doe_error_diffusion).  Plain left-to-right rows (snake: false) run as a
// diagonal wavefront (row j trails row j-1 by two samples); serpentine rows
// (the default) make the whole image one serial chain, which runs in one
// workgroup that stages rows through LDS.
export function errorDiffusionQuantize(phi, levels = 4, { snake = true, wrapFirst = true } = {}) {
  const n = phi.length;
  const nx = Math.round(Math.sqrt(n));
  const ny = n / nx;
  // a non-integer row count passes the reference's check and reads past the
  // array there; here it throws with the reference's message
  if (nx * ny !== n || !Number.isInteger(ny)) throw new Error('errorDiffusionQuantize: expected a square grid');
  const plane = phi instanceof Float64Array ? phi : Float64Array.from(phi);
  const src = native.fieldCreate(nx, ny, 1, 1);
  native.fieldUpload(src, plane, plane);
  const dst = native.fieldCreate(nx, ny, 1, 1);
  native.errorDiffusion(src, dst, levels, snake, wrapFirst);
  const out = new Float64Array(n);
  native.fieldDownload(dst, out, new Float64Array(n));
  return out;
}

// Etch depth in [0, lambda / (n - 1)) (ref: doe.js:967-976), on the GPU
// (doe_phase_to_height); the phase travels as one row.
export function phaseToHeight(phase, lambda, n = 1.52) {
  const len = phase.length;
  const h = new Float64Array(len);
  if (len === 0) return h;
  const plane = phase instanceof Float64Array ? phase : Float64Array.from(phase);
  const src = native.fieldCreate(len, 1, 1, 1);
  native.fieldUpload(src, plane, plane);
  const dst = native.fieldCreate(len, 1, 1, 1);
  native.phaseToHeight(src, dst, lambda, n);
  native.fieldDownload(dst, h, new Float64Array(len));
  return h;
}

// Annular mask (ref: doe.js:288-305), used by illuminationProfile('annulus').
export function ringMask(nx, ny, inner = 0.5, outer = 0.95, { soft = 0.02 } = {}) {
  const m = new Float64Array(nx * ny);
  const lo = inner + soft;
  const hi = outer - soft;
  for (let j = 0; j < ny; j++) {
    const v = (j - ny / 2) / (ny / 2);
    for (let i = 0; i < nx; i++) {
      const r = Math.hypot((i - nx / 2) / (nx / 2), v);
      let val = 0;
      if (r >= lo && r <= hi) val = 1;
      else if (r > inner && r < lo) val = 0.5 * (1 - Math.cos((Math.PI * (r - inner)) / soft));
      else if (r > hi && r < outer) val = 0.5 * (1 + Math.cos((Math.PI * (r - hi)) / soft));
      m[i + j * nx] = val;
    }
  }
  return m;
}

// ---------------------------------------------------------------------------
// device loops
// ---------------------------------------------------------------------------
function geometry(o, illumination) {
  const nx = def(o.nx, Math.round(Math.sqrt(illumination.length)));
  const ny = def(o.ny, nx);
  const dx = def(o.dx, 8e-6);
  const dy = def(o.dy, dx);
  return { nx, ny, dx, dy };
}

function f64(a) {
  return a instanceof Float64Array ? a : Float64Array.from(a);
}

// Target amplitude with caller-supplied weights folded in: the loops only
// ever use targetAmp[k] * weights[k] (ref: doe.js:369, 383, 681, 703).
function weightedTarget(targetAmp, weights) {
  if (!weights) return f64(targetAmp);
  const t = new Float64Array(targetAmp.length);
  for (let k = 0; k < t.length; k++) t[k] = targetAmp[k] * weights[k];
  return t;
}

function replayProgress(o, error, efficiency, n) {
  if (!o.onProgress) return;
  const every = def(o.progressEvery, 10);
  for (let it = 0; it < n; it++) {
    if (it % every === 0 || it === n - 1) o.onProgress({ iteration: it, error: error[it], efficiency: efficiency[it] });
  }
}

function runDevice(mode, algorithm, illumination, targetAmp, o, initPhase) {
  const g = geometry(o, illumination);
  const plan = {
    mode,
    algorithm,
    iterations: o.iterations,
    nx: g.nx,
    ny: g.ny,
    dx: g.dx,
    dy: g.dy,
    lambda: o.lambda,
    z: o.z,
    beta: def(o.beta, 0.9),
    feedback: def(o.feedback, 0),
    escape: def(o.escape, 0),
    dtype: loopDtype(),
  };
  const roi = o.roi ? f64(o.roi) : null;
  const seed = def(o.seed, 12345);
  const r = native.gsRun(plan, f64(illumination), targetAmp, roi, initPhase ? f64(initPhase) : null, seed);
  return { r, g };
}

// Image-plane GS / HIO (ref: doe.js:327-437).
export function gerchbergSaxton(illumination, targetAmp, opts = {}) {
  const o = Object.assign({}, DEFAULT_DESIGN, opts);
  const t = weightedTarget(targetAmp, def(o.weights, null));
  const { r, g } = runDevice(MODE_IMAGE_PLANE, ALG_GS, illumination, t, o, o.initPhase || null);
  const error = Array.from(r.error);
  const efficiency = Array.from(r.efficiency);
  replayProgress(o, error, efficiency, o.iterations);
  // the reference also returns the final forward field
  const field = new ComplexField(g.nx, g.ny, g.dx, g.dy);
  field.setAmplitudePhase(illumination, r.phase);
  propagateAngularSpectrumInPlace(field, o.lambda, o.z, { bandLimit: false });
  return {
    phase: r.phase,
    error,
    efficiency,
    amplitude: r.amplitude,
    intensity: r.intensity,
    iterations: o.iterations,
    field,
  };
}

// Far-field loop (ref: doe.js:639-730).
export function fourierCgh(illumination, targetAmp, opts = {}) {
  const o = Object.assign({}, DEFAULT_DESIGN, opts);
  const t = weightedTarget(targetAmp, def(o.weights, null));
  const { r } = runDevice(MODE_FAR_FIELD, ALG_GS, illumination, t, o, o.initPhase || null);
  const error = Array.from(r.error);
  const efficiency = Array.from(r.efficiency);
  replayProgress(o, error, efficiency, o.iterations);
  return { phase: r.phase, error, efficiency, amplitude: r.amplitude, intensity: r.intensity, iterations: o.iterations, field: null };
}

// Multi-plane GS (ref: doe.js:557-626): one element, a different target at
// each distance.  The whole loop runs on the device (doe_gs_multiplane_run):
// per plane an angular-spectrum step to z, the GS projection on the plane's
// target inside its ROI and the step back, bandLimit false, flat initial
// phase unless initPhase is given.
export function gerchbergSaxtonMultiPlane(illumination, planes, opts = {}) {
  const o = Object.assign({}, DEFAULT_DESIGN, opts);
  const g = geometry(o, illumination);
  if (planes.length === 0) throw new Error('gerchbergSaxtonMultiPlane: no planes');
  const spec = planes.map((p) => ({
    z: p.z,
    targetAmp: f64(p.targetAmp),
    roi: p.roi ? f64(p.roi) : null,
    weight: def(p.weight, 1),
  }));
  const r = native.gsMultiPlane(f64(illumination), spec, g.nx, g.ny, g.dx, g.dy, o.lambda, o.iterations,
    loopDtype(), o.initPhase ? f64(o.initPhase) : null);
  const error = Array.from(r.error);
  const efficiency = Array.from(r.efficiency);
  const perPlaneError = r.perPlaneError.map((a) => Array.from(a));
  if (o.onProgress) {
    const every = def(o.progressEvery, 10);
    for (let it = 0; it < o.iterations; it++) {
      if (it % every === 0 || it === o.iterations - 1) {
        o.onProgress({ iteration: it, error: error[it], perPlane: perPlaneError.map((a) => a[it]) });
      }
    }
  }
  return { phase: r.phase, error, efficiency, perPlaneError, iterations: o.iterations };
}

// Weighted GS (ref: doe.js:455-546).  With the drop-in's own base loops the
// whole weighted iteration runs on the device; any other base function runs
// the reference's host-side wrapper around it.
export function weightedGerchbergSaxton(illumination, targetAmp, opts = {}, baseFn = gerchbergSaxton) {
  const o = Object.assign({}, DEFAULT_DESIGN, opts);
  if (baseFn === gerchbergSaxton || baseFn === fourierCgh) {
    const mode = baseFn === fourierCgh ? MODE_FAR_FIELD : MODE_IMAGE_PLANE;
    const { r } = runDevice(mode, ALG_WGS, illumination, f64(targetAmp), o, o.initPhase || null);
    const error = Array.from(r.error);
    const efficiency = Array.from(r.efficiency);
    replayProgress(o, error, efficiency, o.iterations);
    return {
      phase: r.phase,
      error,
      efficiency,
      amplitude: r.amplitude,
      intensity: r.intensity,
      iterations: o.iterations,
      field: null,
      weightedIterations: o.iterations,
      weights: r.weights,
      rmse: r.rmse,
    };
  }
  return weightedHost(illumination, targetAmp, o, baseFn);
}

function weightedHost(illumination, targetAmp, o, baseFn) {
  const n = illumination.length;
  const inRoi = (k) => (o.roi ? o.roi[k] : 1) > 0;
  const weights = new Float64Array(n);
  for (let k = 0; k < n; k++) if (inRoi(k)) weights[k] = 1;
  let result = null;
  let bestPhase = null;
  let bestRmse = Infinity;
  let best = null;
  const hist = { error: [], efficiency: [] };
  for (let it = 0; it < o.iterations; it++) {
    result = baseFn(illumination, targetAmp, Object.assign({}, o, {
      weights, iterations: 1, initPhase: result ? result.phase : o.initPhase, onProgress: null,
    }));
    let num = 0;
    let den = 0;
    for (let k = 0; k < n; k++) {
      if (!inRoi(k)) continue;
      const d = result.amplitude[k] - targetAmp[k];
      num += d * d;
      den += targetAmp[k] * targetAmp[k];
    }
    const rmse = den > 0 ? Math.sqrt(num / den) : 0;
    if (rmse < bestRmse) {
      bestRmse = rmse;
      bestPhase = Float64Array.from(result.phase);
      best = { amplitude: Float64Array.from(result.amplitude), intensity: Float64Array.from(result.intensity) };
    }
    const beta = def(o.beta, 0.9);
    let wsum = 0;
    let tsum = 0;
    for (let k = 0; k < n; k++) {
      if (!inRoi(k)) continue;
      const a = result.amplitude[k];
      const t = targetAmp[k];
      const ratio = a > 1e-30 ? t / a : 2;
      weights[k] *= Math.pow(Math.min(Math.max(ratio, 0.25), 4), beta);
      weights[k] = Math.min(Math.max(weights[k], 1e-4), 1e4);
      wsum += weights[k] * weights[k] * t * t;
      tsum += t * t;
    }
    const s = wsum > 0 ? Math.sqrt(tsum / wsum) : 1;
    for (let k = 0; k < n; k++) weights[k] *= s;
    hist.error.push(bestRmse);
    hist.efficiency.push(result.efficiency[result.efficiency.length - 1]);
  }
  replayProgress(o, hist.error, hist.efficiency, o.iterations);
  if (bestPhase) {
    result = Object.assign({}, result, {
      iterations: o.iterations,
      error: hist.error,
      efficiency: hist.efficiency,
      phase: bestPhase,
      amplitude: best.amplitude,
      intensity: best.intensity,
      weightedIterations: o.iterations,
    });
  }
  result.weights = weights;
  result.rmse = bestRmse;
  return result;
}

// Separable clamped Gaussian blur (ref: doe.js:761-794) on the GPU
// (doe_raster_blur).  The taps are built here exactly as the reference
// builds them, so the device sums the same products in the same order.
export function blurRaster(src, nx, ny, sigmaPx = 1) {
  if (sigmaPx <= 0) return Float64Array.from(src);
  const r = Math.max(1, Math.ceil(3 * sigmaPx));
  const ker = new Float64Array(2 * r + 1);
  let ks = 0;
  for (let i = -r; i <= r; i++) ks += ker[i + r] = Math.exp(-(i * i) / (2 * sigmaPx * sigmaPx));
  for (let i = 0; i < ker.length; i++) ker[i] /= ks;
  const plane = src instanceof Float64Array ? src : Float64Array.from(src);
  if (plane.length !== nx * ny) throw new Error(`blurRaster: raster length ${plane.length} does not match ${nx}x${ny}`);
  const dev = native.fieldCreate(nx, ny, 1, 1);
  native.fieldUpload(dev, plane, plane);
  native.rasterBlur(dev, dev, sigmaPx, ker);
  const out = new Float64Array(nx * ny);
  native.fieldDownload(dev, out, new Float64Array(nx * ny));
  return out;
}

function discTarget(nx, ny, frac) {
  const t = new Float64Array(nx * ny);
  const r = frac * Math.min(nx, ny) * 0.5;
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) t[i + j * nx] = Math.hypot(i - (nx - 1) / 2, j - (ny - 1) / 2) <= r ? 1 : 0;
  }
  return t;
}

// ref: doe.js:796-896
export function designDOE(cfg) {
  const nx = def(cfg.nx, 256);
  const ny = def(cfg.ny, nx);
  const dx = def(cfg.dx, 8e-6);
  const dy = def(cfg.dy, dx);
  const lambda = def(cfg.lambda, 532e-9);
  const z = def(cfg.z, 0.03);
  const mode = def(cfg.mode, 'image-plane');
  const algorithm = cfg.algorithm && cfg.algorithm !== 'auto' ? cfg.algorithm : mode === 'far-field' ? 'gs' : 'wgs';
  const levels = def(cfg.levels, 0);
  const iterations = def(cfg.iterations, 60);
  const illumination = def(cfg.illumination, discMask(nx, ny, dx, dy, (Math.min(nx * dx, ny * dy) / 2) * 0.95, { soft: 2 * dx }));
  const roi = def(cfg.roi, null);
  let irr = def(cfg.target, discTarget(nx, ny, 0.35));
  if (cfg.smooth > 0) irr = blurRaster(irr, nx, ny, cfg.smooth);
  const targetAmp = new Float64Array(nx * ny);
  for (let k = 0; k < targetAmp.length; k++) targetAmp[k] = Math.sqrt(Math.max(0, irr[k]));
  // match the demanded power to the available power (discrete sums)
  let pin = 0;
  for (let k = 0; k < illumination.length; k++) pin += illumination[k] * illumination[k];
  let pt = 0;
  for (let k = 0; k < targetAmp.length; k++) pt += targetAmp[k] * targetAmp[k];
  if (pt > 0) {
    const s = Math.sqrt(pin / pt);
    for (let k = 0; k < targetAmp.length; k++) targetAmp[k] *= s;
  }
  let initPhase = cfg.initPhase;
  if (!initPhase) {
    const init = def(cfg.init, 'random');
    if (init === 'flat') initPhase = flatPhase(nx, ny);
    else if (init === 'lens') initPhase = kinoformLens(nx, ny, dx, dy, lambda, z);
    else initPhase = randomPhase(nx, ny, def(cfg.seed, 12345));
  }
  const common = {
    nx, ny, dx, dy, lambda, z, roi, iterations,
    seed: def(cfg.seed, 12345),
    feedback: algorithm === 'hio' ? def(cfg.feedback, 0.8) : 0,
    escape: def(cfg.escape, Math.round(iterations * 0.2)),
    beta: def(cfg.beta, 0.9),
    initPhase,
    onProgress: cfg.onProgress,
    progressEvery: def(cfg.progressEvery, 10),
  };
  const base = mode === 'far-field' ? fourierCgh : gerchbergSaxton;
  const result = algorithm === 'wgs'
    ? weightedGerchbergSaxton(illumination, targetAmp, common, base)
    : base(illumination, targetAmp, common);
  result.mode = mode === 'far-field' ? 'far-field' : 'image-plane';
  result.algorithm = algorithm;
  if (levels > 1) {
    result.phaseContinuous = Float64Array.from(result.phase);
    result.phase = errorDiffusionQuantize(result.phase, levels);
    result.levels = levels;
    result.quantized = true;
  }
  Object.assign(result, { illumination, targetAmp, roi, lambda, z, nx, ny, dx, dy });
  return result;
}

// ref: doe.js:924-932
export function simulateDOE(result, { phase = null, z = null, lambda = null, bandLimit = false } = {}) {
  const field = new ComplexField(result.nx, result.ny, result.dx, result.dy);
  field.setAmplitudePhase(result.illumination, phase || result.phase);
  propagateAngularSpectrumInPlace(field, lambda === null ? result.lambda : lambda, z === null ? result.z : z, { bandLimit });
  return field;
}

// ref: doe.js:979-1000
export function illuminationProfile(type, nx, ny, dx, dy, opts = {}) {
  const radius = def(opts.radius, (Math.min(nx * dx, ny * dy) / 2) * 0.95);
  const ap = discMask(nx, ny, dx, dy, radius, { soft: def(opts.soft, 2 * dx) });
  if (type === 'tophat') return ap;
  if (type === 'gaussian') {
    const out = new Float64Array(nx * ny);
    const w = def(opts.waist, radius * 0.6);
    for (let j = 0; j < ny; j++) {
      const y = (j - ny / 2) * dy;
      for (let i = 0; i < nx; i++) {
        const x = (i - nx / 2) * dx;
        out[i + j * nx] = ap[i + j * nx] * Math.exp(-(x * x + y * y) / (w * w));
      }
    }
    return out;
  }
  if (type === 'annulus') return ringMask(nx, ny, def(opts.inner, 0.6), def(opts.outer, 0.98), { soft: 0.02 });
  if (type === 'rect') return rectMask(nx, ny, dx, dy, def(opts.width, nx * dx * 0.9), def(opts.height, ny * dy * 0.9), { soft: 2 * dx });
  return ap;
}
