// ORACLE (test infrastructure only) — FP64 restatement of the reference's
// phase-retrieval loops (ref: src/doe.js:49-86, 327-546, 639-730, 761-905,
// 979-1000).  ES2019.
//
// Pinning: the FFT/propagator parts are pinned closed-form (see propagate.js).
// tests/test_oracle_cpu.py checks the loops against the reference's own
// behavioural tests (ref test/doe.test.js, test/projector.test.js
// thresholds), the published design results (ref README.md:29, 35:
// image-plane 'DOE' rmse 0.0064 / efficiency 0.78, far-field rmse 8.2 %,
// rho 0.991, efficiency 81 %) and first-iteration identities;
// tests/test_oracle_pins_cpu.py checks each step against numpy written from
// its definition: the HIO relaxation and escape period, the far-field WGS
// weight update (clamp, power, clamp, renormalisation) over two iterations,
// the best-iterate bookkeeping over nine, the a = 0 projection branch and
// the illumination profiles; tests/test_oracle_mutants_cpu.py plants slips
// in a copy of this file and checks that those pins fail.

import { ComplexField, discMask } from './field.js';
import { propagateAngularSpectrumInPlace, lensPhase, wrapPhase } from './propagate.js';
import { fft2d, fftshift2d, ifftshift2d } from './fft.js';

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

function pick(v, d) {
  return v === undefined || v === null ? d : v;
}

// xorshift32 with seed>>>0 || 1 (ref: doe.js:65-75)
export function makeRng(seed) {
  let s = (pick(seed, 1) >>> 0) || 1;
  return function next() {
    s ^= s << 13;
    s >>>= 0;
    s ^= s >>> 17;
    s ^= s << 5;
    s >>>= 0;
    return s / 4294967296;
  };
}

// ref: doe.js:81-86
export function randomPhase(nx, ny, seed) {
  const rng = makeRng(pick(seed, 1));
  const p = new Float64Array(nx * ny);
  for (let k = 0; k < p.length; k++) p[k] = (rng() * 2 - 1) * Math.PI;
  return p;
}

export function flatPhase(nx, ny) {
  return new Float64Array(nx * ny);
}

export function kinoformLens(nx, ny, dx, dy, lambda, f) {
  return lensPhase(new ComplexField(nx, ny, dx, dy), lambda, f, { wrap: true });
}

function merged(opts) {
  const o = {};
  Object.keys(DEFAULT_DESIGN).forEach(function (k) { o[k] = DEFAULT_DESIGN[k]; });
  Object.keys(opts || {}).forEach(function (k) { o[k] = opts[k]; });
  return o;
}

// Image-plane GS / HIO loop (ref: doe.js:327-437).
export function gerchbergSaxton(illumination, targetAmp, opts) {
  const o = merged(opts);
  const nx = pick(o.nx, Math.round(Math.sqrt(illumination.length)));
  const ny = pick(o.ny, nx);
  const dx = pick(o.dx, 8e-6);
  const dy = pick(o.dy, dx);
  const lambda = o.lambda;
  const z = o.z;
  const iterations = o.iterations;
  const roi = pick(o.roi, null);
  const weights = pick(o.weights, null);
  const doe = new ComplexField(nx, ny, dx, dy);
  const plane = new ComplexField(nx, ny, dx, dy);
  const phase = o.initPhase ? Float64Array.from(o.initPhase) : flatPhase(nx, ny);
  const aT = new Float64Array(nx * ny);
  const error = [];
  const efficiency = [];
  for (let it = 0; it < iterations; it++) {
    doe.setAmplitudePhase(illumination, phase);
    plane.copyFrom(doe);
    propagateAngularSpectrumInPlace(plane, lambda, z, { bandLimit: false });
    let eN = 0;
    let eD = 0;
    let inRoi = 0;
    let total = 0;
    for (let k = 0; k < plane.size; k++) {
      const a = Math.hypot(plane.re[k], plane.im[k]);
      const i2 = a * a;
      aT[k] = a;
      total += i2;
      const m = roi ? roi[k] : 1;
      const t = targetAmp[k] * (weights ? weights[k] : 1);
      if (m > 0) {
        inRoi += i2 * m;
        if (t > 0 || i2 > 0) {
          const d = a * m - t * m;
          eN += d * d;
          eD += t * t;
        }
      }
    }
    error.push(eD > 0 ? Math.sqrt(eN / eD) : 0);
    efficiency.push(total > 0 ? (inRoi * dx * dy) / (total * dx * dy) : 0);
    const fb = it >= pick(o.escape, 0) ? o.feedback : 0;
    for (let k = 0; k < plane.size; k++) {
      const m = roi ? roi[k] : 1;
      const t = targetAmp[k] * (weights ? weights[k] : 1);
      const a = aT[k];
      let na;
      if (m > 0) {
        na = t;
        if (fb > 0 && a > 0) na = t + fb * (a - t);
      } else {
        na = a * (1 - fb) + fb * a;
      }
      if (a > 0) {
        const s = na / a;
        plane.re[k] *= s;
        plane.im[k] *= s;
      }
    }
    propagateAngularSpectrumInPlace(plane, lambda, -z, { bandLimit: false });
    for (let k = 0; k < plane.size; k++) phase[k] = Math.atan2(plane.im[k], plane.re[k]);
    if (o.onProgress && (it % pick(o.progressEvery, 10) === 0 || it === iterations - 1)) {
      o.onProgress({ iteration: it, error: error[it], efficiency: efficiency[it] });
    }
  }
  // final forward pass for the reported amplitude (ref: doe.js:418-426)
  doe.setAmplitudePhase(illumination, phase);
  propagateAngularSpectrumInPlace(doe, lambda, z, { bandLimit: false });
  const amp = new Float64Array(nx * ny);
  const inten = new Float64Array(nx * ny);
  for (let k = 0; k < doe.size; k++) {
    const a = Math.hypot(doe.re[k], doe.im[k]);
    amp[k] = a;
    inten[k] = a * a;
  }
  return { phase: phase, error: error, efficiency: efficiency, amplitude: amp, intensity: inten, iterations: iterations, field: doe };
}

// Multi-plane GS (ref: doe.js:557-626): per iteration the planes are visited
// in order; each visit propagates the element illum * exp(i phase) to the
// plane's z (no band limit), measures the amplitude error inside the plane's
// ROI, imposes the target amplitude there (the field elsewhere is kept),
// propagates back by -z and takes the phase.  error[it] is the
// weight-averaged plane error and efficiency[it] = 1 - error[it], as the
// reference reports them.  Flat initial phase unless initPhase is given.
export function gerchbergSaxtonMultiPlane(illumination, planes, opts) {
  const o = merged(opts);
  const nx = pick(o.nx, Math.round(Math.sqrt(illumination.length)));
  const ny = pick(o.ny, nx);
  const dx = pick(o.dx, 8e-6);
  const dy = pick(o.dy, dx);
  if (planes.length === 0) throw new Error('gerchbergSaxtonMultiPlane: no planes');
  const field = new ComplexField(nx, ny, dx, dy);
  const phase = o.initPhase ? Float64Array.from(o.initPhase) : flatPhase(nx, ny);
  const amp = new Float64Array(nx * ny);
  const error = [];
  const efficiency = [];
  const perPlaneError = planes.map(function () { return []; });
  for (let it = 0; it < o.iterations; it++) {
    let acc = 0;
    for (let p = 0; p < planes.length; p++) {
      const pl = planes[p];
      const roi = pick(pl.roi, null);
      const t = pl.targetAmp;
      field.setAmplitudePhase(illumination, phase);
      propagateAngularSpectrumInPlace(field, o.lambda, pl.z, { bandLimit: false });
      let num = 0;
      let den = 0;
      for (let k = 0; k < field.size; k++) {
        amp[k] = Math.hypot(field.re[k], field.im[k]);
        const m = roi ? roi[k] : 1;
        if (m > 0) {
          const d = (amp[k] - t[k]) * m;
          num += d * d;
          den += t[k] * t[k];
        }
      }
      const e = den > 0 ? Math.sqrt(num / den) : 0;
      perPlaneError[p].push(e);
      acc += e * pick(pl.weight, 1);
      for (let k = 0; k < field.size; k++) {
        const m = roi ? roi[k] : 1;
        if (m > 0 && amp[k] > 0) {
          const s = t[k] / amp[k];
          field.re[k] *= s;
          field.im[k] *= s;
        }
      }
      propagateAngularSpectrumInPlace(field, o.lambda, -pl.z, { bandLimit: false });
      for (let k = 0; k < field.size; k++) phase[k] = Math.atan2(field.im[k], field.re[k]);
    }
    error.push(acc / planes.length);
    efficiency.push(1 - error[error.length - 1]);
  }
  return { phase: phase, error: error, efficiency: efficiency, perPlaneError: perPlaneError, iterations: o.iterations };
}

// Weighted GS with bounded weights and best-iterate bookkeeping
// (ref: doe.js:455-546).
export function weightedGerchbergSaxton(illumination, targetAmp, opts, baseFn) {
  baseFn = baseFn || gerchbergSaxton;
  const o = merged(opts);
  const nx = pick(o.nx, Math.round(Math.sqrt(illumination.length)));
  const ny = pick(o.ny, nx);
  const n = nx * ny;
  const roiOf = function (k) { return o.roi ? o.roi[k] : 1; };
  const weights = new Float64Array(n);
  for (let k = 0; k < n; k++) if (roiOf(k) > 0) weights[k] = 1;
  let result = null;
  let bestPhase = null;
  let bestRmse = Infinity;
  let bestSnap = null;
  const hist = { error: [], efficiency: [] };
  for (let it = 0; it < o.iterations; it++) {
    const sub = {};
    Object.keys(o).forEach(function (k) { sub[k] = o[k]; });
    sub.weights = weights;
    sub.iterations = 1;
    sub.initPhase = result ? result.phase : o.initPhase;
    sub.onProgress = null;
    result = baseFn(illumination, targetAmp, sub);
    let num = 0;
    let den = 0;
    for (let k = 0; k < n; k++) {
      if (roiOf(k) <= 0) continue;
      const d = result.amplitude[k] - targetAmp[k];
      num += d * d;
      den += targetAmp[k] * targetAmp[k];
    }
    const rmse = den > 0 ? Math.sqrt(num / den) : 0;
    if (rmse < bestRmse) {
      bestRmse = rmse;
      bestPhase = Float64Array.from(result.phase);
      bestSnap = { amplitude: Float64Array.from(result.amplitude), intensity: Float64Array.from(result.intensity) };
    }
    const beta = pick(o.beta, 0.9);
    let wsum = 0;
    let tsum = 0;
    for (let k = 0; k < n; k++) {
      if (roiOf(k) <= 0) continue;
      const measured = result.amplitude[k];
      const desired = targetAmp[k];
      const ratio = measured > 1e-30 ? desired / measured : 2;
      weights[k] *= Math.pow(Math.min(Math.max(ratio, 0.25), 4), beta);
      weights[k] = Math.min(Math.max(weights[k], 1e-4), 1e4);
      wsum += weights[k] * weights[k] * desired * desired;
      tsum += desired * desired;
    }
    const norm = wsum > 0 ? Math.sqrt(tsum / wsum) : 1;
    for (let k = 0; k < n; k++) weights[k] *= norm;
    hist.error.push(bestRmse);
    hist.efficiency.push(result.efficiency[result.efficiency.length - 1]);
    if (o.onProgress && (it % pick(o.progressEvery, 10) === 0 || it === o.iterations - 1)) {
      o.onProgress({ iteration: it, error: bestRmse, efficiency: result.efficiency[result.efficiency.length - 1] });
    }
  }
  if (bestPhase) {
    const r = {};
    Object.keys(result).forEach(function (k) { r[k] = result[k]; });
    r.iterations = o.iterations;
    r.error = hist.error;
    r.efficiency = hist.efficiency;
    r.phase = bestPhase;
    r.amplitude = bestSnap.amplitude;
    r.intensity = bestSnap.intensity;
    r.weightedIterations = o.iterations;
    result = r;
  }
  result.weights = weights;
  result.rmse = bestRmse;
  return result;
}

// Far-field (unitary DFT) loop (ref: doe.js:639-730).
export function fourierCgh(illumination, targetAmp, opts) {
  const o = merged(opts);
  const nx = pick(o.nx, Math.round(Math.sqrt(illumination.length)));
  const ny = pick(o.ny, nx);
  const iterations = o.iterations;
  const roi = pick(o.roi, null);
  const weights = pick(o.weights, null);
  const n = nx * ny;
  const re = new Float64Array(n);
  const im = new Float64Array(n);
  const phase = o.initPhase ? Float64Array.from(o.initPhase) : randomPhase(nx, ny, o.seed);
  const error = [];
  const efficiency = [];
  const inten = new Float64Array(n);
  const amp = new Float64Array(n);
  const norm = 1 / Math.sqrt(nx * ny);
  for (let it = 0; it < iterations; it++) {
    for (let k = 0; k < n; k++) {
      re[k] = illumination[k] * Math.cos(phase[k]);
      im[k] = illumination[k] * Math.sin(phase[k]);
    }
    ifftshift2d(re, im, nx, ny);
    fft2d(re, im, nx, ny, false);
    fftshift2d(re, im, nx, ny, true);
    for (let k = 0; k < n; k++) {
      re[k] *= norm;
      im[k] *= norm;
    }
    let eN = 0;
    let eD = 0;
    let inRoi = 0;
    let total = 0;
    for (let k = 0; k < n; k++) {
      const a = Math.hypot(re[k], im[k]);
      amp[k] = a;
      inten[k] = a * a;
      total += a * a;
      const m = roi ? roi[k] : 1;
      const t = targetAmp[k] * (weights ? weights[k] : 1);
      if (m > 0) {
        inRoi += a * a * m;
        const d = (a - t) * m;
        eN += d * d;
        eD += t * t;
      }
    }
    error.push(eD > 0 ? Math.sqrt(eN / eD) : 0);
    efficiency.push(total > 0 ? inRoi / total : 0);
    for (let k = 0; k < n; k++) {
      const m = roi ? roi[k] : 1;
      const a = Math.hypot(re[k], im[k]);
      const t = m > 0 ? targetAmp[k] * (weights ? weights[k] : 1) : a;
      if (a > 0) {
        const s = t / a;
        re[k] *= s;
        im[k] *= s;
      } else {
        re[k] = t;
        im[k] = 0;
      }
    }
    ifftshift2d(re, im, nx, ny);
    fft2d(re, im, nx, ny, true);
    fftshift2d(re, im, nx, ny, true);
    for (let k = 0; k < n; k++) {
      re[k] *= Math.sqrt(nx * ny);
      im[k] *= Math.sqrt(nx * ny);
      phase[k] = Math.atan2(im[k], re[k]);
    }
    if (o.onProgress && (it % pick(o.progressEvery, 10) === 0 || it === iterations - 1)) {
      o.onProgress({ iteration: it, error: error[it], efficiency: efficiency[it] });
    }
  }
  return { phase: phase, error: error, efficiency: efficiency, amplitude: amp, intensity: inten, iterations: iterations, field: null };
}

// Nearest of `levels` uniform phase levels, wrapped (ref: doe.js:167-174).
export function quantizePhase(phi, levels) {
  levels = pick(levels, 4);
  const step = (2 * Math.PI) / levels;
  const out = new Float64Array(phi.length);
  for (let k = 0; k < phi.length; k++) out[k] = Math.round(phi[k] / step) * step;
  return wrapPhase(out);
}

// Floyd–Steinberg on the circle (ref: doe.js:185-226): each sample is
// rounded to a level, its wrapped error is scattered 7/16 ahead in the row
// and 3/16, 5/16, 1/16 to the row below (mirrored on right-to-left rows).
export function errorDiffusionQuantize(phi, levels, opts) {
  levels = pick(levels, 4);
  const o = opts || {};
  const snake = pick(o.snake, true);
  const wrapFirst = pick(o.wrapFirst, true);
  const n = phi.length;
  const nx = Math.round(Math.sqrt(n));
  const ny = n / nx;
  if (nx * ny !== n) throw new Error('errorDiffusionQuantize: expected a square grid');
  const buf = wrapFirst ? wrapPhase(Float64Array.from(phi)) : Float64Array.from(phi);
  const out = new Float64Array(n);
  const step = (2 * Math.PI) / levels;
  const scatter = [[1, 0, 7 / 16], [-1, 1, 3 / 16], [0, 1, 5 / 16], [1, 1, 1 / 16]];
  for (let j = 0; j < ny; j++) {
    const dir = snake && j % 2 === 1 ? -1 : 1;
    for (let c = 0; c < nx; c++) {
      const i = dir > 0 ? c : nx - 1 - c;
      const k = i + j * nx;
      const q = Math.round(buf[k] / step) * step;
      out[k] = q;
      let e = buf[k] - q;
      if (e > Math.PI) e -= 2 * Math.PI;
      else if (e <= -Math.PI) e += 2 * Math.PI;
      for (let t = 0; t < scatter.length; t++) {
        const ii = i + dir * scatter[t][0];
        const jj = j + scatter[t][1];
        if (ii >= 0 && ii < nx && jj < ny) buf[ii + jj * nx] += e * scatter[t][2];
      }
    }
  }
  return wrapPhase(out);
}

// Etch depth in [0, lambda / (n - 1)) (ref: doe.js:967-976).
export function phaseToHeight(phase, lambda, nIndex) {
  nIndex = pick(nIndex, 1.52);
  const s = lambda / (2 * Math.PI * (nIndex - 1));
  const step = lambda / (nIndex - 1);
  const h = new Float64Array(phase.length);
  for (let k = 0; k < phase.length; k++) h[k] = (((phase[k] * s) % step) + step) % step;
  return h;
}

// Separable Gaussian blur with clamped edges (ref: doe.js:761-794).
export function blurRaster(src, nx, ny, sigmaPx) {
  sigmaPx = pick(sigmaPx, 1);
  if (sigmaPx <= 0) return Float64Array.from(src);
  const r = Math.max(1, Math.ceil(3 * sigmaPx));
  const ker = new Float64Array(2 * r + 1);
  let ks = 0;
  for (let i = -r; i <= r; i++) {
    ker[i + r] = Math.exp(-(i * i) / (2 * sigmaPx * sigmaPx));
    ks += ker[i + r];
  }
  for (let i = 0; i < ker.length; i++) ker[i] /= ks;
  const tmp = new Float64Array(src.length);
  const out = new Float64Array(src.length);
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      let acc = 0;
      for (let d = -r; d <= r; d++) acc += src[Math.min(nx - 1, Math.max(0, i + d)) + j * nx] * ker[d + r];
      tmp[i + j * nx] = acc;
    }
  }
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      let acc = 0;
      for (let d = -r; d <= r; d++) acc += tmp[i + Math.min(ny - 1, Math.max(0, j + d)) * nx] * ker[d + r];
      out[i + j * nx] = acc;
    }
  }
  return out;
}

function defaultDiscTarget(nx, ny, frac) {
  const t = new Float64Array(nx * ny);
  const r = frac * Math.min(nx, ny) * 0.5;
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) t[i + j * nx] = Math.hypot(i - (nx - 1) / 2, j - (ny - 1) / 2) <= r ? 1 : 0;
  }
  return t;
}

function weightedWrapper(fn, illumination, targetAmp, common, algorithm) {
  if (algorithm === 'wgs') return weightedGerchbergSaxton(illumination, targetAmp, common, fn);
  return fn(illumination, targetAmp, common);
}

// Orchestration (ref: doe.js:796-896).  Quantisation (levels > 1) is outside
// the hot path and not restated.
export function designDOE(cfg) {
  const nx = pick(cfg.nx, 256);
  const ny = pick(cfg.ny, nx);
  const dx = pick(cfg.dx, 8e-6);
  const dy = pick(cfg.dy, dx);
  const lambda = pick(cfg.lambda, 532e-9);
  const z = pick(cfg.z, 0.03);
  const mode = pick(cfg.mode, 'image-plane');
  const algorithm = cfg.algorithm && cfg.algorithm !== 'auto' ? cfg.algorithm : mode === 'far-field' ? 'gs' : 'wgs';
  const iterations = pick(cfg.iterations, 60);
  const illumination = pick(cfg.illumination, discMask(nx, ny, dx, dy, (Math.min(nx * dx, ny * dy) / 2) * 0.95, { soft: 2 * dx }));
  const roi = pick(cfg.roi, null);
  let tIrr = pick(cfg.target, defaultDiscTarget(nx, ny, 0.35));
  if (cfg.smooth > 0) tIrr = blurRaster(tIrr, nx, ny, cfg.smooth);
  const targetAmp = new Float64Array(nx * ny);
  for (let k = 0; k < targetAmp.length; k++) targetAmp[k] = Math.sqrt(Math.max(0, tIrr[k]));
  let inP = 0;
  for (let k = 0; k < illumination.length; k++) inP += illumination[k] * illumination[k];
  let tP = 0;
  for (let k = 0; k < targetAmp.length; k++) tP += targetAmp[k] * targetAmp[k];
  if (tP > 0) {
    const s = Math.sqrt(inP / tP);
    for (let k = 0; k < targetAmp.length; k++) targetAmp[k] *= s;
  }
  let initPhase = cfg.initPhase;
  if (!initPhase) {
    const init = pick(cfg.init, 'random');
    if (init === 'flat') initPhase = flatPhase(nx, ny);
    else if (init === 'lens') initPhase = kinoformLens(nx, ny, dx, dy, lambda, z);
    else initPhase = randomPhase(nx, ny, pick(cfg.seed, 12345));
  }
  const common = {
    nx: nx, ny: ny, dx: dx, dy: dy, lambda: lambda, z: z, roi: roi, iterations: iterations,
    seed: pick(cfg.seed, 12345),
    feedback: algorithm === 'hio' ? pick(cfg.feedback, 0.8) : 0,
    escape: pick(cfg.escape, Math.round(iterations * 0.2)),
    beta: pick(cfg.beta, 0.9),
    initPhase: initPhase,
    onProgress: cfg.onProgress,
    progressEvery: pick(cfg.progressEvery, 10),
  };
  let result;
  if (mode === 'far-field') {
    result = weightedWrapper(fourierCgh, illumination, targetAmp, common, algorithm);
    result.mode = 'far-field';
  } else {
    result = weightedWrapper(gerchbergSaxton, illumination, targetAmp, common, algorithm);
    result.mode = 'image-plane';
  }
  result.algorithm = algorithm;
  result.illumination = illumination;
  result.targetAmp = targetAmp;
  result.roi = roi;
  result.lambda = lambda;
  result.z = z;
  result.nx = nx;
  result.ny = ny;
  result.dx = dx;
  result.dy = dy;
  return result;
}

// Illumination profiles (ref: doe.js:979-1000); tophat and gaussian.
export function illuminationProfile(type, nx, ny, dx, dy, opts) {
  const o = opts || {};
  const radius = pick(o.radius, (Math.min(nx * dx, ny * dy) / 2) * 0.95);
  const ap = discMask(nx, ny, dx, dy, radius, { soft: pick(o.soft, 2 * dx) });
  if (type === 'gaussian') {
    const out = new Float64Array(nx * ny);
    const w = pick(o.waist, radius * 0.6);
    for (let j = 0; j < ny; j++) {
      const y = (j - ny / 2) * dy;
      for (let i = 0; i < nx; i++) {
        const x = (i - nx / 2) * dx;
        out[i + j * nx] = ap[i + j * nx] * Math.exp(-(x * x + y * y) / (w * w));
      }
    }
    return out;
  }
  return ap;
}
