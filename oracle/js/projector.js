// ORACLE (test infrastructure only) — FP64 restatement of the projector
// pipeline pieces on the hot path (ref: src/projector.js:48-498, 554-614).
// ES2019.  The 5x7 glyph table is data the target raster must reproduce
// exactly, so its rows are the reference's.

import { ComplexField } from './field.js';
import { propagateAngularSpectrumInPlace, focalPlaneSpectrum } from './propagate.js';
import { designDOE, illuminationProfile } from './doe.js';
import { maskedSum, regionStats, rmsError, correlation, ssim, rmsRadius, zeroOrderFraction, spotUniformity } from './metrics.js';

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

// 1-bit text raster, nearest-neighbour upscaled (ref: projector.js:97-123)
export function renderText(text, opts) {
  const o = opts || {};
  const scale = o.scale === undefined ? 8 : o.scale;
  const spacing = o.spacing === undefined ? 1 : o.spacing;
  const chars = text.toUpperCase().split('');
  const w = chars.length * (5 + spacing) - spacing;
  const h = 7;
  const r = new Float64Array(w * h);
  chars.forEach(function (ch, ci) {
    const g = GLYPHS[ch] === undefined ? GLYPHS[' '] : GLYPHS[ch];
    const rows = g.split('/');
    for (let y = 0; y < 7; y++) {
      for (let x = 0; x < 5; x++) if (rows[y][x] === '1') r[x + ci * (5 + spacing) + y * w] = 1;
    }
  });
  const w2 = w * scale;
  const h2 = h * scale;
  const out = new Float64Array(w2 * h2);
  for (let y = 0; y < h2; y++) {
    for (let x = 0; x < w2; x++) out[x + y * w2] = r[Math.floor(x / scale) + Math.floor(y / scale) * w];
  }
  return { data: out, width: w2, height: h2 };
}

// Bilinear, aspect-preserving resample (ref: projector.js:126-154)
export function resampleRaster(src, sw, sh, dw, dh, opts) {
  const o = opts || {};
  const fit = o.fit === undefined ? 'contain' : o.fit;
  const bg = o.background === undefined ? 0 : o.background;
  const gain = o.gain === undefined ? 1 : o.gain;
  const out = new Float64Array(dw * dh).fill(bg);
  const s = fit === 'contain' ? Math.min(dw / sw, dh / sh) : Math.max(dw / sw, dh / sh);
  const ox = (dw - sw * s) / 2;
  const oy = (dh - sh * s) / 2;
  for (let y = 0; y < dh; y++) {
    const fy = (y - oy) / s;
    const y0 = Math.floor(fy);
    if (y0 < 0 || y0 >= sh) continue;
    const ty = fy - y0;
    const y1 = Math.min(y0 + 1, sh - 1);
    for (let x = 0; x < dw; x++) {
      const fx = (x - ox) / s;
      const x0 = Math.floor(fx);
      if (x0 < 0 || x0 >= sw) continue;
      const tx = fx - x0;
      const x1 = Math.min(x0 + 1, sw - 1);
      const v =
        src[x0 + y0 * sw] * (1 - tx) * (1 - ty) +
        src[x1 + y0 * sw] * tx * (1 - ty) +
        src[x0 + y1 * sw] * (1 - tx) * ty +
        src[x1 + y1 * sw] * tx * ty;
      out[x + y * dw] = v * gain;
    }
  }
  return out;
}

// Soft dilated ROI around the target support (ref: projector.js:262-315)
export function autoRoi(irr, nx, ny, opts) {
  const o = opts || {};
  const dilate = o.dilate === undefined ? 2 : o.dilate;
  const soft = o.soft === undefined ? 2 : o.soft;
  const threshold = o.threshold === undefined ? 0.05 : o.threshold;
  let mx = 0;
  for (let k = 0; k < irr.length; k++) if (irr[k] > mx) mx = irr[k];
  const bin = new Float64Array(irr.length);
  for (let k = 0; k < irr.length; k++) bin[k] = irr[k] > threshold * mx ? 1 : 0;
  const r = Math.max(1, Math.round(dilate));
  const tmp = new Float64Array(irr.length);
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      let v = 0;
      for (let d = -r; d <= r && v === 0; d++) {
        const ii = i + d;
        if (ii >= 0 && ii < nx && bin[ii + j * nx] > 0) v = 1;
      }
      tmp[i + j * nx] = v;
    }
  }
  const dil = new Float64Array(irr.length);
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      let v = 0;
      for (let d = -r; d <= r && v === 0; d++) {
        const jj = j + d;
        if (jj >= 0 && jj < ny && tmp[i + jj * nx] > 0) v = 1;
      }
      dil[i + j * nx] = v;
    }
  }
  if (soft <= 0) return dil;
  const out = new Float64Array(irr.length);
  const s = Math.max(1, Math.round(soft));
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      let acc = 0;
      let n = 0;
      for (let dj = -s; dj <= s; dj++) {
        const jj = j + dj;
        if (jj < 0 || jj >= ny) continue;
        for (let di = -s; di <= s; di++) {
          const ii = i + di;
          if (ii < 0 || ii >= nx) continue;
          acc += dil[ii + jj * nx];
          n++;
        }
      }
      out[i + j * nx] = n > 0 ? acc / n : 0;
    }
  }
  return out;
}

// Every target kind (ref: projector.js:169-255).
export function makeTarget(kind, nx, ny, opts) {
  const o = opts || {};
  const irr = new Float64Array(nx * ny);
  const cx = (nx - 1) / 2;
  const cy = (ny - 1) / 2;
  const dflt = (v, d) => (v === undefined || v === null ? d : v);
  const fill = (fn) => {
    for (let j = 0; j < ny; j++) for (let i = 0; i < nx; i++) irr[i + j * nx] = fn(i, j);
  };
  let name = kind;
  if (kind === 'image') {
    if (!o.raster) throw new Error('makeTarget("image"): pass {raster, width, height}');
    const gamma = dflt(o.gamma, 1);
    const res = resampleRaster(o.raster, o.width, o.height, nx, ny, { fit: 'contain' });
    for (let k = 0; k < irr.length; k++) irr[k] = gamma === 1 ? res[k] : Math.pow(Math.max(0, res[k]), gamma);
    name = dflt(o.name, 'image');
  } else if (kind === 'text') {
    const text = dflt(o.text, 'DOE');
    const g = renderText(text, { scale: dflt(o.scale, 6), spacing: dflt(o.spacing, 1) });
    irr.set(resampleRaster(g.data, g.width, g.height, nx, ny, { fit: 'contain' }));
    name = 'text:' + text;
  } else if (kind === 'disc') {
    const r = dflt(o.radius, 0.35) * Math.min(nx, ny) * 0.5;
    fill((i, j) => (Math.hypot(i - cx, j - cy) <= r ? 1 : 0));
  } else if (kind === 'ring') {
    const r0 = dflt(o.inner, 0.25) * Math.min(nx, ny) * 0.5;
    const r1 = dflt(o.outer, 0.4) * Math.min(nx, ny) * 0.5;
    fill((i, j) => {
      const r = Math.hypot(i - cx, j - cy);
      return r >= r0 && r <= r1 ? 1 : 0;
    });
  } else if (kind === 'dots' || kind === 'grid') {
    const cols = dflt(o.cols, 8);
    const rows = dflt(o.rows, 8);
    const px = nx / cols;
    const py = ny / rows;
    const sigma = dflt(o.sigma, 0.22) * Math.min(px, py);
    const dark = o.missing || [];
    fill((i, j) => {
      const fy = ((j + 0.5) / py) % 1 - 0.5;
      const fx = ((i + 0.5) / px) % 1 - 0.5;
      const ddx = fx * px;
      const ddy = fy * py;
      const ci = (i / px) | 0;
      const cj = (j / py) | 0;
      for (const m of dark) if (m[0] === ci && m[1] === cj) return 0;
      return Math.exp(-(ddx * ddx + ddy * ddy) / (2 * sigma * sigma));
    });
  } else if (kind === 'checker') {
    const cells = dflt(o.cells, 8);
    fill((i, j) => ((Math.floor((i / nx) * cells) + Math.floor((j / ny) * cells)) % 2 === 0 ? 1 : 0));
  } else if (kind === 'lines') {
    const period = dflt(o.period, 12);
    const duty = dflt(o.duty, 0.5);
    fill((i) => ((i % period) / period < duty ? 1 : 0));
  } else if (kind === 'zone') {
    const period = dflt(o.period, 16);
    fill((i, j) => {
      const r = Math.hypot(i - cx, j - cy);
      return Math.sin((Math.PI * r * r) / period) > 0 ? 1 : 0;
    });
  } else {
    throw new Error('makeTarget: unknown kind "' + kind + '"');
  }
  const roi = o.roi ? o.roi : autoRoi(irr, nx, ny, { dilate: dflt(o.dilate, 2), soft: dflt(o.softEdge, 2) });
  const grid = kind === 'dots' || kind === 'grid' ? { cols: dflt(o.cols, 8), rows: dflt(o.rows, 8) } : null;
  return { irradiance: irr, roi: roi, name: name, grid: grid };
}

export const DEFAULT_PROJECTOR = {
  n: 256, pitch: 8e-6, lambda: 532e-9, distance: 0.03, aperture: 0.95, apodization: 2,
  source: 'tophat', waist: 0.6, inner: 0.6, outer: 0.98,
  mode: 'image-plane', algorithm: 'auto', iterations: 60, levels: 0, seed: 12345, quantize: false,
  target: { kind: 'text', text: 'DOE' },
};

export class ProjectorSystem {
  // ref: projector.js:347-366
  constructor(cfg) {
    cfg = cfg || {};
    const c = {};
    Object.keys(DEFAULT_PROJECTOR).forEach(function (k) { c[k] = DEFAULT_PROJECTOR[k]; });
    Object.keys(cfg).forEach(function (k) { c[k] = cfg[k]; });
    const tgt = {};
    Object.keys(DEFAULT_PROJECTOR.target).forEach(function (k) { tgt[k] = DEFAULT_PROJECTOR.target[k]; });
    Object.keys(cfg.target || {}).forEach(function (k) { tgt[k] = cfg.target[k]; });
    c.target = tgt;
    this.cfg = c;
    this.nx = c.n;
    this.ny = c.n;
    this.dx = c.pitch;
    this.dy = c.pitch;
    this.lambda = c.lambda;
    if (cfg.distance === undefined && c.mode === 'far-field') c.distance = (c.n * c.pitch * c.pitch) / c.lambda;
    this.z = c.distance;
    this.target = null;
    this.result = null;
    this.setTarget(c.target);
  }

  setTarget(spec) {
    if (spec && spec.irradiance) {
      this.target = { irradiance: spec.irradiance, roi: spec.roi ? spec.roi : autoRoi(spec.irradiance, this.nx, this.ny), name: spec.name ? spec.name : 'custom' };
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
    const radius = ((Math.min(this.nx * this.dx, this.ny * this.dy) / 2) * this.cfg.aperture) / 1;
    return illuminationProfile(this.cfg.source, this.nx, this.ny, this.dx, this.dy, {
      radius: radius, soft: this.cfg.apodization * this.dx, waist: radius * this.cfg.waist,
    });
  }

  // ref: projector.js:438-467
  design(opts) {
    const o = opts || {};
    const r = designDOE({
      nx: this.nx, ny: this.ny, dx: this.dx, dy: this.dy, lambda: this.lambda, z: this.z,
      mode: this.cfg.mode, algorithm: this.cfg.algorithm,
      iterations: o.iterations === undefined ? this.cfg.iterations : o.iterations,
      levels: o.levels === undefined ? this.cfg.levels : o.levels,
      seed: o.seed === undefined ? this.cfg.seed : o.seed,
      illumination: this.illumination(), target: this.target.irradiance, roi: this.target.roi,
      onProgress: o.onProgress, progressEvery: o.progressEvery === undefined ? 10 : o.progressEvery,
      beta: o.beta === undefined ? 0.9 : o.beta, escape: o.escape, initPhase: o.initPhase,
    });
    r.iterations = o.iterations === undefined ? this.cfg.iterations : o.iterations;
    r.target = this.target;
    this.result = r;
    return r;
  }

  // ref: projector.js:475-498
  simulate(opts) {
    const o = opts || {};
    if (!this.result) throw new Error('ProjectorSystem.simulate: call design() first');
    const dist = o.z === undefined || o.z === null ? this.z : o.z;
    const mode = o.mode ? o.mode : this.cfg.mode;
    const phase = o.phase ? o.phase : this.result.phase;
    const pad = o.pad === undefined ? 1 : o.pad;
    if (mode === 'far-field') {
      if (!(dist > 0)) throw new Error('ProjectorSystem.simulate: far-field mode needs a positive throw');
      const f = new ComplexField(this.nx, this.ny, this.dx, this.dy);
      f.setAmplitudePhase(this.result.illumination, phase);
      return focalPlaneSpectrum(f, this.lambda, dist);
    }
    let field = new ComplexField(this.nx, this.ny, this.dx, this.dy);
    field.setAmplitudePhase(this.result.illumination, phase);
    if (pad > 1) field = field.resample(this.nx * pad, this.ny * pad);
    propagateAngularSpectrumInPlace(field, this.lambda, dist, { bandLimit: !!o.bandLimit });
    if (pad > 1) field = field.resample(this.nx, this.ny);
    return field;
  }

  // ref: projector.js:554-614
  screenMetrics(field, opts) {
    const o = opts || {};
    const tgt = o.target ? o.target : this.target;
    const irr = field.intensity();
    const mask = o.roi ? o.roi : tgt.roi;
    const element = new ComplexField(this.nx, this.ny, this.dx, this.dy);
    element.setAmplitudePhase(this.result.illumination, o.phase ? o.phase : this.result.phase);
    const cell = field.dx * field.dy;
    const it = Float64Array.from(tgt.irradiance);
    const pt = maskedSum(irr, mask) * cell;
    const tt = maskedSum(it, null) * cell;
    const s = tt > 0 ? pt / tt : 0;
    for (let k = 0; k < it.length; k++) it[k] *= s;
    const prov = regionStats(irr, mask);
    const stats = regionStats(irr, mask, { threshold: 0.1 * prov.max });
    const spots = tgt.grid ? spotUniformity(irr, this.nx, this.ny, tgt.grid) : null;
    const incident = maskedSum(this.illumination(), null) * this.dx * this.dy;
    const total = maskedSum(irr, null) * cell;
    const inSignal = maskedSum(irr, mask) * cell;
    const rms = rmsError(irr, it, mask);
    const corr = correlation(irr, it, mask);
    const ss = ssim(irr, it, this.nx, this.ny, { window: 8, weight: 'target' });
    const rr = rmsRadius(irr, this.nx, field.dx, 0, 0, field.dy);
    return {
      efficiency: incident > 0 ? inSignal / incident : 0,
      fluxFraction: total > 0 ? inSignal / total : 0,
      uniformity: spots ? spots.uniformity : stats.uniformity,
      spotUniformity: spots ? spots.uniformity : null,
      flatness: stats.flatness,
      contrast: stats.cv,
      rmse: rms,
      fidelity: (corr + 1) / 2,
      correlation: corr,
      ssim: ss,
      snr: rms > 0 ? 20 * Math.log10(1 / rms) : Infinity,
      zeroOrder: zeroOrderFraction(element),
      rmsRadius: rr.r,
      meanIntensity: stats.mean,
      peakIntensity: stats.max,
      powerInSignal: inSignal,
      powerAtScreen: total,
      powerIncident: incident,
    };
  }

  // ref: projector.js:622-641: screen metrics at evenly spaced distances
  // around the design throw, one simulate() per distance
  defocusCurve(opts) {
    const o = opts || {};
    const points = o.points === undefined ? 21 : o.points;
    const span = o.span === undefined ? 0.6 : o.span;
    const pad = o.pad === undefined ? 2 : o.pad;
    const mode = o.mode === undefined ? null : o.mode;
    const z0 = this.z * (1 - span);
    const z1 = this.z * (1 + span);
    const out = [];
    for (let i = 0; i < points; i++) {
      const z = z0 + ((z1 - z0) * i) / (points - 1);
      const m = this.screenMetrics(this.simulate({ z: z, pad: pad, mode: mode }));
      out.push({ z: z, zMm: z * 1e3, fidelity: m.fidelity, correlation: m.correlation, ssim: m.ssim, rmse: m.rmse, efficiency: m.efficiency, uniformity: m.uniformity });
    }
    return out;
  }

  // ref: projector.js:644-653: screen irradiance along the axis
  focusStack(opts) {
    const o = opts || {};
    const points = o.points === undefined ? 9 : o.points;
    const span = o.span === undefined ? 0.5 : o.span;
    const pad = o.pad === undefined ? 2 : o.pad;
    const mode = o.mode === undefined ? null : o.mode;
    const out = [];
    for (let i = 0; i < points; i++) {
      const z = this.z * (1 - span) + (this.z * 2 * span * i) / (points - 1);
      out.push({ z: z, intensity: this.simulate({ z: z, pad: pad, mode: mode }).intensity() });
    }
    return out;
  }
}
