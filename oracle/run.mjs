// ORACLE job runner (test infrastructure only).  Usage:
//   node oracle/run.mjs <dir>
// <dir>/job.json = {"op": ..., "args": {...}}; array inputs are little-endian
// Float64 files <dir>/in_<name>.f64; array outputs are written as
// <dir>/out_<name>.f64 and scalars/short lists to <dir>/result.json.
import fs from 'fs';
import path from 'path';
import { fft2d, dft2dNaive } from './js/fft.js';
import { ComplexField } from './js/field.js';
import {
  angularSpectrumTransferFunction,
  propagateAngularSpectrumInPlace,
  propagateFresnelFFT,
  propagateFresnelFFTInverse,
  focalPlaneSpectrum,
  focusSweep,
} from './js/propagate.js';
import { gerchbergSaxton, weightedGerchbergSaxton, fourierCgh, designDOE, randomPhase, gerchbergSaxtonMultiPlane, blurRaster, errorDiffusionQuantize, quantizePhase, phaseToHeight, illuminationProfile } from './js/doe.js';
import { ProjectorSystem, makeTarget, autoRoi, resampleRaster, renderText } from './js/projector.js';
import { planeToRGB, planeToGray } from './js/color.js';
import { ssim, rmsError, correlation, regionStats, zeroOrderFraction, rmsRadius, spotUniformity } from './js/metrics.js';

const dir = process.argv[2];
const job = JSON.parse(fs.readFileSync(path.join(dir, 'job.json'), 'utf8'));
const a = job.args || {};

function inp(name) {
  const p = path.join(dir, 'in_' + name + '.f64');
  if (!fs.existsSync(p)) return null;
  const b = fs.readFileSync(p);
  return new Float64Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}
function out(name, arr) {
  fs.writeFileSync(path.join(dir, 'out_' + name + '.f64'), Buffer.from(Float64Array.from(arr).buffer));
}
function fieldOf(nx, ny, dx, dy) {
  const f = new ComplexField(nx, ny, dx, dy);
  f.re.set(inp('re'));
  f.im.set(inp('im'));
  return f;
}
function nowMs() {
  const t = process.hrtime();
  return t[0] * 1e3 + t[1] / 1e6;
}

const res = {};
const op = job.op;
if (op === 'fft2d' || op === 'dft2d') {
  const re = inp('re');
  const im = inp('im');
  if (op === 'fft2d') {
    fft2d(re, im, a.nx, a.ny, !!a.inverse);
    out('re', re);
    out('im', im);
  } else {
    const r = dft2dNaive(re, im, a.nx, a.ny, !!a.inverse);
    out('re', r.re);
    out('im', r.im);
  }
} else if (op === 'asm') {
  const f = fieldOf(a.nx, a.ny, a.dx, a.dy);
  propagateAngularSpectrumInPlace(f, a.lambda, a.z, { bandLimit: a.bandLimit, evanescent: a.evanescent, softEdge: a.softEdge });
  out('re', f.re);
  out('im', f.im);
} else if (op === 'transfer') {
  const H = angularSpectrumTransferFunction(a.nx, a.ny, a.dx, a.dy, a.lambda, a.z, { bandLimit: a.bandLimit, evanescent: a.evanescent, softEdge: a.softEdge });
  out('re', H.re);
  out('im', H.im);
  res.propagatingFraction = H.propagatingFraction;
} else if (op === 'fresnel') {
  const f = fieldOf(a.nx, a.ny, a.dx, a.dy);
  const r = propagateFresnelFFT(f, a.lambda, a.z, {});
  out('re', r.field.re);
  out('im', r.field.im);
  res.dx = r.dx;
  res.dy = r.dy;
  res.undersampled = r.undersampled;
  res.zCritical = r.zCritical;
} else if (op === 'fresnel_inv') {
  const f = fieldOf(a.nx, a.ny, a.in_dx, a.in_dy);
  const r = propagateFresnelFFTInverse(f, a.lambda, a.z, { dx: a.dx, dy: a.dy });
  out('re', r.re);
  out('im', r.im);
} else if (op === 'lens') {
  const f = fieldOf(a.nx, a.ny, a.dx, a.dy);
  const r = focalPlaneSpectrum(f, a.lambda, a.f);
  out('re', r.re);
  out('im', r.im);
  res.dx = r.dx;
  res.dy = r.dy;
} else if (op === 'gs') {
  // one phase-retrieval run with explicit inputs
  const opts = {
    nx: a.nx, ny: a.ny, dx: a.dx, dy: a.dy, lambda: a.lambda, z: a.z, iterations: a.iterations,
    beta: a.beta, feedback: a.feedback || 0, escape: a.escape || 0, seed: a.seed,
    roi: inp('roi'), initPhase: inp('phase0'),
  };
  const base = a.mode === 'far-field' ? fourierCgh : gerchbergSaxton;
  const r = a.algorithm === 'wgs'
    ? weightedGerchbergSaxton(inp('illum'), inp('target'), opts, base)
    : base(inp('illum'), inp('target'), opts);
  out('phase', r.phase);
  out('amplitude', r.amplitude);
  out('intensity', r.intensity);
  if (r.weights) out('weights', r.weights);
  res.error = r.error;
  res.efficiency = r.efficiency;
  res.rmse = r.rmse === undefined ? null : r.rmse;
} else if (op === 'multiplane') {
  // gerchbergSaxtonMultiPlane with explicit inputs: in_target<q>, in_roi<q>
  const planes = a.planes.map(function (p, q) {
    return { z: p.z, targetAmp: inp('target' + q), roi: inp('roi' + q), weight: p.weight };
  });
  const r = gerchbergSaxtonMultiPlane(inp('illum'), planes, {
    nx: a.nx, ny: a.ny, dx: a.dx, dy: a.dy, lambda: a.lambda, iterations: a.iterations, initPhase: inp('phase0'),
  });
  out('phase', r.phase);
  res.error = r.error;
  res.efficiency = r.efficiency;
  res.perPlaneError = r.perPlaneError;
} else if (op === 'focus_sweep') {
  const sw = focusSweep(fieldOf(a.nx, a.ny, a.dx, a.dy), a.lambda, a.z0, a.z1, a.steps, { bandLimit: !!a.bandLimit });
  const n = a.nx * a.ny;
  const re = new Float64Array(n * sw.length);
  const im = new Float64Array(n * sw.length);
  sw.forEach(function (p, s) { re.set(p.field.re, s * n); im.set(p.field.im, s * n); });
  out('re', re);
  out('im', im);
  res.z = sw.map(function (p) { return p.z; });
} else if (op === 'random_phase') {
  out('phase', randomPhase(a.nx, a.ny, a.seed));
} else if (op === 'projector') {
  // ProjectorSystem(cfg).design(opts); simulate(); screenMetrics()
  const sys = new ProjectorSystem(a.cfg);
  const t0 = nowMs();
  const r = sys.design(a.design || {});
  const t1 = nowMs();
  const field = sys.simulate(a.simulate || {});
  const m = sys.screenMetrics(field);
  const t2 = nowMs();
  out('phase', r.phase);
  out('illumination', r.illumination);
  out('target_amp', r.targetAmp);
  out('target', sys.target.irradiance);
  out('roi', sys.target.roi);
  out('screen_re', field.re);
  out('screen_im', field.im);
  res.error = r.error;
  res.efficiency = r.efficiency;
  res.algorithm = r.algorithm;
  res.mode = r.mode;
  res.z = sys.z;
  res.screen_dx = field.dx;
  res.metrics = m;
  res.design_ms = t1 - t0;
  res.simulate_metrics_ms = t2 - t1;
  if (a.defocus) res.defocus = sys.defocusCurve(a.defocus);
  if (a.stack) {
    const st = sys.focusStack(a.stack);
    const n = sys.nx * sys.ny;
    const all = new Float64Array(n * st.length);
    st.forEach(function (p, i) { all.set(p.intensity, i * n); });
    out('stack', all);
    res.stack_z = st.map(function (p) { return p.z; });
  }
} else if (op === 'projector_eval') {
  // screenMetrics(simulate()) of a given element phase (e.g. one designed on
  // the GPU) under ProjectorSystem(cfg)
  const sys = new ProjectorSystem(a.cfg);
  sys.result = { phase: inp('phase'), illumination: sys.illumination() };
  const field = sys.simulate(a.simulate || {});
  res.metrics = sys.screenMetrics(field);
  out('screen_re', field.re);
  out('screen_im', field.im);
} else if (op === 'target') {
  // makeTarget(kind, nx, ny, opts); an image raster comes in as in_raster
  const o = Object.assign({}, a.opts || {});
  if (a.kind === 'image') o.raster = inp('raster');
  const t = makeTarget(a.kind, a.nx, a.ny, o);
  out('irradiance', t.irradiance);
  out('roi', t.roi);
  res.name = t.name;
  res.grid = t.grid;
} else if (op === 'auto_roi') {
  out('roi', autoRoi(inp('irr'), a.nx, a.ny, a.opts || {}));
} else if (op === 'blur') {
  out('out', blurRaster(inp('src'), a.nx, a.ny, a.sigma));
} else if (op === 'resample') {
  out('out', resampleRaster(inp('src'), a.sw, a.sh, a.dw, a.dh, a.opts || {}));
} else if (op === 'render_text') {
  const g = renderText(a.text, a.opts || {});
  out('data', g.data);
  res.width = g.width;
  res.height = g.height;
} else if (op === 'image') {
  // planeToRGB / planeToGray of in_plane; bytes come back as Float64
  const r = a.gray ? planeToGray(inp('plane'), a.opts || {}) : planeToRGB(inp('plane'), a.opts || {});
  out('bytes', a.gray ? r.gray : r.rgb);
  res.lo = r.lo;
  res.hi = r.hi;
} else if (op === 'error_diffusion') {
  out('out', errorDiffusionQuantize(inp('phase'), a.levels, a.opts || {}));
} else if (op === 'quantize') {
  out('out', quantizePhase(inp('phase'), a.levels));
} else if (op === 'phase_to_height') {
  out('out', phaseToHeight(inp('phase'), a.lambda, a.n));
} else if (op === 'metrics') {
  const p = inp('plane');
  const t = inp('target');
  const m = inp('mask');
  res.rmse = rmsError(p, t, m);
  res.correlation = correlation(p, t, m);
  res.ssim = ssim(p, t, a.nx, a.ny, { window: 8, weight: a.weight || null });
  res.stats = regionStats(p, m);
} else if (op === 'metric_fns') {
  // zeroOrderFraction of the field in_re/in_im; rmsRadius and spotUniformity
  // of in_plane (either part may be absent)
  if (inp('re')) res.zeroOrder = zeroOrderFraction(fieldOf(a.nx, a.ny, a.dx || 1, a.dy || 1));
  const p = inp('plane');
  if (p) {
    res.rmsRadius = rmsRadius(p, a.nx, a.dx, a.cx, a.cy, a.dy);
    const su = spotUniformity(p, a.nx, a.ny, a.spots || {});
    res.spotUniformity = su.uniformity;
    res.spotCells = su.cells;
    res.spotCv = su.cv;
  }
} else if (op === 'sample_image') {
  const f = new ComplexField(a.nx, a.ny, a.dx || 1, a.dy || 1);
  f.sampleImage(inp('img'), a.width, a.height, a.scale, a.fill);
  out('re', f.re);
  out('im', f.im);
} else if (op === 'field_resample') {
  const f = fieldOf(a.nx, a.ny, a.dx, a.dy);
  const g = f.resample(a.nx2, a.ny2, a.dx2, a.dy2);
  out('re', g.re);
  out('im', g.im);
} else if (op === 'illumination') {
  out('profile', illuminationProfile(a.type, a.nx, a.ny, a.dx, a.dy, a.opts || {}));
} else if (op === 'time') {
  // CPU baseline timing of one unit of work on a single JS thread,
  // 1 warm-up + reps (the reference's bench.mjs methodology).
  const n = a.n;
  const reps = a.reps || 3;
  let fn;
  if (a.what === 'fft2d') {
    const re = new Float64Array(n * n).map(function (_, k) { return Math.sin(k * 0.37); });
    const im = new Float64Array(n * n).map(function (_, k) { return Math.cos(k * 0.11); });
    fn = function () { fft2d(re, im, n, n, false); };
  } else if (a.what === 'asm') {
    const f = new ComplexField(n, n, a.dx, a.dx);
    for (let k = 0; k < f.size; k++) { f.re[k] = Math.sin(k * 0.37); f.im[k] = Math.cos(k * 0.11); }
    fn = function () { propagateAngularSpectrumInPlace(f, a.lambda, a.z, { bandLimit: true }); };
  } else if (a.what === 'fresnel') {
    const f = new ComplexField(n, n, a.dx, a.dx);
    for (let k = 0; k < f.size; k++) { f.re[k] = Math.sin(k * 0.37); f.im[k] = Math.cos(k * 0.11); }
    fn = function () { propagateFresnelFFT(f, a.lambda, a.z, {}); };
  } else if (a.what === 'wgs_ff') {
    const ill = new Float64Array(n * n).fill(1);
    const tgt = new Float64Array(n * n).fill(1);
    fn = function () {
      weightedGerchbergSaxton(ill, tgt, { nx: n, ny: n, iterations: a.iterations || 1, initPhase: randomPhase(n, n, 1) }, fourierCgh);
    };
  } else if (a.what === 'target') {
    // makeTarget of one kind (its autoRoi included) on an n x n grid
    fn = function () { makeTarget(a.kind || 'text', n, n, a.opts || {}); };
  } else if (a.what === 'image') {
    // planeToRGB of an n x n plane (options as given, e.g. a percentile)
    const plane = new Float64Array(n * n).map(function (_, k) { return Math.abs(Math.sin(k * 0.37)) ** 4; });
    fn = function () { planeToRGB(plane, a.opts || {}); };
  } else {
    throw new Error('unknown timing target ' + a.what);
  }
  fn();
  const t0 = nowMs();
  for (let i = 0; i < reps; i++) fn();
  res.ms = (nowMs() - t0) / reps;
  res.node = process.version;
} else {
  throw new Error('unknown op ' + op);
}
fs.writeFileSync(path.join(dir, 'result.json'), JSON.stringify(res));
