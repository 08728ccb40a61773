// GPU check of the JavaScript drop-in (diffractiveopticalelement_amd/js)
// against the FP64 oracle, run by tests/test_gpu_dropin.py.  Prints one JSON
// object per check and exits non-zero if any check failed.
import * as gpu from '../../diffractiveopticalelement_amd/js/index.js';
import * as offt from '../../oracle/js/fft.js';
import * as oprop from '../../oracle/js/propagate.js';
import { ComplexField as OField } from '../../oracle/js/field.js';
import { ProjectorSystem as OProjector } from '../../oracle/js/projector.js';
import { designDOE as oDesign, gerchbergSaxtonMultiPlane as oMultiPlane, blurRaster as oBlur, errorDiffusionQuantize as oFS, phaseToHeight as oHeight } from '../../oracle/js/doe.js';
import { makeTarget as oMakeTarget, autoRoi as oAutoRoi, resampleRaster as oResample } from '../../oracle/js/projector.js';
import { planeToRGB as oRGB, planeToGray as oGray } from '../../oracle/js/color.js';
import * as om from '../../oracle/js/metrics.js';
import zlib from 'zlib';

// screenMetrics' figures (ref: projector.js:554-614) from the oracle's FP64
// metric loops over the downloaded irradiance of a device screen
function hostScreenMetrics(sys, field) {
  const tgt = sys.target;
  const irr = field.intensity();
  const mask = tgt.roi;
  const element = new OField(sys.nx, sys.ny, sys.dx, sys.dy);
  element.setAmplitudePhase(sys.result.illumination, sys.result.phase);
  const cell = field.dx * field.dy;
  const scaled = Float64Array.from(tgt.irradiance);
  const pt = om.maskedSum(irr, mask) * cell;
  const tt = om.maskedSum(scaled, null) * cell;
  const s = tt > 0 ? pt / tt : 0;
  for (let k = 0; k < scaled.length; k++) scaled[k] *= s;
  const provisional = om.regionStats(irr, mask, {});
  const stats = om.regionStats(irr, mask, { threshold: 0.1 * provisional.max });
  const spots = tgt.grid ? om.spotUniformity(irr, sys.nx, sys.ny, tgt.grid) : null;
  const incident = om.maskedSum(sys.illumination(), null) * sys.dx * sys.dy;
  const total = om.maskedSum(irr, null) * cell;
  const inSignal = om.maskedSum(irr, mask) * cell;
  const rms = om.rmsError(irr, scaled, mask);
  const corr = om.correlation(irr, scaled, mask);
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
    ssim: om.ssim(irr, scaled, sys.nx, sys.ny, { window: 8, weight: 'target' }),
    snr: rms > 0 ? 20 * Math.log10(1 / rms) : Infinity,
    zeroOrder: om.zeroOrderFraction(element),
    rmsRadius: om.rmsRadius(irr, sys.nx, field.dx, 0, 0, field.dy).r,
    meanIntensity: stats.mean,
    peakIntensity: stats.max,
    powerInSignal: inSignal,
    powerAtScreen: total,
    powerIncident: incident,
  };
}

const results = [];
function check(name, ok, info) {
  results.push({ name, ok: !!ok, info });
  console.log(JSON.stringify({ name, ok: !!ok, info }));
}
function rel(ar, ai, br, bi) {
  let n = 0;
  let d = 0;
  for (let k = 0; k < ar.length; k++) {
    n += (ar[k] - br[k]) ** 2 + (ai[k] - bi[k]) ** 2;
    d += br[k] ** 2 + bi[k] ** 2;
  }
  return Math.sqrt(n / d);
}
function lcg(n, seed) {
  let s = seed;
  const a = new Float64Array(n);
  for (let i = 0; i < n; i++) {
    s = (1103515245 * s + 12345) % 2147483648;
    a[i] = s / 2147483648 - 0.5;
  }
  return a;
}
function pair(nx, ny, dx, seed) {
  const g = new gpu.ComplexField(nx, ny, dx, dx);
  const o = new OField(nx, ny, dx, dx);
  const re = lcg(nx * ny, seed);
  const im = lcg(nx * ny, seed + 1);
  g.re.set(re);
  g.im.set(im);
  o.re.set(re);
  o.im.set(im);
  return [g, o];
}
const LAM = 532e-9;

for (const prec of ['f32', 'f64']) {
  gpu.setPrecision(prec);
  const tol = prec === 'f64' ? 1e-12 : 1e-5;
  // fft2d / fft1d (ref: fft.js:86-154)
  {
    const re = lcg(64 * 32, 3);
    const im = lcg(64 * 32, 4);
    const a = [Float64Array.from(re), Float64Array.from(im)];
    const b = [Float64Array.from(re), Float64Array.from(im)];
    gpu.fft2d(a[0], a[1], 64, 32, false);
    offt.fft2d(b[0], b[1], 64, 32, false);
    check(`fft2d ${prec}`, rel(a[0], a[1], b[0], b[1]) < tol, rel(a[0], a[1], b[0], b[1]));
    gpu.fft2d(a[0], a[1], 64, 32, true);
    check(`fft2d round trip ${prec}`, rel(a[0], a[1], re, im) < tol, rel(a[0], a[1], re, im));
    const c = [Float64Array.from(re), Float64Array.from(im)];
    const d = [Float64Array.from(re), Float64Array.from(im)];
    gpu.fft1d(c[0], c[1], 32, 64, 5, false);
    offt.fft1d(d[0], d[1], 32, 64, 5, false);
    check(`fft1d strided ${prec}`, rel(c[0], c[1], d[0], d[1]) < tol, rel(c[0], c[1], d[0], d[1]));
  }
  // propagators (ref: propagate.js)
  for (const band of [false, true]) {
    const [g, o] = pair(128, 128, 8e-6, 11);
    const gr = gpu.propagateAngularSpectrum(g, LAM, 0.04, { bandLimit: band });
    const or = oprop.propagateAngularSpectrum(o, LAM, 0.04, { bandLimit: band });
    const e = rel(gr.re, gr.im, or.re, or.im);
    check(`asm band=${band} ${prec}`, e < tol, e);
    gpu.propagateAngularSpectrumInPlace(g, LAM, 0.04, { bandLimit: band });
    oprop.propagateAngularSpectrumInPlace(o, LAM, 0.04, { bandLimit: band });
    const e2 = rel(g.re, g.im, o.re, o.im);
    check(`asm in place band=${band} ${prec}`, e2 < tol, e2);
  }
  {
    const [g, o] = pair(128, 128, 8e-6, 21);
    const fg = gpu.propagateFresnelFFT(g, LAM, 0.03);
    const fo = oprop.propagateFresnelFFT(o, LAM, 0.03);
    const e = rel(fg.field.re, fg.field.im, fo.field.re, fo.field.im);
    check(`fresnel ${prec}`, e < tol && fg.dx === fo.dx && fg.undersampled === fo.undersampled, e);
    const bg = gpu.propagateFresnelFFTInverse(fg.field, LAM, 0.03, { dx: 8e-6, dy: 8e-6 });
    const bo = oprop.propagateFresnelFFTInverse(fo.field, LAM, 0.03, { dx: 8e-6, dy: 8e-6 });
    const e2 = rel(bg.re, bg.im, bo.re, bo.im);
    check(`fresnel inverse ${prec}`, e2 < tol, e2);
    const lg = gpu.focalPlaneSpectrum(g, LAM, 0.05);
    const lo = oprop.focalPlaneSpectrum(o, LAM, 0.05);
    const e3 = rel(lg.re, lg.im, lo.re, lo.im);
    check(`lens ${prec}`, e3 < tol && lg.dx === lo.dx, e3);
    let threw = false;
    try {
      gpu.propagateFresnelFFTInverse(fg.field, LAM, 0.03, {});
    } catch (err) {
      threw = /pass the DOE-plane pitch/.test(err.message);
    }
    check('fresnel inverse without pitch throws', threw);
  }
  {
    const Hg = gpu.angularSpectrumTransferFunction(64, 64, 8e-6, 8e-6, LAM, 0.05, { softEdge: 0.01 });
    const Ho = oprop.angularSpectrumTransferFunction(64, 64, 8e-6, 8e-6, LAM, 0.05, { softEdge: 0.01 });
    const e = rel(Hg.re, Hg.im, Ho.re, Ho.im);
    check('transfer function', e < 1e-12 && Hg.propagatingFraction === Ho.propagatingFraction, e);
  }
}

// device mirror semantics: host writes after a device op are honoured
{
  gpu.setPrecision('f64');
  const [g, o] = pair(32, 32, 8e-6, 5);
  gpu.propagateAngularSpectrumInPlace(g, LAM, 0.02, { bandLimit: false });
  g.re[7] += 1.0; // host write on the downloaded mirror
  oprop.propagateAngularSpectrumInPlace(o, LAM, 0.02, { bandLimit: false });
  o.re[7] += 1.0;
  gpu.propagateAngularSpectrumInPlace(g, LAM, -0.02, { bandLimit: false });
  oprop.propagateAngularSpectrumInPlace(o, LAM, -0.02, { bandLimit: false });
  check('mirror write-back', rel(g.re, g.im, o.re, o.im) < 1e-12, rel(g.re, g.im, o.re, o.im));
}

// ComplexField resample (bilinear at another pitch, pad / crop) and
// sampleImage on the device: FP64 fields equal the oracle bit for bit
{
  gpu.setPrecision('f64');
  const [g, o] = pair(64, 64, 4e-6, 9);
  for (const [n2, dx2] of [[128, 2.5e-6], [32, 6e-6], [128, 4e-6]]) {
    const rg = g.resample(n2, n2, dx2, dx2);
    const ro = o.resample(n2, n2, dx2, dx2);
    check(`ComplexField.resample ${n2} @ ${dx2}`, sameBits(rg.re, ro.re) && sameBits(rg.im, ro.im) && rg.dx === ro.dx);
  }
  const img = lcg(40 * 30, 17).map((v) => v + 0.5);
  const sg = new gpu.ComplexField(64, 64, 4e-6).sampleImage(img, 40, 30, 0.8, 0.25);
  const so = new OField(64, 64, 4e-6).sampleImage(img, 40, 30, 0.8, 0.25);
  check('ComplexField.sampleImage', sameBits(sg.re, so.re) && sameBits(sg.im, so.im));
}

// ProjectorSystem end to end: design on the GPU, screenMetrics within 1e-3
function metricsClose(a, b, keys, tol) {
  const bad = keys.filter((k) => !(Math.abs(a[k] - b[k]) < tol));
  return { ok: bad.length === 0, bad: bad.map((k) => [k, a[k], b[k]]) };
}
const KEYS = ['rmse', 'efficiency', 'correlation', 'ssim', 'fluxFraction', 'zeroOrder', 'uniformity', 'flatness'];
// image-plane: the small reference configuration itself (n 256, 100 WGS
// iterations, seed 0, text 'DOE'); far-field: 128^2 GS, 40 iterations
const PROJ_CFGS = {
  'image-plane': { n: 256, pitch: 8e-6, lambda: 532e-9, distance: 0.03, mode: 'image-plane', algorithm: 'wgs',
    iterations: 100, seed: 0, target: { kind: 'text', text: 'DOE' } },
  'far-field': { n: 128, pitch: 8e-6, lambda: 532e-9, mode: 'far-field', iterations: 40, seed: 12345,
    target: { kind: 'text', text: 'DOE' } },
};
const oracleProj = {};
for (const mode of Object.keys(PROJ_CFGS)) {
  const os = new OProjector(PROJ_CFGS[mode]);
  const ro = os.design();
  oracleProj[mode] = { ro, mo: os.screenMetrics(os.simulate()) };
}
for (const prec of ['f32', 'f64']) {
  gpu.setPrecision(prec);
  for (const mode of ['image-plane', 'far-field']) {
    const cfg = PROJ_CFGS[mode];
    const gs = new gpu.ProjectorSystem(cfg);
    const rg = gs.design();
    const { ro, mo } = oracleProj[mode];
    const screen = gs.simulate();
    const mg = gs.screenMetrics(screen);
    const c = metricsClose(mg, mo, KEYS, 1e-3);
    check(`projector ${mode} ${prec} screenMetrics`, c.ok, c.bad);
    // device reductions vs the host FP64 loops over the same downloaded screen
    const mh = hostScreenMetrics(gs, screen);
    const rk = Object.keys(mh).filter((k) => mh[k] !== null && Number.isFinite(mh[k]));
    const bad = rk.filter((k) => !(Math.abs(mg[k] - mh[k]) <= 1e-9 * Math.abs(mh[k]) + 1e-15));
    check(`projector ${mode} ${prec} screenMetrics device vs host`, bad.length === 0 && rk.length >= 17,
      bad.map((k) => [k, mg[k], mh[k]]));
    let maxe = 0;
    for (let i = 0; i < 10; i++) maxe = Math.max(maxe, Math.abs(rg.error[i] - ro.error[i]), Math.abs(rg.efficiency[i] - ro.efficiency[i]));
    check(`projector ${mode} ${prec} first 10 iterates`, maxe < 1e-4, maxe);
    check(`projector ${mode} ${prec} algorithm`, rg.algorithm === ro.algorithm && rg.mode === ro.mode, [rg.algorithm, rg.mode]);
    if (mode === 'image-plane') {
      // the reference's own thresholds (ref test/projector.test.js:62-84)
      check(`projector thresholds ${prec}`, mg.rmse < 0.06 && mg.correlation > 0.98 && mg.efficiency > 0.55 &&
        mg.ssim > 0.95 && mg.zeroOrder < 0.01 && mg.fluxFraction > 0.55 && mg.snr > 15, mg);
    }
  }
}

// Parseval through design + simulate (ref test/projector.test.js:86-102), FP64 mode
{
  gpu.setPrecision('f64');
  const sys = new gpu.ProjectorSystem({ n: 64, pitch: 8e-6, lambda: 532e-9, distance: 0.03, iterations: 10, target: { kind: 'disc' } });
  sys.design({ iterations: 10 });
  const irr = sys.simulate().intensity();
  let screen = 0;
  for (const v of irr) screen += v;
  let inc = 0;
  for (const v of sys.illumination()) inc += v * v;
  check('Parseval through design+simulate (f64)', Math.abs(screen / inc - 1) < 1e-9, screen / inc - 1);
}

// designDOE determinism and wrapped phase (ref test/doe.test.js:350-382)
{
  gpu.setPrecision('f32');
  const t = new Float64Array(64 * 64);
  for (let j = 0; j < 64; j++) for (let i = 0; i < 64; i++) t[i + j * 64] = Math.hypot(i - 32, j - 32) <= 20 ? 1 : 0;
  const cfg = { nx: 64, ny: 64, dx: 8e-6, lambda: LAM, z: 0.03, iterations: 10, target: t, mode: 'image-plane' };
  const a = gpu.designDOE(cfg);
  const b = gpu.designDOE(cfg);
  let md = 0;
  let wrapped = true;
  for (let k = 0; k < a.phase.length; k++) {
    md = Math.max(md, Math.abs(a.phase[k] - b.phase[k]));
    if (Math.abs(a.phase[k]) > Math.PI + 1e-6) wrapped = false;
  }
  check('designDOE deterministic', md < 1e-12 && wrapped && a.error.length === 10, md);
  const ref = oDesign(cfg);
  check('designDOE error history vs oracle', Math.abs(a.error[9] - ref.error[9]) < 1e-3, [a.error[9], ref.error[9]]);
}

// Multi-distance callers (ref: propagate.js:427-434, projector.js:622-653,
// doe.js:557-626): batched on the device, equal to the oracle's per-z loops
for (const prec of ['f32', 'f64']) {
  gpu.setPrecision(prec);
  const tol = prec === 'f64' ? 1e-12 : 1e-5;
  {
    const [g, o] = pair(64, 64, 8e-6, 31);
    const sg = gpu.focusSweep(g, LAM, 0.01, 0.03, 4, { bandLimit: true });
    const so = oprop.focusSweep(o, LAM, 0.01, 0.03, 4, { bandLimit: true });
    let worst = 0;
    for (let s = 0; s < so.length; s++) worst = Math.max(worst, rel(sg[s].field.re, sg[s].field.im, so[s].field.re, so[s].field.im));
    check(`focusSweep ${prec}`, sg.length === 5 && worst < tol && sg[4].z === so[4].z, worst);
  }
  {
    const cfg = { n: 64, pitch: 8e-6, lambda: 532e-9, distance: 0.03, mode: 'image-plane', iterations: 8, seed: 3, target: { kind: 'disc' } };
    const gs = new gpu.ProjectorSystem(cfg);
    const os = new OProjector(cfg);
    gs.design();
    os.design();
    // the same element on both sides, so the curves compare the callers alone
    os.result.phase = Float64Array.from(gs.result.phase);
    const dg = gs.defocusCurve({ points: 5, span: 0.5, pad: 2 });
    const dor = os.defocusCurve({ points: 5, span: 0.5, pad: 2 });
    const mtol = prec === 'f64' ? 1e-9 : 1e-3;
    let bad = [];
    for (let i = 0; i < 5; i++) {
      for (const k of ['fidelity', 'correlation', 'ssim', 'rmse', 'efficiency', 'uniformity']) {
        if (!(Math.abs(dg[i][k] - dor[i][k]) <= mtol)) bad.push([i, k, dg[i][k], dor[i][k]]);
      }
      if (dg[i].z !== dor[i].z) bad.push([i, 'z', dg[i].z, dor[i].z]);
    }
    check(`defocusCurve ${prec}`, bad.length === 0, bad);
    const fg = gs.focusStack({ points: 3, span: 0.5, pad: 2 });
    const fo = os.focusStack({ points: 3, span: 0.5, pad: 2 });
    let worst = 0;
    for (let i = 0; i < 3; i++) {
      let num = 0;
      let den = 0;
      for (let k = 0; k < fo[i].intensity.length; k++) {
        num += (fg[i].intensity[k] - fo[i].intensity[k]) ** 2;
        den += fo[i].intensity[k] ** 2;
      }
      worst = Math.max(worst, Math.sqrt(num / den));
    }
    check(`focusStack ${prec}`, worst < (prec === 'f64' ? 1e-12 : 1e-5), worst);
  }
  {
    const n = 64;
    const ill = new Float64Array(n * n);
    const ta = new Float64Array(n * n);
    const tb = new Float64Array(n * n);
    const roi = new Float64Array(n * n);
    for (let j = 0; j < n; j++) {
      for (let i = 0; i < n; i++) {
        const r = Math.hypot(i - n / 2, j - n / 2);
        ill[i + j * n] = r <= 0.45 * n ? 1 : 0;
        ta[i + j * n] = r <= 0.2 * n ? 1.3 : 0;
        tb[i + j * n] = Math.abs(i - n / 2) < 8 && Math.abs(j - n / 2) < 20 ? 1.1 : 0;
        roi[i + j * n] = r <= 0.28 * n ? 1 : 0;
      }
    }
    const planes = [{ z: 0.015, targetAmp: ta, roi }, { z: 0.03, targetAmp: tb, weight: 0.5 }];
    const opts = { nx: n, ny: n, dx: 8e-6, lambda: LAM, iterations: 10 };
    const mg = gpu.gerchbergSaxtonMultiPlane(ill, planes, opts);
    const mo = oMultiPlane(ill, planes, opts);
    let worst = 0;
    for (let q = 0; q < 2; q++) for (let it = 0; it < 10; it++) worst = Math.max(worst, Math.abs(mg.perPlaneError[q][it] - mo.perPlaneError[q][it]));
    check(`gerchbergSaxtonMultiPlane ${prec}`, worst < (prec === 'f64' ? 1e-9 : 1e-4) && mg.error.length === 10, worst);
    let threw = false;
    try {
      gpu.gerchbergSaxtonMultiPlane(ill, [], opts);
    } catch (err) {
      threw = /no planes/.test(err.message);
    }
    check('gerchbergSaxtonMultiPlane without planes throws', threw);
  }
}
{
  gpu.setPrecision('f32');
  const sys = new gpu.ProjectorSystem({ n: 64, pitch: 8e-6, lambda: 532e-9, distance: 0.03, iterations: 6, target: { kind: 'disc' } });
  const r = sys.designMultiPlane([{ z: 0.02, kind: 'disc' }, { z: 0.04, kind: 'text', text: 'A' }], { iterations: 6 });
  check('designMultiPlane', r.mode === 'multi-plane' && r.phase.length === 64 * 64 && r.perPlaneError.length === 2 &&
    r.error.every((e) => Number.isFinite(e)), r.error);
}

// targets, ROIs, blur, resample on the device (ref: projector.js:126-315, doe.js:761-794)
function same(a, b) {
  if (a.length !== b.length) return false;
  for (let k = 0; k < a.length; k++) if (a[k] !== b[k] && !(a[k] !== a[k] && b[k] !== b[k])) return false;
  return true;
}
// bit for bit: -0 and +0 differ, NaNs compare equal
function sameBits(a, b) {
  if (a.length !== b.length) return false;
  for (let k = 0; k < a.length; k++) if (!Object.is(a[k], b[k])) return false;
  return true;
}
function maxRel(a, b) {
  let w = 0;
  for (let k = 0; k < a.length; k++) w = Math.max(w, Math.abs(a[k] - b[k]) / Math.max(1e-300, Math.abs(b[k])));
  return w;
}
{
  const kinds = [['text', { text: 'DOE' }], ['disc', {}], ['ring', { inner: 0.2 }], ['checker', { cells: 7 }],
    ['lines', { period: 10 }], ['zone', {}], ['grid', { cols: 5, rows: 4, missing: [[1, 1], [2.5, 0], [4, 3]] }],
    ['dots', { cols: 3, rows: 3, dilate: 4, softEdge: 0 }]];
  for (const [nx, ny] of [[128, 128], [90, 70]]) {
    for (const [kind, opts] of kinds) {
      const g = gpu.makeTarget(kind, nx, ny, opts);
      const o = oMakeTarget(kind, nx, ny, opts);
      const exactIrr = kind !== 'grid' && kind !== 'dots';
      const irrOk = exactIrr ? same(g.irradiance, o.irradiance) : maxRel(g.irradiance, o.irradiance) < 1e-14;
      check(`makeTarget ${kind} ${nx}x${ny}`, irrOk && same(g.roi, o.roi) && g.name === o.name &&
        JSON.stringify(g.grid) === JSON.stringify(o.grid), { name: g.name });
    }
  }
  const img = lcg(33 * 21, 9).map((v) => v + 0.5);
  const gi = gpu.makeTarget('image', 100, 80, { raster: img, width: 33, height: 21, gamma: 1.8, name: 'pic' });
  const oi = oMakeTarget('image', 100, 80, { raster: img, width: 33, height: 21, gamma: 1.8, name: 'pic' });
  check('makeTarget image gamma', maxRel(gi.irradiance, oi.irradiance) < 1e-14 && gi.name === 'pic', maxRel(gi.irradiance, oi.irradiance));
  const custom = lcg(77 * 61, 21).map((v) => (v + 0.5) ** 5);
  check('autoRoi', same(gpu.autoRoi(custom, 77, 61, { dilate: 3, soft: 1.6, threshold: 0.1 }),
    oAutoRoi(custom, 77, 61, { dilate: 3, soft: 1.6, threshold: 0.1 })));
  const bg = gpu.blurRaster(custom, 77, 61, 1.7);
  const bo = oBlur(custom, 77, 61, 1.7);
  check('blurRaster', same(bg, bo), maxRel(bg, bo));
  check('blurRaster sigma 0 copies', same(gpu.blurRaster(custom, 77, 61, 0), custom));
  for (const fit of ['contain', 'cover']) {
    const rg = gpu.resampleRaster(img, 33, 21, 64, 50, { fit, background: 0.25, gain: 2 });
    const ro = oResample(img, 33, 21, 64, 50, { fit, background: 0.25, gain: 2 });
    check(`resampleRaster ${fit}`, same(rg, ro));
  }
  // a plane shorter than nx * ny is refused before any device copy
  let short = false;
  try {
    gpu.autoRoi(new Float64Array(77 * 61 - 5), 77, 61);
  } catch (err) {
    short = err instanceof RangeError && /shorter than/.test(err.message);
  }
  check('autoRoi short plane throws RangeError', short);
  let threw = false;
  try {
    gpu.makeTarget('star', 16, 16);
  } catch (err) {
    threw = /unknown kind "star"/.test(err.message);
  }
  check('makeTarget unknown kind throws', threw);
  // setTarget through the projector keeps the reference's result
  const sys = new gpu.ProjectorSystem({ n: 64, target: { kind: 'dots', cols: 4, rows: 4 } });
  const osys = new OProjector({ n: 64, target: { kind: 'dots', cols: 4, rows: 4 } });
  check('ProjectorSystem target', maxRel(sys.target.irradiance, osys.target.irradiance) < 1e-14 &&
    same(sys.target.roi, osys.target.roi));
}

// fabrication output (ref: doe.js:185-226, 967-976)
{
  const ph = lcg(128 * 128, 51).map((v) => v * 12);
  for (const snake of [false, true]) {
    for (const wrapFirst of [true, false]) {
      const g = gpu.errorDiffusionQuantize(ph, 4, { snake, wrapFirst });
      const o = oFS(ph, 4, { snake, wrapFirst });
      check(`errorDiffusionQuantize snake=${snake} wrapFirst=${wrapFirst}`, sameBits(g, o));
    }
  }
  check('phaseToHeight', sameBits(gpu.phaseToHeight(ph, 633e-9, 1.5), oHeight(ph, 633e-9, 1.5)));
  // a 3 x 4 'square' grid by the reference's rule (round(sqrt(12)) = 3 divides 12)
  const p12 = lcg(12, 3).map((v) => v * 9);
  check('errorDiffusionQuantize 3x4', sameBits(gpu.errorDiffusionQuantize(p12, 3, { snake: false }), oFS(p12, 3, { snake: false })));
  let threw = false;
  try {
    gpu.errorDiffusionQuantize(new Float64Array(10), 4, { snake: false });  // 10 / round(sqrt(10)) is not a row count
  } catch (err) {
    threw = /expected a square grid/.test(err.message);
  }
  check('errorDiffusionQuantize non-square throws', threw);
}

// figure pixels and PNG output (ref: color.js:87-160, png.js:89-180)
async function figureChecks() {
  const plane = lcg(97 * 53, 31).map((v, k) => (k % 7 === 0 ? v * 10 : v));
  for (const opts of [{}, { map: 'turbo' }, { map: 'inferno', min: -0.2 }, { map: 'phase', percentile: 0.95 }]) {
    const g = gpu.planeToRGB(plane, opts);
    const o = oRGB(plane, opts).rgb;
    check(`planeToRGB ${JSON.stringify(opts)}`, same(g, o));
  }
  for (const opts of [{}, { invert: true }, { min: 0, max: 2 }]) {
    check(`planeToGray ${JSON.stringify(opts)}`, same(gpu.planeToGray(plane, opts), oGray(plane, opts).gray));
  }
  // a plane longer than 65536 samples travels as one long row
  const big = lcg(512 * 512, 61);
  check('planeToRGB 512^2 plane', same(gpu.planeToRGB(big, { percentile: 0.9 }), oRGB(big, { percentile: 0.9 }).rgb));
  const f = new gpu.ComplexField(64, 32, 8e-6, 8e-6);
  f.re.set(lcg(2048, 41));
  f.im.set(lcg(2048, 43));
  const inten = f.intensity();
  check('fieldToRGB intensity', same(gpu.fieldToRGB(f, { map: 'viridis' }), oRGB(inten, { map: 'viridis' }).rgb));
  const bytes = await gpu.writeFieldPNG(null, f, { map: 'viridis' });
  // decode: chunks, CRCs, inflate, scanlines
  let o = 8;
  const u = Buffer.from(bytes);
  let idat = [];
  let crcOk = true;
  while (o < u.length) {
    const len = u.readUInt32BE(o);
    const type = u.toString('latin1', o + 4, o + 8);
    if (type === 'IDAT') idat.push(u.subarray(o + 8, o + 8 + len));
    const c = zlib.crc32 ? zlib.crc32(u.subarray(o + 4, o + 8 + len)) : null;
    if (c !== null && c !== u.readUInt32BE(o + 8 + len)) crcOk = false;
    o += 12 + len;
  }
  const raw = zlib.inflateSync(Buffer.concat(idat));
  const px = oRGB(inten, { map: 'viridis' }).rgb;
  let rowsOk = raw.length === 32 * (1 + 192);
  for (let y = 0; y < 32 && rowsOk; y++) {
    rowsOk = raw[y * 193] === 0 && same(raw.subarray(y * 193 + 1, (y + 1) * 193), px.subarray(y * 192, (y + 1) * 192));
  }
  check('writeFieldPNG decodes to the mapped pixels', rowsOk && crcOk && u.readUInt32BE(16) === 64 && u.readUInt32BE(20) === 32);
  const gpng = await gpu.writePlanePNG(null, Float64Array.from({ length: 16 }, (_, k) => k / 15), 4, 4);
  const graw = zlib.inflateSync(Buffer.from(gpng.subarray(41, 41 + Buffer.from(gpng).readUInt32BE(33))));
  check('writePlanePNG brightest pixel', graw.length === 20 && graw[3 * 5 + 1 + 3] > 240, graw[3 * 5 + 1 + 3]);
}

figureChecks().then(() => {
  const failed = results.filter((r) => !r.ok);
  console.log(JSON.stringify({ summary: true, total: results.length, failed: failed.length }));
  process.exit(failed.length ? 1 : 0);
}, (err) => {
  console.log(JSON.stringify({ name: 'figure checks', ok: false, info: String(err && err.stack) }));
  process.exit(1);
});
