// ORACLE (test infrastructure only) — FP64 restatement of the reference's
// propagators (ref: src/propagate.js:38-374, 395-420).  ES2019.
//
// Pinned by tests/test_oracle_cpu.py: closed-form H against numpy, ASM
// unitarity and round trip, Gaussian w(z), band-limit zeroing, brute-force
// Fresnel integral, Fresnel Parseval/inverse, ASM == Fresnel at z_c, lens
// transform unitarity and distance invariance; tests/test_oracle_pins_cpu.py:
// the raised-cosine band edge (1 up to (1 - softEdge) lim, 0 from lim, the
// cosine between, separable in fx and fy).

import { ComplexField } from './field.js';
import { fft2d, fftshift2d, ifftshift2d } from './fft.js';

// raised-cosine edge (ref: propagate.js:38-42)
function softWindow(v, lim, inner) {
  if (v <= inner) return 1;
  if (v >= lim) return 0;
  return 0.5 * (1 + Math.cos((Math.PI * (v - inner)) / (lim - inner)));
}

// ref: propagate.js:48-51
export function localFrequencyLimit(n, d, lambda, z) {
  const df = 1 / (n * d);
  return 1 / (lambda * Math.sqrt(Math.pow(2 * df * z, 2) + 1));
}

// Transfer function on the centred frequency grid (ref: propagate.js:57-102).
export function angularSpectrumTransferFunction(nx, ny, dx, dy, lambda, z, opts) {
  const o = opts || {};
  const bandLimit = o.bandLimit === undefined ? true : o.bandLimit;
  const evanescent = o.evanescent === undefined ? true : o.evanescent;
  const softEdge = o.softEdge === undefined ? 0.02 : o.softEdge;
  const re = new Float64Array(nx * ny);
  const im = new Float64Array(nx * ny);
  const dfx = 1 / (nx * dx);
  const dfy = 1 / (ny * dy);
  const invL2 = 1 / (lambda * lambda);
  const fxLim = bandLimit ? localFrequencyLimit(nx, dx, lambda, z) : Infinity;
  const fyLim = bandLimit ? localFrequencyLimit(ny, dy, lambda, z) : Infinity;
  const limX = Math.min(fxLim, 1 / (2 * dx));
  const limY = Math.min(fyLim, 1 / (2 * dy));
  let count = 0;
  for (let j = 0; j < ny; j++) {
    const fy = (j - ny / 2) * dfy;
    const wy = fyLim === Infinity ? 1 : softWindow(Math.abs(fy), limY, limY * (1 - softEdge));
    for (let i = 0; i < nx; i++) {
      const fx = (i - nx / 2) * dfx;
      const wx = fxLim === Infinity ? 1 : softWindow(Math.abs(fx), limX, limX * (1 - softEdge));
      const arg = invL2 - fx * fx - fy * fy;
      const k = i + j * nx;
      const w = wx * wy;
      if (arg >= 0) {
        const ph = 2 * Math.PI * z * Math.sqrt(arg);
        re[k] = w * Math.cos(ph);
        im[k] = w * Math.sin(ph);
        if (w > 0.5) count++;
      } else if (evanescent) {
        re[k] = 0;
        im[k] = 0;
      } else {
        const ph = 2 * Math.PI * z * Math.sqrt(-arg);
        re[k] = w * Math.cos(ph);
        im[k] = w * Math.sin(ph);
        count++;
      }
    }
  }
  return { re: re, im: im, propagatingFraction: count / (nx * ny) };
}

// ref: propagate.js:112-133
export function propagateAngularSpectrumInPlace(field, lambda, z, opts) {
  if (z === 0) return field;
  const nx = field.nx;
  const ny = field.ny;
  const re = field.re;
  const im = field.im;
  ifftshift2d(re, im, nx, ny);
  fft2d(re, im, nx, ny, false);
  fftshift2d(re, im, nx, ny, true);
  const H = angularSpectrumTransferFunction(nx, ny, field.dx, field.dy, lambda, z, opts);
  for (let k = 0; k < re.length; k++) {
    const ur = re[k];
    const ui = im[k];
    re[k] = ur * H.re[k] - ui * H.im[k];
    im[k] = ur * H.im[k] + ui * H.re[k];
  }
  ifftshift2d(re, im, nx, ny);
  fft2d(re, im, nx, ny, true);
  fftshift2d(re, im, nx, ny, true);
  return field;
}

export function propagateAngularSpectrum(field, lambda, z, opts) {
  return propagateAngularSpectrumInPlace(field.clone(), lambda, z, opts);
}

// ref: propagate.js:141-143
export function fresnelMinDistance(n, d, lambda) {
  return (n * d * d) / lambda;
}

// Single-FFT Fresnel to the magnified grid (ref: propagate.js:155-239).
export function propagateFresnelFFT(field, lambda, z, opts) {
  const o = opts || {};
  const nx = field.nx;
  const ny = field.ny;
  const dx = field.dx;
  const dy = field.dy;
  const az = Math.abs(z);
  const k = (2 * Math.PI) / lambda;
  const zCritical = fresnelMinDistance(nx, dx, lambda);
  const undersampled = az < zCritical;
  if (undersampled && o.throwOnUndersampled) {
    throw new Error(
      'propagateFresnelFFT: z=' + (az * 1e3).toFixed(3) + ' mm is below the critical distance ' +
        'z_c=' + (zCritical * 1e3).toFixed(3) + ' mm for a ' + nx + 'x' + ny + ' grid at ' +
        (dx * 1e6).toFixed(2) + ' um pitch — the Fresnel chirp is under-sampled. Use the angular spectrum instead.',
    );
  }
  const dux = (lambda * az) / (nx * dx);
  const duy = (lambda * az) / (ny * dy);
  const re = new Float64Array(nx * ny);
  const im = new Float64Array(nx * ny);
  for (let j = 0; j < ny; j++) {
    const y = field.y(j);
    for (let i = 0; i < nx; i++) {
      const x = field.x(i);
      const q = (Math.PI * (x * x + y * y)) / (lambda * az);
      const sg = (i + j) % 2 ? -1 : 1;
      const c = sg * Math.cos(q);
      const s = sg * Math.sin(q);
      const idx = i + j * nx;
      re[idx] = field.re[idx] * c - field.im[idx] * s;
      im[idx] = field.re[idx] * s + field.im[idx] * c;
    }
  }
  fft2d(re, im, nx, ny, false);
  const amp = (dx * dy) / (lambda * az);
  const kz = k * z - (Math.PI * (nx + ny)) / 2;
  const pr = amp * Math.sin(kz);
  const pj = -amp * Math.cos(kz);
  for (let j = 0; j < ny; j++) {
    const v = (j - ny / 2) * duy;
    for (let i = 0; i < nx; i++) {
      const u = (i - nx / 2) * dux;
      const q = (Math.PI * (u * u + v * v)) / (lambda * az);
      const sg = (i + j) % 2 ? -1 : 1;
      const c = sg * Math.cos(q);
      const s = sg * Math.sin(q);
      const idx = i + j * nx;
      const tr = re[idx] * pr - im[idx] * pj;
      const ti = re[idx] * pj + im[idx] * pr;
      re[idx] = tr * c - ti * s;
      im[idx] = tr * s + ti * c;
    }
  }
  const out = new ComplexField(nx, ny, dux, duy);
  out.re = re;
  out.im = im;
  return { field: out, dx: dux, dy: duy, z: z, magnification: dux / dx, undersampled: undersampled, zCritical: zCritical };
}

// Exact inverse using the inverse DFT (ref: propagate.js:246-316).
export function propagateFresnelFFTInverse(field, lambda, z, opts) {
  const o = opts || {};
  const dx = o.dx;
  const dy = o.dy;
  if (!dx || !dy) throw new Error('propagateFresnelFFTInverse: pass the DOE-plane pitch as {dx, dy}');
  const nx = field.nx;
  const ny = field.ny;
  const az = Math.abs(z);
  const k = (2 * Math.PI) / lambda;
  const dux = field.dx;
  const duy = field.dy;
  const re = new Float64Array(nx * ny);
  const im = new Float64Array(nx * ny);
  for (let j = 0; j < ny; j++) {
    const v = (j - ny / 2) * duy;
    for (let i = 0; i < nx; i++) {
      const u = (i - nx / 2) * dux;
      const q = (Math.PI * (u * u + v * v)) / (lambda * az);
      const sg = (i + j) % 2 ? -1 : 1;
      const c = sg * Math.cos(q);
      const s = -sg * Math.sin(q);
      const idx = i + j * nx;
      re[idx] = field.re[idx] * c - field.im[idx] * s;
      im[idx] = field.re[idx] * s + field.im[idx] * c;
    }
  }
  fft2d(re, im, nx, ny, true);
  const amp = (lambda * az) / (dx * dy);
  const kg = k * z - (Math.PI * (nx + ny)) / 2 - Math.PI / 2;
  const pr = amp * Math.cos(kg);
  const pj = -amp * Math.sin(kg);
  for (let j = 0; j < ny; j++) {
    const y = (j - ny / 2) * dy;
    for (let i = 0; i < nx; i++) {
      const x = (i - nx / 2) * dx;
      const q = (Math.PI * (x * x + y * y)) / (lambda * az);
      const sg = (i + j) % 2 ? -1 : 1;
      const c = sg * Math.cos(q);
      const s = -sg * Math.sin(q);
      const idx = i + j * nx;
      const tr = re[idx] * pr - im[idx] * pj;
      const ti = re[idx] * pj + im[idx] * pr;
      re[idx] = tr * c - ti * s;
      im[idx] = tr * s + ti * c;
    }
  }
  const out = new ComplexField(nx, ny, dx, dy);
  out.re = re;
  out.im = im;
  return out;
}

// Lens focal-plane transform (ref: propagate.js:342-374).
export function focalPlaneSpectrum(field, lambda, f) {
  const nx = field.nx;
  const ny = field.ny;
  const dx = field.dx;
  const dy = field.dy;
  const re = Float64Array.from(field.re);
  const im = Float64Array.from(field.im);
  ifftshift2d(re, im, nx, ny);
  fft2d(re, im, nx, ny, false);
  fftshift2d(re, im, nx, ny, true);
  const dux = (lambda * f) / (nx * dx);
  const duy = (lambda * f) / (ny * dy);
  const amp0 = (dx * dy) / (lambda * f);
  for (let j = 0; j < ny; j++) {
    const v = (j - ny / 2) * duy;
    for (let i = 0; i < nx; i++) {
      const u = (i - nx / 2) * dux;
      const idx = i + j * nx;
      const q = (Math.PI * (u * u + v * v)) / (lambda * f) - Math.PI / 2;
      const c = amp0 * Math.cos(q);
      const s = amp0 * Math.sin(q);
      const ur = re[idx];
      const ui = im[idx];
      re[idx] = ur * c - ui * s;
      im[idx] = ur * s + ui * c;
    }
  }
  const out = new ComplexField(nx, ny, dux, duy);
  out.re = re;
  out.im = im;
  return out;
}

// ref: propagate.js:410-420
export function wrapPhase(phi) {
  for (let i = 0; i < phi.length; i++) {
    let p = phi[i] % (2 * Math.PI);
    if (p > Math.PI) p -= 2 * Math.PI;
    else if (p <= -Math.PI) p += 2 * Math.PI;
    phi[i] = p;
  }
  return phi;
}

// ref: propagate.js:395-407
export function lensPhase(field, lambda, f, opts) {
  const o = opts || {};
  const cx = o.cx === undefined ? 0 : o.cx;
  const cy = o.cy === undefined ? 0 : o.cy;
  const wrap = o.wrap === undefined ? true : o.wrap;
  const k = (2 * Math.PI) / lambda;
  const phi = new Float64Array(field.size);
  for (let j = 0; j < field.ny; j++) {
    const y = field.y(j) - cy;
    for (let i = 0; i < field.nx; i++) {
      const x = field.x(i) - cx;
      phi[i + j * field.nx] = (-k * (x * x + y * y)) / (2 * f);
    }
  }
  return wrap ? wrapPhase(phi) : phi;
}

// ref: propagate.js:427-434: the field propagated to steps + 1 evenly spaced
// distances from zStart to zEnd, one propagation per distance.
export function focusSweep(field, lambda, zStart, zEnd, steps, opts) {
  const out = [];
  for (let s = 0; s <= steps; s++) {
    const z = zStart + ((zEnd - zStart) * s) / steps;
    out.push({ z: z, field: propagateAngularSpectrum(field, lambda, z, opts) });
  }
  return out;
}
