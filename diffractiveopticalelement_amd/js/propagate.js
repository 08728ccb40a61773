// Drop-in for src/propagate.js.  The angular spectrum, single-FFT Fresnel
// (forward and exact inverse) and the lens focal-plane transform run on the
// GPU: the transfer function, band limit, chirps and the fftshift pairs
// (which reduce to checkerboards) are fused into the FFT passes of
// libdoehip.so.  Scalar helpers (local frequency limit, critical distance,
// lens phase, wrapping) stay in JavaScript.
import { ComplexField } from './field.js';
import { native, dtypeCode } from './native.js';

export const DEFAULT_WAVELENGTH = 532e-9;

// ref: propagate.js:48-51
export function localFrequencyLimit(n, d, lambda, z) {
  const df = 1 / (n * d);
  return 1 / (lambda * Math.sqrt(Math.pow(2 * df * z, 2) + 1));
}

// Transfer function on the centred grid, built on the device in FP64
// (ref: propagate.js:57-102).
export function angularSpectrumTransferFunction(nx, ny, dx, dy, lambda, z, opts = {}) {
  const { bandLimit = true, evanescent = true, softEdge = 0.02 } = opts;
  return native.asmTransfer(nx, ny, dx, dy, lambda, z, bandLimit, evanescent, softEdge);
}

// ref: propagate.js:112-133
export function propagateAngularSpectrumInPlace(field, lambda, z, opts = {}) {
  if (z === 0) return field;
  const { bandLimit = true, evanescent = true, softEdge = 0.02 } = opts;
  const h = field._device();
  native.asm(h, h, field.dx, field.dy, lambda, z, bandLimit, evanescent, softEdge);
  field._markDevice();
  return field;
}

// ref: propagate.js:105-109
export function propagateAngularSpectrum(field, lambda, z, opts = {}) {
  const out = new ComplexField(field.nx, field.ny, field.dx, field.dy);
  if (z === 0) return field.clone();
  const { bandLimit = true, evanescent = true, softEdge = 0.02 } = opts;
  native.asm(field._device(), out._deviceOut(), field.dx, field.dy, lambda, z, bandLimit, evanescent, softEdge);
  out._markDevice();
  return out;
}

// ref: propagate.js:141-143
export function fresnelMinDistance(n, d, lambda) {
  return (n * d * d) / lambda;
}

// ref: propagate.js:155-239
export function propagateFresnelFFT(field, lambda, z, opts = {}) {
  const { throwOnUndersampled = false } = opts;
  const { nx, ny, dx } = field;
  const az = Math.abs(z);
  const zCritical = fresnelMinDistance(nx, dx, lambda);
  const undersampled = az < zCritical;
  if (undersampled && throwOnUndersampled) {
    throw new Error(
      `propagateFresnelFFT: z=${(az * 1e3).toFixed(3)} mm is below the critical distance ` +
        `z_c=${(zCritical * 1e3).toFixed(3)} mm for a ${nx}x${ny} grid at ${(dx * 1e6).toFixed(2)} um pitch — ` +
        'the Fresnel chirp is under-sampled. Use the angular spectrum instead.',
    );
  }
  const dux = (lambda * az) / (nx * dx);
  const duy = (lambda * az) / (ny * field.dy);
  const out = new ComplexField(nx, ny, dux, duy);
  native.fresnel(field._device(), out._deviceOut(), field.dx, field.dy, lambda, z, false, 0, 0);
  out._markDevice();
  return { field: out, dx: dux, dy: duy, z, magnification: dux / dx, undersampled, zCritical };
}

// ref: propagate.js:246-316
export function propagateFresnelFFTInverse(field, lambda, z, opts = {}) {
  const { dx, dy } = opts;
  if (!dx || !dy) {
    throw new Error('propagateFresnelFFTInverse: pass the DOE-plane pitch as {dx, dy}');
  }
  const out = new ComplexField(field.nx, field.ny, dx, dy);
  native.fresnel(field._device(), out._deviceOut(), dx, dy, lambda, z, true, field.dx, field.dy);
  out._markDevice();
  return out;
}

export function propagateFraunhofer(field, lambda, z, opts = {}) {
  return propagateFresnelFFT(field, lambda, z, opts);
}

// ref: propagate.js:342-374
export function focalPlaneSpectrum(field, lambda, f) {
  const { nx, ny, dx, dy } = field;
  const out = new ComplexField(nx, ny, (lambda * f) / (nx * dx), (lambda * f) / (ny * dy));
  native.lens(field._device(), out._deviceOut(), dx, dy, lambda, f);
  out._markDevice();
  return out;
}

// ref: propagate.js:383-389
export function propagate(field, lambda, z, opts = {}) {
  const { method = 'asm', pad = 1 } = opts;
  const src = pad > 1 ? field.resample(field.nx * pad, field.ny * pad) : field;
  if (method === 'fresnel') return propagateFresnelFFT(src, lambda, z, opts).field;
  return propagateAngularSpectrum(src, lambda, z, { bandLimit: opts.bandLimit === undefined ? true : opts.bandLimit });
}

// ref: propagate.js:395-407
export function lensPhase(field, lambda, f, opts = {}) {
  const { cx = 0, cy = 0, wrap = true } = opts;
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

// ref: propagate.js:410-420
export function wrapPhase(phi) {
  const tau = 2 * Math.PI;
  for (let i = 0; i < phi.length; i++) {
    let p = phi[i] % tau;
    if (p > Math.PI) p -= tau;
    else if (p <= -Math.PI) p += tau;
    phi[i] = p;
  }
  return phi;
}

// ref: propagate.js:427-434.  All distances run as one batch: the field is
// broadcast into steps+1 device slots, one angular-spectrum call propagates
// each slot by its own z, and every plane is handed out as its own field.
export function focusSweep(field, lambda, zStart, zEnd, steps, opts = {}) {
  const count = steps >= 0 ? Math.floor(steps) + 1 : 0;
  if (count === 0) return [];
  const { bandLimit = true, evanescent = true, softEdge = 0.02 } = opts;
  const zs = new Float64Array(count);
  for (let s = 0; s < count; s++) zs[s] = zStart + ((zEnd - zStart) * s) / steps;
  const dt = dtypeCode();
  const batch = native.fieldCreate(field.nx, field.ny, count, dt);
  native.fieldResample(field._device(dt), batch);
  native.asm(batch, batch, field.dx, field.dy, lambda, zs, bandLimit, evanescent, softEdge);
  const planes = [];
  for (let s = 0; s < count; s++) {
    const out = new ComplexField(field.nx, field.ny, field.dx, field.dy);
    native.fieldCopySlices(batch, s, out._deviceOut(dt), 0, 1);
    out._markDevice();
    planes.push({ z: zs[s], field: out });
  }
  return planes;
}
