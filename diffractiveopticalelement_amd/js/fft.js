// Drop-in for src/fft.js.  fft2d / fft1d run on the GPU through the addon
// (in place on the caller's Float64Array planes, same conventions: forward
// unscaled, inverse scaled by 1/N per axis).  The shifts are index
// permutations and dft2dNaive is the reference's O(N^4) test oracle; both
// stay plain loops.
import { native, dtypeCode } from './native.js';

export function clearTwiddleCache() {
  // device twiddle tables live in the library context for its lifetime
}

export function nextPow2(n) {
  let p = 1;
  while (p < n) p *= 2;
  return p;
}

export function isPow2(n) {
  return Number.isInteger(n) && n > 0 && (n & (n - 1)) === 0;
}

// ref: fft.js:86-139
export function fft1d(re, im, n, stride = 1, offset = 0, inverse = false) {
  native.fft1d(re, im, n, stride, offset, !!inverse, dtypeCode());
}

// ref: fft.js:148-154
export function fft2d(re, im, nx, ny, inverse = false) {
  if (re.length !== nx * ny || im.length !== nx * ny) {
    throw new Error('fft2d: plane length ' + re.length + ' does not match ' + nx + 'x' + ny);
  }
  native.fft2d(re, im, nx, ny, !!inverse, dtypeCode());
}

export function fft2dOut(re, im, nx, ny, inverse = false) {
  const or = Float64Array.from(re);
  const oi = Float64Array.from(im);
  fft2d(or, oi, nx, ny, inverse);
  return { re: or, im: oi };
}

// Circular half-length shift along one strided axis (ref: fft.js:170-181).
export function fftshift1d(arr, n, shift = true, stride = 1, offset = 0) {
  const half = n >>> 1;
  const d = shift ? half : n - half;
  for (let i = 0; i < half; i++) {
    const a = offset + i * stride;
    const b = offset + ((i + d) % n) * stride;
    const t = arr[a];
    arr[a] = arr[b];
    arr[b] = t;
  }
  return arr;
}

export function ifftshift1d(arr, n, stride = 1, offset = 0) {
  return fftshift1d(arr, n, false, stride, offset);
}

export function fftshift2d(re, im, nx, ny, shift = true) {
  for (let y = 0; y < ny; y++) {
    fftshift1d(re, nx, shift, 1, y * nx);
    if (im !== null) fftshift1d(im, nx, shift, 1, y * nx);
  }
  for (let x = 0; x < nx; x++) {
    fftshift1d(re, ny, shift, nx, x);
    if (im !== null) fftshift1d(im, ny, shift, nx, x);
  }
}

export function ifftshift2d(re, im, nx, ny) {
  fftshift2d(re, im, nx, ny, false);
}

// The reference's naive DFT oracle (ref: fft.js:208-237), kept for users'
// own tests: direct O(N^4) sum, same sign and scale conventions.
export function dft2dNaive(re, im, nx, ny, inverse = false) {
  const or = new Float64Array(nx * ny);
  const oi = new Float64Array(nx * ny);
  const sgn = inverse ? 1 : -1;
  const scale = inverse ? 1 / (nx * ny) : 1;
  for (let ky = 0; ky < ny; ky++) {
    for (let kx = 0; kx < nx; kx++) {
      let sr = 0;
      let si = 0;
      for (let y = 0; y < ny; y++) {
        const ay = (sgn * 2 * Math.PI * ky * y) / ny;
        for (let x = 0; x < nx; x++) {
          const ax = (sgn * 2 * Math.PI * kx * x) / nx;
          const c = Math.cos(ax) * Math.cos(ay) - Math.sin(ax) * Math.sin(ay);
          const s = Math.sin(ax) * Math.cos(ay) + Math.cos(ax) * Math.sin(ay);
          sr += re[x + y * nx] * c - im[x + y * nx] * s;
          si += re[x + y * nx] * s + im[x + y * nx] * c;
        }
      }
      or[kx + ky * nx] = sr * scale;
      oi[kx + ky * nx] = si * scale;
    }
  }
  return { re: or, im: oi };
}

export function fft2dReal(real, nx, ny) {
  const re = Float64Array.from(real);
  const im = new Float64Array(nx * ny);
  fft2d(re, im, nx, ny, false);
  return { re, im };
}
