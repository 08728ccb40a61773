// ORACLE (test infrastructure only) — FP64 CPU restatement of the reference's
// FFT (ref: src/fft.js).  Only tests/, the smoke check and bench.py's CPU
// baseline may load it, and only as the checker.  ES2019: runs on Node 12.
//
// Pinned by: tests/test_oracle_cpu.py (numpy.fft, naive DFT, Parseval,
// delta -> flat, shift involution — the known answers of ref test/fft.test.js).

const TWO_PI = 2 * Math.PI;
const twCache = new Map();

// exp(-2*pi*i*k/len) for k < len/2, cached per length (ref: fft.js:29-48).
function stageTwiddles(len) {
  let t = twCache.get(len);
  if (t) return t;
  const h = len >>> 1;
  const c = new Float64Array(h);
  const s = new Float64Array(h);
  for (let k = 0; k < h; k++) {
    const ang = (-2 * Math.PI * k) / len;
    c[k] = Math.cos(ang);
    s[k] = Math.sin(ang);
  }
  t = { c: c, s: s };
  if (twCache.size >= 64) twCache.delete(twCache.keys().next().value);
  twCache.set(len, t);
  return t;
}

export function clearTwiddleCache() {
  twCache.clear();
}

export function isPow2(n) {
  return Number.isInteger(n) && n > 0 && (n & (n - 1)) === 0;
}

export function nextPow2(n) {
  let p = 1;
  while (p < n) p *= 2;
  return p;
}

// Radix-2 decimation-in-time over a strided slice (ref: fft.js:86-139):
// bit-reversal permutation, log2(n) butterfly stages, inverse conjugates the
// kernel and scales by 1/n.
export function fft1d(re, im, n, stride, offset, inverse) {
  stride = stride === undefined ? 1 : stride;
  offset = offset === undefined ? 0 : offset;
  if (n <= 1) return;
  if (!isPow2(n)) throw new Error('fft1d: length must be a power of two, got ' + n);
  let j = 0;
  for (let i = 1; i < n; i++) {
    let bit = n >>> 1;
    while (j & bit) {
      j ^= bit;
      bit >>>= 1;
    }
    j ^= bit;
    if (i < j) {
      const a = offset + i * stride;
      const b = offset + j * stride;
      let t = re[a];
      re[a] = re[b];
      re[b] = t;
      t = im[a];
      im[a] = im[b];
      im[b] = t;
    }
  }
  for (let len = 2; len <= n; len *= 2) {
    const half = len >>> 1;
    const tw = stageTwiddles(len);
    for (let base = 0; base < n; base += len) {
      const p0 = offset + base * stride;
      for (let k = 0; k < half; k++) {
        const wr = tw.c[k];
        const wi = inverse ? -tw.s[k] : tw.s[k];
        const p = p0 + k * stride;
        const q = p + half * stride;
        const tr = re[q] * wr - im[q] * wi;
        const ti = re[q] * wi + im[q] * wr;
        re[q] = re[p] - tr;
        im[q] = im[p] - ti;
        re[p] += tr;
        im[p] += ti;
      }
    }
  }
  if (inverse) {
    const sc = 1 / n;
    for (let i = 0; i < n; i++) {
      const p = offset + i * stride;
      re[p] *= sc;
      im[p] *= sc;
    }
  }
}

// Rows then columns, in place (ref: fft.js:148-154).
export function fft2d(re, im, nx, ny, inverse) {
  if (re.length !== nx * ny || im.length !== nx * ny) {
    throw new Error('fft2d: plane length ' + re.length + ' does not match ' + nx + 'x' + ny);
  }
  for (let y = 0; y < ny; y++) fft1d(re, im, nx, 1, y * nx, !!inverse);
  for (let x = 0; x < nx; x++) fft1d(re, im, ny, nx, x, !!inverse);
}

// Quadrant swap along one axis (ref: fft.js:170-181).
export function fftshift1d(arr, n, shift, stride, offset) {
  shift = shift === undefined ? true : shift;
  stride = stride === undefined ? 1 : stride;
  offset = offset === undefined ? 0 : offset;
  const m = n >>> 1;
  const d = shift ? m : n - m;
  for (let i = 0; i < m; i++) {
    const a = offset + i * stride;
    const b = offset + ((i + d) % n) * stride;
    const t = arr[a];
    arr[a] = arr[b];
    arr[b] = t;
  }
  return arr;
}

export function ifftshift1d(arr, n, stride, offset) {
  return fftshift1d(arr, n, false, stride, offset);
}

// ref: fft.js:188-198
export function fftshift2d(re, im, nx, ny, shift) {
  shift = shift === undefined ? true : shift;
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

// O(N^4) DFT with the same sign and scale conventions (ref: fft.js:208-237).
export function dft2dNaive(re, im, nx, ny, inverse) {
  const or = new Float64Array(nx * ny);
  const oi = new Float64Array(nx * ny);
  const sg = inverse ? 1 : -1;
  const sc = inverse ? 1 / (nx * ny) : 1;
  for (let ky = 0; ky < ny; ky++) {
    for (let kx = 0; kx < nx; kx++) {
      let sr = 0;
      let si = 0;
      for (let y = 0; y < ny; y++) {
        const ay = (sg * TWO_PI * ky * y) / ny;
        const cy = Math.cos(ay);
        const sy = Math.sin(ay);
        for (let x = 0; x < nx; x++) {
          const ax = (sg * TWO_PI * kx * x) / nx;
          const c = Math.cos(ax) * cy - Math.sin(ax) * sy;
          const s = Math.sin(ax) * cy + Math.cos(ax) * sy;
          const vr = re[x + y * nx];
          const vi = im[x + y * nx];
          sr += vr * c - vi * s;
          si += vr * s + vi * c;
        }
      }
      or[kx + ky * nx] = sr * sc;
      oi[kx + ky * nx] = si * sc;
    }
  }
  return { re: or, im: oi };
}
