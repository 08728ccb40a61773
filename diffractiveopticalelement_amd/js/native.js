// Loads the N-API addon (../doe_napi.node) that binds libdoehip.so.  There is
// no JavaScript fallback: importing any drop-in module without the built
// addon, or without a HIP device, throws.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
let addon = null;
try {
  addon = require('../doe_napi.node');
} catch (e) {
  throw new Error('doe_napi.node is not built (run build() of the package): ' + e.message);
}

export const native = addon;

// Compute precision of the device path: 'f32' (default), 'f64' (check mode),
// 'f16' / 'bf16' (half-width storage, FP32 arithmetic; FFT and propagators).
const DTYPES = { f32: 0, f64: 1, f16: 2, bf16: 3 };
let precision = process.env.DOE_PRECISION && DTYPES[process.env.DOE_PRECISION] !== undefined
  ? process.env.DOE_PRECISION
  : 'f32';

export function setPrecision(p) {
  if (DTYPES[p] === undefined) throw new Error('setPrecision: expected f32, f64, f16 or bf16, got ' + p);
  precision = p;
}

export function getPrecision() {
  return precision;
}

export function dtypeCode(p) {
  return DTYPES[p === undefined ? precision : p];
}

// Phase retrieval runs in f32 or f64 only.
export function loopDtype() {
  return precision === 'f64' ? 1 : 0;
}
