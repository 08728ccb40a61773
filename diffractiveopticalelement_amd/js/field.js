// Drop-in for src/field.js: ComplexField with a device-resident planar copy.
//
// The public `re` / `im` Float64Array planes of the reference are kept: they
// are a host mirror that is materialised lazily.  A field touched by a device
// operation holds its data in HBM (FP32 by default, FP64 in check mode) and
// downloads only when `re` / `im` are read; reading them hands out the host
// planes for writing, so the device copy is re-uploaded before the next
// device operation.  Layout (ref: field.js:22-40): row-major, idx = i + j*nx,
// sample (i, j) at x = (i - nx/2)*dx, y = (j - ny/2)*dy.
import { native, dtypeCode } from './native.js';
import { isPow2, fft2d, fftshift2d, ifftshift2d } from './fft.js';

export class ComplexField {
  constructor(nx, ny = nx, dx = 1, dy = dx) {
    if (!isPow2(nx) || !isPow2(ny)) {
      throw new Error(`ComplexField: grid must be power-of-two, got ${nx}x${ny}`);
    }
    this.nx = nx;
    this.ny = ny;
    this.dx = dx;
    this.dy = dy;
    this._re = new Float64Array(nx * ny);
    this._im = new Float64Array(nx * ny);
    this._dev = null; // addon field handle
    this._devDtype = -1;
    this._where = 'host'; // 'host' | 'device' | 'both'
  }

  // ---- host mirror ------------------------------------------------------
  _pull() {
    if (this._where === 'device') {
      native.fieldDownload(this._dev, this._re, this._im);
      this._where = 'both';
    }
  }

  get re() {
    this._pull();
    this._where = 'host';
    return this._re;
  }

  set re(v) {
    this._pull();
    this._re = v;
    this._where = 'host';
  }

  get im() {
    this._pull();
    this._where = 'host';
    return this._im;
  }

  set im(v) {
    this._pull();
    this._im = v;
    this._where = 'host';
  }

  // read-only host view (does not invalidate the device copy)
  _hostRe() {
    this._pull();
    return this._re;
  }

  _hostIm() {
    this._pull();
    return this._im;
  }

  // ---- device copy ------------------------------------------------------
  // Device handle holding the current values, in the active precision.
  _device(dtype = dtypeCode()) {
    if (this._dev && this._devDtype !== dtype) {
      this._pull();
      this._where = 'host';
    }
    if (!this._dev || this._devDtype !== dtype) {
      this._dev = native.fieldCreate(this.nx, this.ny, 1, dtype);
      this._devDtype = dtype;
      this._where = 'host';
    }
    if (this._where === 'host') {
      native.fieldUpload(this._dev, this._re, this._im);
      this._where = 'both';
    }
    return this._dev;
  }

  // Empty device buffer to receive a result (contents are overwritten).
  _deviceOut(dtype = dtypeCode()) {
    if (!this._dev || this._devDtype !== dtype) {
      this._dev = native.fieldCreate(this.nx, this.ny, 1, dtype);
      this._devDtype = dtype;
    }
    return this._dev;
  }

  _markDevice() {
    this._where = 'device';
  }

  // ---- reference surface (ref: field.js:43-513) --------------------------
  get size() {
    return this.nx * this.ny;
  }

  get extentX() {
    return this.nx * this.dx;
  }

  get extentY() {
    return this.ny * this.dy;
  }

  x(i) {
    return (i - this.nx / 2) * this.dx;
  }

  y(j) {
    return (j - this.ny / 2) * this.dy;
  }

  idx(i, j) {
    return i + j * this.nx;
  }

  clone() {
    const f = new ComplexField(this.nx, this.ny, this.dx, this.dy);
    f._re.set(this._hostRe());
    f._im.set(this._hostIm());
    return f;
  }

  copyFrom(other) {
    if (other.nx !== this.nx || other.ny !== this.ny) {
      throw new Error('ComplexField.copyFrom: shape mismatch');
    }
    this.re.set(other._hostRe());
    this.im.set(other._hostIm());
    this.dx = other.dx;
    this.dy = other.dy;
    return this;
  }

  fill(re = 0, im = 0) {
    this.re.fill(re);
    this.im.fill(im);
    return this;
  }

  zero() {
    return this.fill(0, 0);
  }

  intensityAt(i, j) {
    const k = i + j * this.nx;
    const r = this._hostRe();
    const m = this._hostIm();
    return r[k] * r[k] + m[k] * m[k];
  }

  amplitudeAt(i, j) {
    const k = i + j * this.nx;
    return Math.hypot(this._hostRe()[k], this._hostIm()[k]);
  }

  phaseAt(i, j) {
    const k = i + j * this.nx;
    return Math.atan2(this._hostIm()[k], this._hostRe()[k]);
  }

  amplitude(out = new Float64Array(this.size)) {
    const r = this._hostRe();
    const m = this._hostIm();
    for (let k = 0; k < r.length; k++) out[k] = Math.hypot(r[k], m[k]);
    return out;
  }

  intensity(out = new Float64Array(this.size)) {
    const r = this._hostRe();
    const m = this._hostIm();
    for (let k = 0; k < r.length; k++) out[k] = r[k] * r[k] + m[k] * m[k];
    return out;
  }

  phase(out = new Float64Array(this.size)) {
    const r = this._hostRe();
    const m = this._hostIm();
    for (let k = 0; k < r.length; k++) out[k] = Math.atan2(m[k], r[k]);
    return out;
  }

  multiplyPhase(p) {
    const r = this.re;
    const m = this.im;
    for (let k = 0; k < r.length; k++) {
      const c = Math.cos(p[k]);
      const s = Math.sin(p[k]);
      const a = r[k];
      r[k] = a * c - m[k] * s;
      m[k] = a * s + m[k] * c;
    }
    return this;
  }

  multiplyConjPhase(p) {
    const r = this.re;
    const m = this.im;
    for (let k = 0; k < r.length; k++) {
      const c = Math.cos(p[k]);
      const s = -Math.sin(p[k]);
      const a = r[k];
      r[k] = a * c - m[k] * s;
      m[k] = a * s + m[k] * c;
    }
    return this;
  }

  setAmplitudePhase(amp, phi = 0) {
    const r = this.re;
    const m = this.im;
    const sa = typeof amp === 'number';
    const sp = typeof phi === 'number';
    for (let k = 0; k < r.length; k++) {
      const a = sa ? amp : amp[k];
      const p = sp ? phi : phi[k];
      r[k] = a * Math.cos(p);
      m[k] = a * Math.sin(p);
    }
    return this;
  }

  setPlaneWave(amp = 1, phi = 0) {
    return this.setAmplitudePhase(amp, phi);
  }

  power() {
    const r = this._hostRe();
    const m = this._hostIm();
    let s = 0;
    for (let k = 0; k < r.length; k++) s += r[k] * r[k] + m[k] * m[k];
    return s * this.dx * this.dy;
  }

  peakIntensity() {
    const r = this._hostRe();
    const m = this._hostIm();
    let best = 0;
    for (let k = 0; k < r.length; k++) {
      const v = r[k] * r[k] + m[k] * m[k];
      if (v > best) best = v;
    }
    return best;
  }

  normalizePeak(peak = 1) {
    const p = this.peakIntensity();
    if (p > 0) this.scale(Math.sqrt(peak / p));
    return this;
  }

  normalizePower(power = 1) {
    const p = this.power();
    if (p > 0) this.scale(Math.sqrt(power / p));
    return this;
  }

  scale(s) {
    const r = this.re;
    const m = this.im;
    for (let k = 0; k < r.length; k++) {
      r[k] *= s;
      m[k] *= s;
    }
    return this;
  }

  globalPhase(phi) {
    const c = Math.cos(phi);
    const s = Math.sin(phi);
    const r = this.re;
    const m = this.im;
    for (let k = 0; k < r.length; k++) {
      const a = r[k];
      r[k] = a * c - m[k] * s;
      m[k] = a * s + m[k] * c;
    }
    return this;
  }

  multiply(other) {
    const r = this.re;
    const m = this.im;
    const or = other._hostRe();
    const oi = other._hostIm();
    for (let k = 0; k < r.length; k++) {
      const a = r[k];
      r[k] = a * or[k] - m[k] * oi[k];
      m[k] = a * oi[k] + m[k] * or[k];
    }
    return this;
  }

  add(other) {
    const r = this.re;
    const m = this.im;
    const or = other._hostRe();
    const oi = other._hostIm();
    for (let k = 0; k < r.length; k++) {
      r[k] += or[k];
      m[k] += oi[k];
    }
    return this;
  }

  static sum(fields) {
    if (fields.length === 0) throw new Error('ComplexField.sum: no fields');
    const out = fields[0].clone();
    for (let i = 1; i < fields.length; i++) out.add(fields[i]);
    return out;
  }

  static sumIncoherent(fields) {
    if (fields.length === 0) throw new Error('ComplexField.sumIncoherent: no fields');
    const out = fields[0].clone();
    const acc = new Float64Array(out.size);
    for (const f of fields) {
      const r = f._hostRe();
      const m = f._hostIm();
      for (let k = 0; k < acc.length; k++) acc[k] += r[k] * r[k] + m[k] * m[k];
    }
    const r = out.re;
    const m = out.im;
    for (let k = 0; k < acc.length; k++) {
      r[k] = Math.sqrt(acc[k]);
      m[k] = 0;
    }
    return out;
  }

  freqX(i) {
    return (i - this.nx / 2) / (this.nx * this.dx);
  }

  freqY(j) {
    return (j - this.ny / 2) / (this.ny * this.dy);
  }

  get dfx() {
    return 1 / (this.nx * this.dx);
  }

  get dfy() {
    return 1 / (this.ny * this.dy);
  }

  // Centred spectrum (ref: field.js:328-335); the transform runs on the GPU.
  spectrum() {
    const re = Float64Array.from(this._hostRe());
    const im = Float64Array.from(this._hostIm());
    ifftshift2d(re, im, this.nx, this.ny);
    fft2d(re, im, this.nx, this.ny, false);
    fftshift2d(re, im, this.nx, this.ny, true);
    return { re, im, nx: this.nx, ny: this.ny, dfx: this.dfx, dfy: this.dfy };
  }

  // Area-averaged raster sampling (ref: field.js:350-400), on the device
  // (doe_field_sample_image, FP64): re = the raster's average over each
  // sample's box, fill outside the raster; im = 0.
  sampleImage(img, width, height, scale = 1, fill = 0) {
    const raster = img instanceof Float64Array ? img : Float64Array.from(img);
    native.fieldSampleImage(raster, width, height, scale, fill, this._deviceOut(1));
    this._markDevice();
    return this;
  }

  // Pad / crop onto a new grid at the same pitch, or bilinear resample at
  // another pitch (ref: field.js:443-495); both on the device
  // (doe_field_resample), in the field's precision.
  resample(nx2, ny2, dx2 = this.dx, dy2 = this.dy) {
    const out = new ComplexField(nx2, ny2, dx2, dy2);
    const dt = this._where === 'host' ? dtypeCode() : this._devDtype;
    native.fieldResample(this._device(dt), out._deviceOut(dt), this.dx, this.dy, dx2, dy2);
    out._markDevice();
    return out;
  }

  cropCentred(width, height) {
    return this.resample(Math.max(2, Math.round(width / this.dx)), Math.max(2, Math.round(height / this.dy)));
  }

  describe() {
    const r = this.intensityRange();
    return `ComplexField ${this.nx}x${this.ny} @ ${(this.dx * 1e6).toFixed(3)}um  extent=${(this.extentX * 1e3).toFixed(3)}x${(
      this.extentY * 1e3
    ).toFixed(3)}mm  I=[${r.min.toExponential(3)}, ${r.max.toExponential(3)}]`;
  }
}

function raisedCosine(r, edge, soft) {
  if (r <= edge - soft) return 1;
  if (r >= edge) return 0;
  return 0.5 * (1 + Math.cos((Math.PI * (r - edge + soft)) / soft));
}

// ref: field.js:516-536
export function discMask(nx, ny, dx, dy, radius, { soft = 0, cx = 0, cy = 0 } = {}) {
  const m = new Float64Array(nx * ny);
  for (let j = 0; j < ny; j++) {
    const y = (j - ny / 2) * dy - cy;
    for (let i = 0; i < nx; i++) {
      const r = Math.hypot((i - nx / 2) * dx - cx, y);
      m[i + j * nx] = soft > 0 ? raisedCosine(r, radius, soft) : r <= radius ? 1 : 0;
    }
  }
  return m;
}

// ref: field.js:539-558
export function rectMask(nx, ny, dx, dy, width, height, { soft = 0 } = {}) {
  const m = new Float64Array(nx * ny);
  for (let j = 0; j < ny; j++) {
    const y = Math.abs((j - ny / 2) * dy);
    const fy = soft > 0 ? raisedCosine(y, height / 2, soft) : y <= height / 2 ? 1 : 0;
    for (let i = 0; i < nx; i++) {
      const x = Math.abs((i - nx / 2) * dx);
      const fx = soft > 0 ? raisedCosine(x, width / 2, soft) : x <= width / 2 ? 1 : 0;
      m[i + j * nx] = fx * fy;
    }
  }
  return m;
}

// ref: field.js:561-573
export function superGaussian(nx, ny, dx, dy, radius, order = 8) {
  const a = new Float64Array(nx * ny);
  for (let j = 0; j < ny; j++) {
    const y = (j - ny / 2) * dy;
    for (let i = 0; i < nx; i++) {
      a[i + j * nx] = Math.exp(-Math.pow(Math.hypot((i - nx / 2) * dx, y) / radius, order));
    }
  }
  return a;
}
