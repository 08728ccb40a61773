// ORACLE (test infrastructure only) — FP64 restatement of the reference's
// ComplexField subset used on the hot path and of its masks
// (ref: src/field.js:22-196, 350-400, 443-536).  ES2019.
//
// Pinned by tests/test_oracle_pins_cpu.py: sampleImage returns an image of
// the grid's size unchanged, averages 2x2 blocks onto a half-size grid and
// conserves the integral (image + fill) with an inset; the bilinear resample
// reproduces affine planes exactly at the mapped fractional indices and is
// zero beyond the source.

import { isPow2 } from './fft.js';

export class ComplexField {
  constructor(nx, ny, dx, dy) {
    ny = ny === undefined ? nx : ny;
    dx = dx === undefined ? 1 : dx;
    dy = dy === undefined ? dx : dy;
    if (!isPow2(nx) || !isPow2(ny)) {
      throw new Error('ComplexField: grid must be power-of-two, got ' + nx + 'x' + ny);
    }
    this.nx = nx;
    this.ny = ny;
    this.dx = dx;
    this.dy = dy;
    this.re = new Float64Array(nx * ny);
    this.im = new Float64Array(nx * ny);
  }

  get size() {
    return this.nx * this.ny;
  }

  // centred sample coordinates (ref: field.js:56-63)
  x(i) {
    return (i - this.nx / 2) * this.dx;
  }

  y(j) {
    return (j - this.ny / 2) * this.dy;
  }

  clone() {
    const f = new ComplexField(this.nx, this.ny, this.dx, this.dy);
    f.re.set(this.re);
    f.im.set(this.im);
    return f;
  }

  copyFrom(o) {
    if (o.nx !== this.nx || o.ny !== this.ny) throw new Error('ComplexField.copyFrom: shape mismatch');
    this.re.set(o.re);
    this.im.set(o.im);
    this.dx = o.dx;
    this.dy = o.dy;
    return this;
  }

  // ref: field.js:168-179
  setAmplitudePhase(amp, phi) {
    phi = phi === undefined ? 0 : phi;
    const sa = typeof amp === 'number';
    const sp = typeof phi === 'number';
    for (let k = 0; k < this.re.length; k++) {
      const a = sa ? amp : amp[k];
      const p = sp ? phi : phi[k];
      this.re[k] = a * Math.cos(p);
      this.im[k] = a * Math.sin(p);
    }
    return this;
  }

  amplitude() {
    const out = new Float64Array(this.size);
    for (let k = 0; k < out.length; k++) out[k] = Math.hypot(this.re[k], this.im[k]);
    return out;
  }

  intensity() {
    const out = new Float64Array(this.size);
    for (let k = 0; k < out.length; k++) out[k] = this.re[k] * this.re[k] + this.im[k] * this.im[k];
    return out;
  }

  phase() {
    const out = new Float64Array(this.size);
    for (let k = 0; k < out.length; k++) out[k] = Math.atan2(this.im[k], this.re[k]);
    return out;
  }

  // sum |U|^2 dx dy (ref: field.js:190-195)
  power() {
    let s = 0;
    for (let k = 0; k < this.re.length; k++) s += this.re[k] * this.re[k] + this.im[k] * this.im[k];
    return s * this.dx * this.dy;
  }

  // Area average of a raster over each grid sample's box (ref:
  // field.js:350-400): the raster spans the central `scale` fraction of the
  // window; fill outside it.
  sampleImage(img, width, height, scale, fill) {
    scale = scale === undefined ? 1 : scale;
    fill = fill === undefined ? 0 : fill;
    const nx = this.nx;
    const ny = this.ny;
    const s = Math.max(scale, 1e-9);
    const off = (1 - s) / 2;
    const boxW = width / (nx * s);
    const boxH = height / (ny * s);
    for (let j = 0; j < ny; j++) {
      const cy = (((j + 0.5) / ny - off) * height) / s;
      for (let i = 0; i < nx; i++) {
        const cx = (((i + 0.5) / nx - off) * width) / s;
        const ax = cx - boxW / 2;
        const bx = cx + boxW / 2;
        const ay = cy - boxH / 2;
        const by = cy + boxH / 2;
        const lox = Math.max(ax, 0);
        const hix = Math.min(bx, width);
        const loy = Math.max(ay, 0);
        const hiy = Math.min(by, height);
        const area = boxW * boxH;
        let acc = 0;
        if (hix > lox && hiy > loy) {
          const i0 = Math.max(0, Math.floor(lox));
          const i1 = Math.min(width - 1, Math.floor(hix - 1e-12));
          const j0 = Math.max(0, Math.floor(loy));
          const j1 = Math.min(height - 1, Math.floor(hiy - 1e-12));
          for (let jj = j0; jj <= j1; jj++) {
            const wy = Math.min(by, jj + 1) - Math.max(ay, jj);
            for (let ii = i0; ii <= i1; ii++) {
              const wx = Math.min(bx, ii + 1) - Math.max(ax, ii);
              acc += img[ii + jj * width] * wx * wy;
            }
          }
        }
        acc += fill * (area - (hix - lox) * (hiy - loy));
        this.re[i + j * nx] = acc / area;
      }
    }
    this.im.fill(0);
    return this;
  }

  // Pad / crop (same pitch) or bilinear resample at another pitch
  // (ref: field.js:443-495).
  resample(nx2, ny2, dx2, dy2) {
    dx2 = dx2 === undefined ? this.dx : dx2;
    dy2 = dy2 === undefined ? this.dy : dy2;
    if (dx2 !== this.dx || dy2 !== this.dy) return this.resampleBilinear(nx2, ny2, dx2, dy2);
    const out = new ComplexField(nx2, ny2, this.dx, this.dy);
    const ox = Math.floor((nx2 - this.nx) / 2);
    const oy = Math.floor((ny2 - this.ny) / 2);
    for (let j = 0; j < this.ny; j++) {
      const jj = j + oy;
      if (jj < 0 || jj >= ny2) continue;
      for (let i = 0; i < this.nx; i++) {
        const ii = i + ox;
        if (ii < 0 || ii >= nx2) continue;
        out.re[ii + jj * nx2] = this.re[i + j * this.nx];
        out.im[ii + jj * nx2] = this.im[i + j * this.nx];
      }
    }
    return out;
  }

  // Output sample (i, j) at x = (i - nx2/2) dx2 reads the source bilinearly
  // at fractional index x / dx + nx/2 - 1/2, zero outside (ref: field.js:466-495).
  resampleBilinear(nx2, ny2, dx2, dy2) {
    const out = new ComplexField(nx2, ny2, dx2, dy2);
    for (let j = 0; j < ny2; j++) {
      const fy = ((j - ny2 / 2) * dy2) / this.dy + this.ny / 2 - 0.5;
      const j0 = Math.floor(fy);
      const ty = fy - j0;
      for (let i = 0; i < nx2; i++) {
        const fx = ((i - nx2 / 2) * dx2) / this.dx + this.nx / 2 - 0.5;
        const i0 = Math.floor(fx);
        const tx = fx - i0;
        let sr = 0;
        let si = 0;
        for (let b = 0; b < 2; b++) {
          const jj = j0 + b;
          if (jj < 0 || jj >= this.ny) continue;
          const wy = b === 0 ? 1 - ty : ty;
          for (let a = 0; a < 2; a++) {
            const ii = i0 + a;
            if (ii < 0 || ii >= this.nx) continue;
            const w = (a === 0 ? 1 - tx : tx) * wy;
            sr += this.re[ii + jj * this.nx] * w;
            si += this.im[ii + jj * this.nx] * w;
          }
        }
        out.re[i + j * nx2] = sr;
        out.im[i + j * nx2] = si;
      }
    }
    return out;
  }
}

// Disc with optional raised-cosine edge (ref: field.js:516-536).
export function discMask(nx, ny, dx, dy, radius, opts) {
  const o = opts || {};
  const soft = o.soft === undefined ? 0 : o.soft;
  const cx = o.cx === undefined ? 0 : o.cx;
  const cy = o.cy === undefined ? 0 : o.cy;
  const m = new Float64Array(nx * ny);
  for (let j = 0; j < ny; j++) {
    const y = (j - ny / 2) * dy - cy;
    for (let i = 0; i < nx; i++) {
      const x = (i - nx / 2) * dx - cx;
      const r = Math.hypot(x, y);
      let v;
      if (soft > 0) {
        if (r <= radius - soft) v = 1;
        else if (r >= radius) v = 0;
        else v = 0.5 * (1 + Math.cos((Math.PI * (r - radius + soft)) / soft));
      } else {
        v = r <= radius ? 1 : 0;
      }
      m[i + j * nx] = v;
    }
  }
  return m;
}
