// ORACLE (test infrastructure only) — FP64 restatement of the reference's
// ComplexField subset used on the hot path and of its masks
// (ref: src/field.js:22-196, 443-536).  ES2019.

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

  // Centred zero-pad / crop at the same pitch (ref: field.js:443-461).
  resample(nx2, ny2) {
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
