// Drop-in for src/metrics.js: the FP64 figures of merit that judge a design
// (ref: metrics.js:10-336).  They run on the host over the downloaded
// irradiance plane, in double precision, with the reference's definitions.

export function maskedSum(plane, mask = null) {
  let s = 0;
  if (mask === null) {
    for (let k = 0; k < plane.length; k++) s += plane[k];
  } else {
    for (let k = 0; k < plane.length; k++) s += plane[k] * mask[k];
  }
  return s;
}

export function powerOf(plane, dx, dy, mask = null) {
  return maskedSum(plane, mask) * dx * dy;
}

// ref: metrics.js:37-67
export function regionStats(plane, mask, { threshold = 0 } = {}) {
  let n = 0;
  let s1 = 0;
  let s2 = 0;
  let lo = Infinity;
  let hi = -Infinity;
  for (let k = 0; k < plane.length; k++) {
    if (mask && mask[k] <= 0) continue;
    const v = plane[k];
    if (v <= threshold) continue;
    n++;
    s1 += v;
    s2 += v * v;
    lo = Math.min(lo, v);
    hi = Math.max(hi, v);
  }
  if (n === 0) return { n: 0, mean: 0, min: 0, max: 0, std: 0, uniformity: 0, flatness: 0, cv: 0 };
  const mean = s1 / n;
  const std = Math.sqrt(Math.max(0, s2 / n - mean * mean));
  return {
    n,
    mean,
    min: lo,
    max: hi,
    std,
    uniformity: hi + lo > 0 ? 1 - (hi - lo) / (hi + lo) : 1,
    flatness: mean > 0 ? 1 - std / mean : 1,
    cv: mean > 0 ? std / mean : 0,
  };
}

// ref: metrics.js:73-84
export function rmsError(plane, target, mask = null) {
  let num = 0;
  let den = 0;
  for (let k = 0; k < plane.length; k++) {
    if ((mask ? mask[k] : 1) <= 0) continue;
    const d = plane[k] - target[k];
    num += d * d;
    den += target[k] * target[k];
  }
  return den > 0 ? Math.sqrt(num / den) : 0;
}

// ref: metrics.js:87-118
export function correlation(plane, target, mask = null) {
  let sp = 0;
  let st = 0;
  let n = 0;
  for (let k = 0; k < plane.length; k++) {
    if (mask && mask[k] <= 0) continue;
    sp += plane[k];
    st += target[k];
    n++;
  }
  if (n === 0) return 0;
  const mp = sp / n;
  const mt = st / n;
  let cov = 0;
  let vp = 0;
  let vt = 0;
  for (let k = 0; k < plane.length; k++) {
    if (mask && mask[k] <= 0) continue;
    const a = plane[k] - mp;
    const b = target[k] - mt;
    cov += a * b;
    vp += a * a;
    vt += b * b;
  }
  return vp > 0 && vt > 0 ? cov / Math.sqrt(vp * vt) : 0;
}

// ref: metrics.js:121-137
export function psnr(plane, target, mask = null) {
  let mse = 0;
  let n = 0;
  let peak = 0;
  for (let k = 0; k < plane.length; k++) {
    if (mask && mask[k] <= 0) continue;
    const d = plane[k] - target[k];
    mse += d * d;
    peak = Math.max(peak, target[k]);
    n++;
  }
  if (n === 0 || mse === 0) return Infinity;
  return 10 * Math.log10((peak * peak) / (mse / n));
}

// ref: metrics.js:145-157 (element-plane field)
export function zeroOrderFraction(field) {
  const re = field._hostRe ? field._hostRe() : field.re;
  const im = field._hostIm ? field._hostIm() : field.im;
  let sr = 0;
  let si = 0;
  let p = 0;
  for (let k = 0; k < re.length; k++) {
    sr += re[k];
    si += im[k];
    p += re[k] * re[k] + im[k] * im[k];
  }
  return p === 0 ? 0 : (sr * sr + si * si) / (re.length * p);
}

// ref: metrics.js:164-185
export function rmsRadius(plane, nx, dx, cx = null, cy = null, dy = dx) {
  const ny = plane.length / nx;
  const x0 = cx === null ? 0 : cx;
  const y0 = cy === null ? 0 : cy;
  let p = 0;
  let sx = 0;
  let sy = 0;
  for (let j = 0; j < ny; j++) {
    const y = (j - ny / 2) * dy - y0;
    for (let i = 0; i < nx; i++) {
      const x = (i - nx / 2) * dx - x0;
      const v = plane[i + j * nx];
      p += v;
      sx += v * x * x;
      sy += v * y * y;
    }
  }
  if (p === 0) return { x: 0, y: 0, r: 0 };
  const rx = Math.sqrt(sx / p);
  const ry = Math.sqrt(sy / p);
  return { x: rx, y: ry, r: Math.hypot(rx, ry) };
}

// Windowed SSIM, L = larger peak, optional target weighting (ref: metrics.js:198-259).
export function ssim(plane, target, nx, ny, { window = 8, dataRange = null, weight = null } = {}) {
  let L = dataRange;
  if (L === null) {
    let mx = 0;
    for (let k = 0; k < plane.length; k++) mx = Math.max(mx, plane[k], target[k]);
    L = mx > 0 ? mx : 1;
  }
  const C1 = 0.01 * L * (0.01 * L);
  const C2 = 0.03 * L * (0.03 * L);
  const n = window * window;
  let total = 0;
  let wsum = 0;
  for (let wy = 0; wy + window <= ny; wy += window) {
    for (let wx = 0; wx + window <= nx; wx += window) {
      let sp = 0;
      let st = 0;
      let spp = 0;
      let stt = 0;
      let spt = 0;
      for (let j = 0; j < window; j++) {
        let k = wx + (wy + j) * nx;
        for (let i = 0; i < window; i++, k++) {
          const p = plane[k];
          const t = target[k];
          sp += p;
          st += t;
          spp += p * p;
          stt += t * t;
          spt += p * t;
        }
      }
      const mp = sp / n;
      const mt = st / n;
      const vp = Math.abs(spp / n - mp * mp);
      const vt = Math.abs(stt / n - mt * mt);
      const cpt = spt / n - mp * mt;
      const num = (2 * mp * mt + C1) * (2 * cpt + C2);
      const den = (mp * mp + mt * mt + C1) * (vp + vt + C2);
      const w = weight === 'target' ? st / n : 1;
      total += (den !== 0 ? num / den : 0) * w;
      wsum += w;
    }
  }
  return wsum > 0 ? total / wsum : 0;
}

export function visibility(plane, mask = null) {
  const st = regionStats(plane, mask);
  return st.max + st.min > 0 ? (st.max - st.min) / (st.max + st.min) : 0;
}

// Per-cell power uniformity (ref: metrics.js:284-321).
export function spotUniformity(plane, nx, ny, { cols = 8, rows = 8, minPeak = 0.05 } = {}) {
  const cells = new Array(cols * rows).fill(0);
  let peak = 0;
  for (let k = 0; k < plane.length; k++) peak = Math.max(peak, plane[k]);
  const cut = minPeak * peak;
  for (let j = 0; j < ny; j++) {
    const cj = Math.min(rows - 1, Math.floor((j / ny) * rows));
    for (let i = 0; i < nx; i++) {
      const v = plane[i + j * nx];
      if (v >= cut) cells[Math.min(cols - 1, Math.floor((i / nx) * cols)) + cj * cols] += v;
    }
  }
  let lo = Infinity;
  let hi = 0;
  let s1 = 0;
  let s2 = 0;
  let n = 0;
  for (const c of cells) {
    if (c <= 0) continue;
    lo = Math.min(lo, c);
    hi = Math.max(hi, c);
    s1 += c;
    s2 += c * c;
    n++;
  }
  if (n === 0) return { uniformity: 0, cells, min: 0, max: 0, mean: 0, cv: 0 };
  const mean = s1 / n;
  const std = Math.sqrt(Math.max(0, s2 / n - mean * mean));
  return { uniformity: hi + lo > 0 ? (2 * lo) / (hi + lo) : 1, cells, min: lo, max: hi, mean, cv: mean > 0 ? std / mean : 0 };
}

export function summarizeMetrics(m) {
  const f = (v, d = 3) => (Number.isFinite(v) ? v.toFixed(d) : String(v));
  return [
    `efficiency      ${f(m.efficiency * 100, 2)} %`,
    `uniformity      ${f(m.uniformity * 100, 2)} %`,
    `flatness (1-cv) ${f(m.flatness * 100, 2)} %`,
    `contrast        ${f(m.contrast * 100, 2)} %`,
    `RMSE            ${f(m.rmse * 100, 2)} %`,
    `SSIM            ${f(m.ssim, 4)}`,
    `SNR             ${f(m.snr, 2)} dB`,
    `zero-order leak ${f(m.zeroOrder * 100, 3)} %`,
  ].join('\n');
}
