// ORACLE (test infrastructure only) — FP64 restatement of the screen-metric
// reductions (ref: src/metrics.js:10-321).  ES2019.
//
// Pinned by tests/test_oracle_cpu.py (identical planes: rmse 0, correlation
// 1, ssim 1; numpy forms of rmse, correlation, regionStats) and
// tests/test_oracle_pins_cpu.py: SSIM on piecewise-constant and affine
// windows (C1, C2, the data range L, target weighting, partial windows
// skipped), zeroOrderFraction of a plane wave (1), a checkerboard (0) and a
// random field, rmsRadius of Gaussian beams (w / 2 per axis, about a given
// centre), spotUniformity of equal and 1:3 spots above a background below
// the cut; tests/test_oracle_mutants_cpu.py checks those pins catch slips.

export function maskedSum(plane, mask) {
  let s = 0;
  if (!mask) {
    for (let k = 0; k < plane.length; k++) s += plane[k];
    return s;
  }
  for (let k = 0; k < plane.length; k++) s += plane[k] * mask[k];
  return s;
}

// ref: metrics.js:37-67
export function regionStats(plane, mask, opts) {
  const threshold = opts && opts.threshold !== undefined ? opts.threshold : 0;
  let n = 0;
  let sum = 0;
  let sq = 0;
  let mn = Infinity;
  let mx = -Infinity;
  for (let k = 0; k < plane.length; k++) {
    if (mask && mask[k] <= 0) continue;
    const v = plane[k];
    if (v <= threshold) continue;
    n++;
    sum += v;
    sq += v * v;
    if (v < mn) mn = v;
    if (v > mx) mx = v;
  }
  if (n === 0) return { n: 0, mean: 0, min: 0, max: 0, std: 0, uniformity: 0, flatness: 0, cv: 0 };
  const mean = sum / n;
  const std = Math.sqrt(Math.max(0, sq / n - mean * mean));
  return {
    n: n, mean: mean, min: mn, max: mx, std: std,
    uniformity: mx + mn > 0 ? 1 - (mx - mn) / (mx + mn) : 1,
    flatness: mean > 0 ? 1 - std / mean : 1,
    cv: mean > 0 ? std / mean : 0,
  };
}

// ref: metrics.js:73-84
export function rmsError(plane, target, mask) {
  let num = 0;
  let den = 0;
  for (let k = 0; k < plane.length; k++) {
    const m = mask ? mask[k] : 1;
    if (m <= 0) continue;
    const d = plane[k] - target[k];
    num += d * d;
    den += target[k] * target[k];
  }
  return den > 0 ? Math.sqrt(num / den) : 0;
}

// ref: metrics.js:87-118
export function correlation(plane, target, mask) {
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
    const dp = plane[k] - mp;
    const dt = target[k] - mt;
    cov += dp * dt;
    vp += dp * dp;
    vt += dt * dt;
  }
  return vp > 0 && vt > 0 ? cov / Math.sqrt(vp * vt) : 0;
}

// |mean U|^2 / mean |U|^2 on the element plane (ref: metrics.js:145-157)
export function zeroOrderFraction(field) {
  let sr = 0;
  let si = 0;
  let p = 0;
  const n = field.size;
  for (let k = 0; k < n; k++) {
    sr += field.re[k];
    si += field.im[k];
    p += field.re[k] * field.re[k] + field.im[k] * field.im[k];
  }
  if (p === 0) return 0;
  return (sr * sr + si * si) / (n * p);
}

// ref: metrics.js:164-185
export function rmsRadius(plane, nx, dx, cx, cy, dy) {
  dy = dy === undefined ? dx : dy;
  const ny = plane.length / nx;
  const x0 = cx === undefined || cx === null ? 0 : cx;
  const y0 = cy === undefined || cy === null ? 0 : cy;
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

// Block SSIM with L = max peak and optional target weighting (ref: metrics.js:198-259).
export function ssim(plane, target, nx, ny, opts) {
  const o = opts || {};
  const win = o.window === undefined ? 8 : o.window;
  let L = o.dataRange === undefined ? null : o.dataRange;
  if (L === null) {
    let mx = 0;
    for (let k = 0; k < plane.length; k++) {
      if (plane[k] > mx) mx = plane[k];
      if (target[k] > mx) mx = target[k];
    }
    L = mx > 0 ? mx : 1;
  }
  const C1 = 0.01 * L * (0.01 * L);
  const C2 = 0.03 * L * (0.03 * L);
  let total = 0;
  let wsum = 0;
  for (let wy = 0; wy + win <= ny; wy += win) {
    for (let wx = 0; wx + win <= nx; wx += win) {
      let sp = 0;
      let st = 0;
      let spp = 0;
      let stt = 0;
      let spt = 0;
      const n = win * win;
      for (let j = 0; j < win; j++) {
        for (let i = 0; i < win; i++) {
          const k = wx + i + (wy + j) * nx;
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
      const wt = o.weight === 'target' ? st / n : 1;
      total += (den !== 0 ? num / den : 0) * wt;
      wsum += wt;
    }
  }
  return wsum > 0 ? total / wsum : 0;
}

// ref: metrics.js:284-321
export function spotUniformity(plane, nx, ny, opts) {
  const o = opts || {};
  const cols = o.cols === undefined ? 8 : o.cols;
  const rows = o.rows === undefined ? 8 : o.rows;
  const minPeak = o.minPeak === undefined ? 0.05 : o.minPeak;
  const cells = new Array(cols * rows).fill(0);
  let peak = 0;
  for (let k = 0; k < plane.length; k++) if (plane[k] > peak) peak = plane[k];
  const cut = minPeak * peak;
  for (let j = 0; j < ny; j++) {
    const cj = Math.min(rows - 1, Math.floor((j / ny) * rows));
    for (let i = 0; i < nx; i++) {
      const ci = Math.min(cols - 1, Math.floor((i / nx) * cols));
      const v = plane[i + j * nx];
      if (v >= cut) cells[ci + cj * cols] += v;
    }
  }
  let mn = Infinity;
  let mx = 0;
  let sum = 0;
  let sq = 0;
  let n = 0;
  for (let c = 0; c < cells.length; c++) {
    const v = cells[c];
    if (v <= 0) continue;
    mn = Math.min(mn, v);
    mx = Math.max(mx, v);
    sum += v;
    sq += v * v;
    n++;
  }
  if (n === 0) return { uniformity: 0, cells: cells, min: 0, max: 0, mean: 0, cv: 0 };
  const mean = sum / n;
  const std = Math.sqrt(Math.max(0, sq / n - mean * mean));
  return { uniformity: mx + mn > 0 ? (2 * mn) / (mx + mn) : 1, cells: cells, min: mn, max: mx, mean: mean, cv: mean > 0 ? std / mean : 0 };
}
