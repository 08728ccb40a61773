// ORACLE (test infrastructure only) — FP64 restatement of the plane-to-pixel
// mapping (ref: src/color.js:7-160).  The colour-map stop tables are data the
// pixels must reproduce exactly, so their entries are the reference's.

// Stops: [position, [r, g, b]] (ref: color.js:28-71).
export const COLORMAP_STOPS = {
  gray: [[0, [0, 0, 0]], [1, [255, 255, 255]]],
  viridis: [[0, [68, 1, 84]], [0.25, [59, 82, 139]], [0.5, [33, 145, 140]], [0.75, [94, 201, 98]], [1, [253, 231, 37]]],
  inferno: [[0, [0, 0, 4]], [0.25, [87, 16, 110]], [0.5, [188, 55, 84]], [0.75, [249, 142, 9]], [1, [252, 255, 164]]],
  turbo: [[0, [48, 18, 59]], [0.2, [65, 122, 246]], [0.4, [26, 219, 172]], [0.6, [175, 240, 91]], [0.8, [251, 126, 34]], [1, [122, 4, 3]]],
  magenta: [[0, [0, 0, 0]], [0.5, [160, 20, 120]], [1, [255, 220, 255]]],
  phase: [[0, [255, 60, 60]], [0.2, [255, 220, 40]], [0.4, [60, 230, 120]], [0.6, [40, 200, 255]], [0.8, [130, 80, 255]], [1, [255, 60, 60]]],
};

// Piecewise-linear map through the stops, clamped to [0, 1]; NaN falls
// through to the last stop (ref: color.js:7-25).
export function rampColor(stops, t) {
  const x = Math.min(1, Math.max(0, t));
  for (let s = 1; s < stops.length; s++) {
    const p0 = stops[s - 1][0];
    const p1 = stops[s][0];
    if (x <= p1) {
      const f = p1 === p0 ? 0 : (x - p0) / (p1 - p0);
      const c0 = stops[s - 1][1];
      const c1 = stops[s][1];
      return [0, 1, 2].map((c) => Math.round(c0[c] + (c1[c] - c0[c]) * f));
    }
  }
  const last = stops[stops.length - 1][1];
  return [last[0], last[1], last[2]];
}

function dataRange(plane) {
  let mn = Infinity;
  let mx = -Infinity;
  for (let k = 0; k < plane.length; k++) {
    if (plane[k] < mn) mn = plane[k];
    if (plane[k] > mx) mx = plane[k];
  }
  return [mn, mx];
}

// ref: color.js:87-137
export function planeToRGB(plane, opts) {
  const o = opts || {};
  const stops = COLORMAP_STOPS[o.map] || COLORMAP_STOPS.viridis;
  const gamma = o.gamma === undefined ? 1 : o.gamma;
  const logFloor = o.logFloor === undefined ? 1e-4 : o.logFloor;
  const pct = o.percentile === undefined ? null : o.percentile;
  let lo = o.min === undefined ? null : o.min;
  let hi = o.max === undefined ? null : o.max;
  if (lo === null || hi === null) {
    if (pct !== null) {
      const sorted = Float64Array.from(plane).sort();
      if (hi === null) hi = sorted[Math.min(sorted.length - 1, Math.floor(sorted.length * pct))];
      if (lo === null) lo = sorted[0];
    } else {
      const [mn, mx] = dataRange(plane);
      if (lo === null) lo = mn;
      if (hi === null) hi = mx;
    }
  }
  const span = hi - lo || 1;
  const lf = Math.log10(Math.max(1e-12, logFloor));
  const out = new Uint8Array(plane.length * 3);
  for (let k = 0; k < plane.length; k++) {
    let t = o.log ? (Math.log10(Math.max(plane[k], hi * logFloor) / hi) - lf) / -lf : (plane[k] - lo) / span;
    if (gamma !== 1) t = Math.pow(Math.max(0, Math.min(1, t)), gamma);
    const c = rampColor(stops, t);
    out[3 * k] = c[0];
    out[3 * k + 1] = c[1];
    out[3 * k + 2] = c[2];
  }
  return { rgb: out, lo: lo, hi: hi };
}

// ref: color.js:140-160
export function planeToGray(plane, opts) {
  const o = opts || {};
  let lo = o.min === undefined ? null : o.min;
  let hi = o.max === undefined ? null : o.max;
  if (lo === null || hi === null) {
    const [mn, mx] = dataRange(plane);
    if (lo === null) lo = mn;
    if (hi === null) hi = mx;
  }
  const span = hi - lo || 1;
  const out = new Uint8Array(plane.length);
  for (let k = 0; k < plane.length; k++) {
    let t = Math.max(0, Math.min(1, (plane[k] - lo) / span));
    if (o.invert) t = 1 - t;
    out[k] = Math.round(t * 255);
  }
  return { gray: out, lo: lo, hi: hi };
}
