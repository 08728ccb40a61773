// Drop-in for src/color.js: colour maps and plane-to-pixel conversion.  The
// per-pixel mapping runs on the GPU (doe_field_image): the data range, the
// percentile (a radix select over the plane), the colour ramp and the 8-bit
// rounding.  planeToRGB / planeToGray take host planes as the reference's do;
// fieldToRGB / fieldToGray map a device-resident ComplexField's quantity
// without downloading it.
import { native } from './native.js';

// Colour-map stops, [position, [r, g, b]] (data: the reference's tables,
// ref: color.js:28-71).
export const COLORMAP_STOPS = {
  gray: [[0, [0, 0, 0]], [1, [255, 255, 255]]],
  viridis: [[0, [68, 1, 84]], [0.25, [59, 82, 139]], [0.5, [33, 145, 140]], [0.75, [94, 201, 98]], [1, [253, 231, 37]]],
  inferno: [[0, [0, 0, 4]], [0.25, [87, 16, 110]], [0.5, [188, 55, 84]], [0.75, [249, 142, 9]], [1, [252, 255, 164]]],
  turbo: [[0, [48, 18, 59]], [0.2, [65, 122, 246]], [0.4, [26, 219, 172]], [0.6, [175, 240, 91]], [0.8, [251, 126, 34]], [1, [122, 4, 3]]],
  magenta: [[0, [0, 0, 0]], [0.5, [160, 20, 120]], [1, [255, 220, 255]]],
  phase: [[0, [255, 60, 60]], [0.2, [255, 220, 40]], [0.4, [60, 230, 120]], [0.6, [40, 200, 255]], [0.8, [130, 80, 255]], [1, [255, 60, 60]]],
};

// Scalar evaluation of a map (ref: color.js:7-25), for callers that colour
// single values (legends, labels).
function ramp(stops) {
  return (t) => {
    const x = Math.min(1, Math.max(0, t));
    for (let s = 1; s < stops.length; s++) {
      const [p0, c0] = stops[s - 1];
      const [p1, c1] = stops[s];
      if (x <= p1) {
        const f = p1 === p0 ? 0 : (x - p0) / (p1 - p0);
        return [0, 1, 2].map((c) => Math.round(c0[c] + (c1[c] - c0[c]) * f));
      }
    }
    const c = stops[stops.length - 1][1];
    return [c[0], c[1], c[2]];
  };
}

export const COLORMAPS = Object.fromEntries(Object.entries(COLORMAP_STOPS).map(([k, v]) => [k, ramp(v)]));

function stopsArray(name) {
  const stops = COLORMAP_STOPS[name] === undefined ? COLORMAP_STOPS.viridis : COLORMAP_STOPS[name];
  const a = new Float64Array(stops.length * 4);
  stops.forEach(([p, c], q) => a.set([p, c[0], c[1], c[2]], 4 * q));
  return a;
}

// A host plane as a one-field FP64 device field (re = the plane).
function planeField(plane) {
  const p = plane instanceof Float64Array ? plane : Float64Array.from(plane);
  const dev = native.fieldCreate(p.length, 1, 1, 1);
  native.fieldUpload(dev, p, p);
  return dev;
}

function rgbOpts(opts, quantity, pngRows) {
  const { map = 'viridis', min = null, max = null, gamma = 1, log = false, logFloor = 1e-4, percentile = null } = opts;
  return { quantity, stops: stopsArray(map), min, max, gamma, log, logFloor, percentile, pngRows };
}

function grayOpts(opts, quantity, pngRows) {
  const { min = null, max = null, invert = false } = opts;
  return { quantity, gray: true, min, max, invert, pngRows };
}

/**
 * ref: color.js:87-137.  opts: map, min, max, gamma, log, logFloor,
 * percentile (width / height are accepted and unused, as there).
 */
export function planeToRGB(plane, opts = {}) {
  return native.fieldImage(planeField(plane), 0, rgbOpts(opts, 're', false)).bytes;
}

/** ref: color.js:140-160.  opts: min, max, invert. */
export function planeToGray(plane, { min = null, max = null, invert = false } = {}) {
  return native.fieldImage(planeField(plane), 0, grayOpts({ min, max, invert }, 're', false)).bytes;
}

/**
 * A ComplexField's quantity ('intensity' default, 'amplitude', 'phase', 're',
 * 'im') as RGB bytes, mapped on the device.  pngRows: PNG scanlines (each row
 * led by filter byte 0) instead of packed pixels.
 */
export function fieldToRGB(field, opts = {}) {
  const { quantity = 'intensity', pngRows = false } = opts;
  return native.fieldImage(field._device(), 0, rgbOpts(opts, quantity, pngRows)).bytes;
}

export function fieldToGray(field, opts = {}) {
  const { quantity = 'intensity', pngRows = false } = opts;
  return native.fieldImage(field._device(), 0, grayOpts(opts, quantity, pngRows)).bytes;
}

/** ref: color.js:163-171 */
export function wavelengthColor(lambda) {
  const nm = lambda * 1e9;
  if (nm < 450) return [90, 60, 255];
  if (nm < 495) return [40, 140, 255];
  if (nm < 570) return [80, 255, 120];
  if (nm < 590) return [255, 220, 60];
  if (nm < 620) return [255, 140, 40];
  return [255, 70, 60];
}
