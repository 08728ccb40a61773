// Drop-in for src/png.js: 8-bit grayscale / RGB PNG output.  The container
// (signature, IHDR / IDAT / IEND chunks with CRC-32, a zlib stream of stored
// deflate blocks or node's zlib) is byte-serial host work; the pixels of a
// device-resident field are mapped and laid out as PNG scanlines on the GPU
// (writeFieldPNG), so only the finished scanlines cross PCIe.
import { planeToGray, fieldToGray, fieldToRGB } from './color.js';

// CRC-32 (ISO-HDLC, reflected polynomial 0xEDB88320), the PNG chunk checksum.
const CRC = new Uint32Array(256);
for (let n = 0; n < 256; n++) {
  let c = n;
  for (let k = 0; k < 8; k++) c = (c >>> 1) ^ (c & 1 ? 0xedb88320 : 0);
  CRC[n] = c >>> 0;
}

function crc32(bytes, c = 0xffffffff) {
  for (let i = 0; i < bytes.length; i++) c = CRC[(c ^ bytes[i]) & 0xff] ^ (c >>> 8);
  return c;
}

function u32be(out, o, v) {
  out[o] = (v >>> 24) & 0xff;
  out[o + 1] = (v >>> 16) & 0xff;
  out[o + 2] = (v >>> 8) & 0xff;
  out[o + 3] = v & 0xff;
}

// length, type, data, CRC over type + data
function chunk(type, data) {
  const out = new Uint8Array(12 + data.length);
  u32be(out, 0, data.length);
  for (let i = 0; i < 4; i++) out[4 + i] = type.charCodeAt(i);
  out.set(data, 8);
  const c = crc32(data, crc32(out.subarray(4, 8))) ^ 0xffffffff;
  u32be(out, 8 + data.length, c >>> 0);
  return out;
}

// zlib stream of stored deflate blocks (RFC 1950 / 1951) with its Adler-32.
function storedZlib(raw) {
  const max = 65535;
  const blocks = Math.max(1, Math.ceil(raw.length / max));
  const out = new Uint8Array(2 + 5 * blocks + raw.length + 4);
  out[0] = 0x78;
  out[1] = 0x01;
  let o = 2;
  for (let b = 0; b < blocks; b++) {
    const start = b * max;
    const len = Math.min(max, raw.length - start);
    out[o] = b === blocks - 1 ? 1 : 0;
    out[o + 1] = len & 0xff;
    out[o + 2] = len >>> 8;
    out[o + 3] = ~len & 0xff;
    out[o + 4] = (~len >>> 8) & 0xff;
    out.set(raw.subarray(start, start + len), o + 5);
    o += 5 + len;
  }
  let s1 = 1;
  let s2 = 0;
  for (let i = 0; i < raw.length; i++) {
    s1 = (s1 + raw[i]) % 65521;
    s2 = (s2 + s1) % 65521;
  }
  u32be(out, o, ((s2 << 16) | s1) >>> 0);
  return out;
}

// PNG from ready scanlines (each row: filter byte 0 + width * channels bytes).
function assemble(width, height, channels, raw, deflate) {
  let z = null;
  if (deflate) {
    try {
      z = deflate(raw);
    } catch (e) {
      z = null;
    }
  }
  if (!z) z = storedZlib(raw);
  const ihdr = new Uint8Array(13);
  u32be(ihdr, 0, width);
  u32be(ihdr, 4, height);
  ihdr[8] = 8;
  ihdr[9] = channels === 3 ? 2 : 0;
  const parts = [Uint8Array.from([0x89, 0x50, 0x4e, 0x47, 0x0d, 0x0a, 0x1a, 0x0a]), chunk('IHDR', ihdr),
    chunk('IDAT', z), chunk('IEND', new Uint8Array(0))];
  const out = new Uint8Array(parts.reduce((s, p) => s + p.length, 0));
  let o = 0;
  for (const p of parts) {
    out.set(p, o);
    o += p.length;
  }
  return out;
}

/** ref: png.js:89-151.  image: {width, height, rgb | gray, deflate?} */
export function encodePNG({ width, height, rgb = null, gray = null, deflate = null }) {
  const channels = rgb ? 3 : 1;
  const src = rgb || gray;
  if (!src) throw new Error('encodePNG: pass rgb or gray');
  if (src.length < width * height * channels) {
    throw new Error(`encodePNG: buffer too small (${src.length} < ${width * height * channels})`);
  }
  const row = width * channels;
  const raw = new Uint8Array(height * (row + 1));
  for (let y = 0; y < height; y++) raw.set(src.subarray(y * row, (y + 1) * row), y * (row + 1) + 1);
  return assemble(width, height, channels, raw, deflate);
}

async function nodeDeflate(level) {
  try {
    const zlib = await import('zlib');
    return (raw) => zlib.deflateSync(raw, { level });
  } catch (e) {
    return null;
  }
}

/** ref: png.js:158-168: node's zlib when present, stored blocks otherwise. */
export async function encodePNGNode(image, { level = 6 } = {}) {
  return encodePNG(Object.assign({}, image, { deflate: await nodeDeflate(level) }));
}

async function writeBytes(path, png) {
  if (path) {
    const fs = await import('fs');
    fs.writeFileSync(path, png);
  }
  return png;
}

/** ref: png.js:170-180.  The gray mapping runs on the GPU (planeToGray). */
export async function writePlanePNG(path, plane, width, height, opts = {}) {
  const gray = opts.gray ? opts.gray : planeToGray(plane, opts);
  return writeBytes(path, await encodePNGNode({ width, height, gray }));
}

/**
 * A device-resident ComplexField's quantity ('intensity' default) straight to
 * PNG: the GPU maps it (colour map when opts.map is set, else gray) and lays
 * out the scanlines; the host deflates and frames them.
 */
export async function writeFieldPNG(path, field, opts = {}) {
  const rgb = opts.map !== undefined;
  const o = Object.assign({}, opts, { pngRows: true });
  const raw = rgb ? fieldToRGB(field, o) : fieldToGray(field, o);
  const png = assemble(field.nx, field.ny, rgb ? 3 : 1, raw, await nodeDeflate(opts.level === undefined ? 6 : opts.level));
  return writeBytes(path, png);
}
