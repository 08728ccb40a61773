// Writes PNGs with the drop-in encoder for tests/test_png_cpu.py (host
// container only: no GPU call).  Usage: node png_write.mjs <dir>
import fs from 'fs';
import path from 'path';
import { encodePNG, encodePNGNode } from '../../diffractiveopticalelement_amd/js/png.js';

const dir = process.argv[2];
function pattern(n, a, b) {
  const out = new Uint8Array(n);
  for (let k = 0; k < n; k++) out[k] = (k * a + ((k / b) | 0)) % 256;
  return out;
}
const cases = [
  ['gray_5x3', { width: 5, height: 3, gray: pattern(15, 17, 1000) }],
  ['rgb_4x2', { width: 4, height: 2, rgb: pattern(24, 31, 1000) }],
  ['gray_1x1', { width: 1, height: 1, gray: Uint8Array.from([200]) }],
  ['rgb_300x250', { width: 300, height: 250, rgb: pattern(300 * 250 * 3, 7, 13) }],
];
const meta = {};
async function main() {
  for (const [name, img] of cases) {
    fs.writeFileSync(path.join(dir, name + '_stored.png'), encodePNG(img));
    fs.writeFileSync(path.join(dir, name + '_zlib.png'), await encodePNGNode(img));
    fs.writeFileSync(path.join(dir, name + '.raw'), img.rgb || img.gray);
    meta[name] = { width: img.width, height: img.height, channels: img.rgb ? 3 : 1 };
  }
  const errs = [];
  for (const bad of [{ width: 2, height: 2 }, { width: 2, height: 2, gray: new Uint8Array(3) }]) {
    try {
      encodePNG(bad);
      errs.push(null);
    } catch (e) {
      errs.push(e.message);
    }
  }
  meta.errors = errs;
  fs.writeFileSync(path.join(dir, 'meta.json'), JSON.stringify(meta));
}
main().catch((e) => {
  console.error(e);
  process.exit(1);
});
