"""The drop-in PNG container (js/png.js; ref: src/png.js:12-151) checked with
Python's zlib: chunk CRC-32s, the zlib stream's Adler-32 and inflate, the
scanline filter bytes and the pixels (ref test/png.test.js:8-89)."""
import json
import shutil
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
ADDON = ROOT / "diffractiveopticalelement_amd" / "doe_napi.node"

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not ADDON.exists(),
                                reason="node and the built addon are required")


def decode(buf: bytes):
    assert buf[:8] == b"\x89PNG\r\n\x1a\n"
    o, chunks, idat, ihdr = 8, [], b"", None
    while o < len(buf):
        (n,) = struct.unpack(">I", buf[o:o + 4])
        typ, data = buf[o + 4:o + 8], buf[o + 8:o + 8 + n]
        (crc,) = struct.unpack(">I", buf[o + 8 + n:o + 12 + n])
        assert crc == zlib.crc32(typ + data), typ
        chunks.append(typ.decode())
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", data)
        elif typ == b"IDAT":
            idat += data
        o += 12 + n
    return ihdr, chunks, zlib.decompress(idat)  # verifies the Adler-32 trailer


def test_png_container_round_trips(tmp_path):
    proc = subprocess.run(["node", str(ROOT / "tests" / "js" / "png_write.mjs"), str(tmp_path)],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr
    meta = json.loads((tmp_path / "meta.json").read_text())
    assert "pass rgb or gray" in meta["errors"][0] and "buffer too small" in meta["errors"][1]
    for name in ("gray_5x3", "rgb_4x2", "gray_1x1", "rgb_300x250"):
        m = meta[name]
        px = np.frombuffer((tmp_path / f"{name}.raw").read_bytes(), np.uint8)
        sizes = {}
        for kind in ("stored", "zlib"):
            buf = (tmp_path / f"{name}_{kind}.png").read_bytes()
            sizes[kind] = len(buf)
            (w, h, depth, ctype, _, _, _), chunks, raw = decode(buf)
            assert (w, h, depth) == (m["width"], m["height"], 8)
            assert ctype == (2 if m["channels"] == 3 else 0)
            assert chunks == ["IHDR", "IDAT", "IEND"]
            rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * m["channels"])
            assert np.all(rows[:, 0] == 0)
            np.testing.assert_array_equal(rows[:, 1:].ravel(), px)
        if name == "rgb_300x250":  # several stored blocks; real compression is smaller
            assert sizes["zlib"] < sizes["stored"]
