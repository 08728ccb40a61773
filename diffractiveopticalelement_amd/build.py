"""Build the native pieces of the package in-tree.

* ``libdoehip.so``  — HIP kernels + C ABI (``include/doehip.h``), gfx950.
* ``doe_napi.node`` — the Node.js N-API addon over the C ABI.

Object files go to ``_build/`` next to this file; every translation unit is
compiled in parallel and the library is always rebuilt when ``build()`` runs.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
NAPI = PKG / "napi"
BUILD = PKG / "_build"
LIB = PKG / "libdoehip.so"
ADDON = PKG / "doe_napi.node"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DOE_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]

# (source, DOE_PART values): each part is its own object so the template
# instantiations for different precisions compile in parallel.
SOURCES = [
    ("api.hip", [None]),
    ("fft.hip", [1, 2, 3, 4]),
    ("prop.hip", [0, 1, 2, 3, 4]),
    ("gs.hip", [0, 1, 2]),
    ("gs_batch.hip", [None]),
    ("nn.hip", [None]),
    ("warmstart.hip", [None]),
    ("metrics.hip", [None]),
    ("targets.hip", [None]),
    ("fab.hip", [None]),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in proc.stdout.splitlines() if "packed-fp32-ops" not in ln and ln.strip()
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _compile(unit: tuple[Path, int | None]) -> Path:
    src, part = unit
    obj = BUILD / (src.stem + ("" if part is None else f"_p{part}") + ".o")
    defs = [] if part is None else [f"-DDOE_PART={part}"]
    _run([HIPCC, *HIP_FLAGS, *defs, "-c", str(src), "-o", str(obj)])
    return obj


def build_lib(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    units = [(CSRC / s, part) for s, parts in SOURCES for part in parts]
    # largest units first so the long poles start early
    units.sort(key=lambda u: (u[0].name != "gs.hip", u[0].name != "prop.hip"))
    jobs = jobs or min(len(units), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, units))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_napi() -> Path | None:
    """Build the N-API addon if Node's headers are present (skipped otherwise)."""
    src = NAPI / "doe_napi.cc"
    hdr = Path("/usr/include/node/node_api.h")
    if not src.exists() or not hdr.exists():
        return None
    tmp = ADDON.with_suffix(".node.tmp")
    _run([
        "g++", "-O2", "-std=c++17", "-shared", "-fPIC",
        "-I/usr/include/node", f"-I{ROOT / 'include'}",
        str(src), "-o", str(tmp),
        f"-L{PKG}", "-ldoehip", "-Wl,-rpath,$ORIGIN",
    ])
    os.replace(tmp, ADDON)
    return ADDON


def build_all() -> None:
    build_lib()
    build_napi()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}" + (f" and {ADDON}" if ADDON.exists() else ""), file=sys.stderr)
