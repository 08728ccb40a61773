"""Call the FP64 JavaScript oracle (oracle/run.mjs) from pytest.

The oracle is the checker only: tests compare the HIP path against it.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
# DOE_ORACLE_DIR: run another copy of the oracle (the mutation tests plant a
# mistake in a temporary copy and check that the pins catch it)
RUNNER = Path(os.environ.get("DOE_ORACLE_DIR", ROOT / "oracle")) / "run.mjs"


def node_available() -> bool:
    return shutil.which("node") is not None


def run(op: str, args: dict | None = None, inputs: dict | None = None, timeout: float = 600,
        pin_cpu: int | None = None):
    """Run one oracle job; returns (arrays: dict[str, ndarray], result: dict).
    pin_cpu: run node on that one host core (taskset), for timings."""
    with tempfile.TemporaryDirectory(prefix="doe_oracle_") as d:
        dp = Path(d)
        (dp / "job.json").write_text(json.dumps({"op": op, "args": args or {}}))
        for name, arr in (inputs or {}).items():
            if arr is None:
                continue
            np.ascontiguousarray(np.asarray(arr, dtype="<f8")).ravel().tofile(dp / f"in_{name}.f64")
        cmd = ["node", str(RUNNER), str(dp)]
        if pin_cpu is not None and shutil.which("taskset"):
            cmd = ["taskset", "-c", str(pin_cpu)] + cmd
        proc = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        if proc.returncode != 0:
            raise RuntimeError(f"oracle {op} failed:\n{proc.stderr[-4000:]}")
        res = json.loads((dp / "result.json").read_text())
        arrays = {p.name[4:-4]: np.fromfile(p, dtype="<f8") for p in dp.glob("out_*.f64")}
    return arrays, res
