"""The JavaScript drop-in (N-API addon over libdoehip.so) against the FP64
oracle, end to end: fft2d/fft1d, propagators, the device-resident field
mirror, and ProjectorSystem.design()/simulate()/screenMetrics()."""
import json
import shutil
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("node") is None, reason="node is required for the JS drop-in")
def test_js_dropin_parity():
    print("node tests/js/dropin_parity.mjs", flush=True)
    proc = subprocess.run(["node", str(ROOT / "tests" / "js" / "dropin_parity.mjs")],
                          capture_output=True, text=True, timeout=600, cwd=str(ROOT))
    lines = [json.loads(ln) for ln in proc.stdout.splitlines() if ln.startswith("{")]
    failed = [ln for ln in lines if not ln.get("summary") and not ln["ok"]]
    print(proc.stdout[-3000:])
    assert proc.returncode == 0 and not failed, (failed, proc.stderr[-2000:])
    assert any(ln.get("summary") for ln in lines)
