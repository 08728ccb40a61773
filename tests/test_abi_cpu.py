"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, and reports errors through status codes."""
import ctypes as C
import subprocess

import numpy as np

from diffractiveopticalelement_amd import native as nat


def test_library_exports_every_header_symbol():
    lib = nat.load()
    out = subprocess.run(["nm", "-D", "--defined-only", str(nat.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    declared = nat.header_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    bound = {name for name, _, _ in nat.SIGNATURES}
    assert set(declared) <= bound, set(declared) - bound
    assert lib.doe_version() >= 100


def test_status_strings_and_argument_errors():
    lib = nat.load()
    assert lib.doe_status_string(0) == b"ok"
    assert b"power" not in lib.doe_status_string(1)
    # null out-pointer is rejected before any device work
    assert lib.doe_ctx_create(0, None, None) == 1  # DOE_ERR_INVALID
    # a length mismatch in the host fft2d drop-in is reported, not crashed on
    re = np.zeros(12)
    im = np.zeros(12)
    rc = lib.doe_fft2d_host(None, re.ctypes.data_as(nat.c_double_p), im.ctypes.data_as(nat.c_double_p),
                            12, 4, 4, 0, 1)
    assert rc == 1  # null context
    n = C.c_int(-1)
    assert lib.doe_device_count(C.byref(n)) == 0 and n.value >= 0
