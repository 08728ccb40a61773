"""MI355X-native FFT / free-space propagation / phase-retrieval path of the
diffractive-optical-element toolkit.

The product boundary is the C ABI in ``include/doehip.h`` (``libdoehip.so``)
and the JavaScript drop-in modules in ``js/`` that call it through the N-API
addon ``doe_napi.node``.  ``native`` is the ctypes binding used by the Python
tests and the benchmark.
"""
from . import native  # noqa: F401

__all__ = ["native"]
