"""ctypes binding of ``libdoehip.so`` (the C ABI declared in ``include/doehip.h``).

This is plumbing for the Python tests and the benchmark; the product's host
interface for users of the toolkit is the JavaScript drop-in in ``js/`` over
the N-API addon.  Loading fails loudly when the library has not been built:
there is no CPU fallback anywhere on this path.
"""
from __future__ import annotations

import ctypes as C
import weakref
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libdoehip.so"

DOE_OK = 0
DOE_F32, DOE_F64, DOE_F16, DOE_BF16 = 0, 1, 2, 3
DTYPES = {"f32": DOE_F32, "f64": DOE_F64, "f16": DOE_F16, "bf16": DOE_BF16}
MODE_IMAGE_PLANE, MODE_FAR_FIELD = 0, 1
ALG_GS, ALG_WGS = 0, 1

c_double_p = C.POINTER(C.c_double)


class DoeError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class GsPlan(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("algorithm", C.c_int),
        ("iterations", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("dx", C.c_double),
        ("dy", C.c_double),
        ("lambda_", C.c_double),
        ("z", C.c_double),
        ("beta", C.c_double),
        ("feedback", C.c_double),
        ("escape", C.c_int),
        ("dtype", C.c_int),
    ]


class ScreenOpts(C.Structure):
    _fields_ = [("sdx", C.c_double), ("sdy", C.c_double), ("edx", C.c_double), ("edy", C.c_double),
                ("grid_cols", C.c_int), ("grid_rows", C.c_int), ("grid_min_peak", C.c_double),
                ("ssim_weight_target", C.c_int)]


METRIC_NAMES = ["efficiency", "flux_fraction", "uniformity", "spot_uniformity", "flatness", "contrast",
                "rmse", "fidelity", "correlation", "ssim", "snr", "zero_order", "rms_radius",
                "mean_intensity", "peak_intensity", "power_in_signal", "power_at_screen",
                "power_incident"]


class Metrics(C.Structure):
    _fields_ = [(n, C.c_double) for n in METRIC_NAMES] + [("has_spots", C.c_int)]


class MpPlane(C.Structure):
    _fields_ = [("z", C.c_double), ("target_amp", C.c_void_p), ("roi", C.c_void_p),
                ("weight", C.c_double)]


TGT_KINDS = {"disc": 0, "ring": 1, "dots": 2, "grid": 2, "checker": 3, "lines": 4, "zone": 5,
             "raster": 6}
QUANTITIES = {"re": 0, "im": 1, "intensity": 2, "amplitude": 3, "phase": 4}


class TargetSpec(C.Structure):
    _fields_ = [("kind", C.c_int), ("radius", C.c_double), ("inner", C.c_double),
                ("outer", C.c_double), ("cols", C.c_double), ("rows", C.c_double),
                ("sigma", C.c_double), ("missing", C.c_void_p), ("n_missing", C.c_int),
                ("cells", C.c_double), ("period", C.c_double), ("duty", C.c_double),
                ("raster", C.c_void_p), ("raster_w", C.c_int), ("raster_h", C.c_int),
                ("cover", C.c_int), ("gamma", C.c_double), ("background", C.c_double),
                ("gain", C.c_double)]


class ImageOpts(C.Structure):
    _fields_ = [("quantity", C.c_int), ("gray", C.c_int), ("nstops", C.c_int),
                ("stop_pos", C.c_double * 8), ("stop_rgb", (C.c_double * 3) * 8),
                ("has_min", C.c_int), ("has_max", C.c_int), ("min", C.c_double),
                ("max", C.c_double), ("gamma", C.c_double), ("log", C.c_int),
                ("log_floor", C.c_double), ("has_percentile", C.c_int), ("percentile", C.c_double),
                ("invert", C.c_int), ("png_rows", C.c_int)]


# Every symbol the header declares: (name, restype, argtypes)
_VP = C.c_void_p
SIGNATURES = [
    ("doe_version", C.c_int, []),
    ("doe_last_error", C.c_char_p, []),
    ("doe_status_string", C.c_char_p, [C.c_int]),
    ("doe_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("doe_ctx_create", C.c_int, [C.c_int, _VP, C.POINTER(_VP)]),
    ("doe_ctx_destroy", C.c_int, [_VP]),
    ("doe_ctx_sync", C.c_int, [_VP]),
    ("doe_ctx_stream", C.c_int, [_VP, C.POINTER(_VP)]),
    ("doe_ctx_set_flags", C.c_int, [_VP, C.c_int]),
    ("doe_ctx_profile_report", C.c_int, [_VP, C.c_char_p, C.c_long]),
    ("doe_ctx_set_chunk_bytes", C.c_int, [_VP, C.c_long]),
    ("doe_tune_set", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_long]),
    ("doe_tune_get", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_long)]),
    ("doe_autotune_fft2d", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                     C.c_long]),
    ("doe_ctx_get_chunk_bytes", C.c_int, [_VP, C.POINTER(C.c_long)]),
    ("doe_timer_start", C.c_int, [_VP]),
    ("doe_timer_stop", C.c_int, [_VP, c_double_p]),
    ("doe_field_create", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_VP)]),
    ("doe_field_wrap", C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_VP)]),
    ("doe_field_free", C.c_int, [_VP]),
    ("doe_field_info", C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(_VP), C.POINTER(_VP)]),
    ("doe_field_upload", C.c_int, [_VP, _VP, c_double_p, c_double_p]),
    ("doe_field_download", C.c_int, [_VP, _VP, c_double_p, c_double_p]),
    ("doe_field_copy", C.c_int, [_VP, _VP, _VP]),
    ("doe_fft2d", C.c_int, [_VP, _VP, C.c_int]),
    ("doe_fft2d_host", C.c_int, [_VP, c_double_p, c_double_p, C.c_long, C.c_int, C.c_int, C.c_int,
                                 C.c_int]),
    ("doe_fft1d_host", C.c_int, [_VP, c_double_p, c_double_p, C.c_long, C.c_long, C.c_long,
                                 C.c_long, C.c_int, C.c_int]),
    ("doe_asm", C.c_int, [_VP, _VP, _VP, C.c_double, C.c_double, C.c_double, c_double_p, C.c_int,
                          C.c_int, C.c_double]),
    ("doe_asm_multi", C.c_int, [_VP, C.c_int, _VP, _VP, c_double_p, c_double_p, C.c_double, _VP,
                                C.c_int, C.c_int, C.c_double]),
    ("doe_asm_transfer", C.c_int, [_VP, _VP, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_int, C.c_double, c_double_p]),
    ("doe_fresnel", C.c_int, [_VP, _VP, _VP, C.c_double, C.c_double, C.c_double, c_double_p,
                              C.c_int, C.c_double, C.c_double, c_double_p, c_double_p]),
    ("doe_lens", C.c_int, [_VP, _VP, _VP, C.c_double, C.c_double, C.c_double, c_double_p,
                           c_double_p, c_double_p]),
    ("doe_gs_run", C.c_int, [_VP, C.POINTER(GsPlan), c_double_p, c_double_p, c_double_p,
                             c_double_p, C.c_uint, c_double_p, c_double_p, c_double_p,
                             c_double_p, c_double_p, c_double_p, c_double_p]),
    ("doe_gs_batch_create", C.c_int, [_VP, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_double,
                                      C.c_int, C.POINTER(_VP)]),
    ("doe_gs_batch_destroy", C.c_int, [_VP]),
    ("doe_gs_batch_set", C.c_int, [_VP, C.c_int, c_double_p, c_double_p, c_double_p, c_double_p]),
    ("doe_gs_batch_random_init", C.c_int, [_VP, C.c_ulonglong]),
    ("doe_gs_batch_run", C.c_int, [_VP, C.c_int, C.c_int]),
    ("doe_gs_batch_get", C.c_int, [_VP, C.c_int, c_double_p, c_double_p, c_double_p, c_double_p,
                                   c_double_p, c_double_p]),
    ("doe_gs_batch_get_screen", C.c_int, [_VP, C.c_int, c_double_p]),
    ("doe_wsnet_create", C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(_VP)]),
    ("doe_wsnet_destroy", C.c_int, [_VP]),
    ("doe_wsnet_info", C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("doe_wsnet_param_count", C.c_int, [_VP, C.POINTER(C.c_long)]),
    ("doe_wsnet_get_params", C.c_int, [_VP, _VP]),
    ("doe_wsnet_set_params", C.c_int, [_VP, _VP]),
    ("doe_wsnet_forward", C.c_int, [_VP, C.c_int, _VP, _VP]),
    ("doe_wsnet_loss_grad", C.c_int, [_VP, C.c_int, _VP, _VP, _VP, _VP, c_double_p, _VP]),
    ("doe_wsnet_adam_step", C.c_int, [_VP, C.c_float]),
    ("doe_gs_batch_design_order", C.c_int, [_VP, C.POINTER(C.c_int)]),
    ("doe_warmstart_inputs", C.c_int, [_VP, C.c_int, _VP]),
    ("doe_warmstart_labels", C.c_int, [_VP, C.c_int, _VP, _VP, _VP]),
    ("doe_warmstart_apply", C.c_int, [_VP, _VP, _VP]),
    ("doe_warmstart_apply_phasor", C.c_int, [_VP, C.c_int, _VP]),
    ("doe_warmstart_apply_quadratic", C.c_int, [_VP, C.c_int, C.c_double, _VP]),
    ("doe_gs_batch_restart_best", C.c_int, [_VP]),
    ("doe_screen_metrics", C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(ScreenOpts), C.POINTER(Metrics)]),
    ("doe_screen_metrics_batch", C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(ScreenOpts),
                                           C.POINTER(Metrics)]),
    ("doe_screen_metrics_noalloc", C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(ScreenOpts),
                                             C.POINTER(Metrics)]),
    ("doe_field_resample_centred", C.c_int, [_VP, _VP, _VP]),
    ("doe_field_resample", C.c_int, [_VP, _VP, _VP, C.c_double, C.c_double, C.c_double, C.c_double]),
    ("doe_field_sample_image", C.c_int, [_VP, c_double_p, C.c_int, C.c_int, C.c_double, C.c_double, _VP]),
    ("doe_field_copy_slices", C.c_int, [_VP, _VP, C.c_int, _VP, C.c_int, C.c_int]),
    ("doe_target_spec_init", C.c_int, [C.POINTER(TargetSpec), C.c_int]),
    ("doe_target_make", C.c_int, [_VP, C.POINTER(TargetSpec), _VP]),
    ("doe_target_auto_roi", C.c_int, [_VP, _VP, C.c_double, C.c_double, C.c_double]),
    ("doe_target_make_roi", C.c_int, [_VP, C.POINTER(TargetSpec), _VP, C.c_double, C.c_double,
                                      C.c_double]),
    ("doe_raster_blur", C.c_int, [_VP, _VP, _VP, C.c_double, c_double_p]),
    ("doe_image_opts_init", C.c_int, [C.POINTER(ImageOpts)]),
    ("doe_random_phase", C.c_int, [_VP, _VP, C.c_uint]),
    ("doe_error_diffusion", C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, C.c_int]),
    ("doe_phase_to_height", C.c_int, [_VP, _VP, _VP, C.c_double, C.c_double]),
    ("doe_field_image", C.c_int, [_VP, _VP, C.c_int, C.POINTER(ImageOpts), _VP, C.c_long,
                                  c_double_p]),
    ("doe_gs_multiplane_run", C.c_int, [_VP, C.POINTER(GsPlan), c_double_p, C.c_int,
                                        C.POINTER(MpPlane), c_double_p, c_double_p, c_double_p,
                                        c_double_p, c_double_p]),
]

GSB_GRAPH = 1
GSB_PER_CLASS = 2

_lib: C.CDLL | None = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load libdoehip.so and declare every entry point.  Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback for the HIP path)"
        )
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != DOE_OK:
        lib = load()
        raise DoeError(status, f"{lib.doe_status_string(status).decode()}: "
                               f"{lib.doe_last_error().decode()}")


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def _darr(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


class Context:
    """One HIP stream on one device (stream=None: the library's own)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.doe_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(h)))
        self.h = h
        # batches and networks hold the native context: closed before it
        self._children = weakref.WeakSet()

    def _adopt(self, obj) -> None:
        self._children.add(obj)

    def close(self) -> None:
        if self.h:
            for child in list(self._children):
                child.close()
            self.lib.doe_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self) -> None:
        check(self.lib.doe_ctx_sync(self.h))

    def set_flags(self, asynchronous: bool = False, profile: bool = False) -> None:
        check(self.lib.doe_ctx_set_flags(self.h, (1 if asynchronous else 0) | (2 if profile else 0)))

    def profile_report(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        check(self.lib.doe_ctx_profile_report(self.h, buf, len(buf)))
        out = {}
        for ln in buf.value.decode().splitlines():
            tag, cnt, ms, pts = ln.split()
            out[tag] = (int(cnt), float(ms), int(pts))
        return out

    def tune_set(self, nx, ny, batch, dtype, n1=0, colvar=0, rowplanes=0, chunk=-1) -> None:
        check(self.lib.doe_tune_set(self.h, nx, ny, batch, DTYPES[dtype], n1, colvar, rowplanes,
                                    chunk))

    def tune_get(self, nx, ny, batch, dtype) -> dict:
        v = [C.c_int(), C.c_int(), C.c_int(), C.c_long()]
        check(self.lib.doe_tune_get(self.h, nx, ny, batch, DTYPES[dtype], *[C.byref(x) for x in v]))
        return {"n1": v[0].value, "colvar": v[1].value, "rowplanes": v[2].value,
                "chunk": v[3].value}

    def autotune_fft2d(self, nx, ny, batch, dtype="f32", reps=3) -> str:
        buf = C.create_string_buffer(1 << 14)
        check(self.lib.doe_autotune_fft2d(self.h, nx, ny, batch, DTYPES[dtype], reps, buf,
                                          len(buf)))
        return buf.value.decode()

    def set_chunk_bytes(self, nbytes: int) -> None:
        check(self.lib.doe_ctx_set_chunk_bytes(self.h, int(nbytes)))

    def chunk_bytes(self) -> int:
        v = C.c_long()
        check(self.lib.doe_ctx_get_chunk_bytes(self.h, C.byref(v)))
        return v.value

    def timer_start(self) -> None:
        check(self.lib.doe_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_double()
        check(self.lib.doe_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def stream(self) -> int:
        s = C.c_void_p()
        check(self.lib.doe_ctx_stream(self.h, C.byref(s)))
        return s.value or 0

    # ---- fields -------------------------------------------------------
    def field(self, nx: int, ny: int, batch: int = 1, dtype: str = "f32") -> "Field":
        h = C.c_void_p()
        check(self.lib.doe_field_create(self.h, nx, ny, batch, DTYPES[dtype], C.byref(h)))
        return Field(self, h, nx, ny, batch, dtype)

    def wrap(self, re, im, nx: int, ny: int, batch: int = 1, dtype: str = "f32") -> "Field":
        """Non-owning view of device memory (e.g. torch tensors' data_ptr())."""
        h = C.c_void_p()
        check(self.lib.doe_field_wrap(self.h, C.c_void_p(re), C.c_void_p(im), nx, ny, batch,
                                      DTYPES[dtype], C.byref(h)))
        return Field(self, h, nx, ny, batch, dtype)

    def upload(self, re, im=None, dtype: str = "f32") -> "Field":
        re = _darr(re)
        im = np.zeros_like(re) if im is None else _darr(im)
        if re.ndim == 2:
            re, im = re[None], im[None]
        b, ny, nx = re.shape
        f = self.field(nx, ny, b, dtype)
        check(self.lib.doe_field_upload(self.h, f.h, _dptr(re), _dptr(np.ascontiguousarray(im))))
        return f

    # ---- operations ---------------------------------------------------
    def fft2d(self, f: "Field", inverse: bool = False) -> None:
        check(self.lib.doe_fft2d(self.h, f.h, int(inverse)))

    def fft2d_host(self, re: np.ndarray, im: np.ndarray, nx: int, ny: int, inverse=False,
                   dtype: str = "f64") -> None:
        check(self.lib.doe_fft2d_host(self.h, _dptr(re), _dptr(im), re.size, nx, ny,
                                      int(inverse), DTYPES[dtype]))

    def asm(self, fin: "Field", fout: "Field", dx, dy, lam, z, band_limit=True, evanescent=True,
            soft_edge=0.02) -> None:
        zz = _darr(np.broadcast_to(np.asarray(z, dtype=np.float64), (fin.batch,)))
        check(self.lib.doe_asm(self.h, fin.h, fout.h, dx, dy, lam, _dptr(zz), int(band_limit),
                               int(evanescent), soft_edge))

    def asm_multi(self, fins, fouts, dxs, dys, lam, zs, band_limit=True, evanescent=True,
                  soft_edge=0.02) -> None:
        """doe_asm_multi: batches of different (square) sides in one launch per pass;
        zs[k] holds batch k's distances."""
        k = len(fins)
        ins = (C.c_void_p * k)(*[f.h.value if hasattr(f.h, "value") else f.h for f in fins])
        outs = (C.c_void_p * k)(*[f.h.value if hasattr(f.h, "value") else f.h for f in fouts])
        dxa = _darr(dxs)
        dya = _darr(dys)
        zarr = [_darr(np.broadcast_to(np.asarray(z, dtype=np.float64), (f.batch,))) for z, f in
                zip(zs, fins)]
        zp = (c_double_p * k)(*[_dptr(a) for a in zarr])
        check(self.lib.doe_asm_multi(self.h, k, ins, outs, _dptr(dxa), _dptr(dya), lam, zp,
                                     int(band_limit), int(evanescent), soft_edge))

    def asm_transfer(self, nx, ny, dx, dy, lam, z, band_limit=True, evanescent=True,
                     soft_edge=0.02, dtype="f64"):
        f = self.field(nx, ny, 1, dtype)
        frac = C.c_double()
        check(self.lib.doe_asm_transfer(self.h, f.h, dx, dy, lam, z, int(band_limit),
                                        int(evanescent), soft_edge, C.byref(frac)))
        re, im = f.download()
        f.free()
        return re[0], im[0], frac.value

    def fresnel(self, fin: "Field", fout: "Field", dx, dy, lam, z, inverse=False, in_dx=0.0,
                in_dy=0.0):
        zz = _darr(np.broadcast_to(np.asarray(z, dtype=np.float64), (fin.batch,)))
        odx = np.zeros(fin.batch)
        ody = np.zeros(fin.batch)
        check(self.lib.doe_fresnel(self.h, fin.h, fout.h, dx, dy, lam, _dptr(zz), int(inverse),
                                   in_dx, in_dy, _dptr(odx), _dptr(ody)))
        return odx, ody

    def lens(self, fin: "Field", fout: "Field", dx, dy, lam, f):
        ff = _darr(np.broadcast_to(np.asarray(f, dtype=np.float64), (fin.batch,)))
        odx = np.zeros(fin.batch)
        ody = np.zeros(fin.batch)
        check(self.lib.doe_lens(self.h, fin.h, fout.h, dx, dy, lam, _dptr(ff), _dptr(odx),
                                _dptr(ody)))
        return odx, ody

    def screen_metrics(self, screen: "Field", target: "Field", element: "Field", sdx: float,
                       sdy: float, edx: float, edy: float, grid=None, min_peak: float = 0.05,
                       weight_target: bool = True, noalloc: bool = False) -> dict:
        """screenMetrics of a device-resident screen field (target: FP64 field with
        re = target irradiance, im = ROI; element: illumination * e^{i phase});
        grid = (cols, rows) of a spot array or None.  noalloc: the entry point
        that allocates no device memory (doe_screen_metrics_noalloc)."""
        o = ScreenOpts(sdx, sdy, edx, edy, grid[0] if grid else 0, grid[1] if grid else 0, min_peak,
                       1 if weight_target else 0)
        m = Metrics()
        fn = self.lib.doe_screen_metrics_noalloc if noalloc else self.lib.doe_screen_metrics
        check(fn(self.h, screen.h, target.h, element.h, C.byref(o), C.byref(m)))
        out = {n: getattr(m, n) for n in METRIC_NAMES}
        if not m.has_spots:
            out["spot_uniformity"] = None
        return out

    def screen_metrics_batch(self, screens: "Field", target: "Field", element: "Field", sdx, sdy,
                             edx, edy, grid=None, min_peak: float = 0.05,
                             weight_target: bool = True) -> list:
        o = ScreenOpts(sdx, sdy, edx, edy, grid[0] if grid else 0, grid[1] if grid else 0, min_peak,
                       1 if weight_target else 0)
        arr = (Metrics * screens.batch)()
        check(self.lib.doe_screen_metrics_batch(self.h, screens.h, target.h, element.h, C.byref(o),
                                                arr))
        out = []
        for m in arr:
            d = {n: getattr(m, n) for n in METRIC_NAMES}
            if not m.has_spots:
                d["spot_uniformity"] = None
            out.append(d)
        return out

    def resample_centred(self, src: "Field", dst: "Field") -> None:
        check(self.lib.doe_field_resample_centred(self.h, src.h, dst.h))

    def resample(self, src: "Field", dst: "Field", src_dx, src_dy, dst_dx, dst_dy) -> None:
        """ComplexField#resample: centred pad/crop at equal pitches, else bilinear"""
        check(self.lib.doe_field_resample(self.h, src.h, dst.h, src_dx, src_dy, dst_dx, dst_dy))

    def sample_image(self, img, dst: "Field", scale: float = 1.0, fill: float = 0.0) -> None:
        """ComplexField#sampleImage of a host raster (height x width) into dst"""
        a = _darr(img)
        h, w = a.shape
        check(self.lib.doe_field_sample_image(self.h, _dptr(a), w, h, scale, fill, dst.h))

    def copy_slices(self, src: "Field", src_first: int, dst: "Field", dst_first: int, count: int) -> None:
        check(self.lib.doe_field_copy_slices(self.h, src.h, src_first, dst.h, dst_first, count))

    def gs_multiplane(self, illumination, planes, nx, ny, dx, dy, lam, iterations, dtype="f32",
                      init_phase=None) -> dict:
        """planes: list of (z, target_amp, roi or None, weight)"""
        keep = []
        arr = (MpPlane * len(planes))()
        for q, (z, t, roi, w) in enumerate(planes):
            t = _darr(t).reshape(-1)
            keep.append(t)
            arr[q].z, arr[q].target_amp, arr[q].weight = z, t.ctypes.data, w
            if roi is not None:
                r = _darr(roi).reshape(-1)
                keep.append(r)
                arr[q].roi = r.ctypes.data
        plan = GsPlan(iterations=iterations, nx=nx, ny=ny, dx=dx, dy=dy, lambda_=lam,
                      dtype=DTYPES[dtype])
        n = nx * ny
        ill = _darr(illumination).reshape(-1)
        ip = None if init_phase is None else _darr(init_phase).reshape(-1)
        out = {"phase": np.zeros(n), "error": np.zeros(iterations), "efficiency": np.zeros(iterations),
               "per_plane": np.zeros((len(planes), iterations))}
        check(self.lib.doe_gs_multiplane_run(self.h, C.byref(plan), _dptr(ill), len(planes), arr,
                                             None if ip is None else _dptr(ip), _dptr(out["phase"]),
                                             _dptr(out["error"]), _dptr(out["efficiency"]),
                                             _dptr(out["per_plane"])))
        return out

    # ---- targets, ROIs, figure images ----------------------------------
    def target_spec(self, kind: str, **kw) -> TargetSpec:
        """A spec with the reference's defaults for `kind`, overridden by kw
        (raster: 2-D array; missing: list of (col, row))."""
        spec = TargetSpec()
        check(self.lib.doe_target_spec_init(C.byref(spec), TGT_KINDS[kind]))
        keep = []
        for k, v in kw.items():
            if k == "raster":
                a = _darr(v)
                keep.append(a)
                spec.raster, spec.raster_h, spec.raster_w = a.ctypes.data, a.shape[0], a.shape[1]
            elif k == "missing":
                m = np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(-1, 2))
                keep.append(m)
                spec.missing, spec.n_missing = (m.ctypes.data if len(m) else None), len(m)
            else:
                setattr(spec, k, v)
        spec._keep = keep
        return spec

    def target_make(self, specs, nx: int, ny: int, out: "Field" = None) -> "Field":
        """makeTarget on the device: one FP64 field per spec (re = irradiance)."""
        specs = list(specs)
        arr = (TargetSpec * len(specs))(*specs)
        if out is None:
            out = self.field(nx, ny, len(specs), "f64")
        check(self.lib.doe_target_make(self.h, arr, out.h))
        return out

    def target_make_roi(self, specs, nx: int, ny: int, out: "Field" = None, dilate=2.0, soft=2.0,
                        threshold=0.05) -> "Field":
        """makeTarget + autoRoi in one call (re = irradiance, im = ROI)."""
        specs = list(specs)
        arr = (TargetSpec * len(specs))(*specs)
        if out is None:
            out = self.field(nx, ny, len(specs), "f64")
        check(self.lib.doe_target_make_roi(self.h, arr, out.h, dilate, soft, threshold))
        return out

    def target_auto_roi(self, t: "Field", dilate=2.0, soft=2.0, threshold=0.05) -> None:
        check(self.lib.doe_target_auto_roi(self.h, t.h, dilate, soft, threshold))

    def raster_blur(self, src: "Field", dst: "Field", sigma: float, taps=None) -> None:
        tp = None if taps is None else _darr(taps)
        check(self.lib.doe_raster_blur(self.h, src.h, dst.h, sigma, None if tp is None else _dptr(tp)))

    def field_image(self, f: "Field", index: int = 0, quantity: str = "re", gray: bool = False,
                    stops=None, vmin=None, vmax=None, gamma=1.0, log=False, log_floor=1e-4,
                    percentile=None, invert=False, png_rows=False):
        """planeToRGB / planeToGray of one field: (bytes as uint8 array, (lo, hi));
        stops: list of (position, (r, g, b))"""
        o = ImageOpts()
        check(self.lib.doe_image_opts_init(C.byref(o)))
        o.quantity, o.gray = QUANTITIES[quantity], int(gray)
        if stops is not None:
            o.nstops = len(stops)
            for q, (pos, rgb) in enumerate(stops):
                o.stop_pos[q] = pos
                for c in range(3):
                    o.stop_rgb[q][c] = rgb[c]
        if vmin is not None:
            o.has_min, o.min = 1, vmin
        if vmax is not None:
            o.has_max, o.max = 1, vmax
        o.gamma, o.log, o.log_floor, o.invert, o.png_rows = gamma, int(log), log_floor, int(invert), int(png_rows)
        if percentile is not None:
            o.has_percentile, o.percentile = 1, percentile
        ch = 1 if gray else 3
        out = np.zeros(f.ny * (int(png_rows) + f.nx * ch), dtype=np.uint8)
        rng = np.zeros(2)
        check(self.lib.doe_field_image(self.h, f.h, index, C.byref(o), out.ctypes.data, out.size,
                                       _dptr(rng)))
        return out, (rng[0], rng[1])

    # ---- random phase ----------------------------------------------
    def random_phase(self, out: "Field", seed: int = 1) -> None:
        """out.re = the reference's randomPhase stream of seed (ToUint32, 0 -> 1)."""
        check(self.lib.doe_random_phase(self.h, out.h, int(seed) & 0xFFFFFFFF))

    # ---- fabrication output ----------------------------------------
    def error_diffusion(self, phase: "Field", out: "Field", levels: int = 4, snake: bool = True,
                        wrap_first: bool = True) -> None:
        check(self.lib.doe_error_diffusion(self.h, phase.h, out.h, levels, int(snake), int(wrap_first)))

    def phase_to_height(self, phase: "Field", out: "Field", lam: float, n_index: float = 1.52) -> None:
        check(self.lib.doe_phase_to_height(self.h, phase.h, out.h, lam, n_index))

    def gs_run(self, plan: GsPlan, illumination, target_amp, roi=None, init_phase=None,
               seed: int = 12345):
        n = plan.nx * plan.ny
        ill = _darr(illumination).ravel()
        tgt = _darr(target_amp).ravel()
        roi_a = None if roi is None else _darr(roi).ravel()
        ip = None if init_phase is None else _darr(init_phase).ravel()
        out = {
            "phase": np.zeros(n), "error": np.zeros(plan.iterations),
            "efficiency": np.zeros(plan.iterations), "amplitude": np.zeros(n),
            "intensity": np.zeros(n), "weights": np.zeros(n), "rmse": np.zeros(1),
        }
        check(self.lib.doe_gs_run(
            self.h, C.byref(plan), _dptr(ill), _dptr(tgt),
            _dptr(roi_a) if roi_a is not None else None,
            _dptr(ip) if ip is not None else None, C.c_uint(seed & 0xFFFFFFFF),
            _dptr(out["phase"]), _dptr(out["error"]), _dptr(out["efficiency"]),
            _dptr(out["amplitude"]), _dptr(out["intensity"]), _dptr(out["weights"]),
            _dptr(out["rmse"])))
        out["rmse"] = float(out["rmse"][0])
        return out


class GsBatch:
    """Batched far-field GS/WGS over designs of uneven (square) sides."""

    def __init__(self, ctx: Context, sides, algorithm: int, beta: float = 0.9,
                 with_roi: bool = False):
        self.ctx = ctx
        self.sides = [int(v) for v in sides]
        arr = (C.c_int * len(self.sides))(*self.sides)
        h = _VP()
        check(ctx.lib.doe_gs_batch_create(ctx.h, len(self.sides), arr, int(algorithm), float(beta),
                                          int(with_roi), C.byref(h)))
        self.h = h
        self.iterations = 0
        ctx._adopt(self)

    def set(self, index: int, illumination, target_amp, roi=None, init_phase=None) -> None:
        ill = _darr(illumination).ravel()
        tgt = _darr(target_amp).ravel()
        roi_a = None if roi is None else _darr(roi).ravel()
        ip = None if init_phase is None else _darr(init_phase).ravel()
        check(self.ctx.lib.doe_gs_batch_set(self.h, int(index), _dptr(ill), _dptr(tgt),
                                            _dptr(roi_a) if roi_a is not None else None,
                                            _dptr(ip) if ip is not None else None))

    def random_init(self, seed: int) -> None:
        check(self.ctx.lib.doe_gs_batch_random_init(self.h, C.c_ulonglong(seed)))

    def run(self, iterations: int, graph: bool = True, per_class: bool = False) -> None:
        """per_class: one launch per pass per size class (DOE_GSB_PER_CLASS)
        instead of one per pass for the whole batch."""
        flags = (GSB_GRAPH if graph else 0) | (GSB_PER_CLASS if per_class else 0)
        check(self.ctx.lib.doe_gs_batch_run(self.h, int(iterations), flags))
        self.iterations = int(iterations)

    def get(self, index: int) -> dict:
        n = self.sides[index] ** 2
        k = self.iterations
        out = {"phase": np.zeros(n), "error": np.zeros(k), "efficiency": np.zeros(k),
               "amplitude": np.zeros(n), "weights": np.zeros(n), "rmse": np.zeros(1)}
        check(self.ctx.lib.doe_gs_batch_get(self.h, int(index), _dptr(out["phase"]),
                                            _dptr(out["error"]), _dptr(out["efficiency"]),
                                            _dptr(out["amplitude"]), _dptr(out["weights"]),
                                            _dptr(out["rmse"])))
        out["rmse"] = float(out["rmse"][0])
        return out

    def design_order(self):
        arr = (C.c_int * len(self.sides))()
        check(self.ctx.lib.doe_gs_batch_design_order(self.h, arr))
        return list(arr)

    def warm_inputs(self, res: int, input_ptr: int) -> None:
        check(self.ctx.lib.doe_warmstart_inputs(self.h, res, input_ptr))

    def warm_labels(self, res: int, lre_ptr: int, lim_ptr: int, w_ptr: int) -> None:
        check(self.ctx.lib.doe_warmstart_labels(self.h, res, lre_ptr, lim_ptr, w_ptr))

    def warm_apply(self, net: "WsNet", scratch_ptr: int) -> None:
        check(self.ctx.lib.doe_warmstart_apply(self.h, net.h, scratch_ptr))

    def warm_apply_quadratic(self, res: int, spread: float, scratch_ptr: int) -> None:
        check(self.ctx.lib.doe_warmstart_apply_quadratic(self.h, res, spread, scratch_ptr))

    def warm_apply_phasor(self, res: int, phasor_ptr: int) -> None:
        check(self.ctx.lib.doe_warmstart_apply_phasor(self.h, res, phasor_ptr))

    def restart_best(self) -> None:
        """next run starts from one GS projection of each design's best iterate"""
        check(self.ctx.lib.doe_gs_batch_restart_best(self.h))

    def get_screen(self, index: int):
        out = np.zeros(self.iterations)
        check(self.ctx.lib.doe_gs_batch_get_screen(self.h, int(index), _dptr(out)))
        return out

    def close(self) -> None:
        if self.h:
            self.ctx.lib.doe_gs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WsNet:
    """The warm-start network (device-pointer interface of doe_wsnet_*)."""

    def __init__(self, ctx: Context, res: int = 256, base: int = 32, max_batch: int = 8,
                 seed: int = 0):
        self.ctx = ctx
        h = _VP()
        check(ctx.lib.doe_wsnet_create(ctx.h, res, base, max_batch, C.c_ulonglong(seed),
                                       C.byref(h)))
        self.h = h
        ctx._adopt(self)
        self.res, self.base, self.max_batch = res, base, max_batch
        n = C.c_long()
        check(ctx.lib.doe_wsnet_param_count(self.h, C.byref(n)))
        self.nparams = n.value

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, dtype=np.float32)
        check(self.ctx.lib.doe_wsnet_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.nparams
        check(self.ctx.lib.doe_wsnet_set_params(self.h, a.ctypes.data))

    def forward(self, batch: int, input_ptr: int, output_ptr: int) -> None:
        check(self.ctx.lib.doe_wsnet_forward(self.h, batch, input_ptr, output_ptr))

    def loss_grad(self, batch, input_ptr, lre_ptr, lim_ptr, w_ptr, want_grad=False):
        loss = C.c_double()
        g = np.zeros(self.nparams, dtype=np.float32) if want_grad else None
        check(self.ctx.lib.doe_wsnet_loss_grad(self.h, batch, input_ptr, lre_ptr, lim_ptr, w_ptr,
                                               C.byref(loss),
                                               g.ctypes.data if g is not None else None))
        return loss.value, g

    def adam_step(self, lr: float) -> None:
        check(self.ctx.lib.doe_wsnet_adam_step(self.h, C.c_float(lr)))

    def close(self) -> None:
        if self.h:
            self.ctx.lib.doe_wsnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Field:
    def __init__(self, ctx: Context, h, nx, ny, batch, dtype):
        self.ctx, self.h, self.nx, self.ny, self.batch, self.dtype = ctx, h, nx, ny, batch, dtype

    def download(self):
        n = self.nx * self.ny * self.batch
        re = np.zeros(n)
        im = np.zeros(n)
        check(self.ctx.lib.doe_field_download(self.ctx.h, self.h, _dptr(re), _dptr(im)))
        shape = (self.batch, self.ny, self.nx)
        return re.reshape(shape), im.reshape(shape)

    def upload(self, re, im) -> None:
        re = _darr(re).reshape(-1)
        im = _darr(im).reshape(-1)
        n = self.nx * self.ny * self.batch
        if re.size != n or im.size != n:
            raise ValueError(f"Field.upload: planes of {re.size} and {im.size} values for a field of {n}")
        check(self.ctx.lib.doe_field_upload(self.ctx.h, self.h, _dptr(re), _dptr(im)))

    def ptrs(self):
        re = C.c_void_p()
        im = C.c_void_p()
        check(self.ctx.lib.doe_field_info(self.h, None, None, None, None, C.byref(re), C.byref(im)))
        return re.value, im.value

    def free(self) -> None:
        if self.h:
            self.ctx.lib.doe_field_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def header_symbols() -> list[str]:
    """Names of every function declared in include/doehip.h."""
    import re as _re
    text = (PKG.parent / "include" / "doehip.h").read_text()
    return sorted(set(_re.findall(r"\b(doe_[a-z0-9_]+)\s*\(", text)))
