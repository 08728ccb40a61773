// doe_napi.node — thin N-API (v8) addon over the C ABI of libdoehip.so.
//
// The JavaScript drop-in modules (../js/*.js) call these functions; every
// non-OK status of the library becomes a thrown Error carrying the library's
// message (which reproduces the reference's messages where the reference
// throws, e.g. "fft2d: plane length ... does not match NxM").
//
// One process-wide library context on device 0 is created on first use.
#include <node_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "doehip.h"

namespace {

doe_ctx* g_ctx = nullptr;

bool throw_status(napi_env env, int rc) {
  if (rc == DOE_OK) return false;
  std::string msg = doe_last_error();
  if (msg.empty()) msg = doe_status_string(rc);
  napi_throw_error(env, nullptr, msg.c_str());
  return true;
}

doe_ctx* ctx(napi_env env) {
  if (!g_ctx) {
    int rc = doe_ctx_create(0, nullptr, &g_ctx);
    if (rc != DOE_OK) {
      throw_status(env, rc);
      g_ctx = nullptr;
    }
  }
  return g_ctx;
}

#define CHECK_NAPI(call)                                        \
  do {                                                          \
    if ((call) != napi_ok) {                                    \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                           \
    }                                                           \
  } while (0)

struct Args {
  napi_value v[16];
  size_t n = 16;
};

bool get_args(napi_env env, napi_callback_info info, Args& a) {
  return napi_get_cb_info(env, info, &a.n, a.v, nullptr, nullptr) == napi_ok;
}

double num(napi_env env, napi_value v, double dflt = 0.0) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

bool boolean(napi_env env, napi_value v, bool dflt) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b = dflt;
    napi_get_value_bool(env, v, &b);
    return b;
  }
  if (t == napi_number) return num(env, v) != 0;
  return dflt;
}

// Float64Array -> pointer + length (nullptr when the value is not one).
double* f64(napi_env env, napi_value v, size_t* len) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return nullptr;
  napi_typedarray_type t;
  size_t n = 0;
  void* data = nullptr;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &t, &n, &data, &ab, &off) != napi_ok) return nullptr;
  if (t != napi_float64_array) return nullptr;
  if (len) *len = n;
  return (double*)data;
}

// A Float64Array of at least `need` elements.  Returns nullptr with a pending
// RangeError when the array is shorter (the library would read or write
// `need` elements through the pointer), and nullptr without an exception
// when the value is not a Float64Array at all.
double* f64_need(napi_env env, napi_value v, size_t need, const char* what) {
  size_t n = 0;
  double* p = f64(env, v, &n);
  if (p && n < need) {
    const std::string msg = std::string(what) + ": Float64Array of length " + std::to_string(n) +
                            " is shorter than the " + std::to_string(need) + " elements required";
    napi_throw_range_error(env, nullptr, msg.c_str());
    return nullptr;
  }
  return p;
}

bool pending(napi_env env) {
  bool p = false;
  napi_is_exception_pending(env, &p);
  return p;
}

napi_value make_f64(napi_env env, size_t n, double** data) {
  napi_value ab, arr;
  void* p = nullptr;
  napi_create_arraybuffer(env, n * sizeof(double), &p, &ab);
  napi_create_typedarray(env, napi_float64_array, n, ab, 0, &arr);
  *data = (double*)p;
  return arr;
}

napi_value make_num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

void set(napi_env env, napi_value obj, const char* k, napi_value v) {
  napi_set_named_property(env, obj, k, v);
}

doe_field* field_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok) return nullptr;
  return (doe_field*)p;
}

// ---------------------------------------------------------------------------
napi_value Version(napi_env env, napi_callback_info) {
  return make_num(env, doe_version());
}

napi_value DeviceCount(napi_env env, napi_callback_info) {
  int c = 0;
  doe_device_count(&c);
  return make_num(env, c);
}

// fft2d(re, im, nx, ny, inverse, dtype) — in place on host Float64Arrays
napi_value Fft2d(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 4) {
    napi_throw_type_error(env, nullptr, "fft2d(re, im, nx, ny, inverse, dtype)");
    return nullptr;
  }
  size_t lr = 0, li = 0;
  double* re = f64(env, a.v[0], &lr);
  double* im = f64(env, a.v[1], &li);
  if (!re || !im) {
    napi_throw_type_error(env, nullptr, "fft2d: re and im must be Float64Array planes");
    return nullptr;
  }
  const int nx = (int)num(env, a.v[2]);
  const int ny = (int)num(env, a.v[3]);
  const int inv = a.n > 4 && boolean(env, a.v[4], false);
  const int dt = a.n > 5 ? (int)num(env, a.v[5], DOE_F64) : DOE_F64;
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  long len = (long)lr;
  if (li != lr) len = -1;  // mismatched planes: reported as a length error
  throw_status(env, doe_fft2d_host(c, re, im, len < 0 ? (long)li : len, nx, ny, inv, dt));
  return nullptr;
}

// fft1d(re, im, n, stride, offset, inverse, dtype) — one strided line
napi_value Fft1d(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 3) {
    napi_throw_type_error(env, nullptr, "fft1d(re, im, n, stride, offset, inverse, dtype)");
    return nullptr;
  }
  size_t lr = 0, li = 0;
  double* re = f64(env, a.v[0], &lr);
  double* im = f64(env, a.v[1], &li);
  if (!re || !im) {
    napi_throw_type_error(env, nullptr, "fft1d: re and im must be Float64Array planes");
    return nullptr;
  }
  const long n = (long)num(env, a.v[2]);
  const long stride = a.n > 3 ? (long)num(env, a.v[3], 1) : 1;
  const long offset = a.n > 4 ? (long)num(env, a.v[4], 0) : 0;
  const int inv = a.n > 5 && boolean(env, a.v[5], false);
  const int dt = a.n > 6 ? (int)num(env, a.v[6], DOE_F64) : DOE_F64;
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  throw_status(env, doe_fft1d_host(c, re, im, (long)std::min(lr, li), n, stride, offset, inv, dt));
  return nullptr;
}

void finalize_field(napi_env, void* data, void*) { doe_field_free((doe_field*)data); }

// fieldCreate(nx, ny, batch, dtype) -> external handle
napi_value FieldCreate(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  doe_field* f = nullptr;
  const int rc = doe_field_create(c, (int)num(env, a.v[0]), (int)num(env, a.v[1]),
                                  a.n > 2 ? (int)num(env, a.v[2], 1) : 1,
                                  a.n > 3 ? (int)num(env, a.v[3], DOE_F32) : DOE_F32, &f);
  if (throw_status(env, rc)) return nullptr;
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, f, finalize_field, nullptr, &ext));
  return ext;
}

// Elements of every plane of a field (nx * ny * batch).
size_t count_of(doe_field* f) {
  int nx = 0, ny = 0, b = 0;
  doe_field_info(f, &nx, &ny, &b, nullptr, nullptr, nullptr);
  return size_t(nx) * size_t(ny) * size_t(b);
}

// fieldUpload / fieldDownload(handle, re, im): both planes must hold at
// least nx * ny * batch elements (RangeError otherwise).
napi_value FieldTransfer(napi_env env, napi_callback_info info, bool up) {
  Args a;
  get_args(env, info, a);
  const char* what = up ? "fieldUpload" : "fieldDownload";
  doe_field* f = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  if (!f || a.n < 3) {
    napi_throw_type_error(env, nullptr, up ? "fieldUpload(handle, re, im)" : "fieldDownload(handle, re, im)");
    return nullptr;
  }
  const size_t need = count_of(f);
  double* re = f64_need(env, a.v[1], need, what);
  double* im = re ? f64_need(env, a.v[2], need, what) : nullptr;
  if (!re || !im) {
    if (!pending(env))
      napi_throw_type_error(env, nullptr, up ? "fieldUpload(handle, re, im)" : "fieldDownload(handle, re, im)");
    return nullptr;
  }
  throw_status(env, up ? doe_field_upload(ctx(env), f, re, im) : doe_field_download(ctx(env), f, re, im));
  return nullptr;
}
napi_value FieldUpload(napi_env env, napi_callback_info info) { return FieldTransfer(env, info, true); }
napi_value FieldDownload(napi_env env, napi_callback_info info) { return FieldTransfer(env, info, false); }

// Per-batch distances: a number (every field) or a Float64Array of at least
// `batch` entries.  Returns false with a pending RangeError when shorter.
bool zlist(napi_env env, napi_value v, int batch, const char* what, std::vector<double>* out) {
  size_t n = 0;
  double* p = f64(env, v, &n);
  if (p) {
    if (n < size_t(batch)) {
      const std::string msg = std::string(what) + ": " + std::to_string(n) +
                              " distances for a batch of " + std::to_string(batch) + " fields";
      napi_throw_range_error(env, nullptr, msg.c_str());
      return false;
    }
    out->assign(p, p + batch);
  } else {
    out->assign(size_t(batch), num(env, v));
  }
  return true;
}

int batch_of(doe_field* f) {
  int b = 1;
  doe_field_info(f, nullptr, nullptr, &b, nullptr, nullptr, nullptr);
  return b;
}

// asm(hin, hout, dx, dy, lambda, z, bandLimit, evanescent, softEdge)
napi_value Asm(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* fi = field_of(env, a.v[0]);
  doe_field* fo = field_of(env, a.v[1]);
  if (!fi || !fo || a.n < 9) {
    napi_throw_type_error(env, nullptr, "asm(hin, hout, dx, dy, lambda, z, bandLimit, evanescent, softEdge)");
    return nullptr;
  }
  std::vector<double> z;
  if (!zlist(env, a.v[5], batch_of(fi), "asm", &z)) return nullptr;
  throw_status(env, doe_asm(ctx(env), fi, fo, num(env, a.v[2]), num(env, a.v[3]), num(env, a.v[4]),
                            z.data(), boolean(env, a.v[6], true), boolean(env, a.v[7], true),
                            num(env, a.v[8], 0.02)));
  return nullptr;
}

// asmTransfer(nx, ny, dx, dy, lambda, z, bandLimit, evanescent, softEdge)
napi_value AsmTransfer(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  const int nx = (int)num(env, a.v[0]), ny = (int)num(env, a.v[1]);
  doe_field* f = nullptr;
  if (throw_status(env, doe_field_create(c, nx, ny, 1, DOE_F64, &f))) return nullptr;
  double frac = 0;
  int rc = doe_asm_transfer(c, f, num(env, a.v[2]), num(env, a.v[3]), num(env, a.v[4]),
                            num(env, a.v[5]), boolean(env, a.v[6], true), boolean(env, a.v[7], true),
                            num(env, a.v[8], 0.02), &frac);
  napi_value obj = nullptr;
  if (rc == DOE_OK) {
    double *re, *im;
    napi_value vre = make_f64(env, (size_t)nx * ny, &re);
    napi_value vim = make_f64(env, (size_t)nx * ny, &im);
    rc = doe_field_download(c, f, re, im);
    napi_create_object(env, &obj);
    set(env, obj, "re", vre);
    set(env, obj, "im", vim);
    set(env, obj, "propagatingFraction", make_num(env, frac));
  }
  doe_field_free(f);
  if (throw_status(env, rc)) return nullptr;
  return obj;
}

// fresnel(hin, hout, dx, dy, lambda, z, inverse, inDx, inDy) -> {dx, dy}
napi_value Fresnel(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* fi = field_of(env, a.v[0]);
  doe_field* fo = field_of(env, a.v[1]);
  if (!fi || !fo || a.n < 7) {
    napi_throw_type_error(env, nullptr, "fresnel(hin, hout, dx, dy, lambda, z, inverse, inDx, inDy)");
    return nullptr;
  }
  const int b = batch_of(fi);
  std::vector<double> z, odx(b), ody(b);
  if (!zlist(env, a.v[5], b, "fresnel", &z)) return nullptr;
  const int rc = doe_fresnel(ctx(env), fi, fo, num(env, a.v[2]), num(env, a.v[3]), num(env, a.v[4]),
                             z.data(), boolean(env, a.v[6], false), a.n > 7 ? num(env, a.v[7]) : 0,
                             a.n > 8 ? num(env, a.v[8]) : 0, odx.data(), ody.data());
  if (throw_status(env, rc)) return nullptr;
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "dx", make_num(env, odx[0]));
  set(env, obj, "dy", make_num(env, ody[0]));
  return obj;
}

// lens(hin, hout, dx, dy, lambda, f) -> {dx, dy}
napi_value Lens(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* fi = field_of(env, a.v[0]);
  doe_field* fo = field_of(env, a.v[1]);
  if (!fi || !fo || a.n < 6) {
    napi_throw_type_error(env, nullptr, "lens(hin, hout, dx, dy, lambda, f)");
    return nullptr;
  }
  const int b = batch_of(fi);
  std::vector<double> f, odx(b), ody(b);
  if (!zlist(env, a.v[5], b, "lens", &f)) return nullptr;
  const int rc = doe_lens(ctx(env), fi, fo, num(env, a.v[2]), num(env, a.v[3]), num(env, a.v[4]),
                          f.data(), odx.data(), ody.data());
  if (throw_status(env, rc)) return nullptr;
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "dx", make_num(env, odx[0]));
  set(env, obj, "dy", make_num(env, ody[0]));
  return obj;
}

double prop(napi_env env, napi_value obj, const char* k, double dflt) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, k, &has);
  if (!has) return dflt;
  napi_get_named_property(env, obj, k, &v);
  return num(env, v, dflt);
}

// gsRun(plan, illumination, targetAmp, roi|null, initPhase|null, seed)
napi_value metrics_object(napi_env env, const doe_metrics& m) {
  napi_value obj;
  if (napi_create_object(env, &obj) != napi_ok) return nullptr;
  napi_value nul;
  napi_get_null(env, &nul);
  set(env, obj, "efficiency", make_num(env, m.efficiency));
  set(env, obj, "fluxFraction", make_num(env, m.flux_fraction));
  set(env, obj, "uniformity", make_num(env, m.uniformity));
  set(env, obj, "spotUniformity", m.has_spots ? make_num(env, m.spot_uniformity) : nul);
  set(env, obj, "flatness", make_num(env, m.flatness));
  set(env, obj, "contrast", make_num(env, m.contrast));
  set(env, obj, "rmse", make_num(env, m.rmse));
  set(env, obj, "fidelity", make_num(env, m.fidelity));
  set(env, obj, "correlation", make_num(env, m.correlation));
  set(env, obj, "ssim", make_num(env, m.ssim));
  set(env, obj, "snr", make_num(env, m.snr));
  set(env, obj, "zeroOrder", make_num(env, m.zero_order));
  set(env, obj, "rmsRadius", make_num(env, m.rms_radius));
  set(env, obj, "meanIntensity", make_num(env, m.mean_intensity));
  set(env, obj, "peakIntensity", make_num(env, m.peak_intensity));
  set(env, obj, "powerInSignal", make_num(env, m.power_in_signal));
  set(env, obj, "powerAtScreen", make_num(env, m.power_at_screen));
  set(env, obj, "powerIncident", make_num(env, m.power_incident));
  return obj;
}

bool screen_opts(napi_env env, const Args& a, doe_screen_opts* o) {
  o->sdx = num(env, a.v[3]);
  o->sdy = num(env, a.v[4]);
  o->edx = num(env, a.v[5]);
  o->edy = num(env, a.v[6]);
  o->grid_cols = a.n > 7 ? (int)num(env, a.v[7], 0) : 0;
  o->grid_rows = a.n > 8 ? (int)num(env, a.v[8], 0) : 0;
  o->grid_min_peak = a.n > 9 ? num(env, a.v[9], 0.05) : 0.05;
  o->ssim_weight_target = a.n > 10 ? (boolean(env, a.v[10], true) ? 1 : 0) : 1;
  return true;
}

// screenMetrics(screen, target, element, sdx, sdy, edx, edy, cols, rows, minPeak, weightTarget)
// -> the reference's screenMetrics object (projector.js:597-613); target is
// an FP64 field handle with re = irradiance, im = ROI.
napi_value ScreenMetrics(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 7) {
    napi_throw_type_error(env, nullptr,
                          "screenMetrics(screen, target, element, sdx, sdy, edx, edy, cols, rows, "
                          "minPeak, weightTarget)");
    return nullptr;
  }
  doe_field* s = field_of(env, a.v[0]);
  doe_field* t = field_of(env, a.v[1]);
  doe_field* e = field_of(env, a.v[2]);
  if (!s || !t || !e) {
    napi_throw_type_error(env, nullptr, "screenMetrics: screen, target and element must be field handles");
    return nullptr;
  }
  doe_screen_opts o;
  screen_opts(env, a, &o);
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  doe_metrics m;
  if (throw_status(env, doe_screen_metrics(c, s, t, e, &o, &m))) return nullptr;
  return metrics_object(env, m);
}

// screenMetricsBatch(screens, target, element, ...same options) -> array of metrics
napi_value ScreenMetricsBatch(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 7) {
    napi_throw_type_error(env, nullptr, "screenMetricsBatch(screens, target, element, sdx, sdy, edx, edy, ...)");
    return nullptr;
  }
  doe_field* s = field_of(env, a.v[0]);
  doe_field* t = field_of(env, a.v[1]);
  doe_field* e = field_of(env, a.v[2]);
  if (!s || !t || !e) {
    napi_throw_type_error(env, nullptr, "screenMetricsBatch: screens, target and element must be field handles");
    return nullptr;
  }
  int batch = 0;
  doe_field_info(s, nullptr, nullptr, &batch, nullptr, nullptr, nullptr);
  doe_screen_opts o;
  screen_opts(env, a, &o);
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  std::vector<doe_metrics> m(batch > 0 ? batch : 1);
  if (throw_status(env, doe_screen_metrics_batch(c, s, t, e, &o, m.data()))) return nullptr;
  napi_value arr;
  CHECK_NAPI(napi_create_array_with_length(env, batch, &arr));
  for (int b = 0; b < batch; ++b) napi_set_element(env, arr, b, metrics_object(env, m[b]));
  return arr;
}

// fieldResample(src, dst[, srcDx, srcDy, dstDx, dstDy]): centred same-pitch
// resample (src batch 1 broadcasts); with pitches that differ, bilinear in
// physical coordinates
napi_value FieldResample(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* s = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  doe_field* d = a.n > 1 ? field_of(env, a.v[1]) : nullptr;
  if (!s || !d) {
    napi_throw_type_error(env, nullptr, "fieldResample(src, dst, srcDx, srcDy, dstDx, dstDy)");
    return nullptr;
  }
  if (a.n >= 6) {
    throw_status(env, doe_field_resample(ctx(env), s, d, num(env, a.v[2], NAN), num(env, a.v[3], NAN),
                                         num(env, a.v[4], NAN), num(env, a.v[5], NAN)));
  } else {
    throw_status(env, doe_field_resample_centred(ctx(env), s, d));
  }
  return nullptr;
}

// fieldSampleImage(img Float64Array, width, height, scale, fill, dst)
napi_value FieldSampleImage(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* d = a.n > 5 ? field_of(env, a.v[5]) : nullptr;
  const int w = a.n > 1 ? (int)num(env, a.v[1], 0) : 0;
  const int h = a.n > 2 ? (int)num(env, a.v[2], 0) : 0;
  const double* img = a.n > 0 ? f64_need(env, a.v[0], size_t(std::max(w, 0)) * size_t(std::max(h, 0)),
                                         "sampleImage: img") : nullptr;
  if (pending(env)) return nullptr;
  if (!img || !d) {
    napi_throw_type_error(env, nullptr, "fieldSampleImage(img, width, height, scale, fill, dst)");
    return nullptr;
  }
  throw_status(env, doe_field_sample_image(ctx(env), img, w, h, num(env, a.v[3], 1.0), num(env, a.v[4], 0.0), d));
  return nullptr;
}

// fieldCopySlices(src, srcFirst, dst, dstFirst, count)
napi_value FieldCopySlices(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* s = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  doe_field* d = a.n > 2 ? field_of(env, a.v[2]) : nullptr;
  if (!s || !d || a.n < 5) {
    napi_throw_type_error(env, nullptr, "fieldCopySlices(src, srcFirst, dst, dstFirst, count)");
    return nullptr;
  }
  throw_status(env, doe_field_copy_slices(ctx(env), s, (int)num(env, a.v[1]), d, (int)num(env, a.v[3]),
                                          (int)num(env, a.v[4])));
  return nullptr;
}

// gsMultiPlane(illum, planes[{z, targetAmp, roi, weight}], nx, ny, dx, dy, lambda,
//              iterations, dtype, initPhase|null) -> {phase, error, efficiency, perPlaneError[]}
napi_value GsMultiPlane(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 9) {
    napi_throw_type_error(env, nullptr,
                          "gsMultiPlane(illum, planes, nx, ny, dx, dy, lambda, iterations, dtype, initPhase)");
    return nullptr;
  }
  const size_t n = size_t(std::max(0, (int)num(env, a.v[2]))) * size_t(std::max(0, (int)num(env, a.v[3])));
  double* illum = f64_need(env, a.v[0], n, "gsMultiPlane: illumination");
  if (pending(env)) return nullptr;
  bool is_arr = false;
  napi_is_array(env, a.v[1], &is_arr);
  if (!illum || !is_arr) {
    napi_throw_type_error(env, nullptr, "gsMultiPlane: illumination must be a Float64Array, planes an array");
    return nullptr;
  }
  uint32_t np = 0;
  napi_get_array_length(env, a.v[1], &np);
  std::vector<doe_mp_plane> planes(np);
  for (uint32_t q = 0; q < np; ++q) {
    napi_value el, v;
    napi_get_element(env, a.v[1], q, &el);
    napi_get_named_property(env, el, "z", &v);
    planes[q].z = num(env, v);
    napi_get_named_property(env, el, "targetAmp", &v);
    planes[q].target_amp = f64_need(env, v, n, "gsMultiPlane: targetAmp");
    if (pending(env)) return nullptr;
    napi_get_named_property(env, el, "roi", &v);
    planes[q].roi = f64_need(env, v, n, "gsMultiPlane: roi");
    if (pending(env)) return nullptr;
    napi_get_named_property(env, el, "weight", &v);
    planes[q].weight = num(env, v, 1.0);
    if (!planes[q].target_amp) {
      napi_throw_type_error(env, nullptr, "gsMultiPlane: every plane needs a Float64Array targetAmp");
      return nullptr;
    }
  }
  doe_gs_plan plan{};
  plan.nx = (int)num(env, a.v[2]);
  plan.ny = (int)num(env, a.v[3]);
  plan.dx = num(env, a.v[4]);
  plan.dy = num(env, a.v[5]);
  plan.lambda = num(env, a.v[6]);
  plan.iterations = (int)num(env, a.v[7]);
  plan.dtype = (int)num(env, a.v[8], DOE_F32);
  const double* init = a.n > 9 ? f64_need(env, a.v[9], n, "gsMultiPlane: initPhase") : nullptr;
  if (pending(env)) return nullptr;
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  const size_t K = plan.iterations > 0 ? size_t(plan.iterations) : 0;
  double *phase, *err, *eff;
  napi_value jphase = make_f64(env, n, &phase);
  napi_value jerr = make_f64(env, K, &err);
  napi_value jeff = make_f64(env, K, &eff);
  std::vector<double> per(np * K + 1);
  if (throw_status(env, doe_gs_multiplane_run(c, &plan, illum, (int)np, planes.data(), init, phase, err,
                                              eff, per.data())))
    return nullptr;
  napi_value jper;
  CHECK_NAPI(napi_create_array_with_length(env, np, &jper));
  for (uint32_t q = 0; q < np; ++q) {
    double* d;
    napi_value arr = make_f64(env, K, &d);
    std::memcpy(d, per.data() + q * K, K * sizeof(double));
    napi_set_element(env, jper, q, arr);
  }
  napi_value obj;
  CHECK_NAPI(napi_create_object(env, &obj));
  set(env, obj, "phase", jphase);
  set(env, obj, "error", jerr);
  set(env, obj, "efficiency", jeff);
  set(env, obj, "perPlaneError", jper);
  return obj;
}

napi_value GsRun(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  if (a.n < 3) {
    napi_throw_type_error(env, nullptr, "gsRun(plan, illumination, targetAmp, roi, initPhase, seed)");
    return nullptr;
  }
  doe_gs_plan p{};
  p.mode = (int)prop(env, a.v[0], "mode", DOE_MODE_IMAGE_PLANE);
  p.algorithm = (int)prop(env, a.v[0], "algorithm", DOE_ALG_GS);
  p.iterations = (int)prop(env, a.v[0], "iterations", 60);
  p.nx = (int)prop(env, a.v[0], "nx", 0);
  p.ny = (int)prop(env, a.v[0], "ny", p.nx);
  p.dx = prop(env, a.v[0], "dx", 8e-6);
  p.dy = prop(env, a.v[0], "dy", p.dx);
  p.lambda = prop(env, a.v[0], "lambda", 532e-9);
  p.z = prop(env, a.v[0], "z", 0.03);
  p.beta = prop(env, a.v[0], "beta", 0.9);
  p.feedback = prop(env, a.v[0], "feedback", 0.0);
  p.escape = (int)prop(env, a.v[0], "escape", 0);
  p.dtype = (int)prop(env, a.v[0], "dtype", DOE_F32);
  const size_t n = (size_t)p.nx * (size_t)p.ny;
  size_t l1 = 0, l2 = 0, l3 = 0, l4 = 0;
  double* ill = f64(env, a.v[1], &l1);
  double* tgt = f64(env, a.v[2], &l2);
  double* roi = a.n > 3 ? f64(env, a.v[3], &l3) : nullptr;
  double* ip = a.n > 4 ? f64(env, a.v[4], &l4) : nullptr;
  if (!ill || !tgt || l1 != n || l2 != n || (roi && l3 != n) || (ip && l4 != n)) {
    napi_throw_error(env, nullptr, "gsRun: planes must be Float64Array of length nx*ny");
    return nullptr;
  }
  const unsigned seed = a.n > 5 ? (unsigned)(long long)num(env, a.v[5], 12345) : 12345u;
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  double *ph, *am, *in, *we, *er, *ef;
  napi_value vph = make_f64(env, n, &ph), vam = make_f64(env, n, &am), vin = make_f64(env, n, &in),
             vwe = make_f64(env, n, &we);
  const size_t K = p.iterations > 0 ? (size_t)p.iterations : 0;
  napi_value ver = make_f64(env, K, &er), vef = make_f64(env, K, &ef);
  double rmse = 0;
  const int rc = doe_gs_run(c, &p, ill, tgt, roi, ip, seed, ph, K ? er : nullptr, K ? ef : nullptr,
                            am, in, we, &rmse);
  if (throw_status(env, rc)) return nullptr;
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "phase", vph);
  set(env, obj, "amplitude", vam);
  set(env, obj, "intensity", vin);
  set(env, obj, "weights", vwe);
  set(env, obj, "error", ver);
  set(env, obj, "efficiency", vef);
  set(env, obj, "rmse", make_num(env, rmse));
  return obj;
}

// ---- targets, ROIs, figure pixels -------------------------------------------
napi_value prop(napi_env env, napi_value obj, const char* k) {
  napi_value v = nullptr;
  if (napi_get_named_property(env, obj, k, &v) != napi_ok) napi_get_undefined(env, &v);
  return v;
}
bool present(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t != napi_undefined && t != napi_null;
}
double opt_num(napi_env env, napi_value obj, const char* k, double dflt) {
  napi_value v = prop(env, obj, k);
  return present(env, v) ? num(env, v, dflt) : dflt;
}
int32_t* i32(napi_env env, napi_value v, size_t* len) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return nullptr;
  napi_typedarray_type t;
  size_t n = 0;
  void* data = nullptr;
  napi_value ab;
  size_t off = 0;
  if (napi_get_typedarray_info(env, v, &t, &n, &data, &ab, &off) != napi_ok || t != napi_int32_array)
    return nullptr;
  if (len) *len = n;
  return (int32_t*)data;
}
std::string str(napi_env env, napi_value v) {
  char buf[64] = {0};
  size_t n = 0;
  if (napi_get_value_string_utf8(env, v, buf, sizeof(buf), &n) != napi_ok) return std::string();
  return std::string(buf, n);
}

// targetMake(specs[], nx, ny[, {dilate, soft, threshold}]) -> FP64 field
// handle (batch = specs.length, re = irradiance; with the ROI options, im =
// the autoRoi of the same call).  spec: {kind: 'disc'|'ring'|'dots'|'grid'|'checker'|
// 'lines'|'zone'|'raster', radius, inner, outer, cols, rows, sigma,
// missing: Int32Array of (col, row) pairs, cells, period, duty,
// raster: Float64Array, width, height, cover, gamma, background, gain};
// absent options take the reference's defaults.
napi_value TargetMake(napi_env env, napi_callback_info info) {
  Args a;
  bool is_arr = false;
  if (!get_args(env, info, a) || a.n < 3 || napi_is_array(env, a.v[0], &is_arr) != napi_ok || !is_arr) {
    napi_throw_type_error(env, nullptr, "targetMake(specs, nx, ny)");
    return nullptr;
  }
  uint32_t count = 0;
  napi_get_array_length(env, a.v[0], &count);
  if (count == 0) {
    napi_throw_type_error(env, nullptr, "targetMake: no specs");
    return nullptr;
  }
  static const char* kinds[] = {"disc", "ring", "dots", "checker", "lines", "zone", "raster"};
  std::vector<doe_target_spec> specs(count);
  for (uint32_t q = 0; q < count; ++q) {
    napi_value o;
    napi_get_element(env, a.v[0], q, &o);
    std::string k = str(env, prop(env, o, "kind"));
    if (k == "grid") k = "dots";
    int kind = -1;
    for (int t = 0; t < 7; ++t)
      if (k == kinds[t]) kind = t;
    if (kind < 0) {
      napi_throw_error(env, nullptr, ("makeTarget: unknown kind \"" + k + "\"").c_str());
      return nullptr;
    }
    doe_target_spec& S = specs[q];
    doe_target_spec_init(&S, kind);
    S.radius = opt_num(env, o, "radius", S.radius);
    S.inner = opt_num(env, o, "inner", S.inner);
    S.outer = opt_num(env, o, "outer", S.outer);
    S.cols = opt_num(env, o, "cols", S.cols);
    S.rows = opt_num(env, o, "rows", S.rows);
    S.sigma = opt_num(env, o, "sigma", S.sigma);
    S.cells = opt_num(env, o, "cells", S.cells);
    S.period = opt_num(env, o, "period", S.period);
    S.duty = opt_num(env, o, "duty", S.duty);
    S.gamma = opt_num(env, o, "gamma", S.gamma);
    S.background = opt_num(env, o, "background", S.background);
    S.gain = opt_num(env, o, "gain", S.gain);
    S.cover = boolean(env, prop(env, o, "cover"), false) ? 1 : 0;
    size_t nm = 0;
    const int32_t* m = i32(env, prop(env, o, "missing"), &nm);
    if (m && nm >= 2) {
      S.missing = m;
      S.n_missing = int(nm / 2);
    }
    if (kind == DOE_TGT_RASTER) {
      size_t len = 0;
      S.raster = f64(env, prop(env, o, "raster"), &len);
      S.raster_w = (int)opt_num(env, o, "width", 0);
      S.raster_h = (int)opt_num(env, o, "height", 0);
      if (S.raster && (long)len < (long)S.raster_w * S.raster_h) {
        napi_throw_range_error(env, nullptr, "makeTarget: raster shorter than width * height");
        return nullptr;
      }
    }
  }
  doe_ctx* c = ctx(env);
  if (!c) return nullptr;
  doe_field* f = nullptr;
  if (throw_status(env, doe_field_create(c, (int)num(env, a.v[1]), (int)num(env, a.v[2]), (int)count, DOE_F64, &f)))
    return nullptr;
  // optional 4th argument {dilate, soft, threshold}: the ROI in the same call
  int rc;
  if (a.n > 3 && present(env, a.v[3])) {
    rc = doe_target_make_roi(c, specs.data(), f, opt_num(env, a.v[3], "dilate", NAN),
                             opt_num(env, a.v[3], "soft", NAN), opt_num(env, a.v[3], "threshold", NAN));
  } else {
    rc = doe_target_make(c, specs.data(), f);
  }
  if (throw_status(env, rc)) {
    doe_field_free(f);
    return nullptr;
  }
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, f, finalize_field, nullptr, &ext));
  return ext;
}

// targetAutoRoi(field, dilate, soft, threshold): ROI of every field into im
napi_value TargetAutoRoi(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* f = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  if (!f || a.n < 4) {
    napi_throw_type_error(env, nullptr, "targetAutoRoi(field, dilate, soft, threshold)");
    return nullptr;
  }
  throw_status(env, doe_target_auto_roi(ctx(env), f, num(env, a.v[1], NAN), num(env, a.v[2], NAN),
                                        num(env, a.v[3], NAN)));
  return nullptr;
}

// rasterBlur(src, dst, sigma, taps|null)
napi_value RasterBlur(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* s = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  doe_field* d = a.n > 1 ? field_of(env, a.v[1]) : nullptr;
  if (!s || !d || a.n < 3) {
    napi_throw_type_error(env, nullptr, "rasterBlur(src, dst, sigma, taps)");
    return nullptr;
  }
  const double sigma = num(env, a.v[2], NAN);
  // the library reads 2r + 1 taps, r = max(1, ceil(3 sigma)) (doe.js:765)
  const size_t ntaps = sigma > 0 && std::isfinite(sigma) ? size_t(2 * std::max(1.0, std::ceil(3 * sigma)) + 1) : 0;
  const double* taps = a.n > 3 ? f64_need(env, a.v[3], ntaps, "rasterBlur: taps") : nullptr;
  if (pending(env)) return nullptr;
  throw_status(env, doe_raster_blur(ctx(env), s, d, sigma, taps));
  return nullptr;
}

// fieldImage(field, index, {quantity: 're'|'im'|'intensity'|'amplitude'|'phase',
//   gray, stops: Float64Array [pos, r, g, b]*, min, max, gamma, log, logFloor,
//   percentile, invert, pngRows}) -> {bytes: Uint8Array, lo, hi}
napi_value FieldImage(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* f = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  if (!f || a.n < 3) {
    napi_throw_type_error(env, nullptr, "fieldImage(field, index, opts)");
    return nullptr;
  }
  napi_value o = a.v[2];
  doe_image_opts io;
  doe_image_opts_init(&io);
  static const char* qs[] = {"re", "im", "intensity", "amplitude", "phase"};
  napi_value qv = prop(env, o, "quantity");
  if (present(env, qv)) {
    const std::string q = str(env, qv);
    int found = -1;
    for (int t = 0; t < 5; ++t)
      if (q == qs[t]) found = t;
    if (found < 0) {
      napi_throw_error(env, nullptr, ("fieldImage: unknown quantity \"" + q + "\"").c_str());
      return nullptr;
    }
    io.quantity = found;
  }
  io.gray = boolean(env, prop(env, o, "gray"), false) ? 1 : 0;
  size_t ns = 0;
  const double* st = f64(env, prop(env, o, "stops"), &ns);
  if (st) {
    if (ns % 4 != 0 || ns / 4 < 1 || ns / 4 > 8) {
      napi_throw_range_error(env, nullptr, "fieldImage: stops hold 1 to 8 (pos, r, g, b) entries");
      return nullptr;
    }
    io.nstops = int(ns / 4);
    for (int q = 0; q < io.nstops; ++q) {
      io.stop_pos[q] = st[4 * q];
      for (int c = 0; c < 3; ++c) io.stop_rgb[q][c] = st[4 * q + 1 + c];
    }
  }
  napi_value v = prop(env, o, "min");
  if (present(env, v)) {
    io.has_min = 1;
    io.min = num(env, v, NAN);
  }
  v = prop(env, o, "max");
  if (present(env, v)) {
    io.has_max = 1;
    io.max = num(env, v, NAN);
  }
  io.gamma = opt_num(env, o, "gamma", 1);
  io.log = boolean(env, prop(env, o, "log"), false) ? 1 : 0;
  io.log_floor = opt_num(env, o, "logFloor", 1e-4);
  v = prop(env, o, "percentile");
  if (present(env, v)) {
    io.has_percentile = 1;
    io.percentile = num(env, v, NAN);
  }
  io.invert = boolean(env, prop(env, o, "invert"), false) ? 1 : 0;
  io.png_rows = boolean(env, prop(env, o, "pngRows"), false) ? 1 : 0;
  int nx = 0, ny = 0;
  doe_field_info(f, &nx, &ny, nullptr, nullptr, nullptr, nullptr);
  const size_t need = size_t(ny) * ((io.png_rows ? 1 : 0) + size_t(nx) * (io.gray ? 1 : 3));
  napi_value ab, arr;
  void* p = nullptr;
  CHECK_NAPI(napi_create_arraybuffer(env, need, &p, &ab));
  CHECK_NAPI(napi_create_typedarray(env, napi_uint8_array, need, ab, 0, &arr));
  double range[2] = {0, 0};
  if (throw_status(env, doe_field_image(ctx(env), f, (int)num(env, a.v[1], 0), &io, (unsigned char*)p,
                                        (long)need, range)))
    return nullptr;
  napi_value obj;
  CHECK_NAPI(napi_create_object(env, &obj));
  set(env, obj, "bytes", arr);
  set(env, obj, "lo", make_num(env, range[0]));
  set(env, obj, "hi", make_num(env, range[1]));
  return obj;
}

// errorDiffusion(src, dst, levels, snake, wrapFirst): re planes, FP64
napi_value ErrorDiffusion(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* s = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  doe_field* d = a.n > 1 ? field_of(env, a.v[1]) : nullptr;
  if (!s || !d || a.n < 5) {
    napi_throw_type_error(env, nullptr, "errorDiffusion(src, dst, levels, snake, wrapFirst)");
    return nullptr;
  }
  throw_status(env, doe_error_diffusion(ctx(env), s, d, (int)num(env, a.v[2], 4),
                                        boolean(env, a.v[3], true) ? 1 : 0, boolean(env, a.v[4], true) ? 1 : 0));
  return nullptr;
}

// randomPhase(dst, seed): re plane, FP64 or FP32; seed already ToUint32
napi_value RandomPhase(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* d = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  if (!d || a.n < 2) {
    napi_throw_type_error(env, nullptr, "randomPhase(dst, seed)");
    return nullptr;
  }
  uint32_t seed = 0;
  CHECK_NAPI(napi_get_value_uint32(env, a.v[1], &seed));
  throw_status(env, doe_random_phase(ctx(env), d, seed));
  return nullptr;
}

// phaseToHeight(src, dst, lambda, n): re planes, FP64
napi_value PhaseToHeight(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  doe_field* s = a.n > 0 ? field_of(env, a.v[0]) : nullptr;
  doe_field* d = a.n > 1 ? field_of(env, a.v[1]) : nullptr;
  if (!s || !d || a.n < 4) {
    napi_throw_type_error(env, nullptr, "phaseToHeight(src, dst, lambda, n)");
    return nullptr;
  }
  throw_status(env, doe_phase_to_height(ctx(env), s, d, num(env, a.v[2], 0), num(env, a.v[3], 1.52)));
  return nullptr;
}

napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"version", nullptr, Version, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"deviceCount", nullptr, DeviceCount, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fft2d", nullptr, Fft2d, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fft1d", nullptr, Fft1d, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldCreate", nullptr, FieldCreate, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldUpload", nullptr, FieldUpload, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldDownload", nullptr, FieldDownload, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"asm", nullptr, Asm, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"asmTransfer", nullptr, AsmTransfer, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fresnel", nullptr, Fresnel, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"lens", nullptr, Lens, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"gsRun", nullptr, GsRun, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"screenMetrics", nullptr, ScreenMetrics, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"screenMetricsBatch", nullptr, ScreenMetricsBatch, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldResample", nullptr, FieldResample, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldSampleImage", nullptr, FieldSampleImage, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldCopySlices", nullptr, FieldCopySlices, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"gsMultiPlane", nullptr, GsMultiPlane, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"targetMake", nullptr, TargetMake, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"targetAutoRoi", nullptr, TargetAutoRoi, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"rasterBlur", nullptr, RasterBlur, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"fieldImage", nullptr, FieldImage, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"errorDiffusion", nullptr, ErrorDiffusion, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"phaseToHeight", nullptr, PhaseToHeight, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"randomPhase", nullptr, RandomPhase, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof(d) / sizeof(d[0]), d);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
