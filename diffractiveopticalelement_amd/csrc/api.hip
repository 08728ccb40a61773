// C-ABI entry points: context, fields, conversions, fft2d.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "rng.hpp"

namespace doe {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// exp(-2*pi*i*m/N) for m in [0, N): evaluated per octant in long double so
// that symmetric entries are exact negations/swaps of each other.
static void twiddle_host(int n, std::vector<long double>& c, std::vector<long double>& s) {
  c.assign(n, 0.0L);
  s.assign(n, 0.0L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    // angle = 2*pi*m/n; reduce to the first octant of [0, 2*pi)
    int q = (int)((8L * m) / n);  // octant 0..7
    long double a;                // angle within the octant
    long double cc, ss;
    // r = m/n in [0,1); work with exact multiples of 1/(8n)
    const long num = 8L * m - (long)q * n;  // in [0, n)
    a = (pi / 4.0L) * (long double)num / (long double)n;
    const long double ca = cosl(a), sa = sinl(a);
    const long double h = 0.70710678118654752440084436210484903928L;
    switch (q) {
      case 0: cc = ca; ss = sa; break;
      case 1: cc = h * (ca - sa); ss = h * (ca + sa); break;
      case 2: cc = -sa; ss = ca; break;
      case 3: cc = -h * (ca + sa); ss = h * (ca - sa); break;
      case 4: cc = -ca; ss = -sa; break;
      case 5: cc = -h * (ca - sa); ss = -h * (ca + sa); break;
      case 6: cc = sa; ss = -ca; break;
      default: cc = h * (ca + sa); ss = -h * (ca - sa); break;
    }
    c[m] = cc;
    s[m] = -ss;  // forward kernel exp(-i*angle)
  }
  // exact values at the quarter points
  c[0] = 1.0L; s[0] = 0.0L;
  if (n % 4 == 0) { c[n / 4] = 0.0L; s[n / 4] = -1.0L; c[3 * n / 4] = 0.0L; s[3 * n / 4] = 1.0L; }
  if (n % 2 == 0 && n >= 2) { c[n / 2] = -1.0L; s[n / 2] = 0.0L; }
}

int get_twiddle(doe_ctx* ctx, int n, int tsize, Twiddle* out) {
  auto key = std::make_pair(n, tsize);
  auto it = ctx->tw.find(key);
  if (it != ctx->tw.end()) {
    *out = it->second;
    return DOE_OK;
  }
  std::vector<long double> c, s;
  twiddle_host(n, c, s);
  Twiddle t;
  const size_t bytes = size_t(n) * tsize;
  DOE_HIP(hipMalloc(&t.re, bytes));
  DOE_HIP(hipMalloc(&t.im, bytes));
  if (tsize == 8) {
    std::vector<double> a(n), b(n);
    for (int m = 0; m < n; ++m) { a[m] = (double)c[m]; b[m] = (double)s[m]; }
    DOE_HIP(hipMemcpy(t.re, a.data(), bytes, hipMemcpyHostToDevice));
    DOE_HIP(hipMemcpy(t.im, b.data(), bytes, hipMemcpyHostToDevice));
  } else {
    std::vector<float> a(n), b(n);
    for (int m = 0; m < n; ++m) { a[m] = (float)c[m]; b[m] = (float)s[m]; }
    DOE_HIP(hipMemcpy(t.re, a.data(), bytes, hipMemcpyHostToDevice));
    DOE_HIP(hipMemcpy(t.im, b.data(), bytes, hipMemcpyHostToDevice));
  }
  ctx->tw[key] = t;
  *out = t;
  return DOE_OK;
}

int make_plan2(doe_ctx* ctx, int nx, int ny, int tsize, int n1, Plan2* p) {
  p->nx = nx;
  p->ny = ny;
  p->n1 = n1;
  p->n2 = ny / n1;
  int rc = get_twiddle(ctx, nx, tsize, &p->tx);
  if (!rc) rc = get_twiddle(ctx, ny, tsize, &p->ty);
  if (!rc) rc = get_twiddle(ctx, p->n2, tsize, &p->tc);
  return rc;
}

hipEvent_t prof_event(doe_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

int upload_params(doe_ctx* ctx, const void* host, size_t bytes, void** dev) {
  auto holds = [&](const doe_ctx::ParamSlot& sl) {
    return sl.host.size() == bytes && std::memcmp(sl.host.data(), host, bytes) == 0;
  };
  // the block handed out last, unchanged: no transfer
  if (ctx->pcur >= 0 && holds(ctx->pslot[ctx->pcur])) {
    *dev = ctx->pslot[ctx->pcur].dev;
    return DOE_OK;
  }
  // the current slot is free again once the work queued so far has ended
  if (ctx->pcur >= 0) {
    doe_ctx::ParamSlot& old = ctx->pslot[ctx->pcur];
    if (!old.done) DOE_HIP(hipEventCreateWithFlags(&old.done, hipEventDisableTiming));
    DOE_HIP(hipEventRecord(old.done, ctx->stream));
    old.pending = true;
  }
  // a block still held by another slot (calls that alternate between a few
  // parameter sets): read it there, no transfer
  for (int k = 0; k < doe_ctx::kParamSlots; ++k) {
    if (k != ctx->pcur && ctx->pslot[k].dev && holds(ctx->pslot[k])) {
      ctx->pcur = k;
      *dev = ctx->pslot[k].dev;
      return DOE_OK;
    }
  }
  int k = ctx->pnext;
  if (k == ctx->pcur) k = (k + 1) % doe_ctx::kParamSlots;
  ctx->pnext = (k + 1) % doe_ctx::kParamSlots;
  doe_ctx::ParamSlot& sl = ctx->pslot[k];
  if (sl.pending) {
    DOE_HIP(hipEventSynchronize(sl.done));  // kernels that read this slot have ended
    sl.pending = false;
  }
  ctx->pcur = -1;  // nothing is handed out until the copy below is queued
  sl.host.clear();
  if (bytes > sl.cap) {
    if (sl.dev) DOE_HIP(hipFree(sl.dev));
    if (sl.pin) DOE_HIP(hipHostFree(sl.pin));
    sl.dev = sl.pin = nullptr;
    sl.cap = 0;
    size_t cap = 4096;
    while (cap < bytes) cap <<= 1;
    DOE_HIP(hipMalloc(&sl.dev, cap));
    if (cap <= doe_ctx::kParamPinMax) DOE_HIP(hipHostMalloc(&sl.pin, cap, hipHostMallocDefault));
    sl.cap = cap;
  }
  if (sl.pin) {
    // pinned staging: the copy is queued in stream order and the host call
    // returns at once; the staging is rewritten only after the slot's event
    std::memcpy(sl.pin, host, bytes);
    DOE_HIP(hipMemcpyAsync(sl.dev, sl.pin, bytes, hipMemcpyHostToDevice, ctx->stream));
  } else {
    // large blocks (rasters, images): from the caller's memory, consumed
    // before return
    DOE_HIP(hipMemcpyAsync(sl.dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    DOE_HIP(hipStreamSynchronize(ctx->stream));
  }
  sl.host.assign((const char*)host, (const char*)host + bytes);
  ctx->pcur = k;
  *dev = sl.dev;
  return DOE_OK;
}

int ensure_reduce(doe_ctx* ctx, size_t n) {
  if (n > ctx->partials_cap) {
    DOE_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->partials) DOE_HIP(hipFree(ctx->partials));
    size_t cap = 1024;
    while (cap < n) cap <<= 1;
    DOE_HIP(hipMalloc((void**)&ctx->partials, cap * sizeof(double)));
    ctx->partials_cap = cap;
  }
  return DOE_OK;
}

int ensure_work(doe_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->work_cap) return DOE_OK;
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->work_re) DOE_HIP(hipFree(ctx->work_re));
  if (ctx->work_im) DOE_HIP(hipFree(ctx->work_im));
  ctx->work_re = ctx->work_im = nullptr;
  ctx->work_cap = 0;
  if (hipMalloc(&ctx->work_re, bytes) != hipSuccess || hipMalloc(&ctx->work_im, bytes) != hipSuccess) {
    (void)hipGetLastError();
    set_error("work buffer allocation failed");
    return DOE_ERR_OOM;
  }
  ctx->work_cap = bytes;
  return DOE_OK;
}

int ensure_stage(doe_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->stage_cap) return DOE_OK;
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->stage) DOE_HIP(hipFree(ctx->stage));
  size_t cap = 1 << 20;
  while (cap < bytes) cap <<= 1;
  DOE_HIP(hipMalloc(&ctx->stage, cap));
  ctx->stage_cap = cap;
  return DOE_OK;
}

template <typename S>
__global__ void k_from_f64(const double* __restrict__ a, S* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) st<double>(b + i, a[i]);
}
template <typename S>
__global__ void k_to_f64(const S* __restrict__ a, double* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = ld<double>(a + i);
}
template <typename S>
__global__ void k_copy(const S* __restrict__ a, S* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Uniform noise in [-amp, amp) from a hash of the index: the autotuner's
// scratch data.  Zero-filled planes run 5-10 % faster than any real field
// (fewer toggling bits, higher sustained clock), which misranks variants.
template <typename S>
__global__ void k_fill_noise(S* __restrict__ p, long n, unsigned seed, double amp) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = unsigned(i) * 0x9E3779B1u ^ (unsigned(i >> 32) + seed) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  st<double>(p + i, amp * (double(h) * (2.0 / 4294967296.0) - 1.0));
}

template <class F>
static int with_dtype(int dt, F&& f) {
  switch (dt) {
    case DOE_F32: return f(float{});
    case DOE_F64: return f(double{});
    case DOE_F16: return f(__half{});
    case DOE_BF16: return f(__hip_bfloat16{});
    default: set_error("unknown dtype"); return DOE_ERR_INVALID;
  }
}

static int convert_in(doe_ctx* ctx, const double* host, void* dst, long n, int dt) {
  const size_t bytes = size_t(n) * sizeof(double);
  if (dt == DOE_F64) {
    DOE_HIP(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    DOE_HIP(hipStreamSynchronize(ctx->stream));
    return DOE_OK;
  }
  int rc = ensure_stage(ctx, bytes);
  if (rc) return rc;
  DOE_HIP(hipMemcpyAsync(ctx->stage, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  const unsigned grid = unsigned((n + 255) / 256);
  return with_dtype(dt, [&](auto tag) -> int {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_from_f64<S>, dim3(grid), dim3(256), 0, ctx->stream, (const double*)ctx->stage,
                       (S*)dst, n);
    DOE_HIP(hipGetLastError());
    DOE_HIP(hipStreamSynchronize(ctx->stream));
    return DOE_OK;
  });
}

static int convert_out(doe_ctx* ctx, const void* src, double* host, long n, int dt) {
  const size_t bytes = size_t(n) * sizeof(double);
  if (dt == DOE_F64) {
    DOE_HIP(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    DOE_HIP(hipStreamSynchronize(ctx->stream));
    return DOE_OK;
  }
  int rc = ensure_stage(ctx, bytes);
  if (rc) return rc;
  const unsigned grid = unsigned((n + 255) / 256);
  rc = with_dtype(dt, [&](auto tag) -> int {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_to_f64<S>, dim3(grid), dim3(256), 0, ctx->stream, (const S*)src,
                       (double*)ctx->stage, n);
    DOE_HIP(hipGetLastError());
    return DOE_OK;
  });
  if (rc) return rc;
  DOE_HIP(hipMemcpyAsync(host, ctx->stage, bytes, hipMemcpyDeviceToHost, ctx->stream));
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  return DOE_OK;
}

// Storage: any grid of at most 65536 rows and 2^32 points per field (a row
// may be long: a host plane travels as one row).  Transforms and propagators
// additionally need power-of-two sides (check_field).
int check_storage(const doe_ctx* ctx, const doe_field* f) {
  DOE_REQUIRE(ctx != nullptr, DOE_ERR_INVALID, "null context");
  DOE_REQUIRE(f != nullptr, DOE_ERR_INVALID, "null field");
  DOE_REQUIRE(f->nx >= 1 && f->ny >= 1 && f->ny <= 65536 && (long)f->nx * f->ny <= (1L << 32), DOE_ERR_SHAPE,
              "grids need 1..65536 rows and at most 2^32 points");
  return DOE_OK;
}

int check_field(const doe_ctx* ctx, const doe_field* f) {
  DOE_REQUIRE(ctx != nullptr, DOE_ERR_INVALID, "null context");
  DOE_REQUIRE(f != nullptr, DOE_ERR_INVALID, "null field");
  DOE_REQUIRE(is_pow2(f->nx) && is_pow2(f->ny) && f->nx >= 2 && f->ny >= 2 && f->nx <= kMaxN &&
                  f->ny <= kMaxN,
              DOE_ERR_UNSUPPORTED, "grid sides must be powers of two in [2, 8192]");
  return DOE_OK;
}

int dispatch_fft2d(doe_ctx* ctx, doe_field* f, int inverse) {
  switch (f->dtype) {
    case DOE_F32: return run_fft2d<float, float>(ctx, f, inverse);
    case DOE_F64: return run_fft2d<double, double>(ctx, f, inverse);
    case DOE_F16: return run_fft2d<float, __half>(ctx, f, inverse);
    case DOE_BF16: return run_fft2d<float, __hip_bfloat16>(ctx, f, inverse);
  }
  set_error("unknown dtype");
  return DOE_ERR_INVALID;
}

int dispatch_fft_rows(doe_ctx* ctx, doe_field* f, int inverse) {
  switch (f->dtype) {
    case DOE_F32: return run_fft_rows<float, float>(ctx, f, inverse);
    case DOE_F64: return run_fft_rows<double, double>(ctx, f, inverse);
    case DOE_F16: return run_fft_rows<float, __half>(ctx, f, inverse);
    case DOE_BF16: return run_fft_rows<float, __hip_bfloat16>(ctx, f, inverse);
  }
  set_error("unknown dtype");
  return DOE_ERR_INVALID;
}

}  // namespace doe

using namespace doe;

extern "C" {

int doe_version(void) { return 100; }

const char* doe_last_error(void) { return g_last_error.c_str(); }

const char* doe_status_string(int s) {
  switch (s) {
    case DOE_OK: return "ok";
    case DOE_ERR_INVALID: return "invalid argument";
    case DOE_ERR_HIP: return "HIP runtime error";
    case DOE_ERR_OOM: return "out of device memory";
    case DOE_ERR_UNSUPPORTED: return "unsupported size or dtype";
    case DOE_ERR_SHAPE: return "plane length does not match grid";
    case DOE_ERR_STATE: return "invalid call order";
    default: return "unknown status";
  }
}

int doe_device_count(int* count) {
  DOE_REQUIRE(count, DOE_ERR_INVALID, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return DOE_OK;
}

int doe_ctx_create(int device, void* stream, doe_ctx** out) {
  DOE_REQUIRE(out, DOE_ERR_INVALID, "null out");
  *out = nullptr;
  int c = 0;
  DOE_HIP(hipGetDeviceCount(&c));
  DOE_REQUIRE(device >= 0 && device < c, DOE_ERR_INVALID, "no such HIP device");
  DOE_HIP(hipSetDevice(device));
  {
    const int rc = metrics_scratch_init(device);  // the no-allocation entry point's scratch
    if (rc) return rc;
  }
  doe_ctx* ctx = new doe_ctx();
  ctx->device = device;
  if (stream) {
    ctx->stream = (hipStream_t)stream;
    ctx->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
      return DOE_ERR_HIP;
    }
    ctx->own_stream = true;
  }
  if (const char* env = std::getenv("DOE_CHUNK_MIB")) ctx->chunk_bytes = std::atol(env) << 20;
  *out = ctx;
  return DOE_OK;
}

int doe_ctx_set_chunk_bytes(doe_ctx* ctx, long bytes) {
  DOE_REQUIRE(ctx, DOE_ERR_INVALID, "null context");
  DOE_REQUIRE(bytes >= 0, DOE_ERR_INVALID, "chunk size must be >= 0");
  ctx->chunk_bytes = bytes;
  return DOE_OK;
}

int doe_ctx_get_chunk_bytes(doe_ctx* ctx, long* bytes) {
  DOE_REQUIRE(ctx && bytes, DOE_ERR_INVALID, "null argument");
  *bytes = ctx->chunk_bytes;
  return DOE_OK;
}

int doe_ctx_destroy(doe_ctx* ctx) {
  if (!ctx) return DOE_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->tw) {
    hipFree(kv.second.re);
    hipFree(kv.second.im);
  }
  for (auto& sl : ctx->pslot) {
    if (sl.dev) hipFree(sl.dev);
    if (sl.pin) hipHostFree(sl.pin);
    if (sl.done) hipEventDestroy(sl.done);
  }
  if (ctx->stage) hipFree(ctx->stage);
  if (ctx->partials) hipFree(ctx->partials);
  if (ctx->work_re) hipFree(ctx->work_re);
  if (ctx->work_im) hipFree(ctx->work_im);
  for (auto& r : ctx->prof) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto e : ctx->ev_pool) hipEventDestroy(e);
  if (ctx->t0) hipEventDestroy(ctx->t0);
  if (ctx->t1) hipEventDestroy(ctx->t1);
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return DOE_OK;
}

int doe_ctx_sync(doe_ctx* ctx) {
  DOE_REQUIRE(ctx, DOE_ERR_INVALID, "null context");
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  return DOE_OK;
}

int doe_ctx_stream(doe_ctx* ctx, void** stream) {
  DOE_REQUIRE(ctx && stream, DOE_ERR_INVALID, "null argument");
  *stream = (void*)ctx->stream;
  return DOE_OK;
}

int doe_ctx_set_flags(doe_ctx* ctx, int flags) {
  DOE_REQUIRE(ctx, DOE_ERR_INVALID, "null context");
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  ctx->flags = flags;
  return DOE_OK;
}

int doe_ctx_profile_report(doe_ctx* ctx, char* buf, long cap) {
  DOE_REQUIRE(ctx && buf && cap > 0, DOE_ERR_INVALID, "null argument");
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  struct Agg {
    long n = 0;
    double ms = 0;
    long pts = 0;
  };
  std::map<std::string, Agg> agg;
  std::vector<std::string> order;
  for (auto& r : ctx->prof) {
    float ms = 0.f;
    DOE_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    auto it = agg.find(r.tag);
    if (it == agg.end()) {
      order.push_back(r.tag);
      agg[r.tag] = Agg{1, ms, r.pts};
    } else {
      it->second.n += 1;
      it->second.ms += ms;
      it->second.pts += r.pts;
    }
    ctx->ev_pool.push_back(r.a);
    ctx->ev_pool.push_back(r.b);
  }
  ctx->prof.clear();
  std::string out;
  for (auto& k : order) {
    char line[256];
    snprintf(line, sizeof line, "%s %ld %.6f %ld\n", k.c_str(), agg[k].n, agg[k].ms, agg[k].pts);
    out += line;
  }
  DOE_REQUIRE((long)out.size() < cap, DOE_ERR_INVALID, "report buffer too small");
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return DOE_OK;
}

int doe_timer_start(doe_ctx* ctx) {
  DOE_REQUIRE(ctx, DOE_ERR_INVALID, "null context");
  if (!ctx->t0) DOE_HIP(hipEventCreate(&ctx->t0));
  if (!ctx->t1) DOE_HIP(hipEventCreate(&ctx->t1));
  DOE_HIP(hipEventRecord(ctx->t0, ctx->stream));
  return DOE_OK;
}

int doe_timer_stop(doe_ctx* ctx, double* ms) {
  DOE_REQUIRE(ctx && ms && ctx->t0 && ctx->t1, DOE_ERR_STATE, "doe_timer_stop without start");
  DOE_HIP(hipEventRecord(ctx->t1, ctx->stream));
  DOE_HIP(hipEventSynchronize(ctx->t1));
  float f = 0.f;
  DOE_HIP(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
  *ms = f;
  return DOE_OK;
}

int doe_field_create(doe_ctx* ctx, int nx, int ny, int batch, int dtype, doe_field** out) {
  DOE_REQUIRE(ctx && out, DOE_ERR_INVALID, "null argument");
  *out = nullptr;
  DOE_REQUIRE(dtype_size(dtype) > 0, DOE_ERR_INVALID, "unknown dtype");
  DOE_REQUIRE(batch >= 1, DOE_ERR_INVALID, "batch must be >= 1");
  doe_field* f = new doe_field();
  f->nx = nx;
  f->ny = ny;
  f->batch = batch;
  f->dtype = dtype;
  int rc = check_storage(ctx, f);
  if (rc) {
    delete f;
    return rc;
  }
  const size_t bytes = f->count() * dtype_size(dtype);
  if (hipMalloc(&f->re, bytes) != hipSuccess || hipMalloc(&f->im, bytes) != hipSuccess) {
    if (f->re) hipFree(f->re);
    delete f;
    (void)hipGetLastError();
    set_error("doe_field_create: device allocation failed");
    return DOE_ERR_OOM;
  }
  f->owned = true;
  *out = f;
  return DOE_OK;
}

int doe_field_wrap(doe_ctx* ctx, void* re, void* im, int nx, int ny, int batch, int dtype,
                   doe_field** out) {
  DOE_REQUIRE(ctx && out && re && im, DOE_ERR_INVALID, "null argument");
  *out = nullptr;
  DOE_REQUIRE(dtype_size(dtype) > 0, DOE_ERR_INVALID, "unknown dtype");
  DOE_REQUIRE(batch >= 1, DOE_ERR_INVALID, "batch must be >= 1");
  doe_field* f = new doe_field();
  f->nx = nx;
  f->ny = ny;
  f->batch = batch;
  f->dtype = dtype;
  f->re = re;
  f->im = im;
  f->owned = false;
  int rc = check_storage(ctx, f);
  if (rc) {
    delete f;
    return rc;
  }
  *out = f;
  return DOE_OK;
}

int doe_field_free(doe_field* f) {
  if (!f) return DOE_OK;
  if (f->owned) {
    hipFree(f->re);
    hipFree(f->im);
  }
  delete f;
  return DOE_OK;
}

int doe_field_info(const doe_field* f, int* nx, int* ny, int* batch, int* dtype, void** re,
                   void** im) {
  DOE_REQUIRE(f, DOE_ERR_INVALID, "null field");
  if (nx) *nx = f->nx;
  if (ny) *ny = f->ny;
  if (batch) *batch = f->batch;
  if (dtype) *dtype = f->dtype;
  if (re) *re = f->re;
  if (im) *im = f->im;
  return DOE_OK;
}

int doe_field_upload(doe_ctx* ctx, doe_field* f, const double* re, const double* im) {
  int rc = check_storage(ctx, f);
  if (rc) return rc;
  DOE_REQUIRE(re && im, DOE_ERR_INVALID, "null host plane");
  DOE_HIP(hipSetDevice(ctx->device));
  rc = convert_in(ctx, re, f->re, (long)f->count(), f->dtype);
  if (rc) return rc;
  return convert_in(ctx, im, f->im, (long)f->count(), f->dtype);
}

int doe_field_download(doe_ctx* ctx, doe_field* f, double* re, double* im) {
  int rc = check_storage(ctx, f);
  if (rc) return rc;
  DOE_REQUIRE(re && im, DOE_ERR_INVALID, "null host plane");
  DOE_HIP(hipSetDevice(ctx->device));
  rc = convert_out(ctx, f->re, re, (long)f->count(), f->dtype);
  if (rc) return rc;
  return convert_out(ctx, f->im, im, (long)f->count(), f->dtype);
}

int doe_field_copy(doe_ctx* ctx, const doe_field* src, doe_field* dst) {
  int rc = check_storage(ctx, src);
  if (rc) return rc;
  rc = check_storage(ctx, dst);
  if (rc) return rc;
  DOE_REQUIRE(src->nx == dst->nx && src->ny == dst->ny && src->batch == dst->batch &&
                  src->dtype == dst->dtype,
              DOE_ERR_SHAPE, "doe_field_copy: shape mismatch");
  const size_t bytes = src->count() * dtype_size(src->dtype);
  DOE_HIP(hipMemcpyAsync(dst->re, src->re, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  DOE_HIP(hipMemcpyAsync(dst->im, src->im, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  return DOE_OK;
}

int doe_random_phase(doe_ctx* ctx, doe_field* out, unsigned seed) {
  int rc = check_storage(ctx, out);
  if (rc) return rc;
  DOE_REQUIRE(out->dtype == DOE_F64 || out->dtype == DOE_F32, DOE_ERR_INVALID,
              "randomPhase: FP64 or FP32 fields");
  const long n = (long)out->count();
  DOE_REQUIRE(n <= (1L << kRngLevels), DOE_ERR_SHAPE, "randomPhase: at most 2^40 samples per call");
  DOE_HIP(hipSetDevice(ctx->device));
  if (out->dtype == DOE_F64)
    DOE_HIP(launch_random_phase(ctx, static_cast<double*>(out->re), n, seed));
  else
    DOE_HIP(launch_random_phase(ctx, static_cast<float*>(out->re), n, seed));
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_fft2d(doe_ctx* ctx, doe_field* f, int inverse) {
  int rc = check_field(ctx, f);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(ctx->device));
  rc = dispatch_fft2d(ctx, f, inverse);
  if (rc) return rc;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_tune_set(doe_ctx* ctx, int nx, int ny, int batch, int dtype, int n1, int colvar,
                 int rowplanes, long chunk_bytes) {
  DOE_REQUIRE(ctx, DOE_ERR_INVALID, "null context");
  DOE_REQUIRE(n1 == 0 || n1 == 1 || n1 == 2 || n1 == 4, DOE_ERR_INVALID, "tune: n1 must be 0, 1, 2 or 4");
  DOE_REQUIRE(colvar >= 0 && colvar <= 5, DOE_ERR_INVALID, "tune: colvar must be 0..5");
  DOE_REQUIRE(dtype != DOE_F64 || colvar <= 1, DOE_ERR_INVALID,
              "tune: colvar of FP64 fields must be 0 or 1");
  DOE_REQUIRE(rowplanes >= 0 && rowplanes <= 2, DOE_ERR_INVALID,
              "tune: rowplanes must be 0 (built-in), 1 or 2");
  doe_tune t;
  t.n1 = n1;
  t.colvar = colvar;
  t.rowplanes = rowplanes;
  t.chunk = chunk_bytes;
  ctx->tune[std::make_tuple(nx, ny, batch, dtype)] = t;
  return DOE_OK;
}

int doe_tune_get(doe_ctx* ctx, int nx, int ny, int batch, int dtype, int* n1, int* colvar,
                 int* rowplanes, long* chunk_bytes) {
  DOE_REQUIRE(ctx, DOE_ERR_INVALID, "null context");
  const doe_tune t = tune_for(ctx, nx, ny, batch, dtype);
  if (n1) *n1 = t.n1;
  if (colvar) *colvar = t.colvar;
  if (rowplanes) *rowplanes = t.rowplanes;
  if (chunk_bytes) *chunk_bytes = t.chunk >= 0 ? t.chunk : ctx->chunk_bytes;
  return DOE_OK;
}

int doe_autotune_fft2d(doe_ctx* ctx, int nx, int ny, int batch, int dtype, int reps, char* report,
                       long cap) {
  DOE_REQUIRE(ctx, DOE_ERR_INVALID, "null context");
  DOE_REQUIRE(reps > 0, DOE_ERR_INVALID, "autotune: reps must be positive");
  doe_field* f = nullptr;
  int rc = doe_field_create(ctx, nx, ny, batch, dtype, &f);
  if (rc) return rc;
  rc = check_field(ctx, f);
  if (rc) {
    doe_field_free(f);
    return rc;
  }
  // noise of a size that no variant overflows in half-width storage
  // (forward and inverse alternate)
  const long cnt = long(f->count());
  rc = with_dtype(dtype, [&](auto tag) -> int {
    using S = decltype(tag);
    const unsigned grid = unsigned((cnt + 255) / 256);
    hipLaunchKernelGGL(k_fill_noise<S>, dim3(grid), dim3(256), 0, ctx->stream, (S*)f->re, cnt,
                       1u, 0.01);
    hipLaunchKernelGGL(k_fill_noise<S>, dim3(grid), dim3(256), 0, ctx->stream, (S*)f->im, cnt,
                       2u, 0.01);
    DOE_HIP(hipGetLastError());
    return DOE_OK;
  });
  if (rc) {
    doe_field_free(f);
    return rc;
  }
  const auto key = std::make_tuple(nx, ny, batch, dtype);
  const bool had = ctx->tune.count(key) > 0;
  const doe_tune saved_tune = had ? ctx->tune[key] : doe_tune{};
  const int saved_flags = ctx->flags;
  ctx->flags = DOE_CTX_ASYNC;
  const bool f32 = dtype != DOE_F64;  // FP32 arithmetic: every column variant
  const long fb = long(nx) * ny * 2 * long(dtype_size(dtype));
  std::vector<long> chunks = {0};
  if (batch > 1 && fb * batch > (128L << 20)) chunks.push_back(128L << 20);
  std::string rep;
  double best_ms = 1e30;
  doe_tune best{};
  // bring the clock up first: after an idle spell the first milliseconds of
  // work run slower and would rank the first candidates low
  if (had) ctx->tune.erase(key);
  for (int k = 0; k < 16; ++k) {
    if (dispatch_fft2d(ctx, f, 0) || dispatch_fft2d(ctx, f, 1)) break;
  }
  (void)hipGetLastError();
  for (int n1 : {1, 2, 4}) {
    if (!split_fits(nx, ny, n1, f32 ? 4 : 8)) continue;
    for (int rp : {2, 1}) {
      for (int cv = 0; cv <= (f32 ? 5 : 1); ++cv) {
        for (long ch : chunks) {
          doe_tune t;
          t.n1 = n1;
          t.colvar = cv;
          t.rowplanes = rp;
          t.chunk = ch;
          ctx->tune[key] = t;
          // one forward + inverse pair untimed (plans, work planes), then `reps`
          int r = dispatch_fft2d(ctx, f, 0);
          if (!r) r = dispatch_fft2d(ctx, f, 1);
          if (r || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            (void)hipGetLastError();
            continue;  // a variant that does not exist for this shape
          }
          // the better of two timed windows of `reps` pairs
          double ms = 1e30;
          for (int w = 0; !r && w < 2; ++w) {
            double wms = 0;
            r = doe_timer_start(ctx);
            for (int k = 0; !r && k < reps; ++k) {
              r = dispatch_fft2d(ctx, f, 0);
              if (!r) r = dispatch_fft2d(ctx, f, 1);
            }
            if (!r) r = doe_timer_stop(ctx, &wms);
            if (!r) ms = std::min(ms, wms);
          }
          if (r) continue;
          ms /= 2.0 * reps;  // per transform
          char line[160];
          snprintf(line, sizeof line, "n1=%d colvar=%d rowplanes=%d chunk=%ld ms=%.4f\n", n1, cv,
                   rp, ch, ms);
          rep += line;
          if (ms < best_ms) {
            best_ms = ms;
            best = t;
          }
        }
      }
    }
  }
  ctx->flags = saved_flags;
  doe_field_free(f);
  if (best_ms >= 1e30) {
    if (had) ctx->tune[key] = saved_tune;
    else ctx->tune.erase(key);
    set_error("autotune: no configuration ran");
    return DOE_ERR_UNSUPPORTED;
  }
  ctx->tune[key] = best;
  char line[160];
  snprintf(line, sizeof line, "best n1=%d colvar=%d rowplanes=%d chunk=%ld ms=%.4f\n", best.n1,
           best.colvar, best.rowplanes, best.chunk, best_ms);
  rep += line;
  if (report && cap > 0) {
    const size_t n = std::min(rep.size(), size_t(cap - 1));
    std::memcpy(report, rep.data(), n);
    report[n] = 0;
  }
  return DOE_OK;
}

int doe_fft2d_host(doe_ctx* ctx, double* re, double* im, long len, int nx, int ny, int inverse,
                   int dtype) {
  DOE_REQUIRE(ctx && re && im, DOE_ERR_INVALID, "null argument");
  if (len != (long)nx * (long)ny) {
    set_error("fft2d: plane length " + std::to_string(len) + " does not match " +
              std::to_string(nx) + "x" + std::to_string(ny));
    return DOE_ERR_SHAPE;
  }
  doe_field* f = nullptr;
  int rc = doe_field_create(ctx, nx, ny, 1, dtype, &f);
  if (rc) return rc;
  rc = doe_field_upload(ctx, f, re, im);
  if (!rc) rc = doe_fft2d(ctx, f, inverse);
  if (!rc) rc = doe_field_download(ctx, f, re, im);
  doe_field_free(f);
  return rc;
}

int doe_fft1d_host(doe_ctx* ctx, double* re, double* im, long len, long n, long stride,
                   long offset, int inverse, int dtype) {
  DOE_REQUIRE(ctx && re && im, DOE_ERR_INVALID, "null argument");
  if (n <= 1) return DOE_OK;  // ref: fft.js:87
  if (!is_pow2(n)) {
    set_error("fft1d: length must be a power of two, got " + std::to_string(n));
    return DOE_ERR_INVALID;
  }
  DOE_REQUIRE(n <= kMaxN, DOE_ERR_UNSUPPORTED, "fft1d: length above 8192 is not built");
  DOE_REQUIRE(stride != 0 && offset >= 0 && offset < len && offset + (n - 1) * stride >= 0 &&
                  offset + (n - 1) * stride < len,
              DOE_ERR_SHAPE, "fft1d: strided slice leaves the planes");
  std::vector<double> br(2 * n, 0.0), bi(2 * n, 0.0);
  for (long i = 0; i < n; ++i) {
    br[i] = re[offset + i * stride];
    bi[i] = im[offset + i * stride];
  }
  doe_field* f = nullptr;
  int rc = doe_field_create(ctx, (int)n, 2, 1, dtype, &f);
  if (rc) return rc;
  rc = doe_field_upload(ctx, f, br.data(), bi.data());
  if (!rc) rc = dispatch_fft_rows(ctx, f, inverse);
  if (!rc) rc = doe_field_download(ctx, f, br.data(), bi.data());
  doe_field_free(f);
  if (rc) return rc;
  for (long i = 0; i < n; ++i) {
    re[offset + i * stride] = br[i];
    im[offset + i * stride] = bi[i];
  }
  return DOE_OK;
}

}  // extern "C"
