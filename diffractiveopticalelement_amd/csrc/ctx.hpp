// Host-side context, field handles and launch helpers shared by the kernel
// translation units.
#pragma once
#include "passes.hpp"

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/doehip.h"

namespace doe {

void set_error(const std::string& msg);

#define DOE_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::doe::set_error(std::string(#call) + ": " + hipGetErrorString(e_));             \
      return DOE_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define DOE_REQUIRE(cond, status, msg)                                                 \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      ::doe::set_error(msg);                                                           \
      return status;                                                                   \
    }                                                                                  \
  } while (0)

inline bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }
inline size_t dtype_size(int dt) {
  switch (dt) {
    case DOE_F32: return 4;
    case DOE_F64: return 8;
    case DOE_F16: return 2;
    case DOE_BF16: return 2;
    default: return 0;
  }
}

// Device twiddle table for one (N, compute precision): tw[m] = exp(-2*pi*i*m/N).
struct Twiddle {
  void* re = nullptr;
  void* im = nullptr;
};

}  // namespace doe

struct doe_prof_rec {
  const char* tag;
  hipEvent_t a, b;
};

struct doe_ctx {
  int device = 0;
  int flags = 0;  // DOE_CTX_ASYNC | DOE_CTX_PROFILE
  std::vector<doe_prof_rec> prof;
  std::vector<hipEvent_t> ev_pool;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::map<std::pair<int, int>, doe::Twiddle> tw;  // (N, sizeof(T)) -> table
  // device parameter block, refilled per call (small)
  void* dparams = nullptr;
  size_t dparams_cap = 0;
  std::vector<char> params_host;  // last uploaded parameter block
  // staging for FP64 host <-> device conversion
  void* stage = nullptr;
  size_t stage_cap = 0;
  // reduction scratch: partials + per-field arrival counters (zeroed once)
  double* partials = nullptr;
  size_t partials_cap = 0;  // doubles
  unsigned* counters = nullptr;
  size_t counters_cap = 0;
};

struct doe_field {
  int nx = 0, ny = 0, batch = 0, dtype = DOE_F32;
  void* re = nullptr;
  void* im = nullptr;
  bool owned = false;
  size_t count() const { return size_t(nx) * size_t(ny) * size_t(batch); }
};

namespace doe {

// Fetch (creating on first use) the twiddle table for N in precision T.
int get_twiddle(doe_ctx* ctx, int n, int tsize, Twiddle* out);
// Ensure a device parameter block of `bytes` and upload `host` into it on the
// context stream (synchronous w.r.t. the host buffer).
int upload_params(doe_ctx* ctx, const void* host, size_t bytes, void** dev);
// Ensure `n` doubles of partials and `nf` zeroed counters.
int ensure_reduce(doe_ctx* ctx, size_t n, size_t nf);
int ensure_stage(doe_ctx* ctx, size_t bytes);

// Compile-time dispatch over the supported power-of-two line lengths.
template <class F>
inline bool with_n(int n, F&& f) {
  switch (n) {
#define DOE_CASE(v)                          \
  case v:                                    \
    f(std::integral_constant<int, v>{});     \
    return true;
    DOE_CASE(2) DOE_CASE(4) DOE_CASE(8) DOE_CASE(16) DOE_CASE(32) DOE_CASE(64) DOE_CASE(128)
    DOE_CASE(256) DOE_CASE(512) DOE_CASE(1024) DOE_CASE(2048) DOE_CASE(4096) DOE_CASE(8192)
#undef DOE_CASE
    default:
      return false;
  }
}
constexpr int kMaxN = 8192;

// Phase-retrieval loops are built for sides 16..4096.
template <class F>
inline bool with_n_gs(int n, F&& f) {
  switch (n) {
#define DOE_CASE(v)                          \
  case v:                                    \
    f(std::integral_constant<int, v>{});     \
    return true;
    DOE_CASE(16) DOE_CASE(32) DOE_CASE(64) DOE_CASE(128) DOE_CASE(256) DOE_CASE(512)
    DOE_CASE(1024) DOE_CASE(2048) DOE_CASE(4096)
#undef DOE_CASE
    default:
      return false;
  }
}

// Optional per-launch event bracketing (DOE_CTX_PROFILE): the measurement
// hook behind doe_ctx_profile_report.
hipEvent_t prof_event(doe_ctx* ctx);
inline void prof_begin(doe_ctx* ctx, const char* tag) {
  if (!(ctx->flags & DOE_CTX_PROFILE)) return;
  doe_prof_rec r{tag, prof_event(ctx), prof_event(ctx)};
  hipEventRecord(r.a, ctx->stream);
  ctx->prof.push_back(r);
}
inline void prof_end(doe_ctx* ctx) {
  if (!(ctx->flags & DOE_CTX_PROFILE) || ctx->prof.empty()) return;
  hipEventRecord(ctx->prof.back().b, ctx->stream);
}
// Synchronise unless the context runs asynchronously.
inline hipError_t finish(doe_ctx* ctx) {
  if (ctx->flags & DOE_CTX_ASYNC) return hipGetLastError();
  return hipStreamSynchronize(ctx->stream);
}

template <typename T, typename S, int N, int MODE, bool INV1, bool INV2, class Op>
inline hipError_t launch_row(doe_ctx* ctx, const char* tag, const Op& op, const S* ire,
                             const S* iim, S* ore, S* oim, int ny, int batch, const Twiddle& tw) {
  using CF = RowCfg<T, N>;
  const long total = (long)ny * batch;
  const long grid = (total + CF::LINES - 1) / CF::LINES;
  prof_begin(ctx, tag);
  hipLaunchKernelGGL((k_row_pass<T, S, N, MODE, INV1, INV2, Op>), dim3(unsigned(grid)),
                     dim3(CF::WG), 0, ctx->stream, op, ire, iim, ore, oim, ny, total,
                     (const T*)tw.re, (const T*)tw.im);
  prof_end(ctx);
  return hipGetLastError();
}

template <typename T, typename S, int N, int MODE, bool INV1, bool INV2, class Op>
inline hipError_t launch_col(doe_ctx* ctx, const char* tag, const Op& op, const S* ire,
                             const S* iim, S* ore, S* oim, int nx, int batch, const Twiddle& tw) {
  using CF = ColCfg<T, N>;
  const int tiles = (nx + CF::C - 1) / CF::C;
  const long grid = (long)tiles * batch;
  prof_begin(ctx, tag);
  hipLaunchKernelGGL((k_col_pass<T, S, N, MODE, INV1, INV2, Op>), dim3(unsigned(grid)),
                     dim3(CF::WG), 0, ctx->stream, op, ire, iim, ore, oim, nx, tiles,
                     (const T*)tw.re, (const T*)tw.im);
  prof_end(ctx);
  return hipGetLastError();
}

template <typename T, int N>
inline int col_tiles(int nx) {
  return (nx + ColCfg<T, N>::C - 1) / ColCfg<T, N>::C;
}

// Family entry points (defined in their own translation units; each
// (T, S) instantiation is compiled as a separate object, DOE_PART).
template <typename T, typename S>
int run_fft2d(doe_ctx* ctx, doe_field* f, int inverse);
template <typename T, typename S>
int run_fft_rows(doe_ctx* ctx, doe_field* f, int inverse);
struct OpAsmH;
struct ChirpField;
template <typename T, typename S>
int run_asm(doe_ctx* ctx, const doe_field* in, doe_field* out, const OpAsmH& op);
template <typename T, typename S>
int run_fresnel_fwd(doe_ctx* ctx, const doe_field* in, doe_field* out, const ChirpField* dfp);
template <typename T, typename S>
int run_fresnel_inv(doe_ctx* ctx, const doe_field* in, doe_field* out, const ChirpField* dfp);
template <typename T, typename S>
int run_lens(doe_ctx* ctx, const doe_field* in, doe_field* out, const ChirpField* dfp);

}  // namespace doe
