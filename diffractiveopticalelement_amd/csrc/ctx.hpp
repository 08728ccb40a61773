// Host-side context, field handles and launch helpers shared by the kernel
// translation units.
#pragma once
#include "passes.hpp"

#include <hip/hip_ext.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
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

// One tuned configuration of a batched fft2d shape (see doe_autotune_fft2d).
struct doe_tune {
  int n1 = 0;         // column split (0: the built-in choice)
  int colvar = 0;     // column tile variant (FP32 arithmetic; 0: the built-in choice)
  int rowplanes = 0;  // LDS planes of the row pass (0: the built-in choice)
  long chunk = -1;    // chunk bytes (-1: the context's)
};

struct doe_prof_rec {
  const char* tag;
  hipEvent_t a, b;
  long pts;  // grid points the launch covers
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
  // device parameter blocks (per-field geometry and the like), a ring of
  // slots filled in stream order: a slot is refilled only after the work
  // queued behind its previous block has ended, so consecutive calls with
  // different parameters do not drain the stream
  struct ParamSlot {
    void* dev = nullptr;
    void* pin = nullptr;  // pinned staging (blocks up to kParamPinMax)
    size_t cap = 0;
    hipEvent_t done = nullptr;  // recorded each time the slot stops being current
    bool pending = false;
    std::vector<char> host;  // the block the slot holds (empty: none)
  };
  static constexpr int kParamSlots = 8;
  static constexpr size_t kParamPinMax = size_t(1) << 20;
  ParamSlot pslot[kParamSlots];
  int pcur = -1;   // slot of the block handed out last
  int pnext = 0;   // next slot to refill
  // staging for FP64 host <-> device conversion
  void* stage = nullptr;
  size_t stage_cap = 0;
  // reduction scratch: per-workgroup partial sums of the reducing passes
  double* partials = nullptr;
  size_t partials_cap = 0;  // doubles
  // work planes for passes that permute rows (split column passes)
  void* work_re = nullptr;
  void* work_im = nullptr;
  size_t work_cap = 0;  // bytes per plane
  // batched transforms run chunk by chunk, a chunk's planes (re + im) being
  // at most this many bytes; 0 = whole batch at once (the default: on the
  // 16 x 2048^2 batch, 128 MiB chunks measured 8% slower than one launch)
  long chunk_bytes = 0;
  // traversal direction of the next pass; flipped by every row/column launch
  int dir = 0;
  // tuned fft2d configurations, key (nx, ny, batch, dtype)
  std::map<std::tuple<int, int, int, int>, doe_tune> tune;
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
// A device copy of the parameter block `host` (`bytes`), queued on the
// context stream; the host buffer may be reused on return.  Kernels only
// read the block: a later call with the same bytes may be handed the same
// device copy.
int upload_params(doe_ctx* ctx, const void* host, size_t bytes, void** dev);
// Ensure `n` doubles of reduction partials (ctx->partials).
int ensure_reduce(doe_ctx* ctx, size_t n);
int ensure_stage(doe_ctx* ctx, size_t bytes);
// Ensure two work planes of `bytes` each (ctx->work_re / work_im).
int ensure_work(doe_ctx* ctx, size_t bytes);
// Resolve the fixed scratch of doe_screen_metrics_noalloc on `device` (current).
int metrics_scratch_init(int device);

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

// Phase-retrieval loops are built for sides 64..4096 (their row-pass
// reductions need every workgroup inside one field).
template <class F>
inline bool with_n_gs(int n, F&& f) {
  switch (n) {
#define DOE_CASE(v)                          \
  case v:                                    \
    f(std::integral_constant<int, v>{});     \
    return true;
    DOE_CASE(64) DOE_CASE(128) DOE_CASE(256) DOE_CASE(512) DOE_CASE(1024) DOE_CASE(2048)
    DOE_CASE(4096)
#undef DOE_CASE
    default:
      return false;
  }
}

// Optional per-launch event bracketing (DOE_CTX_PROFILE): the measurement
// hook behind doe_ctx_profile_report.
hipEvent_t prof_event(doe_ctx* ctx);
inline void prof_begin(doe_ctx* ctx, const char* tag, long pts) {
  if (!(ctx->flags & DOE_CTX_PROFILE)) return;
  doe_prof_rec r{tag, prof_event(ctx), prof_event(ctx), pts};
  hipEventRecord(r.a, ctx->stream);
  ctx->prof.push_back(r);
}
inline void prof_end(doe_ctx* ctx) {
  if (!(ctx->flags & DOE_CTX_PROFILE) || ctx->prof.empty()) return;
  hipEventRecord(ctx->prof.back().b, ctx->stream);
}
// The row and column passes are timed by their own dispatches instead:
// prof_open makes the record, and the pass's first and last kernels carry
// its start and stop events (hipExtLaunchKernelGGL), so profiling puts no
// event packets between the kernels of a timed loop.
// A stream under capture (by the caller) gets no timed launches.
inline bool prof_open(doe_ctx* ctx, const char* tag, long pts) {
  if (!(ctx->flags & DOE_CTX_PROFILE)) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return false;
  ctx->prof.push_back(doe_prof_rec{tag, prof_event(ctx), prof_event(ctx), pts});
  return true;
}
// Synchronise unless the context runs asynchronously.
inline hipError_t finish(doe_ctx* ctx) {
  if (ctx->flags & DOE_CTX_ASYNC) return hipGetLastError();
  return hipStreamSynchronize(ctx->stream);
}

// Consecutive passes walk their grids in opposite directions, so each pass
// starts on the data the previous one touched last (still in the GPU's
// memory-side cache when the passes run in place).
inline int next_dir(doe_ctx* ctx) {
  const int d = ctx->dir;
  ctx->dir ^= 1;
  return d;
}

// Fields per chunk of a batched transform.  MI355X has a 256 MB memory-side
// cache in front of HBM; keeping a chunk's planes (and the work planes that a
// permuting pass writes) to about half of it lets the second and third passes
// of a transform re-read the chunk from that cache instead of HBM.  Chunks
// are balanced so the last one is not a short tail.
inline int chunk_fields(long chunk_bytes, size_t field_bytes, int batch) {
  if (chunk_bytes <= 0 || batch <= 1) return batch;
  long c = long(size_t(chunk_bytes) / field_bytes);
  if (c < 1) c = 1;
  if (c >= batch) return batch;
  const long nchunks = (batch + c - 1) / c;
  return int((batch + nchunks - 1) / nchunks);
}
inline int chunk_fields(const doe_ctx* ctx, size_t field_bytes, int batch) {
  return chunk_fields(ctx->chunk_bytes, field_bytes, batch);
}

// Column split N1 of the four-step column factorisation for an nx-by-ny grid
// (1 = plain passes).  FP32 compute only; the row workgroup must hold N1 rows.
template <typename T>
inline int col_split(int nx, int ny) {
  if (sizeof(T) == 8) return 1;
  int n1 = ny == 1024 ? 2 : ((ny == 2048 || ny == 4096) ? 4 : 1);
  if (nx < 256 || (long)nx * n1 > 16384) n1 = 1;
  return n1;
}
// Ops that declare FP64_SPLIT (the plain transform) take the split in FP64
// too, with N1 FP64 rows of at most 8192 points in all per row group.
template <class Op, class = void>
struct OpFp64Split : std::false_type {};
template <class Op>
struct OpFp64Split<Op, std::void_t<decltype(Op::FP64_SPLIT)>>
    : std::integral_constant<bool, Op::FP64_SPLIT> {};
template <typename T, int NX, int N1, bool FP64 = false>
constexpr bool split_ok() {
  return N1 == 1 || (sizeof(T) == 4 && NX >= 256 && NX * N1 <= 16384) ||
         (sizeof(T) == 8 && FP64 && NX >= 256 && NX * N1 <= 8192);
}
// Whether the plain fft2d of an nx-by-ny grid can take column split n1 with
// compute of `tsize` bytes (the split_ok bounds at run time).
inline bool split_fits(int nx, int ny, int n1, int tsize) {
  if (n1 == 1) return true;
  const long rows_max = tsize == 4 ? 16384 : 8192;  // N1 rows of a row group
  return (n1 == 2 || n1 == 4) && ny % n1 == 0 && ny / n1 >= 2 && nx >= 256 &&
         long(nx) * n1 <= rows_max;
}
// Column split of the plain fft2d: FP32 as col_split; FP64 splits columns
// of 1024..8192 points in two where the row pair fits (16 x 2048^2 FP64:
// 0.855 ms split in two with one-plane rows, 0.858 ms in four, 0.883 ms
// plain, per transform).
template <typename T>
inline int fft_split(int nx, int ny) {
  if (sizeof(T) == 4) return col_split<float>(nx, ny);
  return ny >= 1024 && ny <= 8192 && nx >= 256 && nx <= 4096 ? 2 : 1;
}
template <class F>
inline bool with_n1(int n1, F&& f) {
  switch (n1) {
    case 1: f(std::integral_constant<int, 1>{}); return true;
    case 2: f(std::integral_constant<int, 2>{}); return true;
    case 4: f(std::integral_constant<int, 4>{}); return true;
    default: return false;
  }
}

// Twiddle tables of a 2-D plan: rows (nx), cross-row step (ny), columns (N2).
struct Plan2 {
  int nx = 0, ny = 0, n1 = 1, n2 = 0;
  Twiddle tx, ty, tc;
};
int make_plan2(doe_ctx* ctx, int nx, int ny, int tsize, int n1, Plan2* p);

template <typename T, typename S, int NX, int N1, int MODE, bool INV1, bool INV2, bool XIN,
          bool XOUT, class Op, int RPL = 2>
inline hipError_t launch_row(doe_ctx* ctx, const char* tag, const Op& op, const S* ire,
                             const S* iim, S* ore, S* oim, int batch, const Plan2& pl) {
  constexpr bool LANES_POSSIBLE = N1 > 1 && LinePlan<NX, PChoice<NX>::P>::NST >= 3 &&
                                  (!XIN || !OpHooks<Op>::pre) && (!XOUT || !OpHooks<Op>::post);
  if constexpr (!split_ok<T, NX, N1, OpFp64Split<Op>::value>() ||
                (RPL == 1 && !(LANES_POSSIBLE || N1 == 1))) {
    return hipErrorInvalidValue;  // never selected by col_split / the tuner
  } else {
    using CF = RowCfg<T, NX, N1, RPL>;
    const long groups = (long)(pl.ny / N1) * batch;
    const long grid = (groups + CF::G - 1) / CF::G;
    if constexpr (Op::NACC > 0) {
      if ((pl.ny / N1) % CF::G != 0) return hipErrorInvalidValue;
    }
    const bool timed = prof_open(ctx, tag, long(pl.nx) * pl.ny * batch);
    const hipEvent_t ev0 = timed ? ctx->prof.back().a : nullptr;
    const hipEvent_t ev1 = timed ? ctx->prof.back().b : nullptr;
    hipExtLaunchKernelGGL((k_row_pass<T, S, NX, N1, MODE, INV1, INV2, XIN, XOUT, Op, RPL>),
                          dim3(unsigned(grid)), dim3(CF::WG), 0, ctx->stream, ev0,
                          Op::NACC > 0 ? nullptr : ev1, 0, op, ire, iim, ore, oim, pl.ny, groups,
                          (const T*)pl.tx.re, (const T*)pl.tx.im, (const T*)pl.ty.re,
                          (const T*)pl.ty.im, next_dir(ctx));
    if constexpr (Op::NACC > 0)
      hipExtLaunchKernelGGL((k_finalize<Op, CF::WG>), dim3(unsigned(batch)), dim3(CF::WG), 0,
                            ctx->stream, nullptr, ev1, 0, op, int((pl.ny / N1) / CF::G));
    return hipGetLastError();
  }
}

// Column tile variants the tuner chooses among (FP32 arithmetic): 0 the op's
// default, 1: 16 points x up to 64 columns, both planes; 2: 32 points x up
// to 64 columns, one plane; 3: 32 points x up to 32 columns, one plane;
// 4: 32 points x up to 16 columns, one plane (small workgroups, several per
// CU); 5: 16 points x up to 32 columns, one plane.
template <typename T, int N2, class Op, int CV>
struct ColVarSel {
  using type = typename OpColCfg<T, N2, Op>::type;
};
template <typename T, int N2, class Op>
struct ColVarSel<T, N2, Op, 1> {
  using type = ColGeom<T, N2, 16, 64, 2>;
};
template <typename T, int N2, class Op>
struct ColVarSel<T, N2, Op, 2> {
  using type = ColGeom<T, N2, 32, 64, 1>;
};
template <typename T, int N2, class Op>
struct ColVarSel<T, N2, Op, 3> {
  using type = ColGeom<T, N2, 32, 32, 1>;
};
template <typename T, int N2, class Op>
struct ColVarSel<T, N2, Op, 4> {
  using type = ColGeom<T, N2, 32, 16, 1>;
};
template <typename T, int N2, class Op>
struct ColVarSel<T, N2, Op, 5> {
  using type = ColGeom<T, N2, 16, 32, 1>;
};
template <typename T, int N2, class Op, int CV>
struct ColVar {
  using type = typename ColVarSel<T, N2, Op, CV>::type;
};

template <typename T, typename S, int N2, int MODE, bool INV1, bool INV2, class Op, int CV = 0>
inline hipError_t launch_col(doe_ctx* ctx, const char* tag, const Op& op, const S* ire,
                             const S* iim, S* ore, S* oim, int batch, const Plan2& pl) {
  using CF = typename ColVar<T, N2, Op, CV>::type;
  const int tiles = (pl.nx + CF::C - 1) / CF::C;
  const long grid = (long)tiles * pl.n1 * batch;
  const bool timed = prof_open(ctx, tag, long(pl.nx) * pl.ny * batch);
  const hipEvent_t ev0 = timed ? ctx->prof.back().a : nullptr;
  const hipEvent_t ev1 = timed ? ctx->prof.back().b : nullptr;
  hipExtLaunchKernelGGL((k_col_pass<T, S, N2, MODE, INV1, INV2, Op, CF>), dim3(unsigned(grid)),
                        dim3(CF::WG), 0, ctx->stream, ev0, Op::NACC > 0 ? nullptr : ev1, 0, op,
                        ire, iim, ore, oim, pl.nx, tiles, pl.n1, (const T*)pl.tc.re,
                        (const T*)pl.tc.im, next_dir(ctx));
  if constexpr (Op::NACC > 0)
    hipExtLaunchKernelGGL((k_finalize<Op, CF::WG>), dim3(unsigned(batch)), dim3(CF::WG), 0,
                          ctx->stream, nullptr, ev1, 0, op, tiles * pl.n1);
  return hipGetLastError();
}

// Row pass at the plan's split (instantiates only the valid (NX, N1) pairs).
template <typename T, typename S, int MODE, bool INV1, bool INV2, bool XIN, bool XOUT, class Op,
          int RPL = 2, class NF>
inline hipError_t row_pass(doe_ctx* ctx, const char* tag, const Op& op, const S* ire, const S* iim,
                           S* ore, S* oim, int batch, const Plan2& pl, NF&& with_sizes) {
  hipError_t e = hipErrorInvalidValue;
  with_sizes(pl.nx, [&](auto NN) {
    constexpr int NX = decltype(NN)::value;
    with_n1(pl.n1, [&](auto MM) {
      constexpr int N1 = decltype(MM)::value;
      e = launch_row<T, S, NX, N1, MODE, INV1, INV2, XIN, XOUT, Op, RPL>(ctx, tag, op, ire, iim,
                                                                         ore, oim, batch, pl);
    });
  });
  return e;
}

template <typename T, typename S, int MODE, bool INV1, bool INV2, class Op, int CV = 0, class NF>
inline hipError_t col_pass(doe_ctx* ctx, const char* tag, const Op& op, const S* ire, const S* iim,
                           S* ore, S* oim, int batch, const Plan2& pl, NF&& with_sizes) {
  hipError_t e = hipErrorInvalidValue;
  with_sizes(pl.n2, [&](auto NN) {
    e = launch_col<T, S, decltype(NN)::value, MODE, INV1, INV2, Op, CV>(ctx, tag, op, ire, iim,
                                                                         ore, oim, batch, pl);
  });
  return e;
}

// The tuned configuration for an fft2d shape, or the built-in one; fields
// left at 0 (n1, rowplanes) take the built-in choice, so the result always
// names the split and the row pass that run.
// Half-width storage (FP16 / BF16) of 2048+ lines defaults to plain passes
// with one-plane rows: at 4 B per point the passes are VALU-bound, and the
// plain row pass (58 VGPRs, 8 waves per SIMD) beats the split's cross-row
// network (16 x 2048^2 of noise: 0.248-0.257 ms against 0.26-0.27 ms for
// the split with two-plane rows; the tuner's best, the split with one-plane
// rows in 128 MiB chunks, 0.246 ms).  FP64 split rows of 1024+ points take
// one LDS plane (16 x 2048^2: 0.855 against 0.875 ms per transform).
inline doe_tune tune_for(const doe_ctx* ctx, int nx, int ny, int batch, int dtype) {
  doe_tune t{};
  auto it = ctx->tune.find(std::make_tuple(nx, ny, batch, dtype));
  if (it != ctx->tune.end()) t = it->second;
  const bool half = dtype == DOE_F16 || dtype == DOE_BF16;
  const bool half_big = half && nx >= 2048 && ny >= 2048;
  if (t.n1 == 0)
    t.n1 = half_big ? 1 : (dtype == DOE_F64 ? fft_split<double>(nx, ny) : fft_split<float>(nx, ny));
  if (t.rowplanes == 0)
    t.rowplanes = half_big || (dtype == DOE_F64 && t.n1 > 1 && nx >= 1024) ? 1 : 2;
  return t;
}

// size dispatchers usable as the `with_sizes` argument
struct AllSizes {
  template <class F>
  bool operator()(int n, F&& f) const { return with_n(n, f); }
};
struct GsSizes {
  template <class F>
  bool operator()(int n, F&& f) const { return with_n_gs(n, f); }
};
// sides 256..8192 only (the tuner's variants; smaller sides keep the built-in
// configuration)
struct LargeSizes {
  template <class F>
  bool operator()(int n, F&& f) const { return n >= 256 && with_n(n, f); }
};

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
