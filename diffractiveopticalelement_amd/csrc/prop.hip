// Free-space propagators (ref: src/propagate.js): angular spectrum, single-FFT
// Fresnel forward/inverse, lens focal-plane transform.  Each is two or three
// passes with the transfer function / chirps / checkerboards fused in.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "ctx.hpp"
#include "merged.hpp"
#include "ops_prop.hpp"

namespace doe {

int check_field(const doe_ctx* ctx, const doe_field* f);

// localFrequencyLimit (ref: propagate.js:48-51), same operation order.
static double local_frequency_limit(int n, double d, double lambda, double z) {
  const double df = 1 / (n * d);
  return 1 / (lambda * std::sqrt(std::pow(2 * df * z, 2) + 1));
}

static AsmField asm_field_params(int nx, int ny, double dx, double dy, double lambda, double z,
                                 int band_limit, double soft_edge) {
  AsmField p{};
  const double inf = std::numeric_limits<double>::infinity();
  const double fx_lim = band_limit ? local_frequency_limit(nx, dx, lambda, z) : inf;
  const double fy_lim = band_limit ? local_frequency_limit(ny, dy, lambda, z) : inf;
  const double fxn = 1 / (2 * dx);
  const double fyn = 1 / (2 * dy);
  p.lim_x = std::min(fx_lim, fxn);
  p.lim_y = std::min(fy_lim, fyn);
  p.inner_x = p.lim_x * (1 - soft_edge);
  p.inner_y = p.lim_y * (1 - soft_edge);
  p.band = band_limit ? 1 : 0;
  p.two_pi_z = 2 * kPi * z;
  p.z = z;
  p.zero = (z == 0) ? 1 : 0;
  return p;
}

template <typename T, typename S>
int run_asm(doe_ctx* ctx, const doe_field* in, doe_field* out, const OpAsmH& op) {
  Plan2 pl;
  int rc = make_plan2(ctx, in->nx, in->ny, sizeof(T), col_split<T>(in->nx, in->ny), &pl);
  if (rc) return rc;
  const size_t plane = size_t(in->nx) * size_t(in->ny);
  const int cf = chunk_fields(ctx, plane * 2 * sizeof(S), in->batch);
  const OpNone none{};
  for (int f0 = 0; f0 < in->batch; f0 += cf) {
    const int nb = std::min(cf, in->batch - f0);
    const S* ire = (const S*)in->re + size_t(f0) * plane;
    const S* iim = (const S*)in->im + size_t(f0) * plane;
    S* re = (S*)out->re + size_t(f0) * plane;
    S* im = (S*)out->im + size_t(f0) * plane;
    // rows forward (+ column split) -> column blocks F . H . F^-1 -> rows
    // inverse; every pass after the first runs in place on the chunk
    DOE_HIP((row_pass<T, S, 0, false, false, false, true>(ctx, "asm_row_fwd", none, ire, iim, re,
                                                           im, nb, pl, AllSizes{})));
    DOE_HIP((col_pass<T, S, 1, false, true>(ctx, "asm_col_h", op.at(f0), re, im, re, im, nb, pl,
                                            AllSizes{})));
    DOE_HIP((row_pass<T, S, 0, true, false, true, false>(ctx, "asm_row_inv", none, re, im, re, im,
                                                          nb, pl, AllSizes{})));
  }
  return DOE_OK;
}

// Two-pass chirp transforms.  Forward (Fresnel forward, lens): row pass with a
// pre hook, then the column blocks with a post hook.  Inverse (Fresnel
// inverse): column blocks with a pre hook on the natural-order input, then
// the row pass with a post hook.
template <typename T, typename S, bool INV, int PRE, int POST>
static int run_chirp2(doe_ctx* ctx, const doe_field* in, doe_field* out, const ChirpField* dfp) {
  Plan2 pl;
  int rc = make_plan2(ctx, in->nx, in->ny, sizeof(T), col_split<T>(in->nx, in->ny), &pl);
  if (rc) return rc;
  const size_t plane = size_t(in->nx) * size_t(in->ny);
  const int cf = chunk_fields(ctx, plane * 2 * sizeof(S), in->batch);
  // the split column pass permutes rows: route it through work planes
  S* wre = nullptr;
  S* wim = nullptr;
  if (pl.n1 > 1) {
    rc = ensure_work(ctx, size_t(cf) * plane * sizeof(S));
    if (rc) return rc;
    wre = (S*)ctx->work_re;
    wim = (S*)ctx->work_im;
  }
  const OpChirp<PRE, 0> first{dfp, in->nx, in->ny};
  const OpChirp<0, POST> last{dfp, in->nx, in->ny};
  for (int f0 = 0; f0 < in->batch; f0 += cf) {
    const int nb = std::min(cf, in->batch - f0);
    const S* ire = (const S*)in->re + size_t(f0) * plane;
    const S* iim = (const S*)in->im + size_t(f0) * plane;
    S* re = (S*)out->re + size_t(f0) * plane;
    S* im = (S*)out->im + size_t(f0) * plane;
    S* xre = pl.n1 > 1 ? wre : re;
    S* xim = pl.n1 > 1 ? wim : im;
    if constexpr (!INV) {
      DOE_HIP((row_pass<T, S, 0, false, false, false, true>(ctx, "chirp_row", first.at(f0), ire,
                                                             iim, xre, xim, nb, pl, AllSizes{})));
      DOE_HIP((col_pass<T, S, 0, false, false>(ctx, "chirp_col", last.at(f0), xre, xim, re, im, nb,
                                               pl, AllSizes{})));
    } else {
      DOE_HIP((col_pass<T, S, 0, true, false>(ctx, "chirp_col", first.at(f0), ire, iim, xre, xim,
                                              nb, pl, AllSizes{})));
      DOE_HIP((row_pass<T, S, 0, true, false, true, false>(ctx, "chirp_row", last.at(f0), xre, xim,
                                                            re, im, nb, pl, AllSizes{})));
    }
  }
  return DOE_OK;
}

constexpr int kFresFwdPre = CH_QUAD | CH_CHECKER;
constexpr int kFresFwdPost = CH_PREFAC | CH_QUAD | CH_OUTGRID | CH_CHECKER;
constexpr int kFresInvPre = CH_QUAD | CH_CONJ | CH_OUTGRID | CH_CHECKER;
constexpr int kFresInvPost = CH_PREFAC | CH_QUAD | CH_CONJ | CH_CHECKER;
constexpr int kLensPre = CH_CHECKER;
constexpr int kLensPost = CH_PREFAC | CH_QUAD | CH_OUTGRID | CH_MINUS_HALF_PI | CH_CHECKER;

template <typename T, typename S>
int run_fresnel_fwd(doe_ctx* ctx, const doe_field* in, doe_field* out, const ChirpField* dfp) {
  return run_chirp2<T, S, false, kFresFwdPre, kFresFwdPost>(ctx, in, out, dfp);
}
template <typename T, typename S>
int run_fresnel_inv(doe_ctx* ctx, const doe_field* in, doe_field* out, const ChirpField* dfp) {
  return run_chirp2<T, S, true, kFresInvPre, kFresInvPost>(ctx, in, out, dfp);
}
template <typename T, typename S>
int run_lens(doe_ctx* ctx, const doe_field* in, doe_field* out, const ChirpField* dfp) {
  return run_chirp2<T, S, false, kLensPre, kLensPost>(ctx, in, out, dfp);
}

#define DOE_PROP_INST(T, S)                                                                     \
  template int run_asm<T, S>(doe_ctx*, const doe_field*, doe_field*, const OpAsmH&);           \
  template int run_fresnel_fwd<T, S>(doe_ctx*, const doe_field*, doe_field*, const ChirpField*); \
  template int run_fresnel_inv<T, S>(doe_ctx*, const doe_field*, doe_field*, const ChirpField*); \
  template int run_lens<T, S>(doe_ctx*, const doe_field*, doe_field*, const ChirpField*);
#if DOE_PART == 0
#define DOE_PROP_EXTERN(T, S)                                                                       \
  extern template int run_asm<T, S>(doe_ctx*, const doe_field*, doe_field*, const OpAsmH&);       \
  extern template int run_fresnel_fwd<T, S>(doe_ctx*, const doe_field*, doe_field*, const ChirpField*); \
  extern template int run_fresnel_inv<T, S>(doe_ctx*, const doe_field*, doe_field*, const ChirpField*); \
  extern template int run_lens<T, S>(doe_ctx*, const doe_field*, doe_field*, const ChirpField*);
DOE_PROP_EXTERN(float, float)
DOE_PROP_EXTERN(double, double)
DOE_PROP_EXTERN(float, __half)
DOE_PROP_EXTERN(float, __hip_bfloat16)
#elif DOE_PART == 1
DOE_PROP_INST(float, float)

// Angular-spectrum propagation of several FP32 batches of square fields of
// different sides (64 ... 2048), at most kMergedMax per call: the rows
// forward, the fused column pass and the rows inverse each run as one launch
// for all of them (merged.hpp), the same three passes as run_asm per batch.
int run_asm_multi_f32(doe_ctx* ctx, int count, const doe_field* const* in, doe_field* const* out,
                      const OpAsmH* ops) {
  MTable<MRowEntry<OpNone>> rf{}, ri{};
  MTable<MColEntry<OpAsmH>> cc{};
  rf.n = ri.n = cc.n = count;
  int rb = 0, cb = 0;
  for (int k = 0; k < count; ++k) {
    const int n = in[k]->nx;
    Plan2 pl;
    int rc = make_plan2(ctx, n, n, sizeof(float), col_split<float>(n, n), &pl);
    if (rc) return rc;
    const int n1 = pl.n1, n2 = n / n1;
    const int cols = mcol_cols(n2), g = mrow_groups(n);
    DOE_REQUIRE(cols > 0 && g > 0 && n1 == (n == 1024 ? 2 : (n == 2048 ? 4 : 1)), DOE_ERR_UNSUPPORTED,
                "doe_asm_multi: sides 64 ... 2048");
    const float* ire = (const float*)in[k]->re;
    const float* iim = (const float*)in[k]->im;
    float* re = (float*)out[k]->re;
    float* im = (float*)out[k]->im;
    const long groups = long(n / n1) * in[k]->batch;
    const int tiles = (n + cols - 1) / cols;
    const OpNone none{};
    rf.first[k] = ri.first[k] = rb;
    rf.e[k] = MRowEntry<OpNone>{none, ire, iim, re, im, (const float*)pl.tx.re,
                                (const float*)pl.tx.im, (const float*)pl.ty.re,
                                (const float*)pl.ty.im, groups, n};
    ri.e[k] = MRowEntry<OpNone>{none, re, im, re, im, (const float*)pl.tx.re,
                                (const float*)pl.tx.im, (const float*)pl.ty.re,
                                (const float*)pl.ty.im, groups, n};
    rb += int((groups + g - 1) / g);
    cc.first[k] = cb;
    cc.e[k] = MColEntry<OpAsmH>{ops[k], re, im, re, im, (const float*)pl.tc.re,
                                (const float*)pl.tc.im, n2, n, tiles, n1};
    cb += tiles * n1 * in[k]->batch;
  }
  rf.first[count] = ri.first[count] = rb;
  cc.first[count] = cb;
  long pts = 0;
  for (int k = 0; k < count; ++k) pts += long(in[k]->count());
  prof_begin(ctx, "asm_multi_row_fwd", pts);
  hipLaunchKernelGGL((k_row_merged<OpNone, 0, false, false, false, true>), dim3(unsigned(rb)),
                     dim3(kMWG), 0, ctx->stream, rf, next_dir(ctx));
  prof_end(ctx);
  prof_begin(ctx, "asm_multi_col_h", pts);
  hipLaunchKernelGGL((k_col_merged<OpAsmH, 1, false, true>), dim3(unsigned(cb)), dim3(kMWG), 0,
                     ctx->stream, cc, next_dir(ctx));
  prof_end(ctx);
  prof_begin(ctx, "asm_multi_row_inv", pts);
  hipLaunchKernelGGL((k_row_merged<OpNone, 0, true, false, true, false>), dim3(unsigned(rb)),
                     dim3(kMWG), 0, ctx->stream, ri, next_dir(ctx));
  prof_end(ctx);
  DOE_HIP(hipGetLastError());
  return DOE_OK;
}
#elif DOE_PART == 2
DOE_PROP_INST(double, double)
#elif DOE_PART == 3
DOE_PROP_INST(float, __half)
#elif DOE_PART == 4
DOE_PROP_INST(float, __hip_bfloat16)
#endif

// Dispatch over the storage precision to the per-object instantiations.
#define DOE_DISPATCH(fn, dtype, ...)                                          \
  ((dtype) == DOE_F32    ? fn<float, float>(__VA_ARGS__)                     \
   : (dtype) == DOE_F64  ? fn<double, double>(__VA_ARGS__)                   \
   : (dtype) == DOE_F16  ? fn<float, __half>(__VA_ARGS__)                    \
   : (dtype) == DOE_BF16 ? fn<float, __hip_bfloat16>(__VA_ARGS__)            \
                         : (set_error("unknown dtype"), DOE_ERR_INVALID))

static int check_pair(doe_ctx* ctx, const doe_field* in, const doe_field* out) {
  int rc = check_field(ctx, in);
  if (rc) return rc;
  rc = check_field(ctx, out);
  if (rc) return rc;
  DOE_REQUIRE(in->nx == out->nx && in->ny == out->ny && in->batch == out->batch &&
                  in->dtype == out->dtype,
              DOE_ERR_SHAPE, "input and output fields differ in shape or dtype");
  return DOE_OK;
}

// Centred-grid transfer function + propagating count (ref: propagate.js:57-102).
template <typename T>
struct OpAsmTable {
  OpAsmH h;
  T* re;
  T* im;
  unsigned long long* count;
  __device__ void apply(int, int i, int j, long k) const {
#pragma clang fp contract(off)
    const int kx = (i + (h.nx >> 1)) & (h.nx - 1);
    const int ky = (j + (h.ny >> 1)) & (h.ny - 1);
    T hr, hi;
    h.template h_at<T>(0, kx, ky, hr, hi);
    re[k] = hr;
    im[k] = hi;
    // propagating count: w > 0.5 for real kz; every sample when evanescent
    // waves are kept as a phase (ref: propagate.js:80-96).
    const int sx = kx < (h.nx >> 1) ? kx : kx - h.nx;
    const int sy = ky < (h.ny >> 1) ? ky : ky - h.ny;
    const double fx = double(sx) * h.dfx, fy = double(sy) * h.dfy;
    const double arg = h.inv_l2 - fx * fx - fy * fy;
    const AsmField p = h.fp[0];
    double w = 1.0;
    if (p.band) w = soft_window(fabs(fx), p.lim_x, p.inner_x) * soft_window(fabs(fy), p.lim_y, p.inner_y);
    bool prop = arg >= 0.0 ? (w > 0.5) : (h.evanescent == 0);
    if (prop) atomicAdd(count, 1ull);
  }
};

}  // namespace doe

#if DOE_PART == 0
using namespace doe;

namespace doe {
int run_asm_multi_f32(doe_ctx* ctx, int count, const doe_field* const* in, doe_field* const* out,
                      const OpAsmH* ops);
}

extern "C" {

int doe_asm(doe_ctx* ctx, const doe_field* in, doe_field* out, double dx, double dy, double lambda,
            const double* z, int band_limit, int evanescent, double soft_edge) {
  int rc = check_pair(ctx, in, out);
  if (rc) return rc;
  DOE_REQUIRE(z, DOE_ERR_INVALID, "null z");
  DOE_HIP(hipSetDevice(ctx->device));
  bool all_zero = true;
  std::vector<AsmField> params(in->batch);
  for (int b = 0; b < in->batch; ++b) {
    params[b] = asm_field_params(in->nx, in->ny, dx, dy, lambda, z[b], band_limit, soft_edge);
    all_zero = all_zero && (z[b] == 0);
  }
  if (all_zero) {  // ref: propagate.js:113, z === 0 returns the field unchanged
    if (in != out) return doe_field_copy(ctx, in, out);
    return DOE_OK;
  }
  void* dp = nullptr;
  rc = upload_params(ctx, params.data(), params.size() * sizeof(AsmField), &dp);
  if (rc) return rc;
  OpAsmH op{};
  op.fp = (const AsmField*)dp;
  op.dfx = 1 / (in->nx * dx);
  op.dfy = 1 / (in->ny * dy);
  op.inv_l2 = 1 / (lambda * lambda);
  op.nx = in->nx;
  op.ny = in->ny;
  op.evanescent = evanescent ? 1 : 0;
  rc = DOE_DISPATCH(run_asm, in->dtype, ctx, in, out, op);
  if (rc) return rc;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_asm_multi(doe_ctx* ctx, int count, const doe_field* const* in, doe_field* const* out,
                  const double* dx, const double* dy, double lambda, const double* const* z,
                  int band_limit, int evanescent, double soft_edge) {
  DOE_REQUIRE(ctx && in && out && dx && dy && z, DOE_ERR_INVALID, "doe_asm_multi: null argument");
  DOE_REQUIRE(count >= 0, DOE_ERR_INVALID, "doe_asm_multi: negative count");
  long nparams = 0;
  for (int k = 0; k < count; ++k) {
    int rc = check_pair(ctx, in[k], out[k]);
    if (rc) return rc;
    DOE_REQUIRE(z[k], DOE_ERR_INVALID, "doe_asm_multi: null z");
    DOE_REQUIRE(in[k]->dtype == DOE_F32 && in[k]->nx == in[k]->ny && in[k]->nx >= 64 &&
                    in[k]->nx <= 2048,
                DOE_ERR_UNSUPPORTED, "doe_asm_multi: FP32 square fields of sides 64 ... 2048");
    nparams += in[k]->batch;
  }
  if (count == 0) return DOE_OK;
  DOE_HIP(hipSetDevice(ctx->device));
  // every field's transfer-function parameters in one block
  std::vector<AsmField> params;
  params.reserve(size_t(nparams));
  for (int k = 0; k < count; ++k)
    for (int b = 0; b < in[k]->batch; ++b)
      params.push_back(asm_field_params(in[k]->nx, in[k]->ny, dx[k], dy[k], lambda, z[k][b],
                                        band_limit, soft_edge));
  void* dp = nullptr;
  int rc = upload_params(ctx, params.data(), params.size() * sizeof(AsmField), &dp);
  if (rc) return rc;
  std::vector<OpAsmH> ops(static_cast<size_t>(count));
  long off = 0;
  for (int k = 0; k < count; ++k) {
    OpAsmH& op = ops[k];
    op = OpAsmH{};
    op.fp = (const AsmField*)dp + off;
    op.dfx = 1 / (in[k]->nx * dx[k]);
    op.dfy = 1 / (in[k]->ny * dy[k]);
    op.inv_l2 = 1 / (lambda * lambda);
    op.nx = in[k]->nx;
    op.ny = in[k]->ny;
    op.evanescent = evanescent ? 1 : 0;
    off += in[k]->batch;
  }
  // up to kMergedMax batches per launch triple
  for (int k0 = 0; k0 < count; k0 += kMergedMax) {
    const int nk = std::min(kMergedMax, count - k0);
    rc = run_asm_multi_f32(ctx, nk, in + k0, out + k0, ops.data() + k0);
    if (rc) return rc;
  }
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_asm_transfer(doe_ctx* ctx, doe_field* out, double dx, double dy, double lambda, double z,
                     int band_limit, int evanescent, double soft_edge,
                     double* propagating_fraction) {
  int rc = check_field(ctx, out);
  if (rc) return rc;
  DOE_REQUIRE(out->batch == 1 && (out->dtype == DOE_F32 || out->dtype == DOE_F64),
              DOE_ERR_UNSUPPORTED, "doe_asm_transfer: batch-1 F32/F64 output only");
  DOE_HIP(hipSetDevice(ctx->device));
  AsmField p = asm_field_params(out->nx, out->ny, dx, dy, lambda, z, band_limit, soft_edge);
  p.zero = 0;  // the table is built for any z, including 0
  // parameter blocks are read-only for kernels (a later call may read the
  // same block again); the count goes to the reduction scratch
  void* dp = nullptr;
  rc = upload_params(ctx, &p, sizeof(p), &dp);
  if (rc) return rc;
  rc = ensure_reduce(ctx, 1);
  if (rc) return rc;
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(ctx->partials);
  DOE_HIP(hipMemsetAsync(cnt, 0, sizeof(*cnt), ctx->stream));
  OpAsmH h{};
  h.fp = (const AsmField*)dp;
  h.dfx = 1 / (out->nx * dx);
  h.dfy = 1 / (out->ny * dy);
  h.inv_l2 = 1 / (lambda * lambda);
  h.nx = out->nx;
  h.ny = out->ny;
  h.evanescent = evanescent ? 1 : 0;
  const long total = (long)out->nx * out->ny;
  const unsigned grid = unsigned((total + 255) / 256);
  if (out->dtype == DOE_F64) {
    OpAsmTable<double> op{h, (double*)out->re, (double*)out->im, cnt};
    hipLaunchKernelGGL(k_elementwise<OpAsmTable<double>>, dim3(grid), dim3(256), 0, ctx->stream, op,
                       out->nx, out->ny, total);
  } else {
    OpAsmTable<float> op{h, (float*)out->re, (float*)out->im, cnt};
    hipLaunchKernelGGL(k_elementwise<OpAsmTable<float>>, dim3(grid), dim3(256), 0, ctx->stream, op,
                       out->nx, out->ny, total);
  }
  DOE_HIP(hipGetLastError());
  unsigned long long c = 0;
  DOE_HIP(hipMemcpyAsync(&c, cnt, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  if (propagating_fraction) *propagating_fraction = double(c) / double(total);
  return DOE_OK;
}

int doe_fresnel(doe_ctx* ctx, const doe_field* in, doe_field* out, double dx, double dy,
                double lambda, const double* z, int inverse, double in_dx, double in_dy,
                double* out_dx, double* out_dy) {
#pragma clang fp contract(off)
  int rc = check_pair(ctx, in, out);
  if (rc) return rc;
  DOE_REQUIRE(z, DOE_ERR_INVALID, "null z");
  DOE_HIP(hipSetDevice(ctx->device));
  const int nx = in->nx, ny = in->ny;
  std::vector<ChirpField> params(in->batch);
  for (int b = 0; b < in->batch; ++b) {
    const double az = std::fabs(z[b]);
    DOE_REQUIRE(az > 0, DOE_ERR_INVALID, "fresnel: z must be nonzero");
    const double k = (2 * kPi) / lambda;
    ChirpField& p = params[b];
    p.laz = lambda * az;
    p.inv_laz = 1 / p.laz;
    if (!inverse) {
      // ref: propagate.js:175-205
      p.dx_in = dx;
      p.dy_in = dy;
      p.dx_out = (lambda * az) / (nx * dx);
      p.dy_out = (lambda * az) / (ny * dy);
      const double amp = (dx * dy) / (lambda * az);
      const double kz = k * z[b] - (kPi * (nx + ny)) / 2;
      p.pre_r = amp * std::sin(kz);
      p.pre_i = -amp * std::cos(kz);
      if (out_dx) out_dx[b] = p.dx_out;
      if (out_dy) out_dy[b] = p.dy_out;
    } else {
      // ref: propagate.js:246-316; dx,dy are the element pitch, in_dx the
      // magnified pitch of the incoming field
      DOE_REQUIRE(dx > 0 && dy > 0, DOE_ERR_INVALID,
                  "propagateFresnelFFTInverse: pass the DOE-plane pitch as {dx, dy}");
      p.dx_in = dx;
      p.dy_in = dy;
      p.dx_out = in_dx;
      p.dy_out = in_dy;
      const double amp = (lambda * az) / (dx * dy);
      const double kg = k * z[b] - (kPi * (nx + ny)) / 2 - kPi / 2;
      p.pre_r = amp * std::cos(kg);
      p.pre_i = -amp * std::sin(kg);
      if (out_dx) out_dx[b] = dx;
      if (out_dy) out_dy[b] = dy;
    }
    chirp_revs(p);
  }
  void* dp = nullptr;
  rc = upload_params(ctx, params.data(), params.size() * sizeof(ChirpField), &dp);
  if (rc) return rc;
  if (!inverse)
    rc = DOE_DISPATCH(run_fresnel_fwd, in->dtype, ctx, in, out, (const ChirpField*)dp);
  else
    rc = DOE_DISPATCH(run_fresnel_inv, in->dtype, ctx, in, out, (const ChirpField*)dp);
  if (rc) return rc;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_lens(doe_ctx* ctx, const doe_field* in, doe_field* out, double dx, double dy,
             double lambda, const double* f, double* out_dx, double* out_dy) {
#pragma clang fp contract(off)
  int rc = check_pair(ctx, in, out);
  if (rc) return rc;
  DOE_REQUIRE(f, DOE_ERR_INVALID, "null focal length");
  DOE_HIP(hipSetDevice(ctx->device));
  const int nx = in->nx, ny = in->ny;
  // fftshift . F . ifftshift == g * c .* F(c .* x), c = (-1)^(i+j),
  // g = (-1)^(nx/2 + ny/2): the shifts become checkerboards.
  const double g = (((nx >> 1) + (ny >> 1)) & 1) ? -1.0 : 1.0;
  std::vector<ChirpField> params(in->batch);
  for (int b = 0; b < in->batch; ++b) {
    ChirpField& p = params[b];
    const double fl = f[b];
    p.laz = lambda * fl;
    p.inv_laz = 1 / p.laz;
    p.dx_in = dx;
    p.dy_in = dy;
    p.dx_out = (lambda * fl) / (nx * dx);  // ref: propagate.js:351-352
    p.dy_out = (lambda * fl) / (ny * dy);
    p.pre_r = g * ((dx * dy) / (lambda * fl));
    p.pre_i = 0.0;
    if (out_dx) out_dx[b] = p.dx_out;
    if (out_dy) out_dy[b] = p.dy_out;
    chirp_revs(p);
  }
  void* dp = nullptr;
  rc = upload_params(ctx, params.data(), params.size() * sizeof(ChirpField), &dp);
  if (rc) return rc;
  rc = DOE_DISPATCH(run_lens, in->dtype, ctx, in, out, (const ChirpField*)dp);
  if (rc) return rc;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

}  // extern "C"
#endif  // DOE_PART == 0
