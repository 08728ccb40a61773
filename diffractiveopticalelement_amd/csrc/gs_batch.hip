// Batched far-field phase retrieval over designs of uneven sizes.
//
// ref: src/doe.js:639-730 fourierCgh (far-field GS) and 455-546
// weightedGerchbergSaxton (the WGS weight update, best-iterate tracking).
// The reference designs one element per call; this engine runs a batch of
// independent square designs of mixed sides (64 ... 4096) with the same
// per-iteration operations as the single-design far-field loop (gs.hip):
// pass B on column tiles, pass A on row groups.
//
// Layout.  Designs are grouped into size classes (largest first); a class's
// designs sit back to back in every per-point array, so a class is a batch of
// equal-size fields for the ordinary row/column pass kernels, with a
// per-class table of design states, histories and reduction slots.  One
// iteration of the whole batch is one column-pass, one finalize and one
// row-pass launch: the classes' parameters travel by value in a table in the
// kernel arguments, each workgroup finds its class there and runs that
// side's pass body (every body 512 threads wide).  The per-class launches
// (one column-pass and one row-pass launch per size class) remain for sides
// of 4096 and on request.  The iteration loop is captured once as a hipGraph
// and replayed.  (A first table-driven kernel that loaded per-item operation
// records from memory raised register spills 4x and lost to the per-class
// launches; with the records in kernel arguments the merged launches are
// 3 % faster on the 256-design batch.)
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <vector>

#include "gs_batch.hpp"
#include "merged.hpp"

namespace doe {

// Counter-based initial phase in FP64: phi = (u*2 - 1)*pi with u in [0, 1)
// the top 53 bits of splitmix64(seed*C1 + index*C2 + k).  The reference's
// xorshift stream is sequential per design; a counter-based stream lets
// every point be generated independently.
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void k_gsb_random_phase(float* __restrict__ phi, long n, unsigned long long seed,
                                   unsigned long long index) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const unsigned long long z = splitmix64(seed * 0xD1B54A32D192ED03ull +
                                          index * 0x8CB92BA72F3D8DD7ull + (unsigned long long)k);
  const double u = double(z >> 11) * (1.0 / 9007199254740992.0);
  phi[k] = float((u * 2 - 1) * 3.141592653589793);
}

// launch tags per size class (profiling)
inline const char* gsb_tag(int n, int pass) {
  static const char* col[] = {"gsb_col_64",   "gsb_col_128",  "gsb_col_256", "gsb_col_512",
                              "gsb_col_1024", "gsb_col_2048", "gsb_col_4096"};
  static const char* row[] = {"gsb_row_64",   "gsb_row_128",  "gsb_row_256", "gsb_row_512",
                              "gsb_row_1024", "gsb_row_2048", "gsb_row_4096"};
  static const char* ini[] = {"gsb_init_64",   "gsb_init_128",  "gsb_init_256", "gsb_init_512",
                              "gsb_init_1024", "gsb_init_2048", "gsb_init_4096"};
  const int i = ilog2(n) - 6;
  return pass == 0 ? col[i] : (pass == 1 ? row[i] : ini[i]);
}

// One launch per pass over every size class (the default; DOE_GSB_PER_CLASS
// selects the per-class launches further down): the table-driven pass
// kernels of merged.hpp with the far-field operations, and a finalize
// kernel over every design.
// Fold every design's column-pass partials and finalise (one workgroup per
// design; design d of the batch's class order).
struct GsbFinEntry {
  OpFFB<float> op;
  int per_field;  // column-pass workgroups per design
};
__global__ void __launch_bounds__(256) k_gsb_fin_merged(MTable<GsbFinEntry> tab) {
  constexpr int NA = OpFFB<float>::NACC;
  __shared__ double sh[16 * NA];
  const int d = blockIdx.x;
  const int c = m_entry(tab.first, tab.n, d);
  const GsbFinEntry& E = tab.e[c];
  const int f = d - tab.first[c];
  const int tid = threadIdx.x;
  const long first = (long)f * E.per_field;
  double sm[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) sm[k] = 0.0;
  for (int w = tid; w < E.per_field; w += 256) {
#pragma unroll
    for (int k = 0; k < NA; ++k) sm[k] += E.op.partials[(first + w) * NA + k];
  }
  block_sum<NA>(sm, sh, tid, 256);
  if (tid == 0) E.op.finalize(f, sm);
}

}  // namespace doe

using namespace doe;

namespace {

// Per-class operation buffers for the current history capacity.
int gsb_prepare(doe_gs_batch* B, int K) {
  doe_ctx* ctx = B->ctx;
  if (K > B->hist_k) {
    void* p = nullptr;
    int rc = B->alloc(&p, sizeof(double) * 3 * size_t(K) * B->count);
    if (rc) return rc;
    B->hist = (double*)p;
    B->hist_k = K;
    if (B->graph) {  // its nodes point at the old histories
      hipGraphExecDestroy(B->graph);
      B->graph = nullptr;
      B->graph_k = -1;
    }
  }
  DOE_HIP(hipMemsetAsync(B->hist, 0, sizeof(double) * 3 * size_t(B->hist_k) * B->count,
                         ctx->stream));
  size_t h = 0;
  for (auto& c : B->cls) {
    GsBuf<float>& b = c.b;
    b = GsBuf<float>{};
    b.illum = B->illum + c.off;
    b.tgt = B->tgt + c.off;
    b.roi = B->has_roi ? B->roi + c.off : nullptr;
    b.w = B->algorithm == DOE_ALG_WGS ? B->w + c.off : nullptr;
    b.phase[0] = B->ph[0] + c.off;
    b.phase[1] = B->ph[1] + c.off;
    b.amp[0] = B->am[0] + c.off;
    b.amp[1] = B->am[1] + c.off;
    b.state = c.state;
    c.hist_off = h;
    b.err_hist = B->hist + h;
    b.eff_hist = B->hist + h + B->hist_k;
    b.scr_hist = B->hist + h + 2 * size_t(B->hist_k);
    b.hstride = 3L * B->hist_k;
    h += size_t(3) * B->hist_k * c.users.size();
    b.nx = c.n;
    b.ny = c.n;
    b.fstride = long(c.n) * c.n;
    b.norm = 1 / std::sqrt(double(c.n) * double(c.n));
    b.gsign = (((c.n >> 1) + (c.n >> 1)) & 1) ? -1.0 : 1.0;
    b.beta = B->beta;
    b.weighted = B->algorithm == DOE_ALG_WGS ? 1 : 0;
    int rc = make_plan2(ctx, c.n, c.n, sizeof(float), col_split<float>(c.n, c.n), &c.pl);
    if (rc) return rc;
  }
  return DOE_OK;
}

// Start of a run: states, weights, and x~ = c .* illum .* exp(i phi0) after
// the row FFT, per size class.
int gsb_start(doe_gs_batch* B) {
  doe_ctx* ctx = B->ctx;
  for (auto& c : B->cls) {
    const int nd = int(c.users.size());
    const long n = long(c.n) * c.n * nd;
    hipLaunchKernelGGL(k_gs_init_state<float>, dim3(unsigned((std::max<long>(n, nd) + 255) / 256)),
                       dim3(256), 0, ctx->stream, c.state, nd, c.b.w, c.b.roi, n, c.b.weighted);
    hipLaunchKernelGGL(k_gs_consts<float>, dim3(unsigned(nd)), dim3(1024), 0, ctx->stream, c.state,
                       c.b.tgt, c.b.roi, long(c.n) * c.n, long(c.n) * c.n);
    DOE_HIP(hipGetLastError());
    OpFFInit<float> op0{c.b, B->phi0 + c.off};
    float* re = B->re + c.off;
    float* im = B->im + c.off;
    DOE_HIP((row_pass<float, float, 0, false, false, false, true>(
        ctx, gsb_tag(c.n, 2), op0, re, im, re, im, nd, c.pl, GsSizes{})));
  }
  return DOE_OK;
}

hipError_t gsb_iterations(doe_gs_batch* B, int K) {
  doe_ctx* ctx = B->ctx;
  const bool gs = B->algorithm == DOE_ALG_GS;
  for (int it = 0; it < K; ++it) {
    const int last = (gs && it == K - 1) ? 1 : 0;
    for (auto& c : B->cls) {
      float* re = B->re + c.off;
      float* im = B->im + c.off;
      OpFFB<float> opb{c.b, c.partials, it, last};
      const hipError_t e = col_pass<float, float, 1, false, true>(
          ctx, gsb_tag(c.n, 0), opb, re, im, re, im, int(c.users.size()), c.pl, GsSizes{});
      if (e != hipSuccess) return e;
    }
    for (auto& c : B->cls) {
      float* re = B->re + c.off;
      float* im = B->im + c.off;
      OpFFA<float> opa{c.b, last};
      const hipError_t e = row_pass<float, float, 1, true, false, true, true>(
          ctx, gsb_tag(c.n, 1), opa, re, im, re, im, int(c.users.size()), c.pl, GsSizes{});
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

bool gsb_merged_ok(const doe_gs_batch* B) {
  for (const auto& c : B->cls)
    if (c.n > 2048) return false;
  return int(B->cls.size()) <= kMergedMax;
}

// The iteration loop with one column-pass, one finalize and one row-pass
// launch per iteration for the whole batch.
hipError_t gsb_iterations_merged(doe_gs_batch* B, int K) {
  doe_ctx* ctx = B->ctx;
  const bool gs = B->algorithm == DOE_ALG_GS;
  MTable<MColEntry<OpFFB<float>>> ct{};
  MTable<MRowEntry<OpFFA<float>>> rt{};
  MTable<GsbFinEntry> ft{};
  const int ncls = int(B->cls.size());
  ct.n = rt.n = ft.n = ncls;
  int cb = 0, rb = 0, fb = 0;
  for (int k = 0; k < ncls; ++k) {
    auto& c = B->cls[k];
    const int nd = int(c.users.size());
    const int n1 = c.pl.n1, n2 = c.n / n1;
    const int cols = mcol_cols(n2);
    const int g = mrow_groups(c.n);
    if (cols <= 0 || g <= 0) return hipErrorInvalidValue;
    const int tiles = (c.n + cols - 1) / cols;
    float* re = B->re + c.off;
    float* im = B->im + c.off;
    ct.first[k] = cb;
    ct.e[k] = MColEntry<OpFFB<float>>{OpFFB<float>{c.b, c.partials, 0, 0}, re, im, re, im,
                                      (const float*)c.pl.tc.re, (const float*)c.pl.tc.im, n2,
                                      c.n, tiles, n1};
    cb += tiles * n1 * nd;
    const long groups = long(c.n / n1) * nd;
    rt.first[k] = rb;
    rt.e[k] = MRowEntry<OpFFA<float>>{OpFFA<float>{c.b, 0}, re, im, re, im,
                                      (const float*)c.pl.tx.re, (const float*)c.pl.tx.im,
                                      (const float*)c.pl.ty.re, (const float*)c.pl.ty.im, groups,
                                      c.n};
    rb += int((groups + g - 1) / g);
    ft.first[k] = fb;
    ft.e[k] = GsbFinEntry{OpFFB<float>{c.b, c.partials, 0, 0}, tiles * n1};
    fb += nd;
  }
  ct.first[ncls] = cb;
  rt.first[ncls] = rb;
  ft.first[ncls] = fb;
  for (int it = 0; it < K; ++it) {
    const int last = (gs && it == K - 1) ? 1 : 0;
    for (int k = 0; k < ncls; ++k) {
      ct.e[k].op.it = it;
      ct.e[k].op.write_amp = last;
      ft.e[k].op.it = it;
      ft.e[k].op.write_amp = last;
      rt.e[k].op.write_phase = last;
    }
    prof_begin(ctx, "gsb_col_merged", long(B->total));
    hipLaunchKernelGGL((k_col_merged<OpFFB<float>, 1, false, true>), dim3(unsigned(cb)),
                       dim3(kMWG), 0, ctx->stream, ct, next_dir(ctx));
    hipLaunchKernelGGL(k_gsb_fin_merged, dim3(unsigned(fb)), dim3(256), 0, ctx->stream, ft);
    prof_end(ctx);
    prof_begin(ctx, "gsb_row_merged", long(B->total));
    hipLaunchKernelGGL((k_row_merged<OpFFA<float>, 1, true, false, true, true>),
                       dim3(unsigned(rb)), dim3(kMWG), 0, ctx->stream, rt, next_dir(ctx));
    prof_end(ctx);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

int doe_gs_batch_create(doe_ctx* ctx, int count, const int* sides, int algorithm, double beta,
                        int with_roi, doe_gs_batch** out) {
  DOE_REQUIRE(ctx && sides && out, DOE_ERR_INVALID, "doe_gs_batch_create: null argument");
  *out = nullptr;
  DOE_REQUIRE(count > 0, DOE_ERR_INVALID, "doe_gs_batch_create: empty batch");
  DOE_REQUIRE(algorithm == DOE_ALG_GS || algorithm == DOE_ALG_WGS, DOE_ERR_INVALID,
              "doe_gs_batch_create: unknown algorithm");
  for (int i = 0; i < count; ++i)
    DOE_REQUIRE(sides[i] >= 64 && sides[i] <= 4096 && is_pow2(sides[i]), DOE_ERR_UNSUPPORTED,
                "doe_gs_batch_create: sides must be powers of two in [64, 4096]");
  DOE_HIP(hipSetDevice(ctx->device));
  auto* B = new doe_gs_batch();
  B->ctx = ctx;
  B->count = count;
  B->algorithm = algorithm;
  B->beta = beta;
  B->has_roi = with_roi != 0;
  B->side.assign(sides, sides + count);
  B->where.resize(count);
  std::map<int, std::vector<int>, std::greater<int>> by_side;
  for (int i = 0; i < count; ++i) by_side[sides[i]].push_back(i);
  size_t o = 0;
  for (auto& kv : by_side) {
    doe_gs_batch::Class c;
    c.n = kv.first;
    c.users = kv.second;
    c.off = o;
    for (size_t d = 0; d < c.users.size(); ++d)
      B->where[c.users[d]] = {int(B->cls.size()), int(d)};
    o += size_t(c.n) * c.n * c.users.size();
    B->cls.push_back(c);
  }
  B->total = o;
  const size_t tb = o * sizeof(float);
  void* p = nullptr;
  int rc = DOE_OK;
  float** planes[] = {&B->re, &B->im, &B->illum, &B->tgt, &B->phi0,
                      &B->ph[0], &B->ph[1], &B->am[0], &B->am[1]};
  for (float** pl : planes) {
    if (!rc) rc = B->alloc(&p, tb);
    if (!rc) *pl = (float*)p;
  }
  if (!rc && B->has_roi) {
    rc = B->alloc(&p, tb);
    if (!rc) B->roi = (float*)p;
  }
  if (!rc && algorithm == DOE_ALG_WGS) {
    rc = B->alloc(&p, tb);
    if (!rc) B->w = (float*)p;
  }
  for (auto& c : B->cls) {
    const size_t nd = c.users.size();
    // column-pass workgroups per design bound the partial slots; row passes
    // of the far-field loop keep no sums
    const size_t slots = size_t(std::max(c.n, 4 * c.n)) * nd;
    if (!rc) rc = B->alloc(&p, sizeof(GsState) * nd);
    if (!rc) c.state = (GsState*)p;
    if (!rc) rc = B->alloc(&p, sizeof(double) * R_NFF * slots);
    if (!rc) c.partials = (double*)p;
  }
  if (!rc && (hipMemset(B->phi0, 0, tb) != hipSuccess || hipMemset(B->ph[0], 0, tb) != hipSuccess ||
              hipMemset(B->ph[1], 0, tb) != hipSuccess))
    rc = DOE_ERR_HIP;
  if (rc) {
    if (rc == DOE_ERR_HIP) set_error("doe_gs_batch_create: hipMemset failed");
    delete B;
    return rc;
  }
  *out = B;
  return DOE_OK;
}

int doe_gs_batch_destroy(doe_gs_batch* B) {
  if (!B) return DOE_OK;
  hipSetDevice(B->ctx->device);
  hipStreamSynchronize(B->ctx->stream);
  delete B;
  return DOE_OK;
}

int doe_gs_batch_set(doe_gs_batch* B, int index, const double* illumination,
                     const double* target_amp, const double* roi, const double* init_phase) {
  DOE_REQUIRE(B, DOE_ERR_INVALID, "doe_gs_batch_set: null batch");
  DOE_REQUIRE(index >= 0 && index < B->count, DOE_ERR_INVALID,
              "doe_gs_batch_set: index out of range");
  DOE_REQUIRE(illumination && target_amp, DOE_ERR_INVALID, "doe_gs_batch_set: null input");
  DOE_REQUIRE(!roi || B->has_roi, DOE_ERR_STATE, "doe_gs_batch_set: batch was created without ROI");
  DOE_REQUIRE(roi || !B->has_roi, DOE_ERR_INVALID,
              "doe_gs_batch_set: this batch needs an ROI for every design");
  DOE_HIP(hipSetDevice(B->ctx->device));
  const auto& c = B->cls[B->where[index].first];
  const long n = long(c.n) * c.n;
  const size_t o = c.off + size_t(B->where[index].second) * n;
  int rc = upload_t<float>(B->ctx, illumination, B->illum + o, n);
  if (!rc) rc = upload_t<float>(B->ctx, target_amp, B->tgt + o, n);
  if (!rc && roi) rc = upload_t<float>(B->ctx, roi, B->roi + o, n);
  if (!rc && init_phase) rc = upload_t<float>(B->ctx, init_phase, B->phi0 + o, n);
  return rc;
}

int doe_gs_batch_random_init(doe_gs_batch* B, unsigned long long seed) {
  DOE_REQUIRE(B, DOE_ERR_INVALID, "doe_gs_batch_random_init: null batch");
  DOE_HIP(hipSetDevice(B->ctx->device));
  for (int i = 0; i < B->count; ++i) {
    const auto& c = B->cls[B->where[i].first];
    const long n = long(c.n) * c.n;
    hipLaunchKernelGGL(k_gsb_random_phase, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                       B->ctx->stream, B->phi0 + c.off + size_t(B->where[i].second) * n, n, seed,
                       (unsigned long long)i);
  }
  DOE_HIP(hipGetLastError());
  DOE_HIP(hipStreamSynchronize(B->ctx->stream));
  return DOE_OK;
}

int doe_gs_batch_run(doe_gs_batch* B, int iterations, int flags) {
  DOE_REQUIRE(B, DOE_ERR_INVALID, "doe_gs_batch_run: null batch");
  DOE_REQUIRE(iterations >= 0, DOE_ERR_INVALID, "doe_gs_batch_run: negative iteration count");
  doe_ctx* ctx = B->ctx;
  DOE_HIP(hipSetDevice(ctx->device));
  int rc = gsb_prepare(B, iterations > 0 ? iterations : 1);
  if (rc) return rc;
  rc = gsb_start(B);
  if (rc) return rc;
  // one launch per pass for the whole batch where every side allows it
  const bool merged = !(flags & DOE_GSB_PER_CLASS) && gsb_merged_ok(B);
  auto loop = [&](int k) { return merged ? gsb_iterations_merged(B, k) : gsb_iterations(B, k); };
  if (iterations > 0) {
    if (flags & DOE_GSB_GRAPH) {
      // captured once per iteration count and replayed; the captured launches
      // keep their arguments (buffers are fixed for the batch's lifetime)
      if (!B->graph || B->graph_k != iterations || B->graph_merged != merged) {
        if (B->graph) {
          hipGraphExecDestroy(B->graph);
          B->graph = nullptr;
        }
        const int saved = ctx->flags;
        ctx->flags &= ~DOE_CTX_PROFILE;
        hipGraph_t g = nullptr;
        DOE_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = loop(iterations);
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        ctx->flags = saved;
        DOE_HIP(le);
        DOE_HIP(ce);
        const hipError_t ie = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        DOE_HIP(ie);
        B->graph_k = iterations;
        B->graph_merged = merged;
      }
      prof_begin(ctx, "gsb_graph", long(B->total) * iterations);
      DOE_HIP(hipGraphLaunch(B->graph, ctx->stream));
      prof_end(ctx);
    } else {
      DOE_HIP(loop(iterations));
    }
  }
  B->last_k = iterations;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_gs_batch_get(doe_gs_batch* B, int index, double* phase_out, double* error_out,
                     double* efficiency_out, double* amplitude_out, double* weights_out,
                     double* best_rmse_out) {
  DOE_REQUIRE(B, DOE_ERR_INVALID, "doe_gs_batch_get: null batch");
  DOE_REQUIRE(index >= 0 && index < B->count, DOE_ERR_INVALID,
              "doe_gs_batch_get: index out of range");
  doe_ctx* ctx = B->ctx;
  DOE_HIP(hipSetDevice(ctx->device));
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  const auto& c = B->cls[B->where[index].first];
  const int d = B->where[index].second;
  const long n = long(c.n) * c.n;
  const size_t o = c.off + size_t(d) * n;
  const bool weighted = B->algorithm == DOE_ALG_WGS;
  GsState st{};
  DOE_HIP(hipMemcpy(&st, c.state + d, sizeof(st), hipMemcpyDeviceToHost));
  int rc = DOE_OK;
  if (phase_out) rc = download_t<float>(ctx, B->ph[0] + o, phase_out, n);
  if (!rc && amplitude_out)
    rc = download_t<float>(ctx, B->am[weighted ? st.best_amp_slot : 0] + o, amplitude_out, n);
  if (!rc && weights_out) {
    if (weighted) {
      rc = download_t<float>(ctx, B->w + o, weights_out, n);
      if (!rc)
        for (long i = 0; i < n; ++i) weights_out[i] *= st.nrm;
    } else {
      std::fill(weights_out, weights_out + n, 0.0);
    }
  }
  if (rc) return rc;
  const int K = B->last_k;
  if ((error_out || efficiency_out) && K > 0) {
    std::vector<double> h(3 * size_t(B->hist_k));
    DOE_HIP(hipMemcpy(h.data(), B->hist + c.hist_off + size_t(d) * 3 * B->hist_k,
                      sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    if (error_out) std::copy(h.begin(), h.begin() + K, error_out);
    if (efficiency_out) std::copy(h.begin() + B->hist_k, h.begin() + B->hist_k + K, efficiency_out);
  }
  if (best_rmse_out) *best_rmse_out = weighted ? st.best_rmse : std::nan("");
  return DOE_OK;
}

int doe_gs_batch_get_screen(doe_gs_batch* B, int index, double* screen_rmse_out) {
  DOE_REQUIRE(B && screen_rmse_out, DOE_ERR_INVALID, "doe_gs_batch_get_screen: null argument");
  DOE_REQUIRE(index >= 0 && index < B->count, DOE_ERR_INVALID,
              "doe_gs_batch_get_screen: index out of range");
  DOE_HIP(hipSetDevice(B->ctx->device));
  DOE_HIP(hipStreamSynchronize(B->ctx->stream));
  const auto& c = B->cls[B->where[index].first];
  const int d = B->where[index].second;
  if (B->last_k > 0)
    DOE_HIP(hipMemcpy(screen_rmse_out,
                      B->hist + c.hist_off + size_t(d) * 3 * B->hist_k + 2 * size_t(B->hist_k),
                      sizeof(double) * B->last_k, hipMemcpyDeviceToHost));
  return DOE_OK;
}

}  // extern "C"
