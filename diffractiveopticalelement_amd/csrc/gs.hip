// Phase retrieval: Gerchberg–Saxton / weighted GS in the image-plane
// (angular-spectrum) and far-field (unitary DFT) geometries.
//
// ref: src/doe.js:327-437 gerchbergSaxton, 455-546 weightedGerchbergSaxton,
//      639-730 fourierCgh.
//
// Far-field loop (2 passes per iteration).  With c = (-1)^(x+y) and
// g = (-1)^(nx/2+ny/2), fftshift.F.ifftshift(x) = g*c.*F(c.*x) at the same
// index, so the loop runs on x~ = c.*x: the spectrum F(x~) is already in the
// reference's centred index space (target/ROI/weights need no remap), and the
// update x~' = illum .* z/|z| with z = F^-1(Y~') carries no checkerboard at
// all.  Pass B (column blocks): F -> measure, project, WGS update -> F^-1.
// Pass A (rows): F^-1 -> phase, re-impose illumination -> F.
//
// Image-plane loop.  Each angular-spectrum propagation is a row pass
// (inverse rows -> spatial-domain work -> forward rows, with the column
// split's cross-row steps) and a column pass (F . H or conj H . F^-1 on the
// column blocks).  Under WGS the weight renormalisation is a grid-wide scalar
// needed before the projection, so the measurement (A1a) and projection
// (A1b) are separate row passes.
#include <cmath>
#include <limits>
#include <vector>

#include "gs_ops.hpp"
#include "rng.hpp"

namespace doe {

int check_field(const doe_ctx* ctx, const doe_field* f);

// ---------------------------------------------------------------------------
// Image plane
// ---------------------------------------------------------------------------
// Initial column pass: x = illum .* exp(i*phi0), then column FFT.
template <typename T>
struct OpIPInit {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = true, HAS_POST = false;
  GsBuf<T> b;
  const T* phi0;
  template <typename U>
  __device__ __forceinline__ void pre(int, int x, int y, U& re, U& im, double*) const {
    const long k = (long)y * b.nx + x;
    const T il = b.illum[k];
    T s, c;
    if constexpr (sizeof(T) == 8) sincos(phi0[k], &s, &c);
    else sincosf(phi0[k], &s, &c);
    re = il * c;
    im = il * s;
  }
  template <typename U>
  __device__ __forceinline__ void mid(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void post(int, int, int, U&, U&, double*) const {}
};

// GS measurement + projection of doe.js:359-397 on the target-plane field U
// with weights w = w' * nrm (constant during the pass).
template <typename T>
__device__ __forceinline__ void ip_measure_project(const GsBuf<T>& b, int it, long k, T& re, T& im,
                                                   double* acc) {
  const T a = sqrt(re * re + im * im);
  const T i2 = a * a;
  const T m = b.roi ? b.roi[k] : T(1);
  T w = T(1);
  if (b.weighted) w = b.w[k] * T(b.state->nrm);
  const T t = b.tgt[k] * w;
  acc[R_TOTAL] += double(i2);
  if (m > T(0)) {
    acc[R_INROI] += double(i2) * double(m);
    if (t > T(0) || i2 > T(0)) {
      const double d = double(a) * double(m) - double(t) * double(m);
      acc[R_ERRNUM] += d * d;
      acc[R_ERRDEN] += double(t) * double(t);
    }
  }
  const T fb = it >= b.escape ? T(b.feedback) : T(0);
  T na;
  if (m > T(0)) {
    na = t;
    if (fb > T(0) && a > T(0)) na = t + fb * (a - t);
  } else {
    na = a;  // a*(1-fb) + fb*a: the field is left as it is outside the ROI
  }
  if (a > T(0)) {
    const T s = na / a;
    re *= s;
    im *= s;
  }
}

template <typename T>
__device__ __forceinline__ void ip_finalize_measure(const GsBuf<T>& b, int it, const double* s) {
#pragma clang fp contract(off)
  const double err = s[R_ERRDEN] > 0 ? sqrt(s[R_ERRNUM] / s[R_ERRDEN]) : 0.0;
  const double eff =
      s[R_TOTAL] > 0 ? (s[R_INROI] * b.dxdy_x * b.dxdy_y) / (s[R_TOTAL] * b.dxdy_x * b.dxdy_y) : 0.0;
  if (!b.weighted) b.err_hist[it] = err;
  b.eff_hist[it] = eff;
}

// A1 (GS, or WGS iteration 0): column F^-1 -> measure + project -> column F.
template <typename T>
struct OpIPMP {
  static constexpr int NACC = R_N;
  static constexpr bool HAS_PRE = false, HAS_POST = false;
  GsBuf<T> b;
  double* partials;
  int it;
  template <typename U>
  __device__ __forceinline__ void pre(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void post(int, int, int, U&, U&, double*) const {}
  __device__ __forceinline__ void mid(int, int x, int y, T& re, T& im, double* acc) const {
    ip_measure_project<T>(b, it, (long)y * b.nx + x, re, im, acc);
  }
  __device__ __forceinline__ void finalize(int, const double* s) const { ip_finalize_measure<T>(b, it, s); }
};

// A1b (WGS): the pass loads the stored target-plane field U; measure +
// project with the renormalised weights, then column F.
template <typename T>
struct OpIPProj {
  static constexpr int NACC = R_N;
  static constexpr bool HAS_PRE = true, HAS_POST = false;
  GsBuf<T> b;
  double* partials;
  int it;
  __device__ __forceinline__ void pre(int, int x, int y, T& re, T& im, double* acc) const {
    ip_measure_project<T>(b, it, (long)y * b.nx + x, re, im, acc);
  }
  template <typename U>
  __device__ __forceinline__ void mid(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void post(int, int, int, U&, U&, double*) const {}
  __device__ __forceinline__ void finalize(int, const double* s) const { ip_finalize_measure<T>(b, it, s); }
};

// A1a (WGS, and the closing pass of GS): column F^-1 -> U stored spatially;
// amplitude of the reported iterate; WGS best-iterate rmse of the previous
// iteration (prev_it) and the weight update (doe.js:490-524).
template <typename T>
struct OpIPMeas {
  static constexpr int NACC = R_N;
  static constexpr bool HAS_PRE = false, HAS_POST = true;
  GsBuf<T> b;
  double* partials;
  int prev_it;
  template <typename U>
  __device__ __forceinline__ void pre(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void mid(int, int, int, U&, U&, double*) const {}
  __device__ __forceinline__ void post(int, int x, int y, T& re, T& im, double* acc) const {
    const long k = (long)y * b.nx + x;
    const T a = sqrt(re * re + im * im);
    if (!b.weighted) {
      b.amp[0][k] = a;
      return;
    }
    b.amp[1 - b.state->best_amp_slot][k] = a;
    const T m = b.roi ? b.roi[k] : T(1);
    if (m > T(0)) {
      const T t0 = b.tgt[k];
      const double d0 = double(a) - double(t0);
      acc[R_RNUM] += d0 * d0;
      acc[R_RDEN] += double(t0) * double(t0);
      const T w = b.w[k] * T(b.state->nrm);
      const T wn = wgs_update<T>(w, a, t0, T(b.beta));
      b.w[k] = wn;
      acc[R_WSUM] += double(wn) * double(wn) * double(t0) * double(t0);
      acc[R_TSUM] += double(t0) * double(t0);
    }
  }
  __device__ __forceinline__ void finalize(int, const double* s) const {
    if (!b.weighted) return;
    GsState* st = b.state;
    const double rmse = s[R_RDEN] > 0 ? sqrt(s[R_RNUM] / s[R_RDEN]) : 0.0;
    if (rmse < st->best_rmse) {
      st->best_rmse = rmse;
      st->best_phase_slot = 1 - st->best_phase_slot;
      st->best_amp_slot = 1 - st->best_amp_slot;
    }
    st->nrm = s[R_WSUM] > 0 ? sqrt(s[R_TSUM] / s[R_WSUM]) : 1.0;
    b.err_hist[prev_it] = st->best_rmse;
  }
};

// A2: column F^-1 -> phase -> illumination re-imposed -> column F.
template <typename T>
struct OpIPPhase {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = false, HAS_POST = false;
  GsBuf<T> b;
  template <typename U>
  __device__ __forceinline__ void pre(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void post(int, int, int, U&, U&, double*) const {}
  __device__ __forceinline__ void mid(int, int x, int y, T& re, T& im, double*) const {
    const long k = (long)y * b.nx + x;
    const int slot = b.weighted ? 1 - b.state->best_phase_slot : 0;
    b.phase[slot][k] = atan2(im, re);
    const T r = sqrt(re * re + im * im);
    const T il = b.illum[k];
    if (r > T(0)) {
      const T s = il / r;
      re *= s;
      im *= s;
    } else {
      re = il;  // atan2(0, 0) = 0
      im = T(0);
    }
  }
};

// Device workspace of one design.
template <typename T>
struct GsWork {
  std::vector<void*> allocs;
  GsBuf<T> b{};
  T* fre = nullptr;
  T* fim = nullptr;
  T* phi0 = nullptr;
  T* roi_d = nullptr;
  double* hist = nullptr;
  ~GsWork() {
    for (void* p : allocs) hipFree(p);
  }
  int alloc(void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("doe_gs_run: device allocation failed");
      return DOE_ERR_OOM;
    }
    allocs.push_back(*p);
    return DOE_OK;
  }
};

// Run a design loop's launches (`f` enqueues them on the context stream) as
// one captured graph, so a long loop of small passes is not bound by the
// host's launch rate; eagerly when per-launch profiling is on or the loop is
// short.  The loops capture kernel launches only (no memsets or copies).
template <class F>
static int run_captured(doe_ctx* ctx, long launches, F&& f) {
  if ((ctx->flags & DOE_CTX_PROFILE) || launches < 16) return f();
  hipGraph_t g = nullptr;
  DOE_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = f();
  const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  DOE_HIP(ce);
  hipGraphExec_t ge = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  DOE_HIP(ie);
  const hipError_t le = hipGraphLaunch(ge, ctx->stream);
  const hipError_t se = hipStreamSynchronize(ctx->stream);
  (void)hipGraphExecDestroy(ge);
  DOE_HIP(le);
  DOE_HIP(se);
  return DOE_OK;
}

template <typename T>
int run_gs(doe_ctx* ctx, const doe_gs_plan* plan, const double* illum, const double* tgt,
                  const double* roi, const double* init_phase, unsigned seed, double* phase_out,
                  double* error_out, double* efficiency_out, double* amplitude_out,
                  double* intensity_out, double* weights_out, double* best_rmse_out) {
  const int nx = plan->nx, ny = plan->ny, K = plan->iterations;
  const long n = (long)nx * ny;
  const bool weighted = plan->algorithm == DOE_ALG_WGS;
  const bool far = plan->mode == DOE_MODE_FAR_FIELD;
  GsWork<T> W;
  const size_t tb = size_t(n) * sizeof(T);
  void* p = nullptr;
  int rc;
#define ALLOC(dst, bytes)                         \
  do {                                            \
    rc = W.alloc(&p, (bytes));                    \
    if (rc) return rc;                            \
    dst = (decltype(dst))p;                       \
  } while (0)
  T *illum_d, *tgt_d, *w_d = nullptr;
  ALLOC(W.fre, tb);
  ALLOC(W.fim, tb);
  ALLOC(illum_d, tb);
  ALLOC(tgt_d, tb);
  ALLOC(W.phi0, tb);
  if (roi) ALLOC(W.roi_d, tb);
  if (weighted) ALLOC(w_d, tb);
  T* ph[2];
  T* am[2];
  ALLOC(ph[0], tb);
  ALLOC(ph[1], tb);
  ALLOC(am[0], tb);
  ALLOC(am[1], tb);
  GsState* st_d;
  ALLOC(st_d, sizeof(GsState));
  ALLOC(W.hist, sizeof(double) * 2 * size_t(K > 0 ? K : 1));
#undef ALLOC
  rc = upload_t<T>(ctx, illum, illum_d, n);
  if (!rc) rc = upload_t<T>(ctx, tgt, tgt_d, n);
  if (!rc && roi) rc = upload_t<T>(ctx, roi, W.roi_d, n);
  if (!rc && init_phase) rc = upload_t<T>(ctx, init_phase, W.phi0, n);
  if (rc) return rc;
  if (!init_phase) {
    if (far)
      DOE_HIP(launch_random_phase(ctx, W.phi0, n, seed));  // fourierCgh default (doe.js:651)
    else
      DOE_HIP(hipMemsetAsync(W.phi0, 0, tb, ctx->stream));  // flat start
  }
  DOE_HIP(hipMemsetAsync(ph[0], 0, tb, ctx->stream));
  DOE_HIP(hipMemsetAsync(ph[1], 0, tb, ctx->stream));
  DOE_HIP(hipMemsetAsync(W.hist, 0, sizeof(double) * 2 * size_t(K > 0 ? K : 1), ctx->stream));
  hipLaunchKernelGGL(k_gs_init_state<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     st_d, 1, w_d, (const T*)W.roi_d, n, weighted ? 1 : 0);
  hipLaunchKernelGGL(k_gs_consts<T>, dim3(1), dim3(1024), 0, ctx->stream, st_d, (const T*)tgt_d,
                     (const T*)W.roi_d, n, n);
  DOE_HIP(hipGetLastError());

  GsBuf<T>& b = W.b;
  b.illum = illum_d;
  b.tgt = tgt_d;
  b.roi = W.roi_d;
  b.w = w_d;
  b.phase[0] = ph[0];
  b.phase[1] = ph[1];
  b.amp[0] = am[0];
  b.amp[1] = am[1];
  b.state = st_d;
  b.err_hist = W.hist;
  b.eff_hist = W.hist + (K > 0 ? K : 1);
  b.nx = nx;
  b.ny = ny;
  b.dxdy_x = plan->dx;
  b.dxdy_y = plan->dy;
  b.norm = 1 / std::sqrt(double(nx) * double(ny));
  b.gsign = (((nx >> 1) + (ny >> 1)) & 1) ? -1.0 : 1.0;
  b.beta = plan->beta;
  b.feedback = plan->feedback;
  b.escape = plan->escape;
  b.weighted = weighted ? 1 : 0;
  b.fstride = n;
  b.hstride = 0;

  Plan2 pl;
  rc = make_plan2(ctx, nx, ny, sizeof(T), col_split<T>(nx, ny), &pl);
  if (rc) return rc;
  // reduction scratch: partial sums per workgroup of a reducing row or
  // column pass
  rc = ensure_reduce(ctx, size_t(std::max(ny, 4 * nx)) * R_NFF);
  if (rc) return rc;
  T* re = W.fre;
  T* im = W.fim;
  const GsSizes sz{};

  if (far) {
    // x~ = c .* illum .* exp(i phi0) -> rows forward
    OpFFInit<T> op0{b, W.phi0};
    DOE_HIP((row_pass<T, T, 0, false, false, false, true>(ctx, "ff_init_row", op0, re, im, re, im, 1,
                                                            pl, sz)));
    rc = run_captured(ctx, 3L * K, [&]() -> int {
      for (int it = 0; it < K; ++it) {
        OpFFB<T> opb{b, ctx->partials, it, it == K - 1 ? 1 : 0};
        DOE_HIP((col_pass<T, T, 1, false, true>(ctx, "ff_col_project", opb, re, im, re, im, 1, pl,
                                                sz)));
        OpFFA<T> opa{b, it == K - 1 ? 1 : 0};
        DOE_HIP((row_pass<T, T, 1, true, false, true, true>(ctx, "ff_row_phase", opa, re, im, re, im,
                                                             1, pl, sz)));
      }
      return DOE_OK;
    });
    if (rc) return rc;
  } else {
    OpAsmH h{};
    std::vector<AsmField> hp(1);
    hp[0].two_pi_z = 2 * kPi * plan->z;
    hp[0].z = plan->z;
    hp[0].lim_x = hp[0].lim_y = hp[0].inner_x = hp[0].inner_y = 0;
    hp[0].band = 0;  // the design loop never band-limits (doe.js:352, 400, 419)
    hp[0].zero = plan->z == 0 ? 1 : 0;
    void* dp = nullptr;
    rc = upload_params(ctx, hp.data(), sizeof(AsmField), &dp);
    if (rc) return rc;
    h.fp = (const AsmField*)dp;
    h.dfx = 1 / (nx * plan->dx);
    h.dfy = 1 / (ny * plan->dy);
    h.inv_l2 = 1 / (plan->lambda * plan->lambda);
    h.nx = nx;
    h.ny = ny;
    h.evanescent = 1;
    OpAsmH hc = h;
    hc.conj = 1;
    h.conj = 0;
    // x = illum .* exp(i phi0) -> rows forward -> column blocks F . H . F^-1
    OpIPInit<T> op0{b, W.phi0};
    DOE_HIP((row_pass<T, T, 0, false, false, false, true>(ctx, "ip_init_row", op0, re, im, re, im, 1,
                                                           pl, sz)));
    DOE_HIP((col_pass<T, T, 1, false, true>(ctx, "ip_col_h", h, re, im, re, im, 1, pl, sz)));
    rc = run_captured(ctx, 7L * K, [&]() -> int {
      for (int it = 0; it < K; ++it) {
        if (!weighted || it == 0) {
          OpIPMP<T> op{b, ctx->partials, it};
          DOE_HIP((row_pass<T, T, 1, true, false, true, true>(ctx, "ip_row_project", op, re, im, re,
                                                               im, 1, pl, sz)));
        } else {
          OpIPMeas<T> om{b, ctx->partials, it - 1};
          DOE_HIP((row_pass<T, T, 0, true, false, true, false>(ctx, "ip_row_measure", om, re, im, re,
                                                                im, 1, pl, sz)));
          OpIPProj<T> op{b, ctx->partials, it};
          DOE_HIP((row_pass<T, T, 0, false, false, false, true>(ctx, "ip_row_reproject", op, re, im,
                                                                 re, im, 1, pl, sz)));
        }
        DOE_HIP((col_pass<T, T, 1, false, true>(ctx, "ip_col_hconj", hc, re, im, re, im, 1, pl,
                                                sz)));
        OpIPPhase<T> opp{b};
        DOE_HIP((row_pass<T, T, 1, true, false, true, true>(ctx, "ip_row_phase", opp, re, im, re, im,
                                                             1, pl, sz)));
        DOE_HIP((col_pass<T, T, 1, false, true>(ctx, "ip_col_h", h, re, im, re, im, 1, pl, sz)));
      }
      // closing measurement: GS final forward amplitude (doe.js:418-426); WGS
      // best check + weight update of the last iteration.
      OpIPMeas<T> om{b, ctx->partials, K - 1};
      DOE_HIP((row_pass<T, T, 0, true, false, true, false>(ctx, "ip_row_measure", om, re, im, re, im,
                                                            1, pl, sz)));
      return DOE_OK;
    });
    if (rc) return rc;
  }

  // ---- outputs ----------------------------------------------------------
  GsState st{};
  DOE_HIP(hipMemcpyAsync(&st, st_d, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  const int pslot = (weighted && !far) ? st.best_phase_slot : 0;
  const int aslot = weighted ? st.best_amp_slot : 0;
  rc = download_t<T>(ctx, ph[pslot], phase_out, n);
  if (rc) return rc;
  if (amplitude_out || intensity_out) {
    std::vector<double> a(n);
    rc = download_t<T>(ctx, am[aslot], a.data(), n);
    if (rc) return rc;
    for (long i = 0; i < n; ++i) {
      if (amplitude_out) amplitude_out[i] = a[i];
      if (intensity_out) intensity_out[i] = a[i] * a[i];
    }
  }
  if (K > 0) {
    std::vector<double> hist(2 * size_t(K));
    DOE_HIP(hipMemcpy(hist.data(), W.hist, sizeof(double) * 2 * size_t(K), hipMemcpyDeviceToHost));
    if (error_out) std::copy(hist.begin(), hist.begin() + K, error_out);
    if (efficiency_out) std::copy(hist.begin() + K, hist.end(), efficiency_out);
  }
  if (weights_out) {
    if (weighted) {
      rc = download_t<T>(ctx, w_d, weights_out, n);
      if (rc) return rc;
      for (long i = 0; i < n; ++i) weights_out[i] *= st.nrm;
    } else {
      std::fill(weights_out, weights_out + n, 0.0);
    }
  }
  if (best_rmse_out) *best_rmse_out = weighted ? st.best_rmse : std::numeric_limits<double>::quiet_NaN();
  return DOE_OK;
}

// Multi-plane GS (gerchbergSaxtonMultiPlane, doe.js:557-626): every
// iteration visits the planes in order; at plane p the element
// illum .* exp(i phi) is propagated by z_p, measured and projected on plane
// p's target and ROI, propagated back by -z_p, and its phase becomes phi.
// This is the image-plane GS step above with per-plane transfer functions,
// targets and ROIs, so the same passes run with the plane's parameters: the
// column pass that closes a step already applies H(z_{p+1}) of the next one.
template <typename T>
int run_gs_multiplane(doe_ctx* ctx, const doe_gs_plan* plan, const double* illum, int P,
                      const doe_mp_plane* planes, const double* init_phase, double* phase_out,
                      double* error_out, double* efficiency_out, double* per_plane_out) {
  const int nx = plan->nx, ny = plan->ny, K = plan->iterations;
  const long n = (long)nx * ny;
  GsWork<T> W;
  const size_t tb = size_t(n) * sizeof(T);
  void* p = nullptr;
  int rc;
#define ALLOC(dst, bytes)                         \
  do {                                            \
    rc = W.alloc(&p, (bytes));                    \
    if (rc) return rc;                            \
    dst = (decltype(dst))p;                       \
  } while (0)
  T *illum_d, *tgt_d, *roi_d, *ph, *am;
  ALLOC(W.fre, tb);
  ALLOC(W.fim, tb);
  ALLOC(illum_d, tb);
  ALLOC(W.phi0, tb);
  ALLOC(ph, tb);
  ALLOC(am, tb);
  ALLOC(tgt_d, tb * size_t(P));
  ALLOC(roi_d, tb * size_t(P));
  GsState* st_d;
  ALLOC(st_d, sizeof(GsState));
  const size_t hk = size_t(K > 0 ? K : 1);
  ALLOC(W.hist, sizeof(double) * hk * size_t(P + 1));  // per-plane error, then efficiency scratch
  AsmField* hp_d;
  ALLOC(hp_d, sizeof(AsmField) * size_t(P));
#undef ALLOC
  rc = upload_t<T>(ctx, illum, illum_d, n);
  std::vector<double> ones;
  for (int q = 0; q < P && !rc; ++q) {
    rc = upload_t<T>(ctx, planes[q].target_amp, tgt_d + size_t(q) * n, n);
    if (!rc) {
      if (!planes[q].roi && ones.empty()) ones.assign(n, 1.0);
      rc = upload_t<T>(ctx, planes[q].roi ? planes[q].roi : ones.data(), roi_d + size_t(q) * n, n);
    }
  }
  std::vector<double> phi;
  if (!init_phase) phi.assign(n, 0.0);  // flatPhase (doe.js:572)
  if (!rc) rc = upload_t<T>(ctx, init_phase ? init_phase : phi.data(), W.phi0, n);
  if (rc) return rc;
  std::vector<AsmField> hp(P);
  for (int q = 0; q < P; ++q) {
    hp[q].two_pi_z = 2 * kPi * planes[q].z;
    hp[q].z = planes[q].z;
    hp[q].lim_x = hp[q].lim_y = hp[q].inner_x = hp[q].inner_y = 0;
    hp[q].band = 0;  // bandLimit: false (doe.js:580, 605)
    hp[q].zero = planes[q].z == 0 ? 1 : 0;
  }
  DOE_HIP(hipMemcpyAsync(hp_d, hp.data(), sizeof(AsmField) * P, hipMemcpyHostToDevice, ctx->stream));
  DOE_HIP(hipMemsetAsync(W.hist, 0, sizeof(double) * hk * size_t(P + 1), ctx->stream));
  hipLaunchKernelGGL(k_gs_init_state<T>, dim3(1), dim3(256), 0, ctx->stream, st_d, 1, (T*)nullptr,
                     (const T*)nullptr, 0L, 0);
  DOE_HIP(hipGetLastError());

  GsBuf<T> b{};
  b.illum = illum_d;
  b.w = nullptr;
  b.phase[0] = b.phase[1] = ph;
  b.amp[0] = b.amp[1] = am;
  b.state = st_d;
  b.eff_hist = W.hist + hk * size_t(P);
  b.nx = nx;
  b.ny = ny;
  b.dxdy_x = plan->dx;
  b.dxdy_y = plan->dy;
  b.norm = 1 / std::sqrt(double(nx) * double(ny));
  b.gsign = 1.0;
  b.beta = 0;
  b.feedback = 0;
  b.escape = std::numeric_limits<int>::max();
  b.weighted = 0;
  b.fstride = n;
  b.hstride = 0;

  Plan2 pl;
  rc = make_plan2(ctx, nx, ny, sizeof(T), col_split<T>(nx, ny), &pl);
  if (rc) return rc;
  rc = ensure_reduce(ctx, size_t(std::max(ny, 4 * nx)) * R_NFF);
  if (rc) return rc;
  T* re = W.fre;
  T* im = W.fim;
  const GsSizes sz{};
  OpAsmH h{};
  h.dfx = 1 / (nx * plan->dx);
  h.dfy = 1 / (ny * plan->dy);
  h.inv_l2 = 1 / (plan->lambda * plan->lambda);
  h.nx = nx;
  h.ny = ny;
  h.evanescent = 1;
  auto hfor = [&](int q, int conj) {
    OpAsmH r = h;
    r.fp = hp_d + q;
    r.conj = conj;
    return r;
  };
  // element -> rows forward -> column blocks F . H(z_0) . F^-1
  OpIPInit<T> op0{b, W.phi0};
  DOE_HIP((row_pass<T, T, 0, false, false, false, true>(ctx, "mp_init_row", op0, re, im, re, im, 1, pl,
                                                         sz)));
  if (K > 0) DOE_HIP((col_pass<T, T, 1, false, true>(ctx, "mp_col_h", hfor(0, 0), re, im, re, im, 1, pl, sz)));
  rc = run_captured(ctx, 4L * K * P, [&]() -> int {
    for (int it = 0; it < K; ++it) {
      for (int q = 0; q < P; ++q) {
        GsBuf<T> bq = b;
        bq.tgt = tgt_d + size_t(q) * n;
        bq.roi = roi_d + size_t(q) * n;
        bq.err_hist = W.hist + hk * size_t(q);
        OpIPMP<T> op{bq, ctx->partials, it};
        DOE_HIP((row_pass<T, T, 1, true, false, true, true>(ctx, "mp_row_project", op, re, im, re, im, 1,
                                                             pl, sz)));
        DOE_HIP((col_pass<T, T, 1, false, true>(ctx, "mp_col_hconj", hfor(q, 1), re, im, re, im, 1, pl,
                                                sz)));
        OpIPPhase<T> opp{bq};
        DOE_HIP((row_pass<T, T, 1, true, false, true, true>(ctx, "mp_row_phase", opp, re, im, re, im, 1,
                                                             pl, sz)));
        const bool last = it == K - 1 && q == P - 1;
        if (!last) {
          const int nq = q + 1 < P ? q + 1 : 0;
          DOE_HIP((col_pass<T, T, 1, false, true>(ctx, "mp_col_h", hfor(nq, 0), re, im, re, im, 1, pl,
                                                  sz)));
        }
      }
    }
    return DOE_OK;
  });
  if (rc) return rc;
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  rc = download_t<T>(ctx, K > 0 ? ph : W.phi0, phase_out, n);
  if (rc) return rc;
  if (K > 0) {
    std::vector<double> hist(hk * size_t(P));
    DOE_HIP(hipMemcpy(hist.data(), W.hist, sizeof(double) * hist.size(), hipMemcpyDeviceToHost));
    for (int it = 0; it < K; ++it) {
      double e = 0;
      for (int q = 0; q < P; ++q) e += hist[size_t(q) * hk + it] * planes[q].weight;
      e /= P;
      if (error_out) error_out[it] = e;
      if (efficiency_out) efficiency_out[it] = 1 - e;
    }
    if (per_plane_out)
      for (int q = 0; q < P; ++q)
        std::copy(hist.begin() + size_t(q) * hk, hist.begin() + size_t(q) * hk + K, per_plane_out + size_t(q) * K);
  }
  return DOE_OK;
}

#define DOE_GS_INST(T)                                                                          \
  template int run_gs<T>(doe_ctx*, const doe_gs_plan*, const double*, const double*, const double*, \
                         const double*, unsigned, double*, double*, double*, double*, double*,    \
                         double*, double*);                                                      \
  template int run_gs_multiplane<T>(doe_ctx*, const doe_gs_plan*, const double*, int,            \
                                    const doe_mp_plane*, const double*, double*, double*, double*, \
                                    double*);
#if DOE_PART == 0
extern template int run_gs<float>(doe_ctx*, const doe_gs_plan*, const double*, const double*,
                                  const double*, const double*, unsigned, double*, double*, double*,
                                  double*, double*, double*, double*);
extern template int run_gs<double>(doe_ctx*, const doe_gs_plan*, const double*, const double*,
                                   const double*, const double*, unsigned, double*, double*,
                                   double*, double*, double*, double*, double*);
extern template int run_gs_multiplane<float>(doe_ctx*, const doe_gs_plan*, const double*, int,
                                             const doe_mp_plane*, const double*, double*, double*,
                                             double*, double*);
extern template int run_gs_multiplane<double>(doe_ctx*, const doe_gs_plan*, const double*, int,
                                              const doe_mp_plane*, const double*, double*, double*,
                                              double*, double*);
#elif DOE_PART == 1
DOE_GS_INST(float)
#elif DOE_PART == 2
DOE_GS_INST(double)
#endif

}  // namespace doe

#if DOE_PART == 0
using namespace doe;

extern "C" int doe_gs_run(doe_ctx* ctx, const doe_gs_plan* plan, const double* illumination,
                          const double* target_amp, const double* roi, const double* init_phase,
                          unsigned seed, double* phase_out, double* error_out,
                          double* efficiency_out, double* amplitude_out, double* intensity_out,
                          double* weights_out, double* best_rmse_out) {
  DOE_REQUIRE(ctx && plan && illumination && target_amp && phase_out, DOE_ERR_INVALID,
              "doe_gs_run: null argument");
  doe_field probe;
  probe.nx = plan->nx;
  probe.ny = plan->ny;
  probe.batch = 1;
  int rc = check_field(ctx, &probe);
  if (rc) return rc;
  DOE_REQUIRE(plan->nx >= 64 && plan->ny >= 64 && plan->nx <= 4096 && plan->ny <= 4096,
              DOE_ERR_UNSUPPORTED, "doe_gs_run: grid sides must be in [64, 4096]");
  DOE_REQUIRE(plan->iterations >= 0, DOE_ERR_INVALID, "doe_gs_run: negative iteration count");
  DOE_REQUIRE(plan->mode == DOE_MODE_IMAGE_PLANE || plan->mode == DOE_MODE_FAR_FIELD,
              DOE_ERR_INVALID, "doe_gs_run: unknown mode");
  DOE_REQUIRE(plan->algorithm == DOE_ALG_GS || plan->algorithm == DOE_ALG_WGS, DOE_ERR_INVALID,
              "doe_gs_run: unknown algorithm");
  DOE_HIP(hipSetDevice(ctx->device));
  const int saved = ctx->flags;
  ctx->flags |= DOE_CTX_ASYNC;  // the loop only synchronises at the end
  if (plan->dtype == DOE_F64)
    rc = run_gs<double>(ctx, plan, illumination, target_amp, roi, init_phase, seed, phase_out,
                        error_out, efficiency_out, amplitude_out, intensity_out, weights_out,
                        best_rmse_out);
  else if (plan->dtype == DOE_F32)
    rc = run_gs<float>(ctx, plan, illumination, target_amp, roi, init_phase, seed, phase_out,
                       error_out, efficiency_out, amplitude_out, intensity_out, weights_out,
                       best_rmse_out);
  else {
    set_error("doe_gs_run: dtype must be DOE_F32 or DOE_F64");
    rc = DOE_ERR_UNSUPPORTED;
  }
  ctx->flags = saved;
  hipStreamSynchronize(ctx->stream);
  return rc;
}

extern "C" int doe_gs_multiplane_run(doe_ctx* ctx, const doe_gs_plan* plan,
                                     const double* illumination, int nplanes,
                                     const doe_mp_plane* planes, const double* init_phase,
                                     double* phase_out, double* error_out, double* efficiency_out,
                                     double* per_plane_error_out) {
  DOE_REQUIRE(ctx && plan && illumination && phase_out, DOE_ERR_INVALID,
              "gerchbergSaxtonMultiPlane: null argument");
  DOE_REQUIRE(nplanes > 0 && planes, DOE_ERR_INVALID, "gerchbergSaxtonMultiPlane: no planes");
  for (int q = 0; q < nplanes; ++q)
    DOE_REQUIRE(planes[q].target_amp, DOE_ERR_INVALID, "gerchbergSaxtonMultiPlane: plane without target");
  doe_field probe;
  probe.nx = plan->nx;
  probe.ny = plan->ny;
  probe.batch = 1;
  int rc = check_field(ctx, &probe);
  if (rc) return rc;
  DOE_REQUIRE(plan->nx >= 64 && plan->ny >= 64 && plan->nx <= 4096 && plan->ny <= 4096,
              DOE_ERR_UNSUPPORTED, "gerchbergSaxtonMultiPlane: grid sides must be in [64, 4096]");
  DOE_REQUIRE(plan->iterations >= 0, DOE_ERR_INVALID, "gerchbergSaxtonMultiPlane: negative iteration count");
  DOE_HIP(hipSetDevice(ctx->device));
  const int saved = ctx->flags;
  ctx->flags |= DOE_CTX_ASYNC;
  if (plan->dtype == DOE_F64)
    rc = run_gs_multiplane<double>(ctx, plan, illumination, nplanes, planes, init_phase, phase_out,
                                   error_out, efficiency_out, per_plane_error_out);
  else if (plan->dtype == DOE_F32)
    rc = run_gs_multiplane<float>(ctx, plan, illumination, nplanes, planes, init_phase, phase_out,
                                  error_out, efficiency_out, per_plane_error_out);
  else {
    set_error("gerchbergSaxtonMultiPlane: dtype must be DOE_F32 or DOE_F64");
    rc = DOE_ERR_UNSUPPORTED;
  }
  ctx->flags = saved;
  hipStreamSynchronize(ctx->stream);
  return rc;
}
#endif  // DOE_PART == 0
