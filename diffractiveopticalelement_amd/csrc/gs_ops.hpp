// Shared pieces of the phase-retrieval engines (single design and batched):
// per-design state and buffers, the far-field loop's fused operations, and
// host<->device conversion helpers.
//
// ref: src/doe.js:455-546 weightedGerchbergSaxton, 639-730 fourierCgh.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "ctx.hpp"
#include "ops_prop.hpp"

namespace doe {

// Per-design device state (one per field of a batch).
struct GsState {
  double best_rmse;
  double nrm;  // weight normalisation: w = w' * nrm
  int best_phase_slot;
  int best_amp_slot;
  int improved;
  int pad;
  // sums of the design's target that stay constant over a run (far field):
  // t0^2 and t0^4 inside the ROI, t0^2 over the whole plane
  double t2_roi, t4_roi, t2_all;
};

template <typename T>
struct GsBuf {
  const T* illum;
  const T* tgt;
  const T* roi;  // nullable: whole plane
  T* w;          // unnormalised weights w' (nullable for GS)
  T* phase[2];
  T* amp[2];
  GsState* state;
  double* err_hist;
  double* eff_hist;
  double* scr_hist;  // far field: screen rmse per iteration (nullable)
  int nx, ny;
  double dxdy_x, dxdy_y;  // dx, dy for the image-plane efficiency expression
  double norm;            // far field: 1/sqrt(nx*ny)
  double gsign;           // far field: (-1)^(nx/2+ny/2)
  double beta;
  double feedback;
  int escape;
  int weighted;
  // a batch of equal-size designs: design f's planes start at f*fstride,
  // its state at state + f, its histories at err_hist/eff_hist + f*hstride
  long fstride;
  long hstride;
};

// reduction slots
enum { R_ERRNUM, R_ERRDEN, R_INROI, R_TOTAL, R_RNUM, R_RDEN, R_WSUM, R_TSUM, R_N };
// far field also accumulates the screen metric of the measured plane
// (projector.js:554-614 screenMetrics rmse: target irradiance T power-matched
// to the plane's ROI power, intensity RMS over the ROI): with I = a^2,
// T = t0^2 and s = sum(m*I) / sum(T), rmse^2 = (sum I^2 - 2 s sum I T +
// s^2 sum T^2) / (s^2 sum T^2) over m > 0.
enum { R_SI2 = R_N, R_SIT, R_ST2, R_TALL, R_NFF };
// the far-field column pass accumulates only what changes between
// iterations; the target's constant sums come from GsState
enum { FF_TOTAL, FF_INROI, FF_SI2, FF_SIT, FF_ERRNUM, FF_ERRDEN, FF_RNUM, FF_WSUM, FF_N };

template <typename T>
__device__ __forceinline__ T clampv(T v, T lo, T hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// WGS weight update of doe.js:514-521 for one ROI pixel; returns w'.
// FP32: the power as exp2(beta * log2(r)) on the hardware log2/exp2 (r is
// clamped to [1/4, 4], so log2(r) is in [-2, 2]).
template <typename T>
__device__ __forceinline__ T wgs_update(T w, T measured, T desired, T beta) {
  const T ratio = measured > T(1e-30) ? desired / measured : T(2);
  const T r = clampv(ratio, T(0.25), T(4));
  T f;
  if constexpr (sizeof(T) == 8) {
    f = pow(r, beta);
  } else {
    f = __builtin_amdgcn_exp2f(beta * __builtin_amdgcn_logf(r));
  }
  const T wn = w * f;
  return clampv(wn, T(1e-4), T(1e4));
}

// ---------------------------------------------------------------------------
// Far field
// ---------------------------------------------------------------------------
// Initial row pass: x~ = c .* illum .* exp(i*phi0), then row FFT.
template <typename T>
struct OpFFInit {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = true, HAS_POST = false;
  GsBuf<T> b;
  const T* phi0;
  template <typename U>
  __device__ __forceinline__ void pre(int f, int x, int y, U& re, U& im, double*) const {
    const long k = f * b.fstride + (long)y * b.nx + x;
    const T il = b.illum[k];
    T s, c;
    if constexpr (sizeof(T) == 8) sincos(phi0[k], &s, &c);
    else sincosf(phi0[k], &s, &c);
    const T cb = ((x + y) & 1) ? T(-1) : T(1);
    re = cb * il * c;
    im = cb * il * s;
  }
  template <typename U>
  __device__ __forceinline__ void mid(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void post(int, int, int, U&, U&, double*) const {}
};

// Column pass B: measurement, projection and WGS update (doe.js:675-712 and
// 490-521) on the centred-index spectrum.
template <typename T>
struct OpFFB {
  static constexpr int NACC = FF_N;
  static constexpr bool HAS_PRE = false, HAS_POST = false;
  // FP32 column tiles.  Lines up to 512 points: the plain transform's tile
  // (32 points per thread, 32 columns, one LDS plane) with the per-point
  // inputs (ROI, target, weight) loaded with the column before the first
  // transform (PREFETCH), so their latency overlaps the field loads; longer
  // lines: 32 points x 16 columns with both planes, inputs loaded between the
  // transforms.  Measured on 32 x 1024^2 WGS designs (512-point column
  // lines): 437 us per column pass against 579 us for 16 x 16 tiles loading
  // the inputs between the transforms.
  template <typename TT, int N>
  using ColGeomT = typename std::conditional<(N <= 512), ColGeom<TT, N, 32, 32, 1>,
                                             ColGeom<TT, N, 32, 16, 2>>::type;
  template <typename TT, int N>
  static constexpr bool PREFETCH = sizeof(TT) == 4 && N <= 512;
  GsBuf<T> b;
  double* partials;
  int it;
  int write_amp;  // GS: last iteration writes |Y| to amp[0]
  template <typename U>
  __device__ __forceinline__ void pre(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void post(int, int, int, U&, U&, double*) const {}
  // per-point inputs of the projection
  struct Aux {
    T m, t0, w;
  };
  __device__ __forceinline__ Aux load_aux(int f, int x, int y) const {
    // per-design base pointers (uniform across the workgroup), 32-bit offsets
    const long base = f * b.fstride;
    const int k = y * b.nx + x;
    Aux r;
    r.m = b.roi ? (b.roi + base)[k] : T(1);
    r.t0 = (b.tgt + base)[k];
    r.w = b.weighted ? (b.w + base)[k] : T(1);
    return r;
  }
  __device__ __forceinline__ void mid(int f, int x, int y, T& re, T& im, double* acc) const {
    mid_aux(f, x, y, re, im, acc, load_aux(f, x, y));
  }
  __device__ __forceinline__ void mid_aux(int f, int x, int y, T& re, T& im, double* acc,
                                          const Aux& ax) const {
    const long base = f * b.fstride;
    const int k = y * b.nx + x;
    const GsState* st = b.state + f;
    const T a = sqrt(re * re + im * im) * T(b.norm);
    const T m = ax.m;
    const T t0 = ax.t0;
    T w = T(1);
    if (b.weighted) w = ax.w * T(st->nrm);
    const T t = t0 * w;
    const double I = double(a) * double(a), Tt = double(t0) * double(t0);
    acc[FF_TOTAL] += I;
    if (m > T(0)) {
      acc[FF_INROI] += I * double(m);
      acc[FF_SI2] += I * I;
      acc[FF_SIT] += I * Tt;
      const double d = (double(a) - double(t)) * double(m);
      acc[FF_ERRNUM] += d * d;
      acc[FF_ERRDEN] += double(t) * double(t);
      if (b.weighted) {
        const double d0 = double(a) - double(t0);
        acc[FF_RNUM] += d0 * d0;
        const T wn = wgs_update<T>(w, a, t0, T(b.beta));
        (b.w + base)[k] = wn;
        acc[FF_WSUM] += double(wn) * double(wn) * Tt;
      }
    }
    if (b.weighted) {
      (b.amp[1 - st->best_amp_slot] + base)[k] = a;
    } else if (write_amp) {
      (b.amp[0] + base)[k] = a;
    }
    const T tp = m > T(0) ? t : a;
    if (a > T(0)) {
      // FP32: hardware reciprocal (about 1 ulp) instead of the division
      // sequence, for normal amplitudes only (the raw reciprocal is not
      // specified for denormal inputs)
      T s;
      if constexpr (sizeof(T) == 4)
        s = a >= __FLT_MIN__ ? tp * __builtin_amdgcn_rcpf(a) : tp / a;
      else
        s = tp / a;
      re *= s;
      im *= s;
    } else {
      const T cb = ((x + y) & 1) ? T(-1) : T(1);
      re = T(b.gsign) * cb * tp / T(b.norm);
      im = T(0);
    }
  }
  __device__ __forceinline__ void finalize(int f, const double* s) const {
    GsState* st = b.state + f;
    double* err_hist = b.err_hist + f * b.hstride;
    double* eff_hist = b.eff_hist + f * b.hstride;
    if (b.scr_hist) {
      const double sc = st->t2_all > 0 ? s[FF_INROI] / st->t2_all : 0.0;
      const double den = sc * sc * st->t4_roi;
      const double num = s[FF_SI2] - 2 * sc * s[FF_SIT] + sc * sc * st->t4_roi;
      b.scr_hist[f * b.hstride + it] = den > 0 ? sqrt(fmax(num, 0.0) / den) : 0.0;
    }
    const double err = s[FF_ERRDEN] > 0 ? sqrt(s[FF_ERRNUM] / s[FF_ERRDEN]) : 0.0;
    const double eff = s[FF_TOTAL] > 0 ? s[FF_INROI] / s[FF_TOTAL] : 0.0;
    if (b.weighted) {
      const double rmse = st->t2_roi > 0 ? sqrt(s[FF_RNUM] / st->t2_roi) : 0.0;
      if (rmse < st->best_rmse) {
        st->best_rmse = rmse;
        st->improved = 1;
        st->best_amp_slot = 1 - st->best_amp_slot;
      } else {
        st->improved = 0;
      }
      st->nrm = s[FF_WSUM] > 0 ? sqrt(st->t2_roi / s[FF_WSUM]) : 1.0;
      err_hist[it] = st->best_rmse;
    } else {
      err_hist[it] = err;
    }
    eff_hist[it] = eff;
  }
};

// Row pass A: phase of the inverse transform, illumination re-imposed.
template <typename T>
struct OpFFA {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = false, HAS_POST = false;
  GsBuf<T> b;
  int write_phase;  // GS: last iteration; WGS: when the iterate improved
  template <typename U>
  __device__ __forceinline__ void pre(int, int, int, U&, U&, double*) const {}
  template <typename U>
  __device__ __forceinline__ void post(int, int, int, U&, U&, double*) const {}
  // the illumination is loaded with the row, before the first transform
  struct Aux {
    T il;
  };
  template <typename TT, int N>
  static constexpr bool PREFETCH = sizeof(TT) == 4;
  __device__ __forceinline__ Aux load_aux(int f, int x, int y) const {
    return Aux{(b.illum + f * b.fstride)[y * b.nx + x]};
  }
  __device__ __forceinline__ void mid(int f, int x, int y, T& re, T& im, double* acc) const {
    mid_aux(f, x, y, re, im, acc, load_aux(f, x, y));
  }
  __device__ __forceinline__ void mid_aux(int f, int x, int y, T& re, T& im, double*,
                                          const Aux& ax) const {
    const long k = f * b.fstride + (long)y * b.nx + x;
    const T cb = ((x + y) & 1) ? T(-1) : T(1);
    const bool wp = b.weighted ? (b.state[f].improved != 0) : (write_phase != 0);
    if (wp) b.phase[0][k] = atan2(cb * im, cb * re);
    const T il = ax.il;
    T s = T(0);
    bool nz;
    if constexpr (sizeof(T) == 4) {
      // FP32: one reciprocal square root (about 1 ulp) instead of a
      // correctly rounded square root and division per point
      const T q = re * re + im * im;
      nz = q > T(0);
      if (nz) s = q >= __FLT_MIN__ ? il * rsqrtf(q) : il / sqrtf(q);
    } else {
      const T r = sqrt(re * re + im * im);
      nz = r > T(0);
      if (nz) s = il / r;
    }
    if (nz) {
      re *= s;
      im *= s;
    } else {
      re = cb * il;
      im = T(0);
    }
  }
};

// Reset `nst` design states and (WGS) the weights of their `n` points.
template <typename T>
__global__ void k_gs_init_state(GsState* st, int nst, T* w, const T* roi, long n, int weighted) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nst) {
    st[i].best_rmse = __builtin_inf();
    st[i].nrm = 1.0;
    st[i].best_phase_slot = 0;
    st[i].best_amp_slot = 0;
    st[i].improved = 0;
    st[i].pad = 0;
  }
  if (weighted && i < n) w[i] = (roi ? (roi[i] > T(0) ? T(1) : T(0)) : T(1));
}

// Constant target sums of each design (one 1024-thread block per design,
// fixed-order reduction): t0^2 and t0^4 where the ROI is positive, t0^2
// over the plane.  Designs of `n` points, `fstride` apart.
template <typename T>
__global__ void __launch_bounds__(1024) k_gs_consts(GsState* st, const T* tgt, const T* roi, long n,
                                                     long fstride) {
  const long base = (long)blockIdx.x * fstride;
  double v[3] = {0.0, 0.0, 0.0};
  for (long k = threadIdx.x; k < n; k += 1024) {
    const double t2 = double(tgt[base + k]) * double(tgt[base + k]);
    v[2] += t2;
    if (!roi || roi[base + k] > T(0)) {
      v[0] += t2;
      v[1] += t2 * t2;
    }
  }
  __shared__ double red[16][3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    for (int off = 32; off >= 1; off >>= 1) v[c] += __shfl_xor(v[c], off, 64);
    if (lane == 0) red[wv][c] = v[c];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 16; ++j)
      for (int c = 0; c < 3; ++c) r[c] += red[j][c];
    st[blockIdx.x].t2_roi = r[0];
    st[blockIdx.x].t4_roi = r[1];
    st[blockIdx.x].t2_all = r[2];
  }
}

template <typename T>
__global__ void k_from_f64_t(const double* __restrict__ a, T* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = T(a[i]);
}
template <typename T>
__global__ void k_to_f64_t(const T* __restrict__ a, double* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = double(a[i]);
}

template <typename T>
inline int upload_t(doe_ctx* ctx, const double* host, T* dst, long n) {
  int rc = ensure_stage(ctx, size_t(n) * sizeof(double));
  if (rc) return rc;
  DOE_HIP(hipMemcpyAsync(ctx->stage, host, size_t(n) * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  hipLaunchKernelGGL(k_from_f64_t<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const double*)ctx->stage, dst, n);
  DOE_HIP(hipGetLastError());
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  return DOE_OK;
}

template <typename T>
inline int download_t(doe_ctx* ctx, const T* src, double* host, long n) {
  int rc = ensure_stage(ctx, size_t(n) * sizeof(double));
  if (rc) return rc;
  hipLaunchKernelGGL(k_to_f64_t<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     src, (double*)ctx->stage, n);
  DOE_HIP(hipGetLastError());
  DOE_HIP(hipMemcpyAsync(host, ctx->stage, size_t(n) * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  DOE_HIP(hipStreamSynchronize(ctx->stream));
  return DOE_OK;
}

}  // namespace doe
