// Projection-quality figures of a screen field on the device, in FP64
// (ProjectorSystem#screenMetrics, projector.js:554-614, with the definitions
// of metrics.js:10-321).
//
// The screen irradiance I = re^2 + im^2 is formed from the device-resident
// field (any storage type) in FP64 and never leaves the GPU.  Two passes over
// the plane:
//   A  sums that fix the power-matching scale s and the thresholds:
//      sum(I m), sum(T), sum(I), sum(I x^2), sum(I y^2), count / sums over
//      m > 0, and the maxima (ROI peak, plane peak, target peak);
//   B  everything that needs s or a threshold: rmsError and correlation
//      against s T (two-pass, like metrics.js:87-118), the thresholded
//      region statistics, and the 8 x 8 SSIM windows (one block per band of
//      8 rows; a window's sums are folded over 8 adjacent lanes).
// Spot-array cells (metrics.js:284-321) and the element-plane sums
// (zeroOrderFraction, metrics.js:145-157; incident power) are separate small
// kernels.  Every reduction is per-block partials folded in a fixed order,
// so results are deterministic.  The scalar formulas are evaluated on the
// host in the reference's order.
#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
#include <vector>

#include "ctx.hpp"

namespace doe {
namespace {

constexpr int kMB = 256;  // threads per block
constexpr int kMaxBlocksA = 2048;

enum { OP_SUM = 0, OP_MAX = 1, OP_MIN = 2 };

// pass A components
enum { A_SIM, A_ST, A_SI, A_SIX2, A_SIY2, A_NM, A_SMI, A_SMT, A_PROV, A_PEAK, A_MAXT, A_N };
constexpr unsigned long long kOpsA =
    (unsigned long long)OP_MAX << (2 * A_PROV) | (unsigned long long)OP_MAX << (2 * A_PEAK) |
    (unsigned long long)OP_MAX << (2 * A_MAXT);
// pass B components
enum { B_NUM, B_DEN, B_COV, B_VP, B_VT, B_RN, B_RS1, B_RS2, B_SST, B_SSW, B_RMIN, B_RMAX, B_N };
constexpr unsigned long long kOpsB =
    (unsigned long long)OP_MIN << (2 * B_RMIN) | (unsigned long long)OP_MAX << (2 * B_RMAX);
// element components
enum { E_SR, E_SI, E_P, E_AMP, E_N };

// Fixed scratch of doe_screen_metrics_noalloc: the reduction partials (FP64)
// followed by the spot-cell bounds (int).  One copy per GPU, held until the
// process ends; its device address is resolved once per device by
// metrics_scratch_init and kept in g_scratch (host code never passes the
// array's name to a kernel).  Every partial slot is written by its kernel
// before it is read, so nothing needs clearing between calls.
constexpr size_t kScratchBytes = size_t(DOE_NOALLOC_SCRATCH_BYTES);
constexpr long kNoallocMaxCells = 4096;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

__device__ __forceinline__ double op_apply(int op, double a, double b) {
  return op == OP_SUM ? a + b : (op == OP_MAX ? fmax(a, b) : fmin(a, b));
}
__host__ __device__ constexpr int op_of(unsigned long long ops, int c) { return int((ops >> (2 * c)) & 3); }

// Reduce K values over the block (wave shuffles, then the 4 waves in order)
// and store them at out[0..K).
template <int K>
__device__ __forceinline__ void block_store(double (&v)[K], unsigned long long ops, double* out) {
  __shared__ double red[kMB / 64][K];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    double x = v[c];
    for (int off = 32; off >= 1; off >>= 1) x = op_apply(op_of(ops, c), x, __shfl_xor(x, off, 64));
    if (lane == 0) red[w][c] = x;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int c = threadIdx.x;
    double x = red[0][c];
    for (int j = 1; j < kMB / 64; ++j) x = op_apply(op_of(ops, c), x, red[j][c]);
    out[c] = x;
  }
}

__device__ __forceinline__ double init_of(int op) {
  return op == OP_SUM ? 0.0
                      : (op == OP_MAX ? -std::numeric_limits<double>::infinity()
                                      : std::numeric_limits<double>::infinity());
}

// Fold per-block partials part[b*K + c] over b, in a fixed order.
__global__ __launch_bounds__(kMB) void k_met_fold(const double* __restrict__ part, int nblk, int K,
                                                  unsigned long long ops, double* __restrict__ out) {
  __shared__ double red[kMB];
  for (int c = 0; c < K; ++c) {
    const int op = op_of(ops, c);
    double x = init_of(op);
    for (int b = threadIdx.x; b < nblk; b += kMB) x = op_apply(op, x, part[(long)b * K + c]);
    red[threadIdx.x] = x;
    __syncthreads();
    for (int s = kMB / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = op_apply(op, red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
    __syncthreads();
  }
}

template <class S>
__device__ __forceinline__ double irr(const S* re, const S* im, long k) {
  const double a = ld<double>(re + k), b = ld<double>(im + k);
  return a * a + b * b;
}

// Pass A.  tg = target irradiance, m = ROI (both FP64, the screen's grid).
template <class S>
__global__ __launch_bounds__(kMB) void k_met_a(const S* __restrict__ re, const S* __restrict__ im,
                                               const double* __restrict__ tg,
                                               const double* __restrict__ m, int nx, int ny,
                                               double sdx, double sdy, double* __restrict__ part) {
  double v[A_N];
#pragma unroll
  for (int c = 0; c < A_N; ++c) v[c] = 0.0;  // maxima start at 0, as in the reference loops
  const long n = (long)nx * ny;
  for (long k = (long)blockIdx.x * kMB + threadIdx.x; k < n; k += (long)gridDim.x * kMB) {
    const double I = irr(re, im, k), T = tg[k], mk = m[k];
    const int j = int(k / nx), i = int(k - (long)j * nx);
    const double x = (double(i) - nx / 2.0) * sdx, y = (double(j) - ny / 2.0) * sdy;
    v[A_SIM] += I * mk;
    v[A_ST] += T;
    v[A_SI] += I;
    v[A_SIX2] += I * x * x;
    v[A_SIY2] += I * y * y;
    if (mk > 0) {
      v[A_NM] += 1.0;
      v[A_SMI] += I;
      v[A_SMT] += T;
      if (I > 0) v[A_PROV] = fmax(v[A_PROV], I);
    }
    v[A_PEAK] = fmax(v[A_PEAK], I);
    v[A_MAXT] = fmax(v[A_MAXT], T);
  }
  block_store<A_N>(v, kOpsA, part + (long)blockIdx.x * A_N);
}

// Scalars of pass A that pass B needs (same expressions as the host's).
struct ScalA {
  double s, thr, mp, mt, C1, C2;
};
__device__ __forceinline__ ScalA scal_a(const double* A, double cell) {
  ScalA r;
  const double pt = A[A_SIM] * cell, tt = A[A_ST] * cell;
  r.s = tt > 0 ? pt / tt : 0.0;
  r.thr = 0.1 * A[A_PROV];
  const double nm = A[A_NM];
  r.mp = nm > 0 ? A[A_SMI] / nm : 0.0;
  r.mt = nm > 0 ? r.s * A[A_SMT] / nm : 0.0;
  double L = fmax(A[A_PEAK], r.s * A[A_MAXT]);
  L = L > 0 ? L : 1.0;
  r.C1 = 0.01 * L * (0.01 * L);
  r.C2 = 0.03 * L * (0.03 * L);
  return r;
}

// Pass B: one block per (256 columns, 8 rows).
template <class S>
__global__ __launch_bounds__(kMB) void k_met_b(const S* __restrict__ re, const S* __restrict__ im,
                                               const double* __restrict__ tg,
                                               const double* __restrict__ m, int nx, int ny,
                                               const double* __restrict__ A, double cell,
                                               int weighted, double* __restrict__ part) {
  const ScalA sa = scal_a(A, cell);
  double v[B_N];
#pragma unroll
  for (int c = 0; c < B_N; ++c) v[c] = 0.0;
  v[B_RMIN] = std::numeric_limits<double>::infinity();
  v[B_RMAX] = -std::numeric_limits<double>::infinity();
  const int i = blockIdx.x * kMB + threadIdx.x;
  const int j0 = blockIdx.y * 8;
  double cp = 0, ct = 0, cpp = 0, ctt = 0, cpt = 0;  // this column's share of an SSIM window
  if (i < nx) {
    const int j1 = min(j0 + 8, ny);
    for (int j = j0; j < j1; ++j) {
      const long k = (long)j * nx + i;
      const double I = irr(re, im, k), t = sa.s * tg[k];
      cp += I;
      ct += t;
      cpp += I * I;
      ctt += t * t;
      cpt += I * t;
      if (m[k] > 0) {
        const double d = I - t;
        v[B_NUM] += d * d;
        v[B_DEN] += t * t;
        const double dp = I - sa.mp, dt = t - sa.mt;
        v[B_COV] += dp * dt;
        v[B_VP] += dp * dp;
        v[B_VT] += dt * dt;
        if (I > sa.thr) {
          v[B_RN] += 1.0;
          v[B_RS1] += I;
          v[B_RS2] += I * I;
          v[B_RMIN] = fmin(v[B_RMIN], I);
          v[B_RMAX] = fmax(v[B_RMAX], I);
        }
      }
    }
  }
  // SSIM window = 8 adjacent columns of this band (metrics.js:222-257)
  for (int off = 1; off < 8; off <<= 1) {
    cp += __shfl_xor(cp, off, 64);
    ct += __shfl_xor(ct, off, 64);
    cpp += __shfl_xor(cpp, off, 64);
    ctt += __shfl_xor(ctt, off, 64);
    cpt += __shfl_xor(cpt, off, 64);
  }
  const int wx = i & ~7;
  if ((i & 7) == 0 && wx + 8 <= nx && j0 + 8 <= ny) {
    const double nw = 64.0;
    const double mp = cp / nw, mt = ct / nw;
    const double vp = fabs(cpp / nw - mp * mp), vt = fabs(ctt / nw - mt * mt);
    const double cv = cpt / nw - mp * mt;
    const double num = (2 * mp * mt + sa.C1) * (2 * cv + sa.C2);
    const double den = (mp * mp + mt * mt + sa.C1) * (vp + vt + sa.C2);
    const double w = weighted ? ct / nw : 1.0;
    v[B_SST] += (den != 0 ? num / den : 0.0) * w;
    v[B_SSW] += w;
  }
  block_store<B_N>(v, kOpsB, part + ((long)blockIdx.y * gridDim.x + blockIdx.x) * B_N);
}

// Spot-array cells: block c sums the pixels of cell c at or above `cut`.
template <class S>
__global__ __launch_bounds__(kMB) void k_met_cells(const S* __restrict__ re, const S* __restrict__ im,
                                                   int nx, int cols, const int* __restrict__ bx,
                                                   const int* __restrict__ by, const double* A,
                                                   double min_peak, double* __restrict__ cells) {
  const int ci = blockIdx.x % cols, cj = blockIdx.x / cols;
  const int i0 = bx[ci], i1 = bx[ci + 1], j0 = by[cj], j1 = by[cj + 1];
  const double cut = min_peak * A[A_PEAK];
  const int w = i1 - i0;
  const long area = (long)w * (j1 - j0);
  double v[1] = {0.0};
  for (long q = threadIdx.x; q < area; q += kMB) {
    const int dj = int(q / w);
    const long k = (long)(j0 + dj) * nx + i0 + (q - (long)dj * w);
    const double I = irr(re, im, k);
    if (I >= cut) v[0] += I;
  }
  block_store<1>(v, 0, cells + blockIdx.x);
}

// Element plane: sum re, sum im, sum |E|^2, sum |E|.
template <class S>
__global__ __launch_bounds__(kMB) void k_met_elem(const S* __restrict__ re, const S* __restrict__ im,
                                                  long n, double* __restrict__ part) {
  double v[E_N] = {0.0, 0.0, 0.0, 0.0};
  for (long k = (long)blockIdx.x * kMB + threadIdx.x; k < n; k += (long)gridDim.x * kMB) {
    const double a = ld<double>(re + k), b = ld<double>(im + k);
    v[E_SR] += a;
    v[E_SI] += b;
    v[E_P] += a * a + b * b;
    v[E_AMP] += sqrt(a * a + b * b);
  }
  block_store<E_N>(v, 0, part + (long)blockIdx.x * E_N);
}

// Same-pitch resample (field.js:443-465): dst[b][j][i] = src[b or 0][j - oy][i - ox]
// inside src, else 0, with ox = floor((nx2 - nx) / 2), oy likewise (centred
// zero-padding or crop).  src_batch 1 broadcasts one field to every output.
template <class S>
__global__ __launch_bounds__(kMB) void k_resample_centred(const S* __restrict__ sre, const S* __restrict__ sim,
                                                          S* __restrict__ dre, S* __restrict__ dim_, int nx,
                                                          int ny, int nx2, int ny2, int src_batch, long total,
                                                          int ox, int oy) {
  const long q = (long)blockIdx.x * kMB + threadIdx.x;
  if (q >= total) return;
  const long per = (long)nx2 * ny2;
  const int b = int(q / per);
  const long r = q - (long)b * per;
  const int j2 = int(r / nx2), i2 = int(r - (long)j2 * nx2);
  const int i = i2 - ox, j = j2 - oy;
  S vr = S(0.0f), vi = S(0.0f);
  if (i >= 0 && i < nx && j >= 0 && j < ny) {
    const long k = (long)(src_batch == 1 ? 0 : b) * nx * ny + (long)j * nx + i;
    vr = sre[k];
    vi = sim[k];
  }
  dre[q] = vr;
  dim_[q] = vi;
}

// First pixel index of each of `cells` cells along a side of n pixels, with
// the reference's mapping c(i) = min(cells - 1, floor(i / n * cells)), which
// never decreases with i: cell c is [b[c], b[c+1]), b[cells] = n.
void cell_bounds(int n, int cells, int* b) {
  for (int c = 0; c <= cells; ++c) b[c] = n;
  int next = 0;
  for (int i = 0; i < n; ++i) {
    const int c = std::min(cells - 1, int(std::floor((double(i) / double(n)) * double(cells))));
    while (next <= c) b[next++] = i;
  }
}

int blocks_of(long n) { return int(std::min<long>((n + kMB - 1) / kMB, kMaxBlocksA)); }

// fixed: the no-allocation variant's scratch (device address), or nullptr
// to use the context's reduction buffers (allocated on demand).
template <class S, class SE>
int run_metrics(doe_ctx* ctx, const doe_field* scr, const doe_field* tgt, const doe_field* el,
                const doe_screen_opts* o, doe_metrics* out, unsigned char* fixed) {
  hipStream_t st = ctx->stream;
  const int nx = scr->nx, ny = scr->ny;
  const long n = (long)nx * ny;
  const int cols = o->grid_cols, rows = o->grid_rows;
  const int nA = blocks_of(n);
  const dim3 gB((nx + kMB - 1) / kMB, (ny + 7) / 8);
  const long nBb = (long)gB.x * gB.y;
  const long ne = (long)el->nx * el->ny;
  const int nE = blocks_of(ne);
  const long ncell = cols > 0 ? (long)cols * rows : 0;
  // partials: A | B | E, then folded A | B | E and the cells
  const long oA = 0, oB = oA + (long)nA * A_N, oE = oB + nBb * B_N, oFA = oE + (long)nE * E_N;
  const long oFB = oFA + A_N, oFE = oFB + B_N, oC = oFE + E_N, total = oC + ncell;
  // the cell bounds follow the partials (fixed scratch only)
  const size_t bounds_off = (size_t(total) * sizeof(double) + 255) & ~size_t(255);
  const size_t nbounds = ncell > 0 ? size_t(cols + 1 + rows + 1) : 0;
  double* P;
  int rc;
  if (fixed) {
    DOE_REQUIRE(ncell <= kNoallocMaxCells, DOE_ERR_UNSUPPORTED,
                "doe_screen_metrics_noalloc: at most 4096 spot cells");
    DOE_REQUIRE(bounds_off + nbounds * sizeof(int) <= kScratchBytes, DOE_ERR_UNSUPPORTED,
                "doe_screen_metrics_noalloc: the screen needs " +
                    std::to_string(bounds_off + nbounds * sizeof(int)) + " bytes of scratch, more than the " +
                    std::to_string(kScratchBytes) + " the entry point holds");
    P = reinterpret_cast<double*>(fixed);
  } else {
    rc = ensure_reduce(ctx, size_t(total));
    if (rc) return rc;
    P = ctx->partials;
  }
  const S* sre = static_cast<const S*>(scr->re);
  const S* sim = static_cast<const S*>(scr->im);
  const double* tg = static_cast<const double*>(tgt->re);
  const double* mk = static_cast<const double*>(tgt->im);
  const double cell = o->sdx * o->sdy;
  if (!fixed) prof_begin(ctx, "screen_metrics", n);
  hipLaunchKernelGGL(k_met_a<S>, dim3(nA), dim3(kMB), 0, st, sre, sim, tg, mk, nx, ny, o->sdx, o->sdy,
                     P + oA);
  hipLaunchKernelGGL(k_met_fold, dim3(1), dim3(kMB), 0, st, P + oA, nA, int(A_N), kOpsA, P + oFA);
  hipLaunchKernelGGL(k_met_b<S>, gB, dim3(kMB), 0, st, sre, sim, tg, mk, nx, ny, P + oFA, cell,
                     o->ssim_weight_target ? 1 : 0, P + oB);
  hipLaunchKernelGGL(k_met_fold, dim3(1), dim3(kMB), 0, st, P + oB, int(nBb), int(B_N), kOpsB, P + oFB);
  hipLaunchKernelGGL(k_met_elem<SE>, dim3(nE), dim3(kMB), 0, st, static_cast<const SE*>(el->re),
                     static_cast<const SE*>(el->im), ne, P + oE);
  hipLaunchKernelGGL(k_met_fold, dim3(1), dim3(kMB), 0, st, P + oE, nE, int(E_N), 0ull, P + oFE);
  // host staging: stack arrays for the fixed-scratch variant (bounded cell
  // count), vectors otherwise
  int bounds_fix[2 * (kNoallocMaxCells + 1)];
  double h_fix[A_N + B_N + E_N + kNoallocMaxCells];
  std::vector<int> bounds_vec;
  std::vector<double> h_vec;
  if (ncell > 0) {
    int* both = bounds_fix;
    if (!fixed) {
      bounds_vec.resize(nbounds);
      both = bounds_vec.data();
    }
    cell_bounds(nx, cols, both);
    cell_bounds(ny, rows, both + cols + 1);
    const int* dbx;
    if (fixed) {
      int* db = reinterpret_cast<int*>(fixed + bounds_off);
      DOE_HIP(hipMemcpyAsync(db, both, nbounds * sizeof(int), hipMemcpyHostToDevice, st));
      dbx = db;
    } else {
      void* dp = nullptr;
      rc = upload_params(ctx, both, nbounds * sizeof(int), &dp);
      if (rc) return rc;
      dbx = static_cast<const int*>(dp);
    }
    hipLaunchKernelGGL(k_met_cells<S>, dim3(unsigned(ncell)), dim3(kMB), 0, st, sre, sim, nx, cols, dbx,
                       dbx + cols + 1, P + oFA, o->grid_min_peak, P + oC);
  }
  DOE_HIP(hipGetLastError());
  if (!fixed) prof_end(ctx);
  const size_t nh = size_t(total - oFA);
  double* h = h_fix;
  if (!fixed) {
    h_vec.resize(nh);
    h = h_vec.data();
  }
  DOE_HIP(hipMemcpyAsync(h, P + oFA, nh * sizeof(double), hipMemcpyDeviceToHost, st));
  DOE_HIP(hipStreamSynchronize(st));
  const double* A = h;
  const double* B = A + A_N;
  const double* E = B + B_N;
  const double* C = E + E_N;

  // the reference's scalar formulas (projector.js:572-613, metrics.js)
  const double incident = E[E_AMP] * o->edx * o->edy;
  const double totalp = A[A_SI] * cell;
  const double in_signal = A[A_SIM] * cell;
  const double rms = B[B_DEN] > 0 ? std::sqrt(B[B_NUM] / B[B_DEN]) : 0.0;
  const double corr = (A[A_NM] > 0 && B[B_VP] > 0 && B[B_VT] > 0) ? B[B_COV] / std::sqrt(B[B_VP] * B[B_VT])
                                                                  : 0.0;
  double mean = 0, mn = 0, mx = 0, uni = 0, flat = 0, cv = 0;
  if (B[B_RN] > 0) {
    mean = B[B_RS1] / B[B_RN];
    const double sd = std::sqrt(std::max(0.0, B[B_RS2] / B[B_RN] - mean * mean));
    mn = B[B_RMIN];
    mx = B[B_RMAX];
    uni = mx + mn > 0 ? 1 - (mx - mn) / (mx + mn) : 1;
    flat = mean > 0 ? 1 - sd / mean : 1;
    cv = mean > 0 ? sd / mean : 0;
  }
  double spot = 0;
  out->has_spots = ncell > 0 ? 1 : 0;
  if (ncell > 0) {
    double lo = std::numeric_limits<double>::infinity(), hi = 0;
    long k = 0;
    for (long c = 0; c < ncell; ++c) {
      if (C[c] <= 0) continue;
      lo = std::min(lo, C[c]);
      hi = std::max(hi, C[c]);
      ++k;
    }
    spot = k == 0 ? 0.0 : (hi + lo > 0 ? (2 * lo) / (hi + lo) : 1.0);
  }
  double rr = 0;
  if (A[A_SI] != 0) rr = std::hypot(std::sqrt(A[A_SIX2] / A[A_SI]), std::sqrt(A[A_SIY2] / A[A_SI]));
  out->efficiency = incident > 0 ? in_signal / incident : 0;
  out->flux_fraction = totalp > 0 ? in_signal / totalp : 0;
  out->uniformity = ncell > 0 ? spot : uni;
  out->spot_uniformity = ncell > 0 ? spot : std::numeric_limits<double>::quiet_NaN();
  out->flatness = flat;
  out->contrast = cv;
  out->rmse = rms;
  out->fidelity = (corr + 1) / 2;
  out->correlation = corr;
  out->ssim = B[B_SSW] > 0 ? B[B_SST] / B[B_SSW] : 0;
  out->snr = rms > 0 ? 20 * std::log10(1 / rms) : std::numeric_limits<double>::infinity();
  out->zero_order = E[E_P] == 0 ? 0 : (E[E_SR] * E[E_SR] + E[E_SI] * E[E_SI]) / (double(ne) * E[E_P]);
  out->rms_radius = rr;
  out->mean_intensity = mean;
  out->peak_intensity = mx;
  out->power_in_signal = in_signal;
  out->power_at_screen = totalp;
  out->power_incident = incident;
  return DOE_OK;
}

template <class S>
int run_metrics_el(doe_ctx* ctx, const doe_field* scr, const doe_field* tgt, const doe_field* el,
                   const doe_screen_opts* o, doe_metrics* out, unsigned char* fixed) {
  switch (el->dtype) {
    case DOE_F32: return run_metrics<S, float>(ctx, scr, tgt, el, o, out, fixed);
    case DOE_F64: return run_metrics<S, double>(ctx, scr, tgt, el, o, out, fixed);
    case DOE_F16: return run_metrics<S, __half>(ctx, scr, tgt, el, o, out, fixed);
    case DOE_BF16: return run_metrics<S, __hip_bfloat16>(ctx, scr, tgt, el, o, out, fixed);
  }
  set_error("unknown dtype");
  return DOE_ERR_INVALID;
}

int check_metrics_args(doe_ctx* ctx, const doe_field* screen, const doe_field* target,
                       const doe_field* element, const doe_screen_opts* opts, doe_metrics* out) {
  DOE_REQUIRE(ctx && screen && target && element && opts && out, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(screen->batch == 1 && target->batch == 1 && element->batch == 1, DOE_ERR_SHAPE,
              "screenMetrics: one field per call");
  DOE_REQUIRE(target->dtype == DOE_F64, DOE_ERR_INVALID,
              "screenMetrics: the target field holds FP64 irradiance (re) and ROI (im)");
  DOE_REQUIRE(target->nx == screen->nx && target->ny == screen->ny, DOE_ERR_SHAPE,
              "screenMetrics: target and screen grids differ");
  DOE_REQUIRE(screen->nx >= 1 && screen->ny >= 1 && element->nx >= 1 && element->ny >= 1, DOE_ERR_SHAPE,
              "screenMetrics: empty grid");
  DOE_REQUIRE(opts->grid_cols >= 0 && opts->grid_rows >= 0 && (opts->grid_cols > 0) == (opts->grid_rows > 0),
              DOE_ERR_INVALID, "screenMetrics: grid needs both cols and rows");
  DOE_REQUIRE((long)opts->grid_cols * opts->grid_rows <= 65536, DOE_ERR_INVALID,
              "screenMetrics: at most 65536 spot cells");
  return DOE_OK;
}

int metrics_dispatch(doe_ctx* ctx, const doe_field* screen, const doe_field* target,
                     const doe_field* element, const doe_screen_opts* opts, doe_metrics* out,
                     unsigned char* fixed) {
  switch (screen->dtype) {
    case DOE_F32: return run_metrics_el<float>(ctx, screen, target, element, opts, out, fixed);
    case DOE_F64: return run_metrics_el<double>(ctx, screen, target, element, opts, out, fixed);
    case DOE_F16: return run_metrics_el<__half>(ctx, screen, target, element, opts, out, fixed);
    case DOE_BF16: return run_metrics_el<__hip_bfloat16>(ctx, screen, target, element, opts, out, fixed);
  }
  set_error("unknown dtype");
  return DOE_ERR_INVALID;
}

}  // namespace

// ---------------------------------------------------------------------------
// ComplexField#resample at another pitch (field.js:466-495): output sample
// (i, j) at x = (i - nx2/2) dx2 reads the source bilinearly at the fractional
// index fx = x / dx + nx/2 - 1/2 (zero outside the source grid), in FP64 with
// the reference's operation order (no contraction), so FP64 fields equal it
// bit for bit.
template <class S>
__global__ __launch_bounds__(kMB) void k_resample_bilinear(const S* __restrict__ sre, const S* __restrict__ sim,
                                                           S* __restrict__ dre, S* __restrict__ dim_, int nx,
                                                           int ny, int nx2, int ny2, int src_batch, long total,
                                                           double sdx, double sdy, double ddx, double ddy) {
#pragma clang fp contract(off)
  const long q = (long)blockIdx.x * kMB + threadIdx.x;
  if (q >= total) return;
  const long per = (long)nx2 * ny2;
  const int b = int(q / per);
  const long r = q - (long)b * per;
  const int j = int(r / nx2), i = int(r - (long)j * nx2);
  const S* re = sre + (long)(src_batch == 1 ? 0 : b) * nx * ny;
  const S* im = sim + (long)(src_batch == 1 ? 0 : b) * nx * ny;
  const double y = (j - ny2 / 2.0) * ddy;
  const double fy = y / sdy + ny / 2.0 - 0.5;
  const double j0 = floor(fy);
  const double ty = fy - j0;
  const double x = (i - nx2 / 2.0) * ddx;
  const double fx = x / sdx + nx / 2.0 - 0.5;
  const double i0 = floor(fx);
  const double tx = fx - i0;
  double ar = 0, ai = 0;
  for (int bb = 0; bb < 2; ++bb) {
    const double jj = j0 + bb;
    if (jj < 0 || jj >= ny) continue;
    const double wy = bb == 0 ? 1 - ty : ty;
    for (int aa = 0; aa < 2; ++aa) {
      const double ii = i0 + aa;
      if (ii < 0 || ii >= nx) continue;
      const double wx = aa == 0 ? 1 - tx : tx;
      const double w = wx * wy;
      const long k = (long)jj * nx + (long)ii;
      ar += ld<double>(re + k) * w;
      ai += ld<double>(im + k) * w;
    }
  }
  st(dre + q, ar);
  st(dim_ + q, ai);
}

// ComplexField#sampleImage (field.js:350-400): the raster covers the central
// `scale` fraction of the window; grid sample (i, j) averages the raster over
// the box it covers (area weights), `fill` outside the raster; re = the
// average, im = 0.  FP64, the reference's operation order.
template <class S>
__global__ __launch_bounds__(kMB) void k_sample_image(const double* __restrict__ img, int width, int height,
                                                      double scale, double fill, S* __restrict__ dre,
                                                      S* __restrict__ dim_, int nx, int ny) {
#pragma clang fp contract(off)
  const long q = (long)blockIdx.x * kMB + threadIdx.x;
  if (q >= (long)nx * ny) return;
  const int j = int(q / nx), i = int(q - (long)j * nx);
  const double s = fmax(scale, 1e-9);
  const double off = (1 - s) / 2;
  const double boxW = width / (nx * s), boxH = height / (ny * s);
  const double cy = (((j + 0.5) / ny - off) * height) / s;
  const double cx = (((i + 0.5) / nx - off) * width) / s;
  const double ax = cx - boxW / 2, bx = cx + boxW / 2;
  const double ay = cy - boxH / 2, by = cy + boxH / 2;
  const double lox = fmax(ax, 0.0), hix = fmin(bx, (double)width);
  const double loy = fmax(ay, 0.0), hiy = fmin(by, (double)height);
  const double area = boxW * boxH;
  double acc = 0;
  if (hix > lox && hiy > loy) {
    const long i0 = (long)fmax(0.0, floor(lox)), i1 = (long)fmin(width - 1.0, floor(hix - 1e-12));
    const long j0 = (long)fmax(0.0, floor(loy)), j1 = (long)fmin(height - 1.0, floor(hiy - 1e-12));
    for (long jj = j0; jj <= j1; ++jj) {
      const double wy = fmin(by, double(jj + 1)) - fmax(ay, double(jj));
      for (long ii = i0; ii <= i1; ++ii) {
        const double wx = fmin(bx, double(ii + 1)) - fmax(ax, double(ii));
        acc += img[ii + jj * width] * wx * wy;
      }
    }
  }
  acc += fill * (area - (hix - lox) * (hiy - loy));
  st(dre + q, acc / area);
  st(dim_ + q, 0.0);
}

// Resolve the device address of `scratch` on `device` (the caller made it
// current), once per device.  Called by doe_ctx_create.
int metrics_scratch_init(int device) {
  DOE_REQUIRE(device >= 0 && device < kMaxDevices, DOE_ERR_UNSUPPORTED,
              "doe_screen_metrics_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return DOE_OK;
  void* ptr = nullptr;
  DOE_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  DOE_REQUIRE(ptr, DOE_ERR_HIP, "doe_screen_metrics_noalloc: no address for the scratch array");
  g_scratch[device] = ptr;
  return DOE_OK;
}

}  // namespace doe

using namespace doe;

extern "C" {

int doe_screen_metrics(doe_ctx* ctx, const doe_field* screen, const doe_field* target,
                       const doe_field* element, const doe_screen_opts* opts, doe_metrics* out) {
  int rc = check_metrics_args(ctx, screen, target, element, opts, out);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(ctx->device));
  return metrics_dispatch(ctx, screen, target, element, opts, out, nullptr);
}

int doe_screen_metrics_noalloc(doe_ctx* ctx, const doe_field* screen, const doe_field* target,
                               const doe_field* element, const doe_screen_opts* opts, doe_metrics* out) {
  int rc = check_metrics_args(ctx, screen, target, element, opts, out);
  if (rc) return rc;
  DOE_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices, DOE_ERR_UNSUPPORTED,
              "doe_screen_metrics_noalloc: device index beyond 63");
  unsigned char* fixed;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    fixed = static_cast<unsigned char*>(g_scratch[ctx->device]);
  }
  DOE_REQUIRE(fixed, DOE_ERR_STATE, "doe_screen_metrics_noalloc: scratch not set up for this device");
  DOE_HIP(hipSetDevice(ctx->device));
  return metrics_dispatch(ctx, screen, target, element, opts, out, fixed);
}

int doe_screen_metrics_batch(doe_ctx* ctx, const doe_field* screens, const doe_field* target,
                             const doe_field* element, const doe_screen_opts* opts, doe_metrics* out) {
  DOE_REQUIRE(ctx && screens && out, DOE_ERR_INVALID, "null argument");
  const size_t plane = size_t(screens->nx) * screens->ny * dtype_size(screens->dtype);
  for (int b = 0; b < screens->batch; ++b) {
    doe_field v = *screens;
    v.batch = 1;
    v.owned = false;
    v.re = static_cast<char*>(screens->re) + b * plane;
    v.im = static_cast<char*>(screens->im) + b * plane;
    const int rc = doe_screen_metrics(ctx, &v, target, element, opts, out + b);
    if (rc) return rc;
  }
  return DOE_OK;
}

int doe_field_resample_centred(doe_ctx* ctx, const doe_field* src, doe_field* dst) {
  DOE_REQUIRE(ctx && src && dst, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(src->dtype == dst->dtype, DOE_ERR_INVALID, "resample: source and destination dtypes differ");
  DOE_REQUIRE(src->batch == dst->batch || src->batch == 1, DOE_ERR_SHAPE,
              "resample: source batch must be 1 or the destination's");
  DOE_REQUIRE(src->re != dst->re && src->im != dst->im, DOE_ERR_INVALID, "resample: source and destination overlap");
  DOE_HIP(hipSetDevice(ctx->device));
  const long total = long(dst->count());
  const int ox = int(std::floor((dst->nx - src->nx) / 2.0)), oy = int(std::floor((dst->ny - src->ny) / 2.0));
  const dim3 g(unsigned((total + kMB - 1) / kMB));
  hipStream_t st = ctx->stream;
#define DOE_RS(S)                                                                                      \
  hipLaunchKernelGGL(k_resample_centred<S>, g, dim3(kMB), 0, st, static_cast<const S*>(src->re),      \
                     static_cast<const S*>(src->im), static_cast<S*>(dst->re), static_cast<S*>(dst->im), \
                     src->nx, src->ny, dst->nx, dst->ny, src->batch, total, ox, oy)
  switch (dst->dtype) {
    case DOE_F32: DOE_RS(float); break;
    case DOE_F64: DOE_RS(double); break;
    case DOE_F16: DOE_RS(__half); break;
    case DOE_BF16: DOE_RS(__hip_bfloat16); break;
    default: set_error("unknown dtype"); return DOE_ERR_INVALID;
  }
#undef DOE_RS
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_field_resample(doe_ctx* ctx, const doe_field* src, doe_field* dst, double src_dx, double src_dy,
                       double dst_dx, double dst_dy) {
  DOE_REQUIRE(ctx && src && dst, DOE_ERR_INVALID, "null argument");
  if (src_dx == dst_dx && src_dy == dst_dy) return doe_field_resample_centred(ctx, src, dst);
  DOE_REQUIRE(src->dtype == dst->dtype, DOE_ERR_INVALID, "resample: source and destination dtypes differ");
  DOE_REQUIRE(src->batch == dst->batch || src->batch == 1, DOE_ERR_SHAPE,
              "resample: source batch must be 1 or the destination's");
  DOE_REQUIRE(src->re != dst->re && src->im != dst->im, DOE_ERR_INVALID, "resample: source and destination overlap");
  DOE_REQUIRE(src_dx > 0 && src_dy > 0 && dst_dx > 0 && dst_dy > 0 && std::isfinite(src_dx) &&
                  std::isfinite(src_dy) && std::isfinite(dst_dx) && std::isfinite(dst_dy),
              DOE_ERR_INVALID, "resample: pitches must be positive and finite");
  DOE_HIP(hipSetDevice(ctx->device));
  const long total = long(dst->count());
  const dim3 g(unsigned((total + kMB - 1) / kMB));
  hipStream_t st = ctx->stream;
#define DOE_RB(S)                                                                                        \
  hipLaunchKernelGGL(k_resample_bilinear<S>, g, dim3(kMB), 0, st, static_cast<const S*>(src->re),       \
                     static_cast<const S*>(src->im), static_cast<S*>(dst->re), static_cast<S*>(dst->im), \
                     src->nx, src->ny, dst->nx, dst->ny, src->batch, total, src_dx, src_dy, dst_dx, dst_dy)
  switch (dst->dtype) {
    case DOE_F32: DOE_RB(float); break;
    case DOE_F64: DOE_RB(double); break;
    case DOE_F16: DOE_RB(__half); break;
    case DOE_BF16: DOE_RB(__hip_bfloat16); break;
    default: set_error("unknown dtype"); return DOE_ERR_INVALID;
  }
#undef DOE_RB
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_field_sample_image(doe_ctx* ctx, const double* img, int width, int height, double scale, double fill,
                           doe_field* dst) {
  DOE_REQUIRE(ctx && img && dst, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(width >= 1 && height >= 1, DOE_ERR_SHAPE, "sampleImage: empty raster");
  DOE_REQUIRE(dst->batch == 1, DOE_ERR_SHAPE, "sampleImage: one field per call");
  DOE_HIP(hipSetDevice(ctx->device));
  void* dimg = nullptr;
  int rc = upload_params(ctx, img, size_t(width) * size_t(height) * sizeof(double), &dimg);
  if (rc) return rc;
  const long total = long(dst->count());
  const dim3 g(unsigned((total + kMB - 1) / kMB));
  hipStream_t st = ctx->stream;
  const double* im = static_cast<const double*>(dimg);
#define DOE_SI(S)                                                                                       \
  hipLaunchKernelGGL(k_sample_image<S>, g, dim3(kMB), 0, st, im, width, height, scale, fill,           \
                     static_cast<S*>(dst->re), static_cast<S*>(dst->im), dst->nx, dst->ny)
  switch (dst->dtype) {
    case DOE_F32: DOE_SI(float); break;
    case DOE_F64: DOE_SI(double); break;
    case DOE_F16: DOE_SI(__half); break;
    case DOE_BF16: DOE_SI(__hip_bfloat16); break;
    default: set_error("unknown dtype"); return DOE_ERR_INVALID;
  }
#undef DOE_SI
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_field_copy_slices(doe_ctx* ctx, const doe_field* src, int src_first, doe_field* dst,
                          int dst_first, int count) {
  DOE_REQUIRE(ctx && src && dst, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(src->nx == dst->nx && src->ny == dst->ny && src->dtype == dst->dtype, DOE_ERR_SHAPE,
              "copy_slices: grids or dtypes differ");
  DOE_REQUIRE(count >= 0 && src_first >= 0 && dst_first >= 0 && src_first + count <= src->batch &&
                  dst_first + count <= dst->batch,
              DOE_ERR_SHAPE, "copy_slices: slice range outside the batch");
  DOE_HIP(hipSetDevice(ctx->device));
  const size_t plane = size_t(src->nx) * src->ny * dtype_size(src->dtype);
  const size_t bytes = plane * count;
  if (bytes) {
    DOE_HIP(hipMemcpyAsync(static_cast<char*>(dst->re) + dst_first * plane,
                           static_cast<const char*>(src->re) + src_first * plane, bytes,
                           hipMemcpyDeviceToDevice, ctx->stream));
    DOE_HIP(hipMemcpyAsync(static_cast<char*>(dst->im) + dst_first * plane,
                           static_cast<const char*>(src->im) + src_first * plane, bytes,
                           hipMemcpyDeviceToDevice, ctx->stream));
  }
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

}  // extern "C"
