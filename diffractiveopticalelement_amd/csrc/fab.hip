// Fabrication output on the device: Floyd–Steinberg error diffusion of a
// wrapped phase (errorDiffusionQuantize, doe.js:185-226) and the etch-depth
// map (phaseToHeight, doe.js:967-976), in FP64 with the reference's operation
// order and contraction off, so results equal the reference bit for bit.
//
// Error diffusion is a recurrence.  A sample's input holds, in order, the
// 1/16, 5/16 and 3/16 shares of the three samples above it and then the 7/16
// share of its predecessor in the row.  Two schedules:
//   serpentine rows (the reference's default): a row runs in the direction
//     opposite to the previous one, so its first sample needs the previous
//     row complete and every sample of an image is one step of one chain.
//     Rows of at most 4096 samples: one workgroup per image, the chain on
//     rows staged in LDS while the other waves stream rows in and out and
//     then add each row's shares into the next in parallel;
//     longer rows: one lane per image, on HBM;
//   left-to-right rows: sample (i, j) needs (i+1, j-1) and (i-1, j), so row
//     j can trail row j-1 by two samples — a diagonal wavefront of
//     nx + 2 (ny - 1) steps per image, one workgroup per image, one barrier
//     per step.
// The one-lane and wavefront schedules keep the pending shares of the next
// row in registers (a sample's three shares arrive from three consecutive
// samples of the row above); the workgroup schedule gathers them after each
// row.  Every schedule reads each sample's original value once.
#include <algorithm>
#include <cmath>

#include "ctx.hpp"

#pragma clang fp contract(off)

namespace doe {
namespace {

constexpr int kFB = 256;

// wrapPhase of one value (propagate.js:410-418): p % 2pi, then into (-pi, pi].
__device__ __forceinline__ double wrap1(double p) {
  const double two_pi = 2 * M_PI;
  p = fmod(p, two_pi);
  if (p > M_PI)
    p -= two_pi;
  else if (p <= -M_PI)
    p += two_pi;
  return p;
}

// Math.round: nearest integer, ties toward +Infinity; a zero result keeps
// the sign of x (Math.round returns -0 for x in [-0.5, -0]).  A nonzero
// result already has the sign of x, so copysign(r, x) is exact for every r
// (one bit operation instead of a compare and a select on the chain).
__device__ __forceinline__ double js_round(double x) {
  if (x != x || x == INFINITY || x == -INFINITY) return x;
  const double f = floor(x);
  const double r = (x - f >= 0.5) ? f + 1 : f;
  return copysign(r, x);
}

// Math.round(v / step) with the quotient taken as v * inv (inv = 1 / step):
// the two quotients differ by a few ulps, which changes the rounding only
// when the quotient lies that close to a half-integer (an integer boundary
// gives the same result from either side), so within 2^-40 of one the IEEE
// division decides.  Away from a half-integer, floor(y + 0.5) is the
// round-half-up of y exactly (the sum's rounding error, at most 2^-53 |y|,
// is far inside the 2^-40 margin), and the fraction test runs beside it
// (y - floor(y) is exact, the hardware fraction instruction).
__device__ __forceinline__ double round_quot(double v, double step, double inv) {
  const double y = v * inv;
  const double d = __builtin_amdgcn_fract(y) - 0.5;
  const double r = floor(y + 0.5);
  if (fabs(d) > 0x1p-40 * fmax(1.0, fabs(y))) return copysign(r, y);
  return js_round(v / step);
}

// One sample: quantise v, return the wrapped error; q is the level.  With
// |levels| >= 2 the error is at most |step| / 2 (1 + 2^-50) <= pi / 2 + ulps
// (NaN and infinite inputs give NaN), so WRAP = false skips a wrap that
// could not change it.
template <bool WRAP = true>
__device__ __forceinline__ double quantise(double v, double step, double inv, double* q) {
  *q = round_quot(v, step, inv) * step;
  double e = v - *q;
  if constexpr (WRAP) {
    if (e > M_PI)
      e -= 2 * M_PI;
    else if (e <= -M_PI)
      e += 2 * M_PI;
  }
  return e;
}

// Serpentine (or plain, with snake = 0) scan of one image per lane.  out
// holds the working values: row j+1's sums are written as its samples
// complete, and the levels replace the values as the scan passes; wrapped at
// the end.
template <bool WRAP>
__global__ __launch_bounds__(64) void k_diffuse_lane(const double* __restrict__ in, double* __restrict__ out,
                                                     int nx, int ny, int batch, double step, double inv,
                                                     int snake, int wrap_first) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= batch) return;
  const long n = (long)nx * ny;
  const double* src = in + (long)b * n;
  double* o = out + (long)b * n;
  // out doubles as the working buffer of the row below the current one:
  // o[next row] holds that row's partial sums until the row is processed.
  for (long k = 0; k < nx; ++k) o[k] = wrap_first ? wrap1(src[k]) : src[k];
  for (int j = 0; j < ny; ++j) {
    const int dir = (snake && (j & 1)) ? -1 : 1;
    const bool below = j + 1 < ny;
    double* row = o + (long)j * nx;
    double* nrow = o + (long)(j + 1) * nx;
    const double* nsrc = src + (long)(j + 1) * nx;
    double carry = 0;  // 7/16 of the previous sample of this row
    // pending sums of the row below: p_next (the current sample's: initial
    // value + 1/16 of the sample behind), p_cur (the sample behind's: + 5/16),
    // which completes with the current sample's 3/16
    const int i0 = dir == 1 ? 0 : nx - 1;
    double p_next = below ? (wrap_first ? wrap1(nsrc[i0]) : nsrc[i0]) : 0, p_cur = 0;
    // blocks of 8 samples: the block's inputs are loaded together, then the
    // chain runs on registers
    // the row below, or (on the last row) any valid row: loads stay in bounds
    const double* nsafe = below ? nsrc : src;
    for (int c0 = 0; c0 < nx; c0 += 8) {
      // the block's 16 loads at clamped (valid) positions, all in flight
      // together; entries past the row's end are never used
      double rv[8], av[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = min(c0 + u, nx - 1);
        const int i = dir == 1 ? c : nx - 1 - c;
        rv[u] = row[i];
        av[u] = nsafe[min(max(i + dir, 0), nx - 1)];
      }
      if (wrap_first) {
#pragma unroll
        for (int u = 0; u < 8; ++u) av[u] = wrap1(av[u]);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = c0 + u;
        if (c >= nx) break;
        const int i = dir == 1 ? c : nx - 1 - c;
        // the first sample of a row has no 7/16 share (and keeps a -0)
        const double v = c ? rv[u] + carry : rv[u];
        double q;
        const double e = quantise<WRAP>(v, step, inv, &q);
        row[i] = q;
        carry = e * (7.0 / 16);
        if (below) {
          if (c > 0) nrow[i - dir] = p_cur + e * (3.0 / 16);
          p_cur = p_next + e * (5.0 / 16);
          p_next = av[u] + e * (1.0 / 16);
        }
      }
    }
    if (below) nrow[dir == 1 ? nx - 1 : 0] = p_cur;
  }
  for (long k = 0; k < n; ++k) o[k] = wrap1(o[k]);
}

// Serpentine rows of one image per workgroup, rows staged through LDS: three
// row buffers rotate, plus a row of errors.  At row j, lane 0 runs the chain
// along row j (its sums, replaced by the levels; the errors into the error
// row) while waves 1.. write row j-1's levels (wrapped) out and load row
// j+2's inputs into the buffer that frees; after a barrier every thread adds
// row j's shares into row j+1 — column m gets, in the order the scan would
// deliver them, 1/16 of the error behind m, 5/16 of m's, 3/16 of the one
// ahead (behind and ahead in row j's direction).  Two barriers per row.
constexpr int kRowThreads = 256;
constexpr int kRowMax = 4096;  // longest row staged (4 x 32 KiB of LDS)
constexpr int kRowPad = 8;     // LDS rows padded on both sides: the chain's block
                               // reads past a row's end land in the padding

// The chain along one serpentine row in scan direction DIR; pc / pe point at
// the row's first sample (in scan order) and at its error slot.  The 7/16
// share starts as -0 (x + -0 == x for every x: the first sample keeps its
// value and sign).
template <int DIR, bool WRAP>
__device__ __forceinline__ void chain_row(double* pc, double* pe, int nx, double step, double inv) {
  double carry = -0.0;
  auto sample = [&](int c, double r) {
    double q;
    const double e = quantise<WRAP>(r + carry, step, inv, &q);
    pc[DIR * c] = q;
    pe[DIR * c] = e;
    carry = e * (7.0 / 16);
  };
  int c0 = 0;
  for (; c0 + 8 <= nx; c0 += 8) {
    double rv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) rv[u] = pc[DIR * (c0 + u)];
#pragma unroll
    for (int u = 0; u < 8; ++u) sample(c0 + u, rv[u]);
  }
  for (int c = c0; c < nx; ++c) sample(c, pc[DIR * c]);
}

__global__ __launch_bounds__(kRowThreads) void k_diffuse_rows(const double* __restrict__ in, double* __restrict__ out,
                                                             int nx, int ny, double step, double inv,
                                                             int wrap_first) {
  extern __shared__ double R[];  // 3 rows + the error row, each nx + 2 kRowPad
  const int ld = nx + 2 * kRowPad;
  auto row = [&](int slot) { return R + slot * ld + kRowPad; };
  double* err = row(3);
  const long n = (long)nx * ny;
  const double* src = in + (long)blockIdx.x * n;
  double* o = out + (long)blockIdx.x * n;
  const int tid = threadIdx.x;
  for (int k = tid; k < nx; k += kRowThreads) {
    const double a = src[k];
    row(0)[k] = wrap_first ? wrap1(a) : a;
    if (ny > 1) {
      const double b = src[nx + k];
      row(1)[k] = wrap_first ? wrap1(b) : b;
    }
  }
  __syncthreads();
  for (int j = 0; j < ny; ++j) {
    const int dir = (j & 1) ? -1 : 1;
    if (tid == 0) {
      // lane 0's lane index (0) added to the row pointers: the compiler then
      // keeps the chain's LDS addresses in vector registers instead of
      // copying a scalar address into one before every store
      const int lane = int(__builtin_amdgcn_mbcnt_lo(~0u, 0u));
      double* cur = row(j % 3) + lane;
      double* err_l = err + lane;
      const bool wrap_e = fabs(step) > 3.0;  // |levels| < 2: the error can reach -pi
      if (dir == 1) {
        if (wrap_e)
          chain_row<1, true>(cur, err_l, nx, step, inv);
        else
          chain_row<1, false>(cur, err_l, nx, step, inv);
      } else {
        if (wrap_e)
          chain_row<-1, true>(cur + nx - 1, err_l + nx - 1, nx, step, inv);
        else
          chain_row<-1, false>(cur + nx - 1, err_l + nx - 1, nx, step, inv);
      }
    } else if (tid >= 64) {
      // the buffer of row j-1 (levels) becomes row j+2's
      double* old = row((j + 2) % 3);
      for (int k = tid - 64; k < nx; k += kRowThreads - 64) {
        if (j >= 1) o[(long)(j - 1) * nx + k] = wrap1(old[k]);
        if (j + 2 < ny) {
          const double a = src[(long)(j + 2) * nx + k];
          old[k] = wrap_first ? wrap1(a) : a;
        }
      }
    }
    __syncthreads();
    if (j + 1 < ny) {
      double* nxt = row((j + 1) % 3);
      for (int m = tid; m < nx; m += kRowThreads) {
        double v = nxt[m];
        const int mb = m - dir, ma = m + dir;  // behind and ahead in row j's scan
        if (mb >= 0 && mb < nx) v = v + err[mb] * (1.0 / 16);
        v = v + err[m] * (5.0 / 16);
        if (ma >= 0 && ma < nx) v = v + err[ma] * (3.0 / 16);
        nxt[m] = v;
      }
    }
    __syncthreads();
  }
  const double* last = row((ny - 1) % 3);
  for (int k = tid; k < nx; k += kRowThreads) o[(long)(ny - 1) * nx + k] = wrap1(last[k]);
}

// Plain (left-to-right) rows on a diagonal wavefront: step t processes sample
// i = t - 2j of every row j in flight.  One workgroup per image; row j is
// thread j mod 1024's (R rows per thread).  Sample (i, j) gathers, in the
// order the serial scan delivers them, 1/16 of e(i-1, j-1), 5/16 of
// e(i, j-1), 3/16 of e(i+1, j-1) and 7/16 of e(i-1, j): the thread keeps the
// row above's two older errors in registers and reads the newest, written
// one step earlier, from a per-row LDS slot (two slots by step parity).  out
// holds the wrapped inputs, then the levels, wrapped at the end.
constexpr int kWaveThreads = 1024;
constexpr int kWaveRowsMax = 8192;
template <bool WRAP, int R>
__global__ __launch_bounds__(kWaveThreads) void k_diffuse_wave(const double* __restrict__ in,
                                                               double* __restrict__ out, int nx, int ny,
                                                               double step, double inv, int wrap_first) {
  __shared__ double eb[2][kWaveRowsMax];
  const long n = (long)nx * ny;
  const double* src = in + (long)blockIdx.x * n;
  double* o = out + (long)blockIdx.x * n;
  const int tid = threadIdx.x;
  for (long k = tid; k < n; k += kWaveThreads) o[k] = wrap_first ? wrap1(src[k]) : src[k];
  __syncthreads();
  // x[r][d]: input of column i + d of the thread's row r (four loads in flight)
  double carry[R], h1[R], h2[R], x[R][4];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = tid + r * kWaveThreads;
    carry[r] = -0.0;  // x + -0 == x: the first sample of a row keeps its value and sign
    h1[r] = h2[r] = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) x[r][d] = j < ny && d < nx ? o[(long)j * nx + d] : 0;
  }
  const int steps = nx + 2 * (ny - 1);
  for (int t = 0; t < steps; ++t) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = tid + r * kWaveThreads;
      const int i = t - 2 * j;
      if (j < ny && i >= -1 && i < nx) {
        // e(i+1, j-1), written by row j-1 at step t-1 (t >= 1 whenever j >= 1)
        const bool up = j > 0 && i + 1 < nx;
        const double c = up ? eb[(t - 1) & 1][j - 1] : 0.0;
        if (i >= 0) {
          double v = x[r][0];
          if (j > 0) {
            if (i >= 1) v = v + h1[r] * (1.0 / 16);
            v = v + h2[r] * (5.0 / 16);
            if (up) v = v + c * (3.0 / 16);
          }
          v = v + carry[r];
          double qv;
          const double e = quantise<WRAP>(v, step, inv, &qv);
          eb[t & 1][j] = e;
          o[(long)j * nx + i] = qv;
          carry[r] = e * (7.0 / 16);
          x[r][0] = x[r][1];
          x[r][1] = x[r][2];
          x[r][2] = x[r][3];
          if (i + 4 < nx) x[r][3] = o[(long)j * nx + i + 4];
        }
        h1[r] = h2[r];
        h2[r] = c;
      }
    }
    __syncthreads();
  }
  for (long k = tid; k < n; k += kWaveThreads) o[k] = wrap1(o[k]);
}

// phaseToHeight: ((phase * s) % step + step) % step, s = lambda / (2 pi (n - 1)),
// step = lambda / (n - 1).
__global__ __launch_bounds__(kFB) void k_height(const double* __restrict__ ph, double* __restrict__ h, long total,
                                                double s, double step) {
  const long k = (long)blockIdx.x * kFB + threadIdx.x;
  if (k < total) h[k] = fmod(fmod(ph[k] * s, step) + step, step);
}

}  // namespace
}  // namespace doe

using namespace doe;

extern "C" {

int doe_error_diffusion(doe_ctx* ctx, const doe_field* phase, doe_field* out, int levels, int snake,
                        int wrap_first) {
  DOE_REQUIRE(ctx && phase && out, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(phase->dtype == DOE_F64 && out->dtype == DOE_F64, DOE_ERR_INVALID,
              "errorDiffusionQuantize: FP64 fields");
  DOE_REQUIRE(phase->nx == out->nx && phase->ny == out->ny && phase->batch == out->batch, DOE_ERR_SHAPE,
              "errorDiffusionQuantize: phase and output grids differ");
  {
    // the schedules read input rows ahead of the rows they write
    const char* a = static_cast<const char*>(phase->re);
    const char* b = static_cast<const char*>(out->re);
    const size_t bytes = phase->count() * sizeof(double);
    DOE_REQUIRE(a + bytes <= b || b + bytes <= a, DOE_ERR_INVALID,
                "errorDiffusionQuantize: output overlaps the phase");
  }
  DOE_REQUIRE(levels != 0, DOE_ERR_INVALID, "errorDiffusionQuantize: levels must be non-zero");
  const int nx = phase->nx, ny = phase->ny, B = phase->batch;
  const double step = (2 * M_PI) / levels;
  const double inv = 1 / step;
  const bool wrap_e = std::fabs(step) > 3.0;  // |levels| < 2: the error can reach -pi
  DOE_HIP(hipSetDevice(ctx->device));
  const double* in = static_cast<const double*>(phase->re);
  double* o = static_cast<double*>(out->re);
  if (snake && nx <= kRowMax) {
    hipLaunchKernelGGL(k_diffuse_rows, dim3(B), dim3(kRowThreads), 4 * (nx + 2 * kRowPad) * sizeof(double),
                       ctx->stream, in, o, nx, ny, step, inv, wrap_first ? 1 : 0);
  } else if (snake) {
    auto k = wrap_e ? k_diffuse_lane<true> : k_diffuse_lane<false>;
    hipLaunchKernelGGL(k, dim3((B + 63) / 64), dim3(64), 0, ctx->stream, in, o, nx, ny, B, step, inv, 1,
                       wrap_first ? 1 : 0);
  } else {
    DOE_REQUIRE(ny <= kWaveRowsMax, DOE_ERR_UNSUPPORTED, "errorDiffusionQuantize: at most 8192 rows");
    const int rows = (ny + kWaveThreads - 1) / kWaveThreads;  // rows per thread
    auto k = rows <= 1   ? (wrap_e ? k_diffuse_wave<true, 1> : k_diffuse_wave<false, 1>)
             : rows <= 2 ? (wrap_e ? k_diffuse_wave<true, 2> : k_diffuse_wave<false, 2>)
             : rows <= 4 ? (wrap_e ? k_diffuse_wave<true, 4> : k_diffuse_wave<false, 4>)
                         : (wrap_e ? k_diffuse_wave<true, 8> : k_diffuse_wave<false, 8>);
    hipLaunchKernelGGL(k, dim3(B), dim3(kWaveThreads), 0, ctx->stream, in, o, nx, ny, step, inv,
                       wrap_first ? 1 : 0);
  }
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_phase_to_height(doe_ctx* ctx, const doe_field* phase, doe_field* out, double lambda, double n_index) {
  DOE_REQUIRE(ctx && phase && out, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(phase->dtype == DOE_F64 && out->dtype == DOE_F64, DOE_ERR_INVALID, "phaseToHeight: FP64 fields");
  DOE_REQUIRE(phase->nx == out->nx && phase->ny == out->ny && phase->batch == out->batch, DOE_ERR_SHAPE,
              "phaseToHeight: phase and output grids differ");
  DOE_HIP(hipSetDevice(ctx->device));
  const double s = lambda / (2 * M_PI * (n_index - 1));
  const double step = lambda / (n_index - 1);
  const long total = (long)phase->count();
  hipLaunchKernelGGL(k_height, dim3(unsigned((total + kFB - 1) / kFB)), dim3(kFB), 0, ctx->stream,
                     static_cast<const double*>(phase->re), static_cast<double*>(out->re), total, s, step);
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

}  // extern "C"
