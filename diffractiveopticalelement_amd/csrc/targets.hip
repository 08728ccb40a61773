// Target rasters, regions of interest and figure images on the device.
//
//   makeTarget (projector.js:169-255): every analytic kind evaluated per
//     pixel; the 'text' and 'image' kinds as a host raster resampled
//     bilinearly on the device (resampleRaster, projector.js:126-154);
//   autoRoi (projector.js:262-315): threshold at a fraction of the plane's
//     maximum, square dilation, box-blurred soft edge, as four separable
//     passes over byte / integer planes;
//   blurRaster (doe.js:761-794): separable Gaussian with clamped edges;
//   planeToRGB / planeToGray (color.js:7-160): a field's scalar quantity
//     mapped to 8-bit pixels, optionally laid out as PNG scanlines (each row
//     led by filter byte 0, png.js:96-101), with the data range, or a
//     percentile of the data as the upper end (an 8-pass radix select over
//     order-preserving 64-bit keys), found on the device.
//
// Arithmetic is FP64 in the reference's operation order with floating-point
// contraction off, so masks, dilation, box blur, bilinear resampling and the
// linear colour mapping equal the reference's results bit for bit; exp,
// sin, pow, log10 and atan2 come from the device math library.  Math.hypot
// is evaluated as the ECMAScript engine evaluates it (normalised,
// compensated sum of squares), so hypot-based masks match exactly too.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <tuple>
#include <vector>

#include "ctx.hpp"

#pragma clang fp contract(off)

namespace doe {
namespace {

constexpr int kTB = 256;
constexpr int kMaxTaps = 65536;  // blur half-width limit

// ---- ECMAScript number semantics used by the reference ---------------------

// Math.hypot(x, y) as V8 evaluates it: the larger magnitude m (Infinity wins
// over NaN), then sqrt of the Kahan-compensated sum of (|x|/m)^2, (|y|/m)^2,
// times m.
__device__ inline double js_hypot(double x, double y) {
  const double ax = fabs(x), ay = fabs(y);
  double m = 0;
  if (ax > m) m = ax;
  if (ay > m) m = ay;
  if (m == INFINITY) return INFINITY;
  if (x != x || y != y) return NAN;
  if (m == 0) return 0;
  double sum = 0, comp = 0;
  const double v[2] = {ax / m, ay / m};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double s = v[q] * v[q] - comp;
    const double p = sum + s;
    comp = (p - sum) - s;
    sum = p;
  }
  return sqrt(sum) * m;
}

__host__ __device__ inline double js_max(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == 0 && b == 0) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}
__host__ __device__ inline double js_min(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == 0 && b == 0) return std::signbit(a) ? a : b;
  return a < b ? a : b;
}
// Math.round: nearest integer, ties toward +Infinity.
__host__ __device__ inline double js_round(double x) {
  if (x != x || x == INFINITY || x == -INFINITY) return x;
  const double f = floor(x);
  return (x - f >= 0.5) ? f + 1 : f;
}
// Storing a number into a Uint8Array (ToUint8: truncate, modulo 256).
__device__ inline unsigned char js_u8(double v) {
  if (v != v || v == INFINITY || v == -INFINITY) return 0;
  double m = fmod(trunc(v), 256.0);
  if (m < 0) m += 256.0;
  return (unsigned char)m;
}

// Order-preserving 64-bit key of a double (TypedArray sort order: -0 before
// +0, every NaN after +Infinity).
__device__ inline unsigned long long dkey(double v) {
  unsigned long long u = v != v ? 0x7FF8000000000000ull : (unsigned long long)__double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double dval(unsigned long long k) {
  const unsigned long long u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)u);
}

template <class S>
__device__ __forceinline__ double ld(const S* p, long k) {
  return static_cast<double>(static_cast<float>(p[k]));
}
template <>
__device__ __forceinline__ double ld<float>(const float* p, long k) {
  return p[k];
}
template <>
__device__ __forceinline__ double ld<double>(const double* p, long k) {
  return p[k];
}

// Column and row of a flat index k < nx * ny <= 2^32 (32-bit division: the
// 64-bit one is a long software sequence).
__device__ __forceinline__ void split_ij(long k, int nx, int& i, int& j) {
  const unsigned kk = (unsigned)k, w = (unsigned)nx;
  const unsigned q = kk / w;
  j = (int)q;
  i = (int)(kk - q * w);
}

// ---- makeTarget ------------------------------------------------------------

struct TgtRec {
  int kind;
  int rw, rh;        // raster size
  int mcols, mrows;  // dark-cell table size (0: none)
  int gamma_on;
  long roff;         // raster offset (doubles) in the raster block
  long moff;         // dark-cell table offset (bytes) in the table block
  double p0, p1, p2; // kind parameters (radii, pitches and sigma, cells, period and duty)
  double s, ox, oy;  // raster placement: scale and offsets (disc, ring: squared-radius bounds)
  double gamma, background, gain;  // (ring: squared-radius bounds in gamma, background)
};

// Irradiance of one sample of a target of kind K (projector.js:169-255).
template <int K>
__device__ __forceinline__ double target_at(const TgtRec& R, const double* __restrict__ rasters,
                                            const unsigned char* __restrict__ dark, int i, int j,
                                            int nx, int ny, double cx, double cy) {
  // Disc, ring, zone: a = i - cx and b = j - cy are multiples of 1/2, so
  // d2 = a^2 + b^2 is exact; away from a boundary it decides the sample,
  // near one (relative 1e-9) the engine's hypot does, as in the reference.
  if constexpr (K == DOE_TGT_DISC) {
    const double a = i - cx, b = j - cy, d2 = a * a + b * b;
    if (d2 < R.s) return 1;
    if (d2 > R.ox) return 0;
    return js_hypot(a, b) <= R.p0 ? 1 : 0;
  } else if constexpr (K == DOE_TGT_RING) {
    const double a = i - cx, b = j - cy, d2 = a * a + b * b;
    if (d2 > R.oy && d2 < R.gamma) return 1;  // clearly between r0 and r1
    if (d2 < R.s || d2 > R.background) return 0;  // clearly inside r0 or outside r1
    const double r = js_hypot(a, b);
    return (r >= R.p0 && r <= R.p1) ? 1 : 0;
  } else if constexpr (K == DOE_TGT_DOTS) {
    const double px = R.p0, py = R.p1, sg = R.p2;
    // x % 1 of a positive finite x is x - floor(x), exactly
    const double qy = (j + 0.5) / py, qx = (i + 0.5) / px;
    const double fy = (qy - floor(qy)) - 0.5;
    const double fx = (qx - floor(qx)) - 0.5;
    const double ddx = fx * px, ddy = fy * py;
    double v = exp(-(ddx * ddx + ddy * ddy) / (2 * sg * sg));
    if (R.mcols > 0) {
      const long ci = (long)(i / px), cj = (long)(j / py);
      if (ci < R.mcols && cj < R.mrows && dark[R.moff + cj * R.mcols + ci]) v = 0;
    }
    return v;
  } else if constexpr (K == DOE_TGT_CHECKER) {
    const double u = floor(((double)i / nx) * R.p0);
    const double w = floor(((double)j / ny) * R.p0);
    const double h = (u + w) * 0.5;  // (u + w) % 2 === 0: an even integer
    return (isfinite(h) && h == floor(h)) ? 1 : 0;
  } else if constexpr (K == DOE_TGT_LINES) {
    return (fmod((double)i, R.p0) / R.p0) < R.p1 ? 1 : 0;
  } else if constexpr (K == DOE_TGT_ZONE) {
    // sin(pi t) > 0 with t = r^2 / period iff floor(t) is even; decided from
    // the exact d2 unless t is within 1e-6 of an integer
    const double a = i - cx, b = j - cy, d2 = a * a + b * b;
    const double t = d2 / R.p0;
    if (fabs(t) < 1e7) {
      const double fl = floor(t), fr = t - fl;
      if (fr > 1e-6 && fr < 1 - 1e-6) return (fl * 0.5 == floor(fl * 0.5)) ? 1 : 0;
    }
    const double r = js_hypot(a, b);
    return sin((M_PI * r * r) / R.p0) > 0 ? 1 : 0;
  } else {
    double v = R.background;
    const double fy = (j - R.oy) / R.s;
    const double y0 = floor(fy);
    if (y0 >= 0 && y0 < R.rh) {
      const double fx = (i - R.ox) / R.s;
      const double x0 = floor(fx);
      if (x0 >= 0 && x0 < R.rw) {
        const double ty = fy - y0, tx = fx - x0;
        const long y1 = (long)fmin(y0 + 1, (double)(R.rh - 1));
        const long x1 = (long)fmin(x0 + 1, (double)(R.rw - 1));
        const double* src = rasters + R.roff;
        const long r0 = (long)y0 * R.rw, r1 = y1 * R.rw;
        const double a = src[(long)x0 + r0] * (1 - tx) * (1 - ty);
        const double c = src[x1 + r0] * tx * (1 - ty);
        const double d = src[(long)x0 + r1] * (1 - tx) * ty;
        const double e = src[x1 + r1] * tx * ty;
        v = (((a + c) + d) + e) * R.gain;
      }
    }
    if (R.gamma_on) v = pow(js_max(0, v), R.gamma);
    return v;
  }
}

// Every sample of field b, walked by the block's threads; four samples per
// thread per step, so their independent FP64 chains interleave.  Returns the
// thread's maximum over v > m from +0 (NaN never selected), for the ROI.
template <int K>
__device__ __forceinline__ double target_field(const TgtRec& R, const double* __restrict__ rasters,
                                               const unsigned char* __restrict__ dark, double* __restrict__ o,
                                               int nx, int ny) {
  const long n = (long)nx * ny;
  const double cx = (nx - 1) / 2.0, cy = (ny - 1) / 2.0;
  const long step = (long)gridDim.x * kTB;
  long k = (long)blockIdx.x * kTB + threadIdx.x;
  double m = 0;
  for (; k + 3 * step < n; k += 4 * step) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int i, j;
      split_ij(k + u * step, nx, i, j);
      v[u] = target_at<K>(R, rasters, dark, i, j, nx, ny, cx, cy);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      o[k + u * step] = v[u];
      if (v[u] > m) m = v[u];
    }
  }
  for (; k < n; k += step) {
    int i, j;
    split_ij(k, nx, i, j);
    const double v = target_at<K>(R, rasters, dark, i, j, nx, ny, cx, cy);
    o[k] = v;
    if (v > m) m = v;
  }
  return m;
}

// part (nullable): per-block maxima of the irradiance, part[b * gridDim.x +
// block], the first half of autoRoi's maximum.
__global__ __launch_bounds__(kTB) void k_target(const TgtRec* __restrict__ recs,
                                                const double* __restrict__ rasters,
                                                const unsigned char* __restrict__ dark,
                                                double* __restrict__ out, int nx, int ny,
                                                double* __restrict__ part) {
  const int b = blockIdx.y;
  const TgtRec R = recs[b];
  double* o = out + (long)b * nx * ny;
  double m;
  switch (R.kind) {
    case DOE_TGT_DISC: m = target_field<DOE_TGT_DISC>(R, rasters, dark, o, nx, ny); break;
    case DOE_TGT_RING: m = target_field<DOE_TGT_RING>(R, rasters, dark, o, nx, ny); break;
    case DOE_TGT_DOTS: m = target_field<DOE_TGT_DOTS>(R, rasters, dark, o, nx, ny); break;
    case DOE_TGT_CHECKER: m = target_field<DOE_TGT_CHECKER>(R, rasters, dark, o, nx, ny); break;
    case DOE_TGT_LINES: m = target_field<DOE_TGT_LINES>(R, rasters, dark, o, nx, ny); break;
    case DOE_TGT_ZONE: m = target_field<DOE_TGT_ZONE>(R, rasters, dark, o, nx, ny); break;
    default: m = target_field<DOE_TGT_RASTER>(R, rasters, dark, o, nx, ny); break;
  }
  if (!part) return;
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double red[kTB / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kTB / 64; ++w) m = fmax(m, red[w]);
    part[(long)b * gridDim.x + blockIdx.x] = m;
  }
}

// ---- autoRoi ---------------------------------------------------------------

// Per-block maxima of irr over irr > m (NaN never selected, start at +0),
// part[b * gridDim.x + block]; eight independent loads in flight per thread.
__global__ __launch_bounds__(kTB) void k_roi_max(const double* __restrict__ irr, long n,
                                                 double* __restrict__ part) {
  const int b = blockIdx.y;
  const double* p = irr + (long)b * n;
  double m = 0;
  const long step = (long)gridDim.x * kTB;
  long k = (long)blockIdx.x * kTB + threadIdx.x;
  for (; k + 7 * step < n; k += 8 * step) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[k + u * step];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (v[u] > m) m = v[u];
  }
  for (; k < n; k += step) {
    const double v = p[k];
    if (v > m) m = v;
  }
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double red[kTB / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kTB / 64; ++w) m = fmax(m, red[w]);
    part[(long)b * gridDim.x + blockIdx.x] = m;
  }
}

// thr[b] = threshold * (maximum of field b's partials).
__global__ void k_roi_thr(const double* __restrict__ part, int nblk, double threshold,
                          double* __restrict__ thr) {
  const int b = blockIdx.x;
  double m = 0;
  for (int q = threadIdx.x; q < nblk; q += 64) m = fmax(m, part[(long)b * nblk + q]);
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if (threadIdx.x == 0) thr[b] = threshold * m;
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// The whole ROI of a 64 x 32 output tile, bit-packed.  The tile plus a halo
// of h = r + s is thresholded into 128-bit row masks (two wave ballots per
// row); the square dilation is an OR of 2r+1 row masks followed by 2r+1
// shifted copies (the max filter is separable, the order is free), cleared
// outside the plane; the box sum along x is a popcount of a (2s+1)-bit
// window; the sum along y a running window down each column; interior
// samples divide through a table of acc / (2s+1)^2 (the same correctly
// rounded quotients).  Integer-exact throughout, so equal to the separable
// passes below, for r + s <= kRoiHalo.
constexpr int kRoiW = 64, kRoiH = 32, kRoiHalo = 16;
constexpr int kRoiRows = kRoiH + 2 * kRoiHalo;  // 64: one lane per halo'd row

struct Bits128 {
  unsigned long long lo, hi;
};
__device__ __forceinline__ Bits128 shl128(Bits128 v, int d) {  // toward higher bit indices
  if (d == 0) return v;
  return {v.lo << d, (v.hi << d) | (v.lo >> (64 - d))};
}
__device__ __forceinline__ Bits128 shr128(Bits128 v, int d) {
  if (d == 0) return v;
  return {(v.lo >> d) | (v.hi << (64 - d)), v.hi >> d};
}
// bits [p, p + w) of v, w <= 64, p + w <= 128
__device__ __forceinline__ unsigned long long window128(Bits128 v, int p, int w) {
  const unsigned long long m = w >= 64 ? ~0ull : ((1ull << w) - 1);
  if (p >= 64) return (v.hi >> (p - 64)) & m;
  const unsigned long long x = p ? ((v.lo >> p) | (v.hi << (64 - p))) : v.lo;
  return x & m;
}
// mask of bits [a, b) of a 128-bit row, 0 <= a <= b <= 128
__device__ __forceinline__ Bits128 range128(int a, int b) {
  auto upto = [](int k) -> Bits128 {  // bits [0, k)
    if (k <= 0) return {0, 0};
    if (k < 64) return {(1ull << k) - 1, 0};
    if (k == 64) return {~0ull, 0};
    if (k < 128) return {~0ull, (1ull << (k - 64)) - 1};
    return {~0ull, ~0ull};
  };
  const Bits128 hb = upto(b), ha = upto(a);
  return {hb.lo & ~ha.lo, hb.hi & ~ha.hi};
}

__global__ __launch_bounds__(kTB) void k_roi_tile(const double* __restrict__ irr,
                                                  const double* __restrict__ thrs,
                                                  double* __restrict__ roi, int nx, int ny, int r,
                                                  int s) {
  __shared__ Bits128 binw[kRoiRows];
  __shared__ Bits128 dilw[kRoiRows];
  __shared__ unsigned char rs[kRoiRows][kRoiW];
  __shared__ double quot[(2 * kRoiHalo + 1) * (2 * kRoiHalo + 1) + 1];
  const int b = blockIdx.z;
  const long n = (long)nx * ny;
  const double* p = irr + (long)b * n;
  const double thr = thrs[b];
  const int x0 = blockIdx.x * kRoiW, y0 = blockIdx.y * kRoiH;
  const int h = r + s;
  const int BW = kRoiW + 2 * h, BH = kRoiH + 2 * h, DH = kRoiH + 2 * s;
  const int tx = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int win = 2 * s + 1, full = win * win;
  // exact interior quotients acc / (2s+1)^2
  for (int q = threadIdx.x; q <= full; q += kTB) quot[q] = (double)q / (double)full;
  // thresholds: all loads of the wave's rows in flight, then two ballots a row
  constexpr int kPer = kRoiRows / (kTB / 64);
  double va[kPer], vb[kPer];
  const int Xa = x0 - h + tx, Xb = Xa + 64;
  const bool ina = Xa >= 0 && Xa < nx, inb = tx + 64 < BW && Xb >= 0 && Xb < nx;
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int yy = w + (kTB / 64) * it;
    const int Y = y0 - h + yy;
    const bool rowin = yy < BH && Y >= 0 && Y < ny;
    const double* prow = p + (long)Y * nx;
    va[it] = (rowin && ina) ? prow[Xa] : NAN;
    vb[it] = (rowin && inb) ? prow[Xb] : NAN;
  }
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int yy = w + (kTB / 64) * it;
    const unsigned long long lo = __ballot(va[it] > thr), hi = __ballot(vb[it] > thr);
    if (tx == 0 && yy < BH) binw[yy] = {lo, hi};
  }
  __syncthreads();
  // dilation, one lane per row of the dilated band (rows y0 - s ..)
  if (w == 0 && tx < DH) {
    const int Y = y0 - s + tx;
    Bits128 v = {0, 0};
    if (Y >= 0 && Y < ny) {
      for (int d = 0; d <= 2 * r; ++d) {  // bin rows tx .. tx + 2r
        const Bits128 u = binw[tx + d];
        v.lo |= u.lo;
        v.hi |= u.hi;
      }
      Bits128 dd = v;
      for (int d = 1; d <= r; ++d) {
        const Bits128 a = shl128(v, d), c = shr128(v, d);
        dd.lo |= a.lo | c.lo;
        dd.hi |= a.hi | c.hi;
      }
      // bits xx with x0 - h + xx inside the plane
      const Bits128 m = range128(max(0, h - x0), min(BW, nx - x0 + h));
      v = {dd.lo & m.lo, dd.hi & m.hi};
    }
    dilw[tx] = v;
  }
  __syncthreads();
  const int X = x0 + tx;
  if (s == 0) {
    if (X < nx)
      for (int yy = w; yy < kRoiH && y0 + yy < ny; yy += kTB / 64)
        roi[(long)b * n + (long)(y0 + yy) * nx + X] = (double)window128(dilw[yy], tx + h, 1);
    return;
  }
  // box sums along x: popcount of bits [tx + r, tx + r + 2s] of each dilated row
  for (int yy = w; yy < DH; yy += kTB / 64) rs[yy][tx] = (unsigned char)__popcll(window128(dilw[yy], tx + r, win));
  __syncthreads();
  if (X >= nx) return;
  // sums along y: a running window over the wave's 8 output rows
  const int cx = min(nx - 1, X + s) - max(0, X - s) + 1;
  const int ya = w * (kRoiH / (kTB / 64));
  int acc = 0;
  for (int d = 0; d < win; ++d) acc += rs[ya + d][tx];
  for (int yy = ya; yy < ya + kRoiH / (kTB / 64); ++yy) {
    const int Y = y0 + yy;
    if (Y >= ny) break;
    if (yy > ya) acc += rs[yy + 2 * s][tx] - rs[yy - 1][tx];
    const int cy = min(ny - 1, Y + s) - max(0, Y - s) + 1;
    roi[(long)b * n + (long)Y * nx + X] =
        (cx == win && cy == win) ? quot[acc] : (double)acc / (double)(cx * cy);
  }
}

// Separable passes for any r, s (the tile kernel's halo is bounded).
// tmp = row-wise dilation (half-width r) of irr > thr.
__global__ __launch_bounds__(kTB) void k_roi_row(const double* __restrict__ irr,
                                                 const double* __restrict__ thrs,
                                                 unsigned char* __restrict__ tmp, int nx, int ny,
                                                 int r) {
  const int b = blockIdx.y;
  const long n = (long)nx * ny;
  const long k = (long)blockIdx.x * kTB + threadIdx.x;
  if (k >= n) return;
  const double thr = thrs[b];
  int i, j;
  split_ij(k, nx, i, j);
  (void)j;
  const double* row = irr + (long)b * n + (k - i);
  const int i0 = max(0, i - r), i1 = min(nx - 1, i + r);
  unsigned char v = 0;
  for (int ii = i0; ii <= i1 && !v; ++ii) v = row[ii] > thr;
  tmp[(long)b * n + k] = v;
}

// Column-wise dilation of tmp; the dilated mask goes to dil (bytes) or,
// without a soft edge, straight to the ROI plane.
__global__ __launch_bounds__(kTB) void k_roi_col(const unsigned char* __restrict__ tmp,
                                                 unsigned char* __restrict__ dil,
                                                 double* __restrict__ roi, int nx, int ny, int r) {
  const int b = blockIdx.y;
  const long n = (long)nx * ny;
  const long k = (long)blockIdx.x * kTB + threadIdx.x;
  if (k >= n) return;
  int i, j;
  split_ij(k, nx, i, j);
  const unsigned char* col = tmp + (long)b * n + i;
  const int j0 = max(0, j - r), j1 = min(ny - 1, j + r);
  unsigned char v = 0;
  for (int jj = j0; jj <= j1 && !v; ++jj) v = col[(long)jj * nx];
  if (roi)
    roi[(long)b * n + k] = v;
  else
    dil[(long)b * n + k] = v;
}

// Row sums of dil over the clipped window [i-s, i+s].
__global__ __launch_bounds__(kTB) void k_box_row(const unsigned char* __restrict__ dil,
                                                 int* __restrict__ rs, int nx, int ny, int s) {
  const int b = blockIdx.y;
  const long n = (long)nx * ny;
  const long k = (long)blockIdx.x * kTB + threadIdx.x;
  if (k >= n) return;
  int i, j;
  split_ij(k, nx, i, j);
  (void)j;
  const unsigned char* row = dil + (long)b * n + (k - i);
  const int i0 = max(0, i - s), i1 = min(nx - 1, i + s);
  int acc = 0;
  for (int ii = i0; ii <= i1; ++ii) acc += row[ii];
  rs[(long)b * n + k] = acc;
}

// Column sums of the row sums over [j-s, j+s], divided by the clipped
// window's sample count: exact integers, one rounding in the division.
__global__ __launch_bounds__(kTB) void k_box_col(const int* __restrict__ rs, double* __restrict__ roi,
                                                 int nx, int ny, int s) {
  const int b = blockIdx.y;
  const long n = (long)nx * ny;
  const long k = (long)blockIdx.x * kTB + threadIdx.x;
  if (k >= n) return;
  int i, j;
  split_ij(k, nx, i, j);
  const int* col = rs + (long)b * n + i;
  const int j0 = max(0, j - s), j1 = min(ny - 1, j + s);
  const int i0 = max(0, i - s), i1 = min(nx - 1, i + s);
  long acc = 0;
  for (int jj = j0; jj <= j1; ++jj) acc += col[(long)jj * nx];
  roi[(long)b * n + k] = (double)acc / (double)((long)(j1 - j0 + 1) * (i1 - i0 + 1));
}

// ---- blurRaster --------------------------------------------------------------

// Row pass: tmp = sum_d src[clamp(i + d)] * ker[d + r], d = -r..r in order.
__global__ __launch_bounds__(kTB) void k_blur_row(const double* __restrict__ src, double* __restrict__ tmp,
                                                  const double* __restrict__ ker, int nx, int ny, int r) {
  const int b = blockIdx.y;
  const long n = (long)nx * ny;
  const long k = (long)blockIdx.x * kTB + threadIdx.x;
  if (k >= n) return;
  int i, j;
  split_ij(k, nx, i, j);
  (void)j;
  const double* row = src + (long)b * n + (k - i);
  double acc = 0;
  for (int d = -r; d <= r; ++d) acc += row[min(nx - 1, max(0, i + d))] * ker[d + r];
  tmp[(long)b * n + k] = acc;
}

__global__ __launch_bounds__(kTB) void k_blur_col(const double* __restrict__ tmp, double* __restrict__ dst,
                                                  const double* __restrict__ ker, int nx, int ny, int r) {
  const int b = blockIdx.y;
  const long n = (long)nx * ny;
  const long k = (long)blockIdx.x * kTB + threadIdx.x;
  if (k >= n) return;
  int i, j;
  split_ij(k, nx, i, j);
  const double* col = tmp + (long)b * n + i;
  double acc = 0;
  for (int d = -r; d <= r; ++d) acc += col[(long)min(ny - 1, max(0, j + d)) * nx] * ker[d + r];
  dst[(long)b * n + k] = acc;
}

// ---- planeToRGB / planeToGray ------------------------------------------------

// The scalar quantity of sample k (ComplexField re / im / intensity /
// amplitude / phase, field.js:98-135).
template <class S>
__device__ __forceinline__ double quantity(const S* re, const S* im, long k, int q) {
  const double a = ld(re, k);
  if (q == DOE_Q_RE) return a;
  const double b = ld(im, k);
  switch (q) {
    case DOE_Q_IM: return b;
    case DOE_Q_INTENSITY: return a * a + b * b;
    case DOE_Q_AMPLITUDE: return js_hypot(a, b);
    default: return atan2(b, a);
  }
}

// Per-block min / max keys of the plane (NaN keys skipped unless with_nan:
// the sort order's first element), part[2 * block], part[2 * block + 1].
template <class S>
__global__ __launch_bounds__(kTB) void k_img_minmax(const S* __restrict__ re, const S* __restrict__ im,
                                                    long n, int q, int with_nan,
                                                    unsigned long long* __restrict__ part) {
  unsigned long long lo = ~0ull, hi = 0;
  for (long k = (long)blockIdx.x * kTB + threadIdx.x; k < n; k += (long)gridDim.x * kTB) {
    const double v = quantity(re, im, k, q);
    if (v != v && !with_nan) continue;
    const unsigned long long key = dkey(v);
    lo = key < lo ? key : lo;
    hi = key > hi ? key : hi;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long a = __shfl_xor(lo, off, 64), c = __shfl_xor(hi, off, 64);
    lo = a < lo ? a : lo;
    hi = c > hi ? c : hi;
  }
  __shared__ unsigned long long red[2][kTB / 64];
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = lo;
    red[1][threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kTB / 64; ++w) {
      lo = red[0][w] < lo ? red[0][w] : lo;
      hi = red[1][w] > hi ? red[1][w] : hi;
    }
    part[2 * blockIdx.x] = lo;
    part[2 * blockIdx.x + 1] = hi;
  }
}

// One radix-select pass: histogram of the 8-bit digit at `shift` over the
// keys that match the selected prefix so far.
template <class S>
__global__ __launch_bounds__(kTB) void k_sel_hist(const S* __restrict__ re, const S* __restrict__ im,
                                                  long n, int q, const unsigned long long* __restrict__ st,
                                                  unsigned* __restrict__ hist, int shift) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long prefix = st[0], mask = st[1];
  for (long k = (long)blockIdx.x * kTB + threadIdx.x; k < n; k += (long)gridDim.x * kTB) {
    const unsigned long long key = dkey(quantity(re, im, k, q));
    if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// st = {prefix, mask, rank}: pick the digit that holds the rank, clear the
// histogram for the next pass.
__global__ __launch_bounds__(256) void k_sel_pick(unsigned long long* __restrict__ st,
                                                  unsigned* __restrict__ hist, int shift) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = hist[threadIdx.x];
  hist[threadIdx.x] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long rank = st[2], cum = 0;
    int d = 255;
    for (int q = 0; q < 256; ++q) {
      if (cum + h[q] > rank) {
        d = q;
        break;
      }
      cum += h[q];
    }
    st[2] = rank - cum;
    st[0] |= (unsigned long long)d << shift;
    st[1] |= 255ull << shift;
  }
}

struct ImgParams {
  int q, gray, nstops, log, invert, png, has_min, has_max, pct, gamma_on;
  double min, max, gamma, log_floor, lf;
  double pos[8];
  double rgb[8][3];
};

// range[0] = lo, range[1] = hi as planeToRGB / planeToGray settle them, from
// the given bounds, the folded block min / max keys, or the selected key.
__global__ __launch_bounds__(256) void k_img_range(const ImgParams P, const unsigned long long* __restrict__ part,
                                                   int nblk, int scanned,
                                                   const unsigned long long* __restrict__ st,
                                                   double* __restrict__ range) {
  __shared__ unsigned long long red[2][256];
  unsigned long long lo = ~0ull, hi = 0;
  if (scanned) {
    for (int q = threadIdx.x; q < nblk; q += 256) {
      lo = part[2 * q] < lo ? part[2 * q] : lo;
      hi = part[2 * q + 1] > hi ? part[2 * q + 1] : hi;
    }
  }
  red[0][threadIdx.x] = lo;
  red[1][threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w) {
      const unsigned long long a = red[0][threadIdx.x + w], c = red[1][threadIdx.x + w];
      if (a < red[0][threadIdx.x]) red[0][threadIdx.x] = a;
      if (c > red[1][threadIdx.x]) red[1][threadIdx.x] = c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    range[0] = P.has_min ? P.min : dval(red[0][0]);
    range[1] = P.has_max ? P.max : (P.pct ? dval(st[0]) : dval(red[1][0]));
  }
}

template <class S>
__global__ __launch_bounds__(kTB) void k_image(const S* __restrict__ re, const S* __restrict__ im, int nx,
                                               long n, const ImgParams P, const double* __restrict__ range,
                                               unsigned char* __restrict__ out) {
  const long k = (long)blockIdx.x * kTB + threadIdx.x;
  if (k >= n) return;
  const double lo = range[0], hi = range[1];
  double span = hi - lo;
  if (span == 0 || span != span) span = 1;
  const double p = quantity(re, im, k, P.q);
  const int ch = P.gray ? 1 : 3;
  int xi, yi;
  split_ij(k, nx, xi, yi);
  const long x = xi, y = yi;
  const long stride = (long)P.png + (long)nx * ch;
  unsigned char* o = out + y * stride + P.png + x * ch;
  if (P.png && x == 0) out[y * stride] = 0;
  if (P.gray) {
    double t = js_max(0, js_min(1, (p - lo) / span));
    if (P.invert) t = 1 - t;
    o[0] = js_u8(js_round(t * 255));
    return;
  }
  double t;
  if (P.log) {
    const double v = js_max(p, hi * P.log_floor);
    t = (log10(v / hi) - P.lf) / -P.lf;
  } else {
    t = (p - lo) / span;
  }
  if (P.gamma_on) t = pow(js_max(0, js_min(1, t)), P.gamma);
  const double xx = js_min(1, js_max(0, t));
  for (int s = 1; s < P.nstops; ++s) {
    const double p0 = P.pos[s - 1], p1 = P.pos[s];
    if (xx <= p1) {
      const double f = p1 == p0 ? 0 : (xx - p0) / (p1 - p0);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double c0 = P.rgb[s - 1][c], c1 = P.rgb[s][c];
        o[c] = js_u8(js_round(c0 + (c1 - c0) * f));
      }
      return;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) o[c] = js_u8(P.rgb[P.nstops - 1][c]);
}

inline unsigned blocks_for(long n) { return unsigned((n + kTB - 1) / kTB); }
inline unsigned red_blocks(long n) {
  const long b = (n + kTB * 8 - 1) / (kTB * 8);
  return unsigned(std::max(1L, std::min(b, 2048L)));
}

template <class F>
int with_dt(int dt, F&& f) {
  switch (dt) {
    case DOE_F32: return f(float{});
    case DOE_F64: return f(double{});
    case DOE_F16: return f(__half{});
    case DOE_BF16: return f(__hip_bfloat16{});
  }
  set_error("unknown dtype");
  return DOE_ERR_INVALID;
}

}  // namespace
}  // namespace doe

using namespace doe;

extern "C" {

int doe_target_spec_init(doe_target_spec* s, int kind) {
  DOE_REQUIRE(s, DOE_ERR_INVALID, "null argument");
  std::memset(s, 0, sizeof(*s));
  s->kind = kind;
  s->radius = 0.35;
  s->inner = 0.25;
  s->outer = 0.4;
  s->cols = 8;
  s->rows = 8;
  s->sigma = 0.22;
  s->cells = 8;
  s->period = kind == DOE_TGT_LINES ? 12 : 16;
  s->duty = 0.5;
  s->gamma = 1;
  s->gain = 1;
  return DOE_OK;
}

}  // extern "C"

namespace doe {
namespace {

// makeTarget for every field of out; with `maxima`, the kernel also leaves
// per-block irradiance maxima in ctx->partials ([b * per + block], *per set).
int make_impl(doe_ctx* ctx, const doe_target_spec* specs, doe_field* out, unsigned* maxima) {
  DOE_REQUIRE(ctx && specs && out, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(out->dtype == DOE_F64, DOE_ERR_INVALID, "makeTarget: the target field is FP64");
  DOE_REQUIRE(out->nx >= 1 && out->ny >= 1 && out->batch >= 1, DOE_ERR_SHAPE, "makeTarget: empty grid");
  DOE_REQUIRE(out->batch <= 65535, DOE_ERR_SHAPE, "makeTarget: at most 65535 targets per call");
  const int nx = out->nx, ny = out->ny, B = out->batch;
  std::vector<TgtRec> recs(B);
  std::vector<double> rasters;
  // host raster (pointer, width, height) -> its offset in `rasters`
  std::map<std::tuple<const double*, int, int>, long> raster_refs;
  std::vector<unsigned char> dark;
  const double mn = std::min((double)nx, (double)ny);
  for (int b = 0; b < B; ++b) {
    const doe_target_spec& S = specs[b];
    TgtRec& R = recs[b];
    std::memset(&R, 0, sizeof(R));
    R.kind = S.kind;
    switch (S.kind) {
      case DOE_TGT_DISC:
        R.p0 = S.radius * mn * 0.5;
        // squared radius with a relative margin: below s inside, above ox outside
        // (NaN bounds send every sample to the exact test)
        R.s = R.p0 >= 0 ? R.p0 * R.p0 * (1 - 1e-9) : -1;
        R.ox = R.p0 >= 0 ? R.p0 * R.p0 * (1 + 1e-9) : -1;
        break;
      case DOE_TGT_RING: {
        R.p0 = S.inner * mn * 0.5;
        R.p1 = S.outer * mn * 0.5;
        // fast decisions only for 0 <= r0 <= r1; otherwise every sample takes the exact test
        const bool ok = R.p0 >= 0 && R.p1 >= R.p0;
        const double q0 = R.p0 * R.p0, q1 = R.p1 * R.p1;
        R.s = ok ? q0 * (1 - 1e-9) : -1;                  // d2 below: inside r0
        R.oy = ok ? q0 * (1 + 1e-9) : INFINITY;           // d2 above: beyond r0
        R.gamma = ok ? q1 * (1 - 1e-9) : -INFINITY;       // d2 below: within r1
        R.background = ok ? q1 * (1 + 1e-9) : INFINITY;   // d2 above: outside r1
        break;
      }
      case DOE_TGT_DOTS: {
        DOE_REQUIRE(S.cols > 0 && S.rows > 0 && std::isfinite(S.cols) && std::isfinite(S.rows),
                    DOE_ERR_INVALID, "makeTarget: dots/grid need positive cols and rows");
        DOE_REQUIRE(S.n_missing == 0 || S.missing, DOE_ERR_INVALID, "makeTarget: missing cells pointer is null");
        R.p0 = nx / S.cols;
        R.p1 = ny / S.rows;
        R.p2 = S.sigma * std::min(R.p0, R.p1);
        if (S.n_missing > 0) {
          R.mcols = int((long)((nx - 1) / R.p0) + 1);
          R.mrows = int((long)((ny - 1) / R.p1) + 1);
          DOE_REQUIRE((long)R.mcols * R.mrows <= (1L << 24), DOE_ERR_INVALID, "makeTarget: too many cells");
          R.moff = (long)dark.size();
          dark.resize(dark.size() + size_t(R.mcols) * R.mrows, 0);
          for (int m = 0; m < S.n_missing; ++m) {
            const int mi = S.missing[2 * m], mj = S.missing[2 * m + 1];
            if (mi >= 0 && mi < R.mcols && mj >= 0 && mj < R.mrows) dark[R.moff + (long)mj * R.mcols + mi] = 1;
          }
        }
        break;
      }
      case DOE_TGT_CHECKER: R.p0 = S.cells; break;
      case DOE_TGT_LINES:
        R.p0 = S.period;
        R.p1 = S.duty;
        break;
      case DOE_TGT_ZONE: R.p0 = S.period; break;
      case DOE_TGT_RASTER: {
        DOE_REQUIRE(S.raster && S.raster_w >= 1 && S.raster_h >= 1, DOE_ERR_INVALID,
                    "makeTarget: a raster target needs {raster, width, height}");
        R.rw = S.raster_w;
        R.rh = S.raster_h;
        const double sx = (double)nx / S.raster_w, sy = (double)ny / S.raster_h;
        R.s = S.cover ? std::max(sx, sy) : std::min(sx, sy);
        R.ox = (nx - S.raster_w * R.s) / 2;
        R.oy = (ny - S.raster_h * R.s) / 2;
        R.gamma_on = S.gamma != 1;
        R.gamma = S.gamma;
        R.background = S.background;
        R.gain = S.gain;
        // specs that share a host raster (same array and size) share one copy
        const auto key = std::make_tuple(S.raster, S.raster_w, S.raster_h);
        const auto hit = raster_refs.find(key);
        if (hit != raster_refs.end()) {
          R.roff = hit->second;
        } else {
          R.roff = (long)rasters.size();
          raster_refs.emplace(key, R.roff);
          rasters.insert(rasters.end(), S.raster, S.raster + (size_t)S.raster_w * S.raster_h);
        }
        break;
      }
      default: set_error("makeTarget: unknown kind"); return DOE_ERR_INVALID;
    }
  }
  DOE_HIP(hipSetDevice(ctx->device));
  // one parameter block: records | rasters | dark-cell tables
  const size_t rb = recs.size() * sizeof(TgtRec), ab = rasters.size() * sizeof(double);
  std::vector<char> blk(rb + ab + dark.size() + 8);
  std::memcpy(blk.data(), recs.data(), rb);
  if (ab) std::memcpy(blk.data() + rb, rasters.data(), ab);
  if (!dark.empty()) std::memcpy(blk.data() + rb + ab, dark.data(), dark.size());
  void* dev = nullptr;
  int rc = upload_params(ctx, blk.data(), blk.size(), &dev);
  if (rc) return rc;
  char* base = static_cast<char*>(dev);
  const long n = (long)nx * ny;
  // a few thousand workgroups in all, each walking its field's grid
  const unsigned per = std::max(1u, std::min(blocks_for(n), (4096u + B - 1) / B));
  double* part = nullptr;
  if (maxima) {
    // room for the maxima and, after them, the per-field thresholds of the ROI
    rc = ensure_reduce(ctx, size_t(B) * (per + 1));
    if (rc) return rc;
    part = ctx->partials;
    *maxima = per;
  }
  prof_begin(ctx, "target_make", n * B);
  hipLaunchKernelGGL(k_target, dim3(per, B), dim3(kTB), 0, ctx->stream, (const TgtRec*)base,
                     (const double*)(base + rb), (const unsigned char*)(base + rb + ab), (double*)out->re, nx,
                     ny, part);
  prof_end(ctx);
  DOE_HIP(hipGetLastError());
  return DOE_OK;
}

// autoRoi of every field of t; `known` > 0: ctx->partials already holds
// `known` per-block maxima per field (left by make_impl).
int roi_impl(doe_ctx* ctx, doe_field* t, double dilate, double soft, double threshold, unsigned known) {
  const int nx = t->nx, ny = t->ny, B = t->batch;
  const long n = (long)nx * ny;
  const int big = std::max(nx, ny);
  const int r = (int)std::min((double)big, js_max(1, js_round(dilate)));
  const bool has_soft = !(soft <= 0);
  const int s = has_soft ? (int)std::min((double)big, js_max(1, js_round(soft))) : 0;
  const long nblk = known ? long(known) : std::max(1L, std::min(64L, (n + kTB * 16 - 1) / (kTB * 16)));
  int rc = ensure_reduce(ctx, size_t(B) * (nblk + 1));  // no reallocation when `known`
  if (rc) return rc;
  double* part = ctx->partials;
  double* thr = ctx->partials + size_t(B) * nblk;
  double* irr = static_cast<double*>(t->re);
  double* roi = static_cast<double*>(t->im);
  hipStream_t st = ctx->stream;
  prof_begin(ctx, "target_roi", n * B);
  if (!known) hipLaunchKernelGGL(k_roi_max, dim3(unsigned(nblk), B), dim3(kTB), 0, st, irr, n, part);
  hipLaunchKernelGGL(k_roi_thr, dim3(B), dim3(64), 0, st, (const double*)part, int(nblk), threshold, thr);
  if (r + s <= kRoiHalo) {
    const dim3 g((nx + kRoiW - 1) / kRoiW, (ny + kRoiH - 1) / kRoiH, B);
    hipLaunchKernelGGL(k_roi_tile, g, dim3(kTB), 0, st, irr, (const double*)thr, roi, nx, ny, r, s);
  } else {
    rc = ensure_work(ctx, size_t(n) * B * sizeof(int));
    if (rc) return rc;
    unsigned char* tmp = static_cast<unsigned char*>(ctx->work_re);
    unsigned char* dil = tmp + size_t(n) * B;
    int* rsum = static_cast<int*>(ctx->work_im);
    const dim3 g(blocks_for(n), B);
    hipLaunchKernelGGL(k_roi_row, g, dim3(kTB), 0, st, irr, (const double*)thr, tmp, nx, ny, r);
    hipLaunchKernelGGL(k_roi_col, g, dim3(kTB), 0, st, tmp, dil, has_soft ? nullptr : roi, nx, ny, r);
    if (has_soft) {
      hipLaunchKernelGGL(k_box_row, g, dim3(kTB), 0, st, dil, rsum, nx, ny, s);
      hipLaunchKernelGGL(k_box_col, g, dim3(kTB), 0, st, rsum, roi, nx, ny, s);
    }
  }
  prof_end(ctx);
  DOE_HIP(hipGetLastError());
  return DOE_OK;
}

int check_roi_args(const doe_field* t, double dilate, double soft, double threshold) {
  DOE_REQUIRE(t->dtype == DOE_F64, DOE_ERR_INVALID, "autoRoi: the target field is FP64");
  DOE_REQUIRE(t->nx >= 1 && t->ny >= 1 && t->batch >= 1, DOE_ERR_SHAPE, "autoRoi: empty grid");
  DOE_REQUIRE(t->batch <= 65535, DOE_ERR_SHAPE, "autoRoi: at most 65535 fields per call");
  DOE_REQUIRE(dilate == dilate && soft == soft && threshold == threshold, DOE_ERR_INVALID,
              "autoRoi: dilate, soft and threshold must be numbers");
  return DOE_OK;
}

}  // namespace
}  // namespace doe

extern "C" {

int doe_target_make(doe_ctx* ctx, const doe_target_spec* specs, doe_field* out) {
  DOE_REQUIRE(ctx && specs && out, DOE_ERR_INVALID, "null argument");
  int rc = make_impl(ctx, specs, out, nullptr);
  if (rc) return rc;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_target_auto_roi(doe_ctx* ctx, doe_field* t, double dilate, double soft, double threshold) {
  DOE_REQUIRE(ctx && t, DOE_ERR_INVALID, "null argument");
  int rc = check_roi_args(t, dilate, soft, threshold);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(ctx->device));
  rc = roi_impl(ctx, t, dilate, soft, threshold, 0);
  if (rc) return rc;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_target_make_roi(doe_ctx* ctx, const doe_target_spec* specs, doe_field* out, double dilate,
                        double soft, double threshold) {
  DOE_REQUIRE(ctx && specs && out, DOE_ERR_INVALID, "null argument");
  int rc = check_roi_args(out, dilate, soft, threshold);
  if (rc) return rc;
  unsigned per = 0;
  rc = make_impl(ctx, specs, out, &per);
  if (rc) return rc;
  rc = roi_impl(ctx, out, dilate, soft, threshold, per);
  if (rc) return rc;
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_raster_blur(doe_ctx* ctx, const doe_field* src, doe_field* dst, double sigma, const double* taps) {
  DOE_REQUIRE(ctx && src && dst, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(src->dtype == DOE_F64 && dst->dtype == DOE_F64, DOE_ERR_INVALID, "blurRaster: FP64 fields");
  DOE_REQUIRE(src->nx == dst->nx && src->ny == dst->ny && src->batch == dst->batch, DOE_ERR_SHAPE,
              "blurRaster: source and destination grids differ");
  DOE_REQUIRE(sigma == sigma, DOE_ERR_INVALID, "blurRaster: sigma must be a number");
  DOE_REQUIRE(src->batch <= 65535, DOE_ERR_SHAPE, "blurRaster: at most 65535 fields per call");
  const int nx = src->nx, ny = src->ny, B = src->batch;
  const long n = (long)nx * ny;
  DOE_HIP(hipSetDevice(ctx->device));
  if (sigma <= 0) {
    if (src->re != dst->re)
      DOE_HIP(hipMemcpyAsync(dst->re, src->re, size_t(n) * B * sizeof(double), hipMemcpyDeviceToDevice,
                             ctx->stream));
    DOE_HIP(finish(ctx));
    return DOE_OK;
  }
  const double rr = std::max(1.0, std::ceil(3 * sigma));
  DOE_REQUIRE(rr <= kMaxTaps, DOE_ERR_INVALID, "blurRaster: sigma too large (half-width above 65536)");
  const int r = (int)rr;
  std::vector<double> ker(2 * r + 1);
  if (taps) {
    std::copy(taps, taps + ker.size(), ker.begin());
  } else {
    double ks = 0;
    for (int i = -r; i <= r; ++i) {
      ker[i + r] = std::exp(-(double)(i * (long)i) / (2 * sigma * sigma));
      ks += ker[i + r];
    }
    for (double& k : ker) k /= ks;
  }
  void* dker = nullptr;
  int rc = upload_params(ctx, ker.data(), ker.size() * sizeof(double), &dker);
  if (rc) return rc;
  rc = ensure_work(ctx, size_t(n) * B * sizeof(double));
  if (rc) return rc;
  double* tmp = static_cast<double*>(ctx->work_re);
  const dim3 g(blocks_for(n), B);
  hipLaunchKernelGGL(k_blur_row, g, dim3(kTB), 0, ctx->stream, (const double*)src->re, tmp,
                     (const double*)dker, nx, ny, r);
  hipLaunchKernelGGL(k_blur_col, g, dim3(kTB), 0, ctx->stream, (const double*)tmp, (double*)dst->re,
                     (const double*)dker, nx, ny, r);
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(ctx));
  return DOE_OK;
}

int doe_image_opts_init(doe_image_opts* o) {
  DOE_REQUIRE(o, DOE_ERR_INVALID, "null argument");
  std::memset(o, 0, sizeof(*o));
  o->quantity = DOE_Q_RE;
  o->nstops = 2;
  o->stop_pos[1] = 1;
  for (int c = 0; c < 3; ++c) o->stop_rgb[1][c] = 255;
  o->gamma = 1;
  o->log_floor = 1e-4;
  return DOE_OK;
}

int doe_field_image(doe_ctx* ctx, const doe_field* f, int index, const doe_image_opts* o, unsigned char* out,
                    long out_len, double* range_out) {
  DOE_REQUIRE(ctx && f && o && out, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(index >= 0 && index < f->batch, DOE_ERR_SHAPE, "image: field index outside the batch");
  DOE_REQUIRE(o->quantity >= DOE_Q_RE && o->quantity <= DOE_Q_PHASE, DOE_ERR_INVALID, "image: unknown quantity");
  DOE_REQUIRE(o->gray || (o->nstops >= 1 && o->nstops <= 8), DOE_ERR_INVALID,
              "image: a colour map has 1 to 8 stops");
  DOE_REQUIRE(!o->has_percentile || (o->percentile >= 0 && o->percentile == o->percentile), DOE_ERR_INVALID,
              "image: percentile must be a non-negative number");
  const int nx = f->nx, ny = f->ny;
  const long n = (long)nx * ny;
  DOE_REQUIRE(n >= 1, DOE_ERR_SHAPE, "image: empty grid");
  const int ch = o->gray ? 1 : 3;
  const long need = (long)ny * ((o->png_rows ? 1 : 0) + (long)nx * ch);
  DOE_REQUIRE(out_len >= need, DOE_ERR_SHAPE, "image: output buffer too small");
  ImgParams P;
  std::memset(&P, 0, sizeof(P));
  P.q = o->quantity;
  P.gray = o->gray ? 1 : 0;
  P.nstops = o->nstops;
  P.log = o->log ? 1 : 0;
  P.invert = o->invert ? 1 : 0;
  P.png = o->png_rows ? 1 : 0;
  P.has_min = o->has_min ? 1 : 0;
  P.has_max = o->has_max ? 1 : 0;
  // planeToGray has no percentile option (color.js:140-160)
  P.pct = (!P.gray && o->has_percentile) ? 1 : 0;
  P.gamma_on = o->gamma != 1;
  P.min = o->min;
  P.max = o->max;
  P.gamma = o->gamma;
  P.log_floor = o->log_floor;
  P.lf = std::log10(js_max(1e-12, o->log_floor));
  for (int s = 0; s < 8; ++s) {
    P.pos[s] = o->stop_pos[s];
    for (int c = 0; c < 3; ++c) P.rgb[s][c] = o->stop_rgb[s][c];
  }
  DOE_HIP(hipSetDevice(ctx->device));
  int rc = ensure_work(ctx, size_t(need));
  if (rc) return rc;
  // reduction scratch: block min / max keys (2 per block), st[3], range[2], hist[256] (u32)
  const unsigned nblk = std::max(1u, std::min(1024u, (unsigned)((n + kTB * 8 - 1) / (kTB * 8))));
  const size_t o_st = 2 * size_t(nblk), o_rg = o_st + 3, o_h = o_rg + 3;
  rc = ensure_reduce(ctx, o_h + 128);
  if (rc) return rc;
  unsigned long long* part = reinterpret_cast<unsigned long long*>(ctx->partials);
  unsigned long long* st = part + o_st;
  double* range = ctx->partials + o_rg;
  unsigned* hist = reinterpret_cast<unsigned*>(ctx->partials + o_h);
  unsigned char* img = static_cast<unsigned char*>(ctx->work_re);
  hipStream_t sm = ctx->stream;
  const size_t plane = size_t(n) * dtype_size(f->dtype);
  const void* re = static_cast<const char*>(f->re) + plane * index;
  const void* im = static_cast<const char*>(f->im) + plane * index;
  const bool need_scan = !(P.has_min && P.has_max);
  rc = with_dt(f->dtype, [&](auto tag) -> int {
    using S = decltype(tag);
    const S* r_ = static_cast<const S*>(re);
    const S* i_ = static_cast<const S*>(im);
    if (need_scan) {
      hipLaunchKernelGGL(k_img_minmax<S>, dim3(nblk), dim3(kTB), 0, sm, r_, i_, n, P.q, P.pct, part);
      if (P.pct && !P.has_max) {
        const double kk = std::floor((double)n * o->percentile);
        const unsigned long long rank = (unsigned long long)std::min((double)(n - 1), kk);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, sm, st, 2, 0ull);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, sm, st + 2, 1, rank);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(128), 0, sm,
                           reinterpret_cast<unsigned long long*>(hist), 128, 0ull);
        const unsigned sblk = std::min(nblk, 512u);
        for (int shift = 56; shift >= 0; shift -= 8) {
          hipLaunchKernelGGL(k_sel_hist<S>, dim3(sblk), dim3(kTB), 0, sm, r_, i_, n, P.q,
                             (const unsigned long long*)st, hist, shift);
          hipLaunchKernelGGL(k_sel_pick, dim3(1), dim3(256), 0, sm, st, hist, shift);
        }
      }
    }
    hipLaunchKernelGGL(k_img_range, dim3(1), dim3(256), 0, sm, P, (const unsigned long long*)part, int(nblk),
                       need_scan ? 1 : 0, (const unsigned long long*)st, range);
    hipLaunchKernelGGL(k_image<S>, dim3(blocks_for(n)), dim3(kTB), 0, sm, r_, i_, nx, n, P,
                       (const double*)range, img);
    return DOE_OK;
  });
  if (rc) return rc;
  DOE_HIP(hipGetLastError());
  DOE_HIP(hipMemcpyAsync(out, img, size_t(need), hipMemcpyDeviceToHost, sm));
  if (range_out) DOE_HIP(hipMemcpyAsync(range_out, range, 2 * sizeof(double), hipMemcpyDeviceToHost, sm));
  DOE_HIP(hipStreamSynchronize(sm));
  return DOE_OK;
}

}  // extern "C"
