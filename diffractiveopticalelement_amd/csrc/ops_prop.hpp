// Elementwise hooks for the free-space propagators (src/propagate.js).
//
// Phase arguments that reach large magnitudes (kz*z ~ 1e5 rad, Fresnel chirps
// ~ 1e4 rad) are formed in FP64 with the same operation order as the
// reference and without FMA contraction, so the FP64 check mode reproduces
// the reference's arguments bit for bit.  The FP32 path converts the FP64
// argument to revolutions, keeps the fractional part (exact in FP64) and
// evaluates the hardware sine/cosine, whose argument is in revolutions (max
// abs error 1.2e-7 on [-0.5, 0.5] revolutions, measured on gfx950 against
// FP64; FP32 sincosf measures 7e-8 on the same grid, at ~4x the instructions).
#pragma once
#include "passes.hpp"

namespace doe {

constexpr double kPi = 3.141592653589793;
constexpr double kInvTwoPi = 0.15915494309189535;    // 1 / (2*pi)

// exp(i*phi) for an FP64 phase, evaluated in precision T.
template <typename T>
__device__ __forceinline__ void cis(double phi, T& c, T& s) {
  if constexpr (sizeof(T) == 8) {
    double sd, cd;
    sincos(phi, &sd, &cd);
    c = cd;
    s = sd;
  } else {
    const double rev = phi * kInvTwoPi;
    const float fr = float(rev - rint(rev));
    c = __builtin_amdgcn_cosf(fr);
    s = __builtin_amdgcn_sinf(fr);
  }
}

template <typename T>
__device__ __forceinline__ void cmul(T& xr, T& xi, T br, T bi) {
  const T r = xr * br - xi * bi;
  xi = xr * bi + xi * br;
  xr = r;
}

struct OpNone {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = false, HAS_POST = false;
  static constexpr bool COL_WIDE = true;
  static constexpr bool FP64_SPLIT = true;
  // the op for a chunk of the batch starting at field f0
  OpNone at(int) const { return *this; }
  template <typename T>
  __device__ __forceinline__ void pre(int, int, int, T&, T&, double*) const {}
  template <typename T>
  __device__ __forceinline__ void mid(int, int, int, T&, T&, double*) const {}
  template <typename T>
  __device__ __forceinline__ void post(int, int, int, T&, T&, double*) const {}
};

// Raised-cosine window of propagate.js:38-42.
__device__ __forceinline__ double soft_window(double v, double lim, double inner) {
  if (v <= inner) return 1.0;
  if (v >= lim) return 0.0;
  return 0.5 * (1.0 + cos((kPi * (v - inner)) / (lim - inner)));
}

// Per-field parameters for the angular-spectrum transfer function.
struct AsmField {
  double two_pi_z;  // (2*pi)*z, formed on the host as the reference does
  double z;         // z itself: the phase in revolutions is z * sqrt(arg)
  double lim_x, inner_x, lim_y, inner_y;
  int band;  // band limit active (finite local frequency limit)
  int zero;  // z == 0: identity (propagate.js:113)
};

// sqrt(x) for a positive normal double: the hardware reciprocal square-root
// estimate refined by Newton-Raphson steps for 1/sqrt and sqrt together
// (x*y0, y0/2; one coupled correction; one residual correction), good to an
// ulp or two; the FP32 transfer function only needs ~1e-14 relative.
__device__ __forceinline__ double sqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  return fma(d, h, g);
}

// Multiply the 2-D spectrum (natural FFT order) by the angular-spectrum
// transfer function H of propagate.js:57-102.  The reference builds H on the
// centred grid and brackets the product with fftshift pairs; because a
// circular shift commutes with the circular convolution, that equals
// F^-1 diag(ifftshift H) F, i.e. H sampled at the signed natural frequency
// index, which is what is evaluated here.  `conj` gives H(-z) = conj(H(z)).
//
// FP64 fields evaluate H per point exactly in the reference's operation
// order (h_at).  FP32 fields split the work by axis: the fused column pass
// (one workgroup = 32 columns x one column block of the grid) first fills an
// LDS table with fy^2 and the y window for each of its rows, each thread
// takes its column's fx terms (inv_l2 - fx^2, the x window, 1/N of the
// inverse transform) once, and a point costs a table read, a subtraction,
// the square root and the phase reduction in FP64, and the hardware sine and
// cosine.  The argument keeps the reference's order, (inv_l2 - fx*fx) -
// fy*fy, so it equals the FP64 path's bit for bit; the windows are rounded
// to FP32 after the reference's FP64 formula.
struct OpAsmH {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = false, HAS_POST = false;
  // FP32 column tiles: 32 points per thread x 32 columns, one LDS plane (two
  // 512-thread workgroups per CU for 512-point lines); 32-column tiles
  // address LDS without padding.  Longer lines keep 32 points per thread
  // with both planes.
  template <typename TT, int N>
  using ColGeomT = typename std::conditional<(N <= 512), ColGeom<TT, N, 32, 32, 1>,
                                             ColGeom<TT, N, 32, 64, 2>>::type;
  // FP32 lines up to 2048 points (the table has to fit the LDS beside the
  // line planes): per-line table of fy^2 (N doubles) and the y window (N
  // floats); longer lines evaluate h_at per point
  template <typename TT, int N>
  static constexpr bool TABLED = sizeof(TT) == 4 && N <= 2048;
  template <typename TT, int N>
  static constexpr int LINE_TAB = TABLED<TT, N> ? N + (N + 1) / 2 : 0;
  template <typename TT, int N>
  static constexpr bool SCALES_INV = TABLED<TT, N>;
  template <typename TT, int N>
  static constexpr bool COL_CTX = TABLED<TT, N>;
  const AsmField* fp;
  double dfx, dfy, inv_l2;
  int nx, ny, evanescent;
  int conj;  // 1: multiply by conj(H) = H(-z)
  OpAsmH at(int f0) const {
    OpAsmH o = *this;
    o.fp += f0;
    return o;
  }
  template <typename T>
  __device__ __forceinline__ void pre(int, int, int, T&, T&, double*) const {}
  template <typename T>
  __device__ __forceinline__ void post(int, int, int, T&, T&, double*) const {}
  template <typename T>
  __device__ __forceinline__ void mid(int f, int kx, int ky, T& re, T& im, double*) const {
    T hr, hi;
    h_at(f, kx, ky, hr, hi);
    if (conj) hi = -hi;
    cmul(re, im, hr, hi);
  }

  // ---- FP32 column pass: per-line table, per-column constants ----------
  template <typename T>
  struct ColCtx {
    double ax;   // inv_l2 - fx*fx
    double zr;   // +-z: revolutions per unit of sqrt(arg); the sign conjugates H
    float wxs;   // x window / N (the inverse column transform's scale)
    float invn;  // 1 / N
    int zero;    // z == 0: H = 1
  };
  template <typename T, int N>
  __device__ __forceinline__ void line_table(double* tab, int f, int k1, int n1cnt, int tid,
                                             int nthr) const {
#pragma clang fp contract(off)
    const AsmField p = fp[f];
    float* wtab = reinterpret_cast<float*>(tab + N);
    for (int q = tid; q < N; q += nthr) {
      const int ky = k1 + n1cnt * q;
      const int sy = ky < (ny >> 1) ? ky : ky - ny;
      const double fy = double(sy) * dfy;
      tab[q] = fy * fy;
      wtab[q] = p.band ? float(soft_window(fabs(fy), p.lim_y, p.inner_y)) : 1.0f;
    }
  }
  template <typename T, int N>
  __device__ __forceinline__ ColCtx<T> col_ctx(int f, int kx) const {
#pragma clang fp contract(off)
    const AsmField p = fp[f];
    const int sx = kx < (nx >> 1) ? kx : kx - nx;
    const double fx = double(sx) * dfx;
    ColCtx<T> c;
    c.ax = inv_l2 - fx * fx;
    c.zr = conj ? -p.z : p.z;
    const double wx = p.band ? soft_window(fabs(fx), p.lim_x, p.inner_x) : 1.0;
    c.invn = 1.0f / float(N);
    c.wxs = float(wx) * c.invn;
    c.zero = p.zero;
    return c;
  }
  // H at line position q of the column: branch-free (selects only), so the
  // caller's registers are never updated conditionally
  template <typename T, int N, class A, class X>
  __device__ __forceinline__ void mid_line(const ColCtx<T>& c, const double* tab, int q, int, int,
                                           int, T& re, T& im, A*, bool, const X&) const {
    static_assert(sizeof(T) == 4, "the table-driven transfer function is the FP32 path");
    const double arg = c.ax - tab[q];
    const float wv = c.wxs * reinterpret_cast<const float*>(tab + N)[q];
    // outside the band limit (w = 0) and for evanescent waves H is 0; the
    // square root's argument is kept normal (1e-300: a phase below 1e-140)
    const float w = (arg >= 0.0 || !evanescent) ? wv : 0.0f;
    const double sq = sqrt_pos(fmax(fabs(arg), 1e-300));
    const double rev = c.zr * sq;
    const float fr = float(rev - rint(rev));
    const float hr = c.zero ? c.invn : w * __builtin_amdgcn_cosf(fr);
    const float hi = c.zero ? 0.0f : w * __builtin_amdgcn_sinf(fr);
    cmul(re, im, hr, hi);
  }

  template <typename T>
  __device__ __forceinline__ void h_at(int f, int kx, int ky, T& hr, T& hi) const {
#pragma clang fp contract(off)
    const AsmField p = fp[f];
    if (p.zero) {
      hr = T(1);
      hi = T(0);
      return;
    }
    const int sx = kx < (nx >> 1) ? kx : kx - nx;
    const int sy = ky < (ny >> 1) ? ky : ky - ny;
    const double fx = double(sx) * dfx;
    const double fy = double(sy) * dfy;
    double w = 1.0;
    if (p.band) w = soft_window(fabs(fx), p.lim_x, p.inner_x) * soft_window(fabs(fy), p.lim_y, p.inner_y);
    // Outside the band limit H is 0 whatever the phase: leave before the
    // square root and the trig (at long distances most of the spectrum).
    if (w == 0.0) {
      hr = T(0);
      hi = T(0);
      return;
    }
    const double arg = inv_l2 - fx * fx - fy * fy;
    double phase;
    if (arg >= 0.0) {
      phase = p.two_pi_z * sqrt(arg);
    } else if (evanescent) {
      hr = T(0);
      hi = T(0);
      return;
    } else {
      phase = p.two_pi_z * sqrt(-arg);
    }
    T c, s;
    cis<T>(phase, c, s);
    hr = T(w) * c;
    hi = T(w) * s;
  }
};

// Single-FFT Fresnel / lens-transform chirp factors.
struct ChirpField {
  double inv_laz;   // 1 / (lambda*|z|) or 1/(lambda*f)
  double laz;       // lambda*|z|  (divisor kept for the FP64 check mode)
  double dx_in, dy_in, dx_out, dy_out;
  double pre_r, pre_i;  // complex prefactor applied on the output side
  // FP32 fields: the chirp phase in revolutions per squared grid index,
  // dx^2 / (2 lambda |z|) on the input and output grids (set by chirp_revs)
  double rx_in, ry_in, rx_out, ry_out;
};
inline void chirp_revs(ChirpField& p) {
  p.rx_in = 0.5 * p.inv_laz * p.dx_in * p.dx_in;
  p.ry_in = 0.5 * p.inv_laz * p.dy_in * p.dy_in;
  p.rx_out = 0.5 * p.inv_laz * p.dx_out * p.dx_out;
  p.ry_out = 0.5 * p.inv_laz * p.dy_out * p.dy_out;
}

// mode bits for OpChirp
enum : int {
  CH_QUAD = 1,      // multiply by exp(+-i*pi*(x^2+y^2)/(lambda*z))
  CH_CONJ = 2,      // conjugate chirp (exp(-i*...))
  CH_CHECKER = 4,   // multiply by (-1)^(x+y)
  CH_PREFAC = 8,    // multiply by the per-field complex prefactor
  CH_OUTGRID = 16,  // coordinates use the output pitch (u,v) instead of (x,y)
  CH_MINUS_HALF_PI = 32,  // lens: phase q - pi/2 folded into the trig argument
};

// Elementwise chirp/checkerboard/prefactor applied as a pre hook (first FFT
// side) or post hook (last FFT side); FIELD_SIDE selects which hook is live.
template <int PRE_MODE, int POST_MODE>
struct OpChirp {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = PRE_MODE != 0, HAS_POST = POST_MODE != 0;
  // FP32 column tiles: 512-point lines use the plain transform's tile (32
  // points x 32 columns, one LDS plane, two workgroups per CU): 228 us per
  // chirp column pass on 16 x 2048^2 against 280 us for the generic 16 x 64
  // two-plane tile; longer lines spill with it and keep the generic tile.
  template <typename TT, int N>
  using ColGeomT = typename std::conditional<N == 512, ColGeom<TT, N, 32, 32, 1>,
                                             ColCfg<TT, N, false>>::type;
  const ChirpField* fp;
  int nx, ny;
  OpChirp at(int f0) const {
    OpChirp o = *this;
    o.fp += f0;
    return o;
  }
  template <typename T>
  __device__ __forceinline__ void apply(int mode, int f, int i, int j, T& re, T& im) const {
#pragma clang fp contract(off)
    if (mode == 0) return;
    const ChirpField p = fp[f];
    T cr = T(1), ci = T(0);
    if (mode & CH_QUAD) {
      if constexpr (sizeof(T) == 8) {
        // FP64 fields: propagate.js:189's operation order
        const double px = (mode & CH_OUTGRID) ? p.dx_out : p.dx_in;
        const double py = (mode & CH_OUTGRID) ? p.dy_out : p.dy_in;
        const double x = double(i - (nx >> 1)) * px;  // (i - nx/2) * dx
        const double y = double(j - (ny >> 1)) * py;
        double q = (kPi * (x * x + y * y)) / p.laz;
        if (mode & CH_MINUS_HALF_PI) q = q - kPi / 2;
        cis<T>(q, cr, ci);
      } else {
        // FP32 fields: the phase in revolutions from the exact integer
        // squares of the grid indices, rx*ix^2 + ry*iy^2 (one FMA; the
        // per-row or per-column term is loop-invariant), then the fraction
        // and the hardware sine / cosine.  Within 1e-11 rad of the
        // reference's argument at the largest phase of a 4096^2 grid.
        const int ix = i - (nx >> 1), iy = j - (ny >> 1);
        const double rx = (mode & CH_OUTGRID) ? p.rx_out : p.rx_in;
        const double ry = (mode & CH_OUTGRID) ? p.ry_out : p.ry_in;
        double rev = fma(rx, double(ix * ix), ry * double(iy * iy));
        if (mode & CH_MINUS_HALF_PI) rev = rev - 0.25;
        const float fr = float(rev - rint(rev));
        cr = __builtin_amdgcn_cosf(fr);
        ci = __builtin_amdgcn_sinf(fr);
      }
      if (mode & CH_CONJ) ci = -ci;
    }
    if (mode & CH_PREFAC) {
      // output-side order of the reference: t = U * prefactor, then * chirp
      cmul(re, im, T(p.pre_r), T(p.pre_i));
    }
    if ((mode & CH_CHECKER) && ((i + j) & 1)) {
      cr = -cr;
      ci = -ci;
    }
    cmul(re, im, cr, ci);
  }
  template <typename T>
  __device__ __forceinline__ void pre(int f, int x, int y, T& re, T& im, double*) const {
    apply(PRE_MODE, f, x, y, re, im);
  }
  template <typename T>
  __device__ __forceinline__ void mid(int, int, int, T&, T&, double*) const {}
  template <typename T>
  __device__ __forceinline__ void post(int f, int x, int y, T& re, T& im, double*) const {
    apply(POST_MODE, f, x, y, re, im);
  }
};

}  // namespace doe
