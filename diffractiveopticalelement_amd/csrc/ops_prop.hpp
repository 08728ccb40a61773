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
  double lim_x, inner_x, lim_y, inner_y;
  int band;  // band limit active (finite local frequency limit)
  int zero;  // z == 0: identity (propagate.js:113)
};

// Multiply the 2-D spectrum (natural FFT order) by the angular-spectrum
// transfer function H of propagate.js:57-102.  The reference builds H on the
// centred grid and brackets the product with fftshift pairs; because a
// circular shift commutes with the circular convolution, that equals
// F^-1 diag(ifftshift H) F, i.e. H sampled at the signed natural frequency
// index, which is what is evaluated here.  `conj` gives H(-z) = conj(H(z)).
struct OpAsmH {
  static constexpr int NACC = 0;
  static constexpr bool HAS_PRE = false, HAS_POST = false;
  // FP32 column tiles.  Lines of 256 and 512 points: 16 points per thread x 16 columns,
  // both LDS planes (512 threads at 128 registers, two workgroups per CU);
  // measured on 16 x 2048^2: 367-371 us per pass against 478 us for 32
  // points x 32 columns at one workgroup per CU (472 us with one plane), 410
  // us for 8-column tiles and 375 us for 16 x 16 with one plane.  Longer
  // lines would spill at 16 points per thread (1024-thread workgroups) and
  // keep 32 points per thread with both planes (the FP64 transfer-function
  // argument fits the 256-register budget unspilled).
  template <typename TT, int N>
  using ColGeomT = typename std::conditional<(N <= 512), ColGeom<TT, N, 16, 16, 2>,
                                             ColGeom<TT, N, 32, 64, 2>>::type;
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
};

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
      const double px = (mode & CH_OUTGRID) ? p.dx_out : p.dx_in;
      const double py = (mode & CH_OUTGRID) ? p.dy_out : p.dy_in;
      const double x = double(i - (nx >> 1)) * px;  // (i - nx/2) * dx
      const double y = double(j - (ny >> 1)) * py;
      // FP64 fields keep propagate.js:189's division; FP32 fields multiply by
      // the reciprocal (one rounding apart: below 1e-11 rad at the largest
      // phase of a 2048^2 grid, far under FP32's resolution) and skip the
      // division's reciprocal-and-refine sequence at every point.
      double q;
      if constexpr (sizeof(T) == 8) {
        q = (kPi * (x * x + y * y)) / p.laz;
      } else {
        q = (kPi * (x * x + y * y)) * p.inv_laz;
      }
      if (mode & CH_MINUS_HALF_PI) q = q - kPi / 2;
      cis<T>(q, cr, ci);
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
