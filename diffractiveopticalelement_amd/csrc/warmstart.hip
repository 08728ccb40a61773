// Hand-off between the batched far-field engine and the warm-start network:
// network inputs from the designs' targets, training labels from converged
// designs, and the initial element phase from a prediction.
//
// Index spaces.  The far-field loop works on x~ = c .* x (c = (-1)^(x+y));
// the reference's centred target-plane field is Y = fftshift(F(ifftshift x))
// = g c .* F(x~) with g = (-1)^(nx/2 + ny/2) (fourierCgh, doe.js:639-730),
// and its inverse is x = g c .* F^-1(c .* Y).  Targets, ROI and the
// network's target-plane phasors live in Y's (centred) index space.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gs_batch.hpp"

struct doe_wsnet;
extern "C" int doe_wsnet_forward(doe_wsnet* N, int batch, const float* input, float* output);
extern "C" int doe_wsnet_info(doe_wsnet* N, int* res, int* max_batch);

namespace doe {
namespace {

// inputs: pooled target irradiance (f x f blocks of t^2), later scaled to max 1
__global__ void k_ws_pool_input(const float* __restrict__ t, float* __restrict__ in, int n, int res,
                                int nd) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)res * res;
  if (i >= per * nd) return;
  const int d = int(i / per);
  const int r = int(i - (long)d * per);
  const int v = r / res, u = r - v * res;
  const int f = n / res;
  const float* td = t + (long)d * n * n;
  float s = 0.f;
  for (int y = 0; y < f; ++y)
    for (int x = 0; x < f; ++x) {
      const float a = td[(long)(v * f + y) * n + u * f + x];
      s += a * a;
    }
  in[i] = s / float(f * f);
}
// per image: divide by the maximum (one workgroup per image)
__global__ void __launch_bounds__(256) k_ws_norm_max(float* __restrict__ in, long per) {
  __shared__ float sh[256];
  float* x = in + (long)blockIdx.x * per;
  float m = 0.f;
  for (long k = threadIdx.x; k < per; k += 256) m = fmaxf(m, x[k]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + st]);
    __syncthreads();
  }
  const float mx = sh[0];
  const float s = mx > 0.f ? 1.f / mx : 0.f;
  for (long k = threadIdx.x; k < per; k += 256) x[k] *= s;
}

// x~ = c .* illum .* exp(i phi) for every point of a class
__global__ void k_ws_element(const float* __restrict__ illum, const float* __restrict__ phi,
                             float* __restrict__ re, float* __restrict__ im, int n, long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long r = i % ((long)n * n);
  const int y = int(r / n), x = int(r - (long)y * n);
  float s, c;
  sincosf(phi[i], &s, &c);
  const float cb = ((x + y) & 1) ? -1.f : 1.f;
  re[i] = cb * illum[i] * c;
  im[i] = cb * illum[i] * s;
}

// labels from F(x~): Y = g c F(x~); pooled over f x f blocks (complex sum),
// label = phasor of the sum, weight = mean target amplitude of the block
__global__ void k_ws_labels(const float* __restrict__ Fre, const float* __restrict__ Fim,
                            const float* __restrict__ t, float* __restrict__ lre,
                            float* __restrict__ lim, float* __restrict__ w, int n, int res, int nd,
                            float g) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)res * res;
  if (i >= per * nd) return;
  const int d = int(i / per);
  const int r = int(i - (long)d * per);
  const int v = r / res, u = r - v * res;
  const int f = n / res;
  const long base = (long)d * n * n;
  double sr = 0, si = 0, sa = 0;
  for (int y = 0; y < f; ++y)
    for (int x = 0; x < f; ++x) {
      const int yy = v * f + y, xx = u * f + x;
      const long k = base + (long)yy * n + xx;
      const float cb = ((xx + yy) & 1) ? -g : g;
      sr += cb * Fre[k];
      si += cb * Fim[k];
      sa += t[k];
    }
  const double m = sqrt(sr * sr + si * si);
  lre[i] = m > 0 ? float(sr / m) : 1.f;
  lim[i] = m > 0 ? float(si / m) : 0.f;
  w[i] = float(sa / (f * f));
}

// Y = t .* phasor(bilinear(z)) at full resolution, then c .* Y (the input of
// the inverse transform)
__global__ void k_ws_pred_field(const float* __restrict__ z, const float* __restrict__ t,
                                float* __restrict__ re, float* __restrict__ im, int n, int res,
                                long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long nn = (long)n * n;
  const int d = int(i / nn);
  const long r = i - (long)d * nn;
  const int y = int(r / n), x = int(r - (long)y * n);
  const float f = float(n) / float(res);
  const float uf = (x + 0.5f) / f - 0.5f, vf = (y + 0.5f) / f - 0.5f;
  const int u0 = max(0, min(res - 1, int(floorf(uf)))), v0 = max(0, min(res - 1, int(floorf(vf))));
  const int u1 = min(res - 1, u0 + 1), v1 = min(res - 1, v0 + 1);
  const float au = fminf(fmaxf(uf - u0, 0.f), 1.f), av = fminf(fmaxf(vf - v0, 0.f), 1.f);
  const float* zd = z + (long)d * res * res * 2;
  auto at = [&](int uu, int vv, int ch) { return zd[((long)vv * res + uu) * 2 + ch]; };
  const float zr = (1 - av) * ((1 - au) * at(u0, v0, 0) + au * at(u1, v0, 0)) +
                   av * ((1 - au) * at(u0, v1, 0) + au * at(u1, v1, 0));
  const float zi = (1 - av) * ((1 - au) * at(u0, v0, 1) + au * at(u1, v0, 1)) +
                   av * ((1 - au) * at(u0, v1, 1) + au * at(u1, v1, 1));
  const float m = sqrtf(zr * zr + zi * zi);
  const float pr = m > 0 ? zr / m : 1.f, pi = m > 0 ? zi / m : 0.f;
  const float cb = ((x + y) & 1) ? -1.f : 1.f;
  re[i] = cb * t[i] * pr;
  im[i] = cb * t[i] * pi;
}

// In place on the forward far field F of the best element field: Y = t .*
// phasor(F) at full resolution, times the checkerboard c (the input of the
// inverse transform) — one GS projection of the best iterate.
__global__ void k_ws_restart(float* __restrict__ re, float* __restrict__ im, const float* __restrict__ t,
                             int n, long total, float g) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long r = i % ((long)n * n);
  const int y = int(r / n), x = int(r - (long)y * n);
  const float cb = ((x + y) & 1) ? -1.f : 1.f;
  const float zr = cb * g * re[i], zi = cb * g * im[i];
  const float m = sqrtf(zr * zr + zi * zi);
  const float pr = m > 0 ? zr / m : 1.f, pi = m > 0 ? zi / m : 0.f;
  re[i] = cb * t[i] * pr;
  im[i] = cb * t[i] * pi;
}

// phi0 = arg(c .* F^-1(c .* Y)) (the global sign g is a global phase)
__global__ void k_ws_phase(const float* __restrict__ re, const float* __restrict__ im,
                           float* __restrict__ phi, int n, long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long r = i % ((long)n * n);
  const int y = int(r / n), x = int(r - (long)y * n);
  const float cb = ((x + y) & 1) ? -1.f : 1.f;
  phi[i] = atan2f(cb * im[i], cb * re[i]);
}

// Per image d of the pooled inputs x (res x res): the centroid (cu, cv) and
// the RMS radius of x about it, as weights; mom[d] = {cu, cv, r}.  One
// workgroup per image, FP64 sums folded in a fixed order.
__global__ void __launch_bounds__(256) k_ws_moments(const float* __restrict__ x, int res,
                                                    double* __restrict__ mom) {
  __shared__ double sh[4][256];
  const long per = (long)res * res;
  const float* xd = x + (long)blockIdx.x * per;
  double s0 = 0, su = 0, sv = 0, s2 = 0;
  for (long k = threadIdx.x; k < per; k += 256) {
    const double w = xd[k];
    const int v = int(k / res), u = int(k - (long)v * res);
    s0 += w;
    su += w * u;
    sv += w * v;
    s2 += w * (double(u) * u + double(v) * v);
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = su;
  sh[2][threadIdx.x] = sv;
  sh[3][threadIdx.x] = s2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double w0 = sh[0][0] > 0 ? sh[0][0] : 1e-30;
    const double cu = sh[1][0] / w0, cv = sh[2][0] / w0;
    const double r2 = fmax(sh[3][0] / w0 - cu * cu - cv * cv, 0.0);
    mom[3 * blockIdx.x] = cu;
    mom[3 * blockIdx.x + 1] = cv;
    mom[3 * blockIdx.x + 2] = sqrt(r2);
  }
}

// Quadratic target-plane phasors exp(i a f^2 |(u, v) - (cu, cv)|^2) per image,
// f = n / res full-resolution pixels per pooled pixel, a = spread pi / r with
// r the RMS radius in full-resolution pixels (at least 4): a chirp of that
// curvature moves a target point at radius r to element radius ~spread n.
__global__ void k_ws_quadratic(const double* __restrict__ mom, float* __restrict__ z, int res, int n,
                               double spread, long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long per = (long)res * res;
  const int d = int(i / per);
  const long r = i - (long)d * per;
  const int v = int(r / res), u = int(r - (long)v * res);
  const double f = double(n) / res;
  const double cu = mom[3 * d], cv = mom[3 * d + 1];
  const double rr = fmax(mom[3 * d + 2] * f, 4.0);
  const double a = spread * M_PI / rr;
  const double ph = a * f * f * ((u - cu) * (u - cu) + (v - cv) * (v - cv));
  double sn, cs;
  sincos(ph, &sn, &cs);
  z[2 * i] = float(cs);
  z[2 * i + 1] = float(sn);
}

unsigned blk(long n) { return unsigned((n + 255) / 256); }

int check_sides(const doe_gs_batch* B, int res) {
  for (const auto& c : B->cls)
    DOE_REQUIRE(c.n >= res && c.n % res == 0, DOE_ERR_SHAPE,
                "warm start: every design side must be a multiple of the network resolution");
  return DOE_OK;
}

// network inputs of all designs in schedule order (class by class)
int ws_inputs(doe_gs_batch* B, int res, float* in) {
  hipStream_t st = B->ctx->stream;
  long o = 0;
  for (const auto& c : B->cls) {
    const int nd = int(c.users.size());
    const long n = long(res) * res * nd;
    hipLaunchKernelGGL(k_ws_pool_input, dim3(blk(n)), dim3(256), 0, st, B->tgt + c.off, in + o, c.n,
                       res, nd);
    hipLaunchKernelGGL(k_ws_norm_max, dim3(unsigned(nd)), dim3(256), 0, st, in + o, long(res) * res);
    o += n;
  }
  DOE_HIP(hipGetLastError());
  return DOE_OK;
}

int fft_class(doe_gs_batch* B, const doe_gs_batch::Class& c, int inverse) {
  doe_field f;
  f.nx = f.ny = c.n;
  f.batch = int(c.users.size());
  f.dtype = DOE_F32;
  f.re = B->re + c.off;
  f.im = B->im + c.off;
  f.owned = false;
  return run_fft2d<float, float>(B->ctx, &f, inverse);
}

}  // namespace
}  // namespace doe

using namespace doe;

extern "C" {

// Per design (schedule order = the order doe_gs_batch_design_order reports):
// input[d] (res^2, pooled target irradiance scaled to max 1).
int doe_warmstart_inputs(doe_gs_batch* B, int res, float* input) {
  DOE_REQUIRE(B && input, DOE_ERR_INVALID, "null argument");
  int rc = check_sides(B, res);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(B->ctx->device));
  rc = ws_inputs(B, res, input);
  if (rc) return rc;
  DOE_HIP(finish(B->ctx));
  return DOE_OK;
}

// Training labels from the designs' current (best) phases: target-plane
// phasor pooled to res^2, and the pooled target amplitude as its weight.
int doe_warmstart_labels(doe_gs_batch* B, int res, float* label_re, float* label_im,
                         float* weight) {
  DOE_REQUIRE(B && label_re && label_im && weight, DOE_ERR_INVALID, "null argument");
  int rc = check_sides(B, res);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(B->ctx->device));
  hipStream_t st = B->ctx->stream;
  long o = 0;
  for (const auto& c : B->cls) {
    const int nd = int(c.users.size());
    const long total = long(c.n) * c.n * nd;
    hipLaunchKernelGGL(k_ws_element, dim3(blk(total)), dim3(256), 0, st, B->illum + c.off,
                       B->ph[0] + c.off, B->re + c.off, B->im + c.off, c.n, total);
    DOE_HIP(hipGetLastError());
    rc = fft_class(B, c, 0);
    if (rc) return rc;
    const float g = (((c.n >> 1) + (c.n >> 1)) & 1) ? -1.f : 1.f;
    const long n = long(res) * res * nd;
    hipLaunchKernelGGL(k_ws_labels, dim3(blk(n)), dim3(256), 0, st, B->re + c.off, B->im + c.off,
                       B->tgt + c.off, label_re + o, label_im + o, weight + o, c.n, res, nd, g);
    o += n;
  }
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(B->ctx));
  return DOE_OK;
}

}  // extern "C"

namespace doe {
namespace {
// initial phases from target-plane phasors z (count * res^2 * 2, schedule order)
int ws_apply_phasor(doe_gs_batch* B, int res, const float* z) {
  hipStream_t st = B->ctx->stream;
  const long per = long(res) * res;
  int rc = DOE_OK;
  long o = 0;
  for (const auto& c : B->cls) {
    const int nd = int(c.users.size());
    const long total = long(c.n) * c.n * nd;
    hipLaunchKernelGGL(k_ws_pred_field, dim3(blk(total)), dim3(256), 0, st, z + per * 2 * o,
                       B->tgt + c.off, B->re + c.off, B->im + c.off, c.n, res, total);
    DOE_HIP(hipGetLastError());
    rc = fft_class(B, c, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ws_phase, dim3(blk(total)), dim3(256), 0, st, B->re + c.off, B->im + c.off,
                       B->phi0 + c.off, c.n, total);
    o += nd;
  }
  DOE_HIP(hipGetLastError());
  return DOE_OK;
}
}  // namespace
}  // namespace doe

extern "C" {

// Initial phases of every design from the network's prediction (replaces the
// designs' initial phases; the next doe_gs_batch_run starts from them).
// `scratch` holds count * res^2 * 3 floats (inputs and predictions).
int doe_warmstart_apply(doe_gs_batch* B, doe_wsnet* net, float* scratch) {
  DOE_REQUIRE(B && net && scratch, DOE_ERR_INVALID, "null argument");
  int res = 0, mb = 0;
  int rc = doe_wsnet_info(net, &res, &mb);
  if (rc) return rc;
  rc = check_sides(B, res);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(B->ctx->device));
  const long per = long(res) * res;
  float* in = scratch;
  float* z = scratch + per * B->count;
  rc = ws_inputs(B, res, in);
  if (rc) return rc;
  for (int d0 = 0; d0 < B->count; d0 += mb) {
    const int nb = std::min(mb, B->count - d0);
    rc = doe_wsnet_forward(net, nb, in + per * d0, z + per * 2 * d0);
    if (rc) return rc;
  }
  rc = ws_apply_phasor(B, res, z);
  if (rc) return rc;
  DOE_HIP(finish(B->ctx));
  return DOE_OK;
}

// Initial phases from given target-plane phasors (device, count*res^2*2
// floats interleaved re, im, schedule order): the same hand-off without a
// network (e.g. a quadratic target-plane phase).
int doe_warmstart_apply_phasor(doe_gs_batch* B, int res, const float* z) {
  DOE_REQUIRE(B && z, DOE_ERR_INVALID, "null argument");
  int rc = check_sides(B, res);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(B->ctx->device));
  rc = ws_apply_phasor(B, res, z);
  if (rc) return rc;
  DOE_HIP(finish(B->ctx));
  return DOE_OK;
}

// Initial phases from the quadratic target-plane phase of every design's
// pooled target (k_ws_quadratic), computed on the device.  `scratch` holds
// count * res^2 * 3 floats; the moments take 3 doubles per design in the
// context's reduction buffer.
int doe_warmstart_apply_quadratic(doe_gs_batch* B, int res, double spread, float* scratch) {
  DOE_REQUIRE(B && scratch, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(spread > 0 && std::isfinite(spread), DOE_ERR_INVALID, "warm start: spread must be positive");
  int rc = check_sides(B, res);
  if (rc) return rc;
  DOE_HIP(hipSetDevice(B->ctx->device));
  hipStream_t st = B->ctx->stream;
  const long per = long(res) * res;
  float* in = scratch;
  float* z = scratch + per * B->count;
  rc = ws_inputs(B, res, in);
  if (rc) return rc;
  rc = ensure_reduce(B->ctx, size_t(3) * B->count);
  if (rc) return rc;
  double* mom = B->ctx->partials;
  hipLaunchKernelGGL(k_ws_moments, dim3(unsigned(B->count)), dim3(256), 0, st, in, res, mom);
  long o = 0;
  for (const auto& c : B->cls) {
    const int nd = int(c.users.size());
    const long total = per * nd;
    hipLaunchKernelGGL(k_ws_quadratic, dim3(blk(total)), dim3(256), 0, st, mom + 3 * o, z + 2 * per * o, res,
                       c.n, spread, total);
    o += nd;
  }
  DOE_HIP(hipGetLastError());
  rc = ws_apply_phasor(B, res, z);
  if (rc) return rc;
  DOE_HIP(finish(B->ctx));
  return DOE_OK;
}

// Restart from the best iterate: every design's initial phase becomes
// arg(F^-1(t * phasor(F(best element field)))); the next doe_gs_batch_run
// starts from it with fresh weights.
int doe_gs_batch_restart_best(doe_gs_batch* B) {
  DOE_REQUIRE(B, DOE_ERR_INVALID, "null argument");
  DOE_HIP(hipSetDevice(B->ctx->device));
  hipStream_t st = B->ctx->stream;
  for (const auto& c : B->cls) {
    const int nd = int(c.users.size());
    const long total = long(c.n) * c.n * nd;
    hipLaunchKernelGGL(k_ws_element, dim3(blk(total)), dim3(256), 0, st, B->illum + c.off,
                       B->ph[0] + c.off, B->re + c.off, B->im + c.off, c.n, total);
    DOE_HIP(hipGetLastError());
    int rc = fft_class(B, c, 0);
    if (rc) return rc;
    const float g = (((c.n >> 1) + (c.n >> 1)) & 1) ? -1.f : 1.f;
    hipLaunchKernelGGL(k_ws_restart, dim3(blk(total)), dim3(256), 0, st, B->re + c.off, B->im + c.off,
                       B->tgt + c.off, c.n, total, g);
    DOE_HIP(hipGetLastError());
    rc = fft_class(B, c, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ws_phase, dim3(blk(total)), dim3(256), 0, st, B->re + c.off, B->im + c.off,
                       B->phi0 + c.off, c.n, total);
  }
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(B->ctx));
  return DOE_OK;
}

// Schedule order of the designs (the order of warm-start inputs and labels).
int doe_gs_batch_design_order(doe_gs_batch* B, int* order) {
  DOE_REQUIRE(B && order, DOE_ERR_INVALID, "null argument");
  int k = 0;
  for (const auto& c : B->cls)
    for (int u : c.users) order[k++] = u;
  return DOE_OK;
}

}  // extern "C"
