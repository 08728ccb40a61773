// Learned warm start for far-field WGS: network kernels, training step and
// the phase hand-off to the batched phase-retrieval engine.
//
// No counterpart in the reference (its designs always start from a random
// or flat phase, doe.js:651, 850-870).  The network maps a target
// irradiance (pooled to res x res) to a phasor field (c, s) in the TARGET
// plane; the element's initial phase is then arg(F^-1(t * (c + i s))) at
// full resolution, one inverse FFT (the far-field operator of fourierCgh,
// doe.js:639-730, inverted), and WGS refines from there.
//
// Architecture (additive-skip U-Net, 3x3 convolutions, ReLU except the
// last, base width C = 32): res^2: 1->C, C->C [s1] | pool | res/2: C->2C,
// 2C->2C [s2] | pool | res/4: 2C->4C, 4C->4C [s3] | pool | res/8: 4C->8C,
// 8C->8C, 8C->4C | up + s3 | 4C->4C, 4C->2C | up + s2 | 2C->2C, 2C->C |
// up + s1 | C->C, C->2.  1.75 M parameters at C = 32.
//
// Training loss per image, phase-aligned and twin-aware: with weights w
// (pooled target amplitude), label phasor y and prediction z,
//   L = (sum w |z|^2 + sum w - 2 max(|sum w z conj(y)|, |sum w z y|)) / sum w,
// the minimum over a global phase of sum w |z - y e^(i theta)|^2 (and of its
// conjugate label: a far-field design and its conjugate-inverted element give
// the same irradiance).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "nn.hpp"

struct doe_gs_batch;

namespace doe {
namespace nn {

// ---------------------------------------------------------------- kernels --
__global__ void k_dw_reduce(const float* __restrict__ part, float* __restrict__ dw, long n,
                            int splits) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[(long)z * n + i];
  dw[i] = s;
}

// db[co] = sum over pixels of dz[p][co]; one workgroup per channel
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ dz, float* __restrict__ db,
                                                   long npix, int C) {
  __shared__ float sh[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (long p = threadIdx.x; p < npix; p += 256) s += dz[p * C + co];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = sh[0];
}

// dz = dy * (y > 0), in place
__global__ void k_relu_bwd(float* __restrict__ dy, const bf16_t* __restrict__ y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(bf2f(y[i]) > 0.f)) dy[i] = 0.f;
}

// 2x2 average pooling (NHWC, H and W of the input even)
__global__ void k_pool2(const bf16_t* __restrict__ x, bf16_t* __restrict__ y, int B, int H, int W,
                        int C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  const long n = (long)B * Ho * Wo * C;
  if (i >= n) return;
  const int c = int(i % C);
  long r = i / C;
  const int wo = int(r % Wo);
  r /= Wo;
  const int ho = int(r % Ho);
  const int b = int(r / Ho);
  const long base = (((long)b * H + 2 * ho) * W + 2 * wo) * C + c;
  const float s = bf2f(x[base]) + bf2f(x[base + C]) + bf2f(x[base + (long)W * C]) +
                  bf2f(x[base + (long)W * C + C]);
  y[i] = f2bf(0.25f * s);
}
// gradient of the pooling: dx[2h+i][2w+j] += dy[h][w] / 4
__global__ void k_pool2_bwd(const float* __restrict__ dy, float* __restrict__ dx, int B, int H,
                            int W, int C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over the input (H x W)
  const long n = (long)B * H * W * C;
  if (i >= n) return;
  const int c = int(i % C);
  long r = i / C;
  const int w = int(r % W);
  r /= W;
  const int h = int(r % H);
  const int b = int(r / H);
  dx[i] += 0.25f * dy[(((long)b * (H / 2) + h / 2) * (W / 2) + w / 2) * C + c];
}
// y = nearest-upsample(x) + s (x at H/2 x W/2)
__global__ void k_up2_add(const bf16_t* __restrict__ x, const bf16_t* __restrict__ s,
                          bf16_t* __restrict__ y, int B, int H, int W, int C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)B * H * W * C;
  if (i >= n) return;
  const int c = int(i % C);
  long r = i / C;
  const int w = int(r % W);
  r /= W;
  const int h = int(r % H);
  const int b = int(r / H);
  y[i] = f2bf(bf2f(x[(((long)b * (H / 2) + h / 2) * (W / 2) + w / 2) * C + c]) + bf2f(s[i]));
}
// gradient of the upsampling: dx[h][w] = sum of dy over the 2x2 block
__global__ void k_up2_bwd(const float* __restrict__ dy, float* __restrict__ dx, int B, int H, int W,
                          int C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over the coarse grid
  const int Hc = H / 2, Wc = W / 2;
  const long n = (long)B * Hc * Wc * C;
  if (i >= n) return;
  const int c = int(i % C);
  long r = i / C;
  const int w = int(r % Wc);
  r /= Wc;
  const int h = int(r % Hc);
  const int b = int(r / Hc);
  const long base = (((long)b * H + 2 * h) * W + 2 * w) * C + c;
  dx[i] = dy[base] + dy[base + C] + dy[base + (long)W * C] + dy[base + (long)W * C + C];
}
__global__ void k_add_inplace(float* __restrict__ a, const float* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}
__global__ void k_f32_to_bf16(const float* __restrict__ a, bf16_t* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// Loss sums per image, two levels (fixed order): block partials of
// (W, Z2, Sr, Si, Tr, Ti) with S = sum w z conj(y), T = sum w z y.
constexpr int kLossBlocks = 64;
__global__ void __launch_bounds__(256) k_loss_part(const float* __restrict__ z,
                                                   const float* __restrict__ yr,
                                                   const float* __restrict__ yi,
                                                   const float* __restrict__ w, long npix,
                                                   double* __restrict__ part) {
  __shared__ double sh[6][256];
  const int b = blockIdx.y;
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < npix; p += 256L * kLossBlocks) {
    const long q = (long)b * npix + p;
    const double wc = w[q], zr = z[2 * q], zi = z[2 * q + 1], lr = yr[q], li = yi[q];
    a[0] += wc;
    a[1] += wc * (zr * zr + zi * zi);
    a[2] += wc * (zr * lr + zi * li);
    a[3] += wc * (zi * lr - zr * li);
    a[4] += wc * (zr * lr - zi * li);
    a[5] += wc * (zi * lr + zr * li);
  }
  for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 6) part[((long)b * kLossBlocks + blockIdx.x) * 6 + threadIdx.x] = sh[threadIdx.x][0];
}
// per image: fold partials, loss, and the gradient coefficients
// out[b*8 + {0: loss, 1: W, 2: dir_r, 3: dir_i, 4: conj flag}]
__global__ void k_loss_fold(const double* __restrict__ part, double* __restrict__ out, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int j = 0; j < kLossBlocks; ++j)
    for (int k = 0; k < 6; ++k) s[k] += part[((long)b * kLossBlocks + j) * 6 + k];
  const double W = s[0] > 0 ? s[0] : 1.0;
  const double aS = sqrt(s[2] * s[2] + s[3] * s[3]);
  const double aT = sqrt(s[4] * s[4] + s[5] * s[5]);
  const bool conj = aT > aS;
  const double m = conj ? aT : aS;
  out[b * 8 + 0] = (s[1] + s[0] - 2 * m) / W;
  out[b * 8 + 1] = W;
  out[b * 8 + 2] = m > 0 ? (conj ? s[4] : s[2]) / m : 0.0;
  out[b * 8 + 3] = m > 0 ? (conj ? s[5] : s[3]) / m : 0.0;
  out[b * 8 + 4] = conj ? 1.0 : 0.0;
}
// dL/dz for the batch-mean loss: (2 w z - 2 w d(|S|)/dz) / (W B)
__global__ void k_loss_grad(const float* __restrict__ z, const float* __restrict__ yr,
                            const float* __restrict__ yi, const float* __restrict__ w,
                            const double* __restrict__ fold, float* __restrict__ dz, long npix,
                            int B) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix * B) return;
  const int b = int(q / npix);
  const double* f = fold + b * 8;
  const double W = f[1], dr = f[2], di = f[3];
  const bool conj = f[4] != 0.0;
  const double wc = w[q], lr = yr[q], li = conj ? -yi[q] : yi[q];
  // |S| with S = sum w z conj(l): d|S|/dzr = w (dr lr - di li), d|S|/dzi = w (dr li + di lr)
  const double gs_r = wc * (dr * lr - di * li);
  const double gs_i = wc * (dr * li + di * lr);
  const double sc = 1.0 / (W * B);
  dz[2 * q] = float((2 * wc * z[2 * q] - 2 * gs_r) * sc);
  dz[2 * q + 1] = float((2 * wc * z[2 * q + 1] - 2 * gs_i) * sc);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16_t* __restrict__ pb, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1 - b1) * gi;
  const float vi = b2 * v[i] + (1 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  pb[i] = f2bf(pi);
}

// ------------------------------------------------------------ host class --
struct Layer {
  int kind;  // 0 conv, 1 pool, 2 up+add
  int cin, cout, relu;
  int res;   // spatial side of this layer's output
  int in, out, skip;  // activation slots
  long woff, boff;    // parameter offsets (conv)
};

}  // namespace nn
}  // namespace doe

using namespace doe;
using namespace doe::nn;

struct doe_wsnet {
  doe_ctx* ctx = nullptr;
  int res = 256, base = 32, max_batch = 0;
  std::vector<Layer> layers;
  std::vector<int> slot_c, slot_res;  // channels and side of each activation slot
  long nparams = 0;
  std::vector<void*> allocs;
  float* p = nullptr;     // FP32 master parameters
  bf16_t* pb = nullptr;   // BF16 copy used by the kernels
  float* g = nullptr;     // gradients
  float* am = nullptr;    // Adam moments
  float* av = nullptr;
  long adam_t = 0;
  std::vector<bf16_t*> act;   // activations per slot
  std::vector<float*> grad;   // activation gradients per slot
  float* out = nullptr;       // FP32 output of the last layer (B x res^2 x 2)
  float* dout = nullptr;
  float* dwp = nullptr;       // MODE 2 partial sums
  long dwp_cap = 0;
  double* lpart = nullptr;
  double* lfold = nullptr;

  ~doe_wsnet() {
    for (void* q : allocs) hipFree(q);
  }
  int alloc(void** q, size_t bytes) {
    if (hipMalloc(q, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("doe_wsnet: device allocation failed");
      return DOE_ERR_OOM;
    }
    allocs.push_back(*q);
    return DOE_OK;
  }
};

namespace {

unsigned blocks(long n) { return unsigned((n + 255) / 256); }

void build_layers(doe_wsnet* N) {
  const int C = N->base, R = N->res;
  auto slot = [&](int c, int r) {
    N->slot_c.push_back(c);
    N->slot_res.push_back(r);
    return int(N->slot_c.size()) - 1;
  };
  const int in = slot(1, R);
  int cur = in;
  auto conv = [&](int cin, int cout, int r, bool relu) {
    Layer L{0, cin, cout, relu ? 1 : 0, r, cur, -1, -1, 0, 0};
    L.out = relu || cout != 2 ? slot(cout, r) : -1;
    N->layers.push_back(L);
    cur = L.out;
    return L.out;
  };
  auto pool = [&](int c, int r) {
    Layer L{1, c, c, 0, r / 2, cur, slot(c, r / 2), -1, 0, 0};
    N->layers.push_back(L);
    cur = L.out;
  };
  auto up = [&](int c, int r, int skip) {
    Layer L{2, c, c, 0, r, cur, slot(c, r), skip, 0, 0};
    N->layers.push_back(L);
    cur = L.out;
  };
  conv(1, C, R, true);
  const int s1 = conv(C, C, R, true);
  pool(C, R);
  conv(C, 2 * C, R / 2, true);
  const int s2 = conv(2 * C, 2 * C, R / 2, true);
  pool(2 * C, R / 2);
  conv(2 * C, 4 * C, R / 4, true);
  const int s3 = conv(4 * C, 4 * C, R / 4, true);
  pool(4 * C, R / 4);
  conv(4 * C, 8 * C, R / 8, true);
  conv(8 * C, 8 * C, R / 8, true);
  conv(8 * C, 4 * C, R / 8, true);
  up(4 * C, R / 4, s3);
  conv(4 * C, 4 * C, R / 4, true);
  conv(4 * C, 2 * C, R / 4, true);
  up(2 * C, R / 2, s2);
  conv(2 * C, 2 * C, R / 2, true);
  conv(2 * C, C, R / 2, true);
  up(C, R, s1);
  conv(C, C, R, true);
  conv(C, 2, R, false);  // output: FP32 buffer, no slot
  long off = 0;
  for (auto& L : N->layers)
    if (L.kind == 0) {
      L.woff = off;
      off += long(L.cout) * 9 * L.cin;
      L.boff = off;
      off += L.cout;
    }
  N->nparams = off;
}

hipError_t conv_fwd(doe_wsnet* N, const Layer& L, int B, const bf16_t* x, bf16_t* y_bf, float* y_f) {
  ConvCall c{};
  c.B = B;
  c.H = c.W = L.res;
  c.Cin = L.cin;
  c.Cout = L.cout;
  c.x = x;
  c.w = N->pb + L.woff;
  c.bias = N->p + L.boff;
  c.y_bf = y_bf;
  c.y_f32 = y_f;
  c.relu = L.relu;
  const long M = long(B) * L.res * L.res;
  hipLaunchKernelGGL(k_conv<0>, dim3(unsigned((M + BM - 1) / BM), unsigned((L.cout + BN - 1) / BN)),
                     dim3(256), 0, N->ctx->stream, c);
  return hipGetLastError();
}

int forward(doe_wsnet* N, int B, const float* input) {
  hipStream_t st = N->ctx->stream;
  const long n0 = long(B) * N->res * N->res;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(n0)), dim3(256), 0, st, input, N->act[0], n0);
  for (const Layer& L : N->layers) {
    if (L.kind == 0) {
      DOE_HIP(conv_fwd(N, L, B, N->act[L.in], L.out >= 0 ? N->act[L.out] : nullptr,
                       L.out >= 0 ? nullptr : N->out));
    } else if (L.kind == 1) {
      const long n = long(B) * L.res * L.res * L.cout;
      hipLaunchKernelGGL(k_pool2, dim3(blocks(n)), dim3(256), 0, st, N->act[L.in], N->act[L.out], B,
                         2 * L.res, 2 * L.res, L.cout);
    } else {
      const long n = long(B) * L.res * L.res * L.cout;
      hipLaunchKernelGGL(k_up2_add, dim3(blocks(n)), dim3(256), 0, st, N->act[L.in], N->act[L.skip],
                         N->act[L.out], B, L.res, L.res, L.cout);
    }
  }
  DOE_HIP(hipGetLastError());
  return DOE_OK;
}

int backward(doe_wsnet* N, int B) {
  hipStream_t st = N->ctx->stream;
  for (size_t s = 0; s < N->grad.size(); ++s) {
    const long n = long(B) * N->slot_res[s] * N->slot_res[s] * N->slot_c[s];
    DOE_HIP(hipMemsetAsync(N->grad[s], 0, n * sizeof(float), st));
  }
  for (int li = int(N->layers.size()) - 1; li >= 0; --li) {
    const Layer& L = N->layers[li];
    if (L.kind == 0) {
      const long npix = long(B) * L.res * L.res;
      float* dz = L.out >= 0 ? N->grad[L.out] : N->dout;
      if (L.relu)
        hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(npix * L.cout)), dim3(256), 0, st, dz,
                           N->act[L.out], npix * L.cout);
      // weights: split the pixel sum so each split is a few hundred MFMA k-steps
      const int ksplit = 4096;
      const int splits = int((npix + ksplit - 1) / ksplit);
      const long nw = long(L.cout) * 9 * L.cin;
      ConvCall c{};
      c.B = B;
      c.H = c.W = L.res;
      c.Cin = L.cin;
      c.Cout = L.cout;
      c.x = N->act[L.in];
      c.w = N->pb + L.woff;
      c.dy = dz;
      c.dwp = N->dwp;
      c.ksplit = ksplit;
      DOE_REQUIRE(long(splits) * nw <= N->dwp_cap, DOE_ERR_STATE, "doe_wsnet: partial buffer");
      hipLaunchKernelGGL(k_conv<2>, dim3(unsigned((L.cout + BM - 1) / BM),
                                         unsigned((9 * L.cin + BN - 1) / BN), unsigned(splits)),
                         dim3(256), 0, st, c);
      hipLaunchKernelGGL(k_dw_reduce, dim3(blocks(nw)), dim3(256), 0, st, N->dwp, N->g + L.woff, nw,
                         splits);
      hipLaunchKernelGGL(k_bias_grad, dim3(unsigned(L.cout)), dim3(256), 0, st, dz, N->g + L.boff,
                         npix, L.cout);
      if (L.in > 0) {  // the network input needs no gradient
        // dX accumulates into the input slot's gradient
        float* dxs = N->dwp + long(splits) * nw;  // scratch after the partials
        DOE_REQUIRE(long(splits) * nw + npix * L.cin <= N->dwp_cap, DOE_ERR_STATE,
                    "doe_wsnet: scratch buffer");
        c.dx = dxs;
        c.mask = nullptr;
        hipLaunchKernelGGL(k_conv<1>, dim3(unsigned((npix + BM - 1) / BM), unsigned((L.cin + BN - 1) / BN)),
                           dim3(256), 0, st, c);
        hipLaunchKernelGGL(k_add_inplace, dim3(blocks(npix * L.cin)), dim3(256), 0, st,
                           N->grad[L.in], dxs, npix * L.cin);
      }
    } else if (L.kind == 1) {
      const long n = long(B) * 4 * L.res * L.res * L.cout;
      hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks(n)), dim3(256), 0, st, N->grad[L.out],
                         N->grad[L.in], B, 2 * L.res, 2 * L.res, L.cout);
    } else {
      const long nfine = long(B) * L.res * L.res * L.cout;
      hipLaunchKernelGGL(k_add_inplace, dim3(blocks(nfine)), dim3(256), 0, st, N->grad[L.skip],
                         N->grad[L.out], nfine);
      float* tmp = N->dwp;
      const long ncoarse = nfine / 4;
      DOE_REQUIRE(ncoarse <= N->dwp_cap, DOE_ERR_STATE, "doe_wsnet: scratch buffer");
      hipLaunchKernelGGL(k_up2_bwd, dim3(blocks(ncoarse)), dim3(256), 0, st, N->grad[L.out], tmp, B,
                         L.res, L.res, L.cout);
      hipLaunchKernelGGL(k_add_inplace, dim3(blocks(ncoarse)), dim3(256), 0, st, N->grad[L.in], tmp,
                         ncoarse);
    }
  }
  DOE_HIP(hipGetLastError());
  return DOE_OK;
}

int loss_and_grad(doe_wsnet* N, int B, const float* yr, const float* yi, const float* w,
                  double* loss_out) {
  hipStream_t st = N->ctx->stream;
  const long npix = long(N->res) * N->res;
  hipLaunchKernelGGL(k_loss_part, dim3(kLossBlocks, unsigned(B)), dim3(256), 0, st, N->out, yr, yi,
                     w, npix, N->lpart);
  hipLaunchKernelGGL(k_loss_fold, dim3(unsigned((B + 63) / 64)), dim3(64), 0, st, N->lpart, N->lfold,
                     B);
  hipLaunchKernelGGL(k_loss_grad, dim3(blocks(npix * B)), dim3(256), 0, st, N->out, yr, yi, w,
                     N->lfold, N->dout, npix, B);
  DOE_HIP(hipGetLastError());
  if (loss_out) {
    std::vector<double> f(8 * size_t(B));
    DOE_HIP(hipMemcpyAsync(f.data(), N->lfold, sizeof(double) * f.size(), hipMemcpyDeviceToHost, st));
    DOE_HIP(hipStreamSynchronize(st));
    double s = 0;
    for (int b = 0; b < B; ++b) s += f[8 * b];
    *loss_out = s / B;
  }
  return DOE_OK;
}

}  // namespace

extern "C" {

int doe_wsnet_create(doe_ctx* ctx, int res, int base, int max_batch, unsigned long long seed,
                     doe_wsnet** out) {
  DOE_REQUIRE(ctx && out, DOE_ERR_INVALID, "doe_wsnet_create: null argument");
  *out = nullptr;
  DOE_REQUIRE(res >= 16 && res % 8 == 0, DOE_ERR_INVALID, "doe_wsnet_create: res must be a multiple of 8, >= 16");
  DOE_REQUIRE(base >= 2 && base <= 64, DOE_ERR_INVALID, "doe_wsnet_create: base width must be in [2, 64]");
  DOE_REQUIRE(max_batch >= 1, DOE_ERR_INVALID, "doe_wsnet_create: max_batch must be positive");
  DOE_HIP(hipSetDevice(ctx->device));
  auto* N = new doe_wsnet();
  N->ctx = ctx;
  N->res = res;
  N->base = base;
  N->max_batch = max_batch;
  build_layers(N);
  void* q = nullptr;
  int rc = DOE_OK;
  float** f32s[] = {&N->p, &N->g, &N->am, &N->av};
  for (float** a : f32s) {
    if (!rc) rc = N->alloc(&q, sizeof(float) * N->nparams);
    if (!rc) *a = (float*)q;
  }
  if (!rc) rc = N->alloc(&q, sizeof(bf16_t) * N->nparams);
  if (!rc) N->pb = (bf16_t*)q;
  for (size_t s = 0; !rc && s < N->slot_c.size(); ++s) {
    const long n = long(max_batch) * N->slot_res[s] * N->slot_res[s] * N->slot_c[s];
    rc = N->alloc(&q, sizeof(bf16_t) * n);
    if (!rc) N->act.push_back((bf16_t*)q);
    if (!rc) rc = N->alloc(&q, sizeof(float) * n);
    if (!rc) N->grad.push_back((float*)q);
  }
  const long nout = long(max_batch) * res * res * 2;
  if (!rc) rc = N->alloc(&q, sizeof(float) * nout);
  if (!rc) N->out = (float*)q;
  if (!rc) rc = N->alloc(&q, sizeof(float) * nout);
  if (!rc) N->dout = (float*)q;
  // partial dW sums + dX scratch: bounded by the largest layer
  long cap = 0;
  for (const Layer& L : N->layers) {
    const long npix = long(max_batch) * L.res * L.res;
    const long splits = (npix + 4095) / 4096;
    cap = std::max(cap, splits * long(L.cout) * 9 * L.cin + npix * std::max(L.cin, L.cout));
  }
  N->dwp_cap = cap;
  if (!rc) rc = N->alloc(&q, sizeof(float) * cap);
  if (!rc) N->dwp = (float*)q;
  if (!rc) rc = N->alloc(&q, sizeof(double) * 6 * kLossBlocks * max_batch);
  if (!rc) N->lpart = (double*)q;
  if (!rc) rc = N->alloc(&q, sizeof(double) * 8 * max_batch);
  if (!rc) N->lfold = (double*)q;
  if (rc) {
    delete N;
    return rc;
  }
  // He-normal weights (std sqrt(2 / fan_in)), zero biases
  std::vector<float> h(N->nparams, 0.f);
  std::mt19937_64 gen(seed);
  for (const Layer& L : N->layers) {
    if (L.kind != 0) continue;
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * L.cin)));
    for (long i = 0; i < long(L.cout) * 9 * L.cin; ++i) h[L.woff + i] = nd(gen);
  }
  rc = doe_wsnet_set_params(N, h.data());
  if (rc) {
    delete N;
    return rc;
  }
  *out = N;
  return DOE_OK;
}

int doe_wsnet_destroy(doe_wsnet* N) {
  if (!N) return DOE_OK;
  hipSetDevice(N->ctx->device);
  hipStreamSynchronize(N->ctx->stream);
  delete N;
  return DOE_OK;
}

int doe_wsnet_info(doe_wsnet* N, int* res, int* max_batch) {
  DOE_REQUIRE(N, DOE_ERR_INVALID, "null argument");
  if (res) *res = N->res;
  if (max_batch) *max_batch = N->max_batch;
  return DOE_OK;
}

int doe_wsnet_param_count(doe_wsnet* N, long* n) {
  DOE_REQUIRE(N && n, DOE_ERR_INVALID, "null argument");
  *n = N->nparams;
  return DOE_OK;
}

int doe_wsnet_set_params(doe_wsnet* N, const float* host) {
  DOE_REQUIRE(N && host, DOE_ERR_INVALID, "null argument");
  hipStream_t st = N->ctx->stream;
  DOE_HIP(hipMemcpyAsync(N->p, host, sizeof(float) * N->nparams, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(N->nparams)), dim3(256), 0, st, N->p, N->pb,
                     N->nparams);
  DOE_HIP(hipMemsetAsync(N->am, 0, sizeof(float) * N->nparams, st));
  DOE_HIP(hipMemsetAsync(N->av, 0, sizeof(float) * N->nparams, st));
  N->adam_t = 0;
  DOE_HIP(hipStreamSynchronize(st));
  return DOE_OK;
}

int doe_wsnet_get_params(doe_wsnet* N, float* host) {
  DOE_REQUIRE(N && host, DOE_ERR_INVALID, "null argument");
  DOE_HIP(hipMemcpyAsync(host, N->p, sizeof(float) * N->nparams, hipMemcpyDeviceToHost,
                         N->ctx->stream));
  DOE_HIP(hipStreamSynchronize(N->ctx->stream));
  return DOE_OK;
}

int doe_wsnet_forward(doe_wsnet* N, int batch, const float* input, float* output) {
  DOE_REQUIRE(N && input, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(batch >= 1 && batch <= N->max_batch, DOE_ERR_INVALID,
              "doe_wsnet_forward: batch exceeds the network's max_batch");
  DOE_HIP(hipSetDevice(N->ctx->device));
  int rc = forward(N, batch, input);
  if (rc) return rc;
  if (output)
    DOE_HIP(hipMemcpyAsync(output, N->out, sizeof(float) * batch * N->res * N->res * 2,
                           hipMemcpyDeviceToDevice, N->ctx->stream));
  DOE_HIP(finish(N->ctx));
  return DOE_OK;
}

int doe_wsnet_loss_grad(doe_wsnet* N, int batch, const float* input, const float* label_re,
                        const float* label_im, const float* weight, double* loss,
                        float* grad_host) {
  DOE_REQUIRE(N && input && label_re && label_im && weight, DOE_ERR_INVALID, "null argument");
  DOE_REQUIRE(batch >= 1 && batch <= N->max_batch, DOE_ERR_INVALID,
              "doe_wsnet_loss_grad: batch exceeds the network's max_batch");
  DOE_HIP(hipSetDevice(N->ctx->device));
  int rc = forward(N, batch, input);
  if (!rc) rc = loss_and_grad(N, batch, label_re, label_im, weight, loss);
  if (!rc) rc = backward(N, batch);
  if (rc) return rc;
  if (grad_host) {
    DOE_HIP(hipMemcpyAsync(grad_host, N->g, sizeof(float) * N->nparams, hipMemcpyDeviceToHost,
                           N->ctx->stream));
  }
  DOE_HIP(hipStreamSynchronize(N->ctx->stream));
  return DOE_OK;
}

int doe_wsnet_adam_step(doe_wsnet* N, float lr) {
  DOE_REQUIRE(N, DOE_ERR_INVALID, "null argument");
  DOE_HIP(hipSetDevice(N->ctx->device));
  const float b1 = 0.9f, b2 = 0.999f;
  N->adam_t += 1;
  const float c1 = 1.f - std::pow(b1, float(N->adam_t));
  const float c2 = 1.f - std::pow(b2, float(N->adam_t));
  hipLaunchKernelGGL(k_adam, dim3(blocks(N->nparams)), dim3(256), 0, N->ctx->stream, N->p, N->g,
                     N->am, N->av, N->pb, N->nparams, lr, b1, b2, 1e-8f, c1, c2);
  DOE_HIP(hipGetLastError());
  DOE_HIP(finish(N->ctx));
  return DOE_OK;
}

}  // extern "C"
