// Kernels of the learned warm start: a small convolutional encoder-decoder
// (BF16 weights and activations, FP32 accumulation on the matrix cores).
//
// No counterpart in the reference.  Convolutions are 3x3, stride 1, zero
// padding, NHWC activations, weights [Cout][3][3][Cin].  All three products
// of a convolution layer run as one implicit-GEMM kernel on
// v_mfma_f32_16x16x16_bf16:
//   MODE 0 forward:  Y[p][co]  = act(sum_k X~[p][k] W[co][k] + b[co]),  K = 9 Cin
//   MODE 1 dX:       dX[p][ci] = sum_k dY~[p][k] W~[k][ci],             K = 9 Cout
//   MODE 2 dW:       dW[co][k] = sum_p dY[p][co] X~[p][k],              K = pixels
// where ~ is the shifted (im2col) view.  MODE 2 splits its pixel sum over
// blockIdx.z into partial sums reduced in a fixed order (deterministic).
#pragma once
#include "common.hpp"

namespace doe {
namespace nn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;  // raw bfloat16 bits

__host__ __device__ inline bf16_t f2bf(float x) {
  // round to nearest even; NaN stays NaN
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return bf16_t(0x7fc0);
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16_t(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) {
  const uint32_t u = uint32_t(h) << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

struct ConvCall {
  int B, H, W, Cin, Cout;
  const bf16_t* x;     // MODE 0, 2: input activations (NHWC)
  const bf16_t* w;     // MODE 0, 1: weights [Cout][9 Cin]
  const float* bias;   // MODE 0
  const float* dy;     // MODE 1, 2: output gradient (NHWC, FP32)
  bf16_t* y_bf;        // MODE 0 output, BF16 (or)
  float* y_f32;        // MODE 0 output, FP32
  int relu;            // MODE 0
  float* dx;           // MODE 1 output
  const bf16_t* mask;  // MODE 1: multiply by (mask > 0) (the layer input after its ReLU)
  float* dwp;          // MODE 2: partial sums [split][Cout][9 Cin]
  int ksplit;          // MODE 2: pixels per split
};

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int AS = BK + 8;  // LDS row strides (shorts)
constexpr int BS = BN + 8;

template <int MODE>
__device__ __forceinline__ float fetch_a(const ConvCall& c, int m, int k, int M, int K) {
  if (m >= M || k >= K) return 0.f;
  if constexpr (MODE == 2) {
    return c.dy[(long)k * c.Cout + m];
  } else {
    const int hw = c.H * c.W;
    const int b = m / hw;
    const int r = m - b * hw;
    const int h = r / c.W;
    const int w = r - h * c.W;
    const int cc = MODE == 0 ? c.Cin : c.Cout;
    const int t = k / cc;
    const int ch = k - t * cc;
    const int kh = t / 3, kw = t - 3 * (t / 3);
    const int hh = MODE == 0 ? h + kh - 1 : h + 1 - kh;
    const int ww = MODE == 0 ? w + kw - 1 : w + 1 - kw;
    if (hh < 0 || hh >= c.H || ww < 0 || ww >= c.W) return 0.f;
    const long q = ((long)(b * c.H + hh) * c.W + ww) * cc + ch;
    if constexpr (MODE == 0) return bf2f(c.x[q]);
    else return c.dy[q];
  }
}

template <int MODE>
__device__ __forceinline__ float fetch_b(const ConvCall& c, int k, int n, int K, int N) {
  if (k >= K || n >= N) return 0.f;
  if constexpr (MODE == 0) {
    return bf2f(c.w[(long)n * K + k]);
  } else if constexpr (MODE == 1) {
    const int t = k / c.Cout;
    const int co = k - t * c.Cout;
    return bf2f(c.w[((long)co * 9 + t) * c.Cin + n]);
  } else {
    // k: pixel, n: (tap, ci)
    const int hw = c.H * c.W;
    const int b = k / hw;
    const int r = k - b * hw;
    const int h = r / c.W;
    const int w = r - h * c.W;
    const int t = n / c.Cin;
    const int ci = n - t * c.Cin;
    const int kh = t / 3, kw = t - 3 * (t / 3);
    const int hh = h + kh - 1, ww = w + kw - 1;
    if (hh < 0 || hh >= c.H || ww < 0 || ww >= c.W) return 0.f;
    return bf2f(c.x[((long)(b * c.H + hh) * c.W + ww) * c.Cin + ci]);
  }
}

// One 64x64 output tile per 256-thread workgroup; wave w computes the 32x32
// quarter (w / 2, w % 2) as 2x2 blocks of 16x16 MFMA products.
template <int MODE>
__global__ void __launch_bounds__(256) k_conv(ConvCall c) {
  __shared__ short As[BM * AS];
  __shared__ short Bs[BK * BS];
  const int npix = c.B * c.H * c.W;
  const int M = MODE == 2 ? c.Cout : npix;
  const int N = MODE == 0 ? c.Cout : (MODE == 1 ? c.Cin : 9 * c.Cin);
  const int Ktot = MODE == 0 ? 9 * c.Cin : (MODE == 1 ? 9 * c.Cout : npix);
  const int m0 = blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  const int kbeg = MODE == 2 ? blockIdx.z * c.ksplit : 0;
  const int kend = MODE == 2 ? min(Ktot, kbeg + c.ksplit) : Ktot;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    {  // A tile: row tid/4, 8 consecutive k
      const int r = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + kk + e;
        As[r * AS + kk + e] = short(f2bf(k < kend ? fetch_a<MODE>(c, m0 + r, k, M, Ktot) : 0.f));
      }
    }
    {  // B tile: k row tid/8, 8 consecutive n
      const int kr = tid >> 3, nn = (tid & 7) * 8;
      const int k = k0 + kr;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        Bs[kr * BS + nn + e] = short(f2bf(k < kend ? fetch_b<MODE>(c, k, n0 + nn + e, Ktot, N) : 0.f));
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK / 16; ++ks) {
      const int kq = ks * 16 + 4 * (lane >> 4);
      bf16x4 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int r = wm + i * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) a[i][j] = As[r * AS + kq + j];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int col = wn + i * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[i][j] = Bs[(kq + j) * BS + col];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  // C layout of the 16x16 product: lane holds rows 4*(lane/16) + q, column lane % 16
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        const float v = acc[i][j][q];
        if constexpr (MODE == 0) {
          float o = v + (c.bias ? c.bias[n] : 0.f);
          if (c.relu) o = fmaxf(o, 0.f);
          const long idx = (long)m * N + n;
          if (c.y_f32) c.y_f32[idx] = o;
          else c.y_bf[idx] = f2bf(o);
        } else if constexpr (MODE == 1) {
          const long idx = (long)m * N + n;
          float o = v;
          if (c.mask && !(bf2f(c.mask[idx]) > 0.f)) o = 0.f;
          c.dx[idx] = o;
        } else {
          c.dwp[((long)blockIdx.z * M + m) * N + n] = v;
        }
      }
}

}  // namespace nn
}  // namespace doe
