// The reference's random phase on the device (makeRng / randomPhase,
// doe.js:65-86): xorshift32 with shifts 13, 17, 5 from seed>>>0 || 1, sample k
// = (s_{k+1} / 2^32 * 2 - 1) * pi with s_{k+1} the state after k+1 steps.
//
// The step is linear over GF(2): s_{k} = M^k s_0 for a 32 x 32 bit matrix M.
// A workgroup builds the jump tables M^(2^i) in LDS (column b of M^(2^i) is
// the image of bit b), each thread jumps to the state before its chunk of 64
// samples with one matrix-vector product per set bit of the chunk start and
// then steps sequentially, so every sample equals the reference's stream bit
// for bit.
#pragma once
#include <cmath>
#include <cstdint>

#include "ctx.hpp"

namespace doe {

constexpr int kRngThreads = 256;
constexpr int kRngChunk = 64;   // samples per thread
constexpr int kRngLevels = 40;  // jump tables: streams of up to 2^40 samples

__device__ __forceinline__ uint32_t xs32_step(uint32_t s) {
  s ^= s << 13;
  s ^= s >> 17;
  s ^= s << 5;
  return s;
}

// tab: 32 columns of a GF(2) matrix; returns tab * s.
__device__ __forceinline__ uint32_t gf2_apply(const uint32_t* tab, uint32_t s) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 32; ++b) r ^= (s >> b & 1u) ? tab[b] : 0u;
  return r;
}

template <typename T>
__global__ __launch_bounds__(kRngThreads) void k_random_phase(T* __restrict__ out, long n, uint32_t seed) {
#pragma clang fp contract(off)
  __shared__ uint32_t tab[kRngLevels][32];
  const int tid = threadIdx.x;
  if (tid < 32) tab[0][tid] = xs32_step(1u << tid);
  __syncthreads();
  for (int i = 1; i < kRngLevels; ++i) {
    if (tid < 32) tab[i][tid] = gf2_apply(tab[i - 1], tab[i - 1][tid]);  // M^(2^i) = (M^(2^(i-1)))^2
    __syncthreads();
  }
  const long k0 = ((long)blockIdx.x * kRngThreads + tid) * kRngChunk;
  if (k0 >= n) return;
  uint32_t s = seed ? seed : 1u;
  for (int i = 0; i < kRngLevels; ++i)
    if ((k0 >> i) & 1) s = gf2_apply(tab[i], s);
  const long k1 = k0 + kRngChunk < n ? k0 + kRngChunk : n;
  for (long k = k0; k < k1; ++k) {
    s = xs32_step(s);
    out[k] = T((double(s) / 4294967296.0 * 2 - 1) * M_PI);
  }
}

// Fill n samples of out with the stream of `seed` on the context's stream.
template <typename T>
inline hipError_t launch_random_phase(doe_ctx* ctx, T* out, long n, uint32_t seed) {
  const long threads = (n + kRngChunk - 1) / kRngChunk;
  const unsigned blocks = unsigned((threads + kRngThreads - 1) / kRngThreads);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_random_phase<T>, dim3(blocks), dim3(kRngThreads), 0, ctx->stream, out, n, seed);
  return hipGetLastError();
}

}  // namespace doe
