// Row and column pass kernels built on the line engine.
//
// A pass loads whole lines (rows or column tiles) of a planar complex batch,
// runs an optional elementwise `pre` hook, one FFT, and either stores (MODE 0)
// or runs a `mid` hook and a second FFT in the opposite stage order (MODE 1),
// then a `post` hook, and stores.  Hooks see (field, line, position) so the
// propagators' transfer functions, chirps and the phase-retrieval projections
// are applied where the data already sits in registers.
//
// Ops with NACC > 0 accumulate FP64 partial sums per thread; the column pass
// reduces them per workgroup into `partials` and the last workgroup of each
// field (per-field arrival counter) folds that field's partials in a fixed
// order and calls op.finalize(field, sums) — deterministic, graph-safe (the
// counter is reset by the finalizing block; no memset needed).
#pragma once
#include "fft_engine.hpp"

namespace doe {

struct NoAcc {};

// Block-wide deterministic reduction of NACC doubles; result valid in tid 0.
template <int NACC>
__device__ __forceinline__ void block_sum(double (&acc)[NACC], double* sh, int tid, int nthr) {
  // wave-level tree with shuffles (64-wide wavefronts)
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    acc[k] = v;
  }
  const int lane = tid & 63, wave = tid >> 6, nw = (nthr + 63) >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NACC; ++k) sh[wave * NACC + k] = acc[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NACC; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sh[w * NACC + k];
      acc[k] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Row pass.  Lines are rows (contiguous); `lines_per_field` = ny.
// ---------------------------------------------------------------------------
template <typename T, typename S, int N, int MODE, bool INV1, bool INV2, class Op>
__global__ void __launch_bounds__((RowCfg<T, N>::WG))
    k_row_pass(Op op, const S* __restrict__ ire, const S* __restrict__ iim, S* __restrict__ ore,
               S* __restrict__ oim, int lines_per_field, long total_lines,
               const T* __restrict__ twr, const T* __restrict__ twi) {
  using CF = RowCfg<T, N>;
  constexpr int P = CF::P;
  using PL = LinePlan<N, P>;
  __shared__ T lre[CF::LDS_ELEMS];
  __shared__ T lim[CF::LDS_ELEMS];
  const int tid = threadIdx.x;
  const int lw = tid / CF::T_LINE;
  const int t = tid - lw * CF::T_LINE;
  const long line = (long)blockIdx.x * CF::LINES + lw;
  const bool active = line < total_lines;
  const int f = active ? int(line / lines_per_field) : 0;
  const int y = active ? int(line - (long)f * lines_per_field) : 0;
  const long base = line * (long)N;
  T vr[P], vi[P];
  double acc[Op::NACC > 0 ? Op::NACC : 1];
#pragma unroll
  for (int k = 0; k < (Op::NACC > 0 ? Op::NACC : 1); ++k) acc[k] = 0.0;
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pos = PL::template in_pos<false>(t, i);
      vr[i] = ld<T>(ire + base + pos);
      vi[i] = ld<T>(iim + base + pos);
      op.pre(f, pos, y, vr[i], vi[i], acc);
    }
  } else {
#pragma unroll
    for (int i = 0; i < P; ++i) vr[i] = vi[i] = T(0);
  }
  const RowLIdx li{lw * CF::LSTRIDE};
  fft_line<T, N, P, INV1, false>(vr, vi, t, lre, lim, li, twr, twi);
  constexpr T invN = T(1) / T(N);
  if constexpr (MODE == 1) {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pos = PL::template out_pos<false>(t, i);
        if constexpr (INV1) {
          vr[i] *= invN;
          vi[i] *= invN;
        }
        op.mid(f, pos, y, vr[i], vi[i], acc);
      }
    }
    fft_line<T, N, P, INV2, true>(vr, vi, t, lre, lim, li, twr, twi);
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pos = PL::template out_pos<true>(t, i);
        if constexpr (INV2) {
          vr[i] *= invN;
          vi[i] *= invN;
        }
        op.post(f, pos, y, vr[i], vi[i], acc);
        st(ore + base + pos, vr[i]);
        st(oim + base + pos, vi[i]);
      }
    }
  } else {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pos = PL::template out_pos<false>(t, i);
        if constexpr (INV1) {
          vr[i] *= invN;
          vi[i] *= invN;
        }
        op.post(f, pos, y, vr[i], vi[i], acc);
        st(ore + base + pos, vr[i]);
        st(oim + base + pos, vi[i]);
      }
    }
  }
  static_assert(Op::NACC == 0, "row passes carry no reductions");
}

// ---------------------------------------------------------------------------
// Column pass.  A workgroup owns C adjacent columns of one field (all ny rows)
// — `tiles` = ceil(nx / C) per field.
// ---------------------------------------------------------------------------
template <typename T, typename S, int N, int MODE, bool INV1, bool INV2, class Op>
__global__ void __launch_bounds__((ColCfg<T, N>::WG))
    k_col_pass(Op op, const S* __restrict__ ire, const S* __restrict__ iim, S* __restrict__ ore,
               S* __restrict__ oim, int nx, int tiles, const T* __restrict__ twr,
               const T* __restrict__ twi) {
  using CF = ColCfg<T, N>;
  constexpr int P = CF::P;
  constexpr int C = CF::C;
  using PL = LinePlan<N, P>;
  __shared__ T lre[CF::LDS_ELEMS];
  __shared__ T lim[CF::LDS_ELEMS];
  const int tid = threadIdx.x;
  const int col = tid % C;
  const int t = tid / C;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int f = lb / tiles;
  const int tile = lb - f * tiles;
  const int x = tile * C + col;
  const bool active = x < nx;
  const long fbase = (long)f * nx * N;
  T vr[P], vi[P];
  double acc[Op::NACC > 0 ? Op::NACC : 1];
#pragma unroll
  for (int k = 0; k < (Op::NACC > 0 ? Op::NACC : 1); ++k) acc[k] = 0.0;
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pos = PL::template in_pos<false>(t, i);
      const long a = fbase + (long)pos * nx + x;
      vr[i] = ld<T>(ire + a);
      vi[i] = ld<T>(iim + a);
      op.pre(f, x, pos, vr[i], vi[i], acc);
    }
  } else {
#pragma unroll
    for (int i = 0; i < P; ++i) vr[i] = vi[i] = T(0);
  }
  const ColLIdx<C> li{col};
  fft_line<T, N, P, INV1, false>(vr, vi, t, lre, lim, li, twr, twi);
  constexpr T invN = T(1) / T(N);
  if constexpr (MODE == 1) {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pos = PL::template out_pos<false>(t, i);
        if constexpr (INV1) {
          vr[i] *= invN;
          vi[i] *= invN;
        }
        op.mid(f, x, pos, vr[i], vi[i], acc);
      }
    }
    fft_line<T, N, P, INV2, true>(vr, vi, t, lre, lim, li, twr, twi);
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pos = PL::template out_pos<true>(t, i);
        if constexpr (INV2) {
          vr[i] *= invN;
          vi[i] *= invN;
        }
        op.post(f, x, pos, vr[i], vi[i], acc);
        const long a = fbase + (long)pos * nx + x;
        st(ore + a, vr[i]);
        st(oim + a, vi[i]);
      }
    }
  } else {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pos = PL::template out_pos<false>(t, i);
        if constexpr (INV1) {
          vr[i] *= invN;
          vi[i] *= invN;
        }
        op.post(f, x, pos, vr[i], vi[i], acc);
        const long a = fbase + (long)pos * nx + x;
        st(ore + a, vr[i]);
        st(oim + a, vi[i]);
      }
    }
  }
  if constexpr (Op::NACC > 0) {
    constexpr int NA = Op::NACC;
    __shared__ double sh[16 * NA];
    __shared__ int last;
    block_sum<NA>(acc, sh, tid, CF::WG);
    if (tid == 0) {
#pragma unroll
      for (int k = 0; k < NA; ++k) op.partials[(long)lb * NA + k] = acc[k];
      __threadfence();
      const unsigned prev = atomicAdd(op.counters + f, 1u);
      last = (prev == unsigned(tiles - 1));
    }
    __syncthreads();
    if (last) {
      __threadfence();
      // fold this field's tiles in a fixed order: thread k sums component k
      // over tile strides, then a fixed tree over strides.
      double s[NA];
#pragma unroll
      for (int k = 0; k < NA; ++k) s[k] = 0.0;
      for (int tl = tid; tl < tiles; tl += CF::WG) {
#pragma unroll
        for (int k = 0; k < NA; ++k) s[k] += op.partials[(long)(f * tiles + tl) * NA + k];
      }
      block_sum<NA>(s, sh, tid, CF::WG);
      if (tid == 0) {
        op.finalize(f, s);
        op.counters[f] = 0u;
      }
    }
  }
}

// Elementwise pass over a batch (no FFT): op.apply(field, x, y, k).
template <class Op>
__global__ void __launch_bounds__(256) k_elementwise(Op op, int nx, int ny, long total) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const long per = (long)nx * ny;
  const int f = int(k / per);
  const long r = k - (long)f * per;
  const int y = int(r / nx);
  const int x = int(r - (long)y * nx);
  op.apply(f, x, y, k);
}

}  // namespace doe
