// Row and column pass kernels built on the line engine.
//
// The 2-D transform of an nx-by-ny grid is split into two passes over HBM.
// The column transform of length ny = N1*N2 uses the four-step
// factorisation
//     X[k1 + N1*k2] = sum_n2 W_N2^(n2*k2) * W_ny^(n2*k1) * sum_n1 x[N2*n1 + n2] W_N1^(n1*k1)
// so that its first two steps (an N1-point DFT across the rows
// {n2 + N2*n1} and the twiddle W_ny^(n2*k1)) run inside the ROW pass, on
// rows a workgroup already holds, and the COLUMN pass only performs N2-point
// transforms over contiguous blocks of N2 rows.  That keeps the column pass's
// tiles short enough to be C = 32 columns wide (128-byte row segments) while
// each point is still read and written once per pass.  Between the passes
// the column axis is in "four-step order": memory row N2*k1 + n2 holds the
// twiddled sub-transform value for (k1, n2).  N1 = 1 degenerates to the plain
// row/column passes.
//
// A pass loads whole lines, runs an optional elementwise `pre` hook, one FFT,
// and either stores (MODE 0) or runs a `mid` hook and a second FFT in the
// opposite stage order (MODE 1), then a `post` hook, and stores.  Hooks see
// (field, x, y) in natural order.  Ops with NACC > 0 accumulate FP64 partial
// sums per thread; each workgroup reduces them in a fixed order and stores
// them into its own slot of `partials` (plain stores), and k_finalize,
// launched right after the pass with one workgroup per field, folds that
// field's slots in slot order and calls op.finalize(field, sums):
// deterministic, no memset between launches and no cross-workgroup
// synchronisation inside a pass.
#pragma once
#include "fft_engine.hpp"

namespace doe {

// Block-wide deterministic reduction of NACC doubles; result valid in tid 0.
template <int NACC>
__device__ __forceinline__ void block_sum(double (&acc)[NACC], double* sh, int tid, int nthr) {
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

// Per-workgroup partial sums, folded per field by a kernel of their own.
// A pass's workgroup sums its per-thread accumulators (fixed order) and
// thread 0 publishes them to partials[slot] with plain stores; k_finalize,
// launched right after the pass, folds each field's partials in slot order
// and calls op.finalize.  No cross-workgroup synchronisation inside the
// pass: an arrival counter needs a release fence per workgroup, which on
// this GPU is an agent-scope L2 writeback (measured: 18-24 % of a far-field
// WGS column pass).
template <class Op, int NTHR>
__device__ __forceinline__ void publish_sums(const Op& op, double (&acc)[Op::NACC], int tid,
                                             long slot) {
  constexpr int NA = Op::NACC;
  __shared__ double sh[16 * NA];
  block_sum<NA>(acc, sh, tid, NTHR);
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NA; ++k) op.partials[slot * NA + k] = acc[k];
  }
}

// One workgroup per field: fold its `per_field` partials (from slot
// f * per_field on) and finalise.
template <class Op, int NTHR>
__global__ void __launch_bounds__(NTHR) k_finalize(Op op, int per_field) {
  constexpr int NA = Op::NACC;
  __shared__ double sh[16 * NA];
  const int f = blockIdx.x;
  const int tid = threadIdx.x;
  const long first = (long)f * per_field;
  double s[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) s[k] = 0.0;
  for (int w = tid; w < per_field; w += NTHR) {
#pragma unroll
    for (int k = 0; k < NA; ++k) s[k] += op.partials[(first + w) * NA + k];
  }
  block_sum<NA>(s, sh, tid, NTHR);
  if (tid == 0) op.finalize(f, s);
}

// ---------------------------------------------------------------------------
// Row pass geometry: G groups of N1 rows, T_LINE = NX/P threads per row.
// ---------------------------------------------------------------------------
template <typename T, int NX, int N1, int PLANES_ = 2, int WGT = 256>
struct RowCfg {
  static constexpr int PLANES = PLANES_;
  static constexpr int P = PChoice<NX>::P;
  static constexpr int T_LINE = NX / P;
  static constexpr int G = cmax(1, WGT / (T_LINE * N1));  // groups per workgroup
  static constexpr int LINES = G * N1;
  static constexpr int WG = LINES * T_LINE;
  // Padded line stride in LDS.  With N1 > 1 the cross-row stages put the N1
  // rows of a group in adjacent lanes, so the stride is 32/N1 mod 32 banks:
  // lanes (t, n1) of a half-wave then fall in distinct banks.
  static constexpr int LBASE = NX + NX / 16;
  static constexpr int LSTRIDE =
      N1 == 1 ? LBASE + 1 : LBASE + ((32 / N1 - LBASE % 32) % 32 + 32) % 32;
  static constexpr bool NEED_LDS = LinePlan<NX, P>::NST > 1 || N1 > 1;
  static constexpr int LDS_ELEMS = NEED_LDS ? LINES * LSTRIDE : 1;
  static constexpr int LDS_IM = PLANES == 2 ? LDS_ELEMS : 1;
  static_assert(WG <= 1024, "row workgroup too large");
  static_assert(N1 == 1 || N1 == 2 || N1 == 4, "cross-row DFT is a lane network of 2 or 4");
};

// Whether a column pass with this op may use the wide tile (ops say so with
// COL_WIDE; the default is no).
template <class Op, class = void>
struct OpWide {
  static constexpr bool value = false;
};
template <class Op>
struct OpWide<Op, std::void_t<decltype(Op::COL_WIDE)>> {
  static constexpr bool value = Op::COL_WIDE;
};

// Column geometry an op asks for.  Three forms, by priority: ColGeomT<T, N>
// (the op names its FP32 geometry per line length, lines of 256+ points);
// COL_P / COL_PLANES (and optionally COL_CMAX: widest tile, default 64
// columns) for FP32 lines of 512+ points; otherwise the generic choice.
// Field plane loads of the passes are non-temporal: each plane is read once
// per pass, and a non-temporal load still hits what the previous pass left
// in the memory-side cache (the alternating pass direction relies on that)
// while not filling the caches with read-once data.  Stores stay ordinary,
// so the next pass finds them there.  Measured on 16 x 2048^2 FP32 fft2d:
// 412 us against 430 us with ordinary loads; non-temporal stores as well:
// 433 us.  Column tiles whose row segment is narrower than a 128-byte line
// (16 FP32 columns) load ordinarily: the neighbouring tile reads the other
// half of each line from L2 (the 16 x 16 ASM column pass fetched 1.15x its
// planes with non-temporal loads, 1.00x with ordinary ones, in the same
// time).
template <typename T, typename S>
__device__ __forceinline__ T ld_field(const S* p) {
  if constexpr (std::is_same_v<S, float> || std::is_same_v<S, double>)
    return T(__builtin_nontemporal_load(p));
  else
    return ld<T>(p);
}

template <class Op, class = void>
struct HasColGeomT : std::false_type {};
template <class Op>
struct HasColGeomT<Op, std::void_t<typename Op::template ColGeomT<float, 256>>> : std::true_type {};
template <class Op, class = void>
struct HasColP : std::false_type {};
template <class Op>
struct HasColP<Op, std::void_t<decltype(Op::COL_P), decltype(Op::COL_PLANES)>> : std::true_type {};
template <class Op, class = void>
struct OpCmax {
  static constexpr int value = 64;
};
template <class Op>
struct OpCmax<Op, std::void_t<decltype(Op::COL_CMAX)>> {
  static constexpr int value = Op::COL_CMAX;
};
template <typename T, int N, class Op,
          int KIND = HasColGeomT<Op>::value ? 2 : (HasColP<Op>::value ? 1 : 0)>
struct OpColCfgK {
  using type = ColCfg<T, N, OpWide<Op>::value>;
};
template <typename T, int N, class Op>
struct OpColCfgK<T, N, Op, 1> {
  using type = typename std::conditional<(sizeof(T) == 4 && N >= 512),
                                         ColGeom<T, N, Op::COL_P, OpCmax<Op>::value, Op::COL_PLANES>,
                                         ColCfg<T, N, OpWide<Op>::value>>::type;
};
template <typename T, int N, class Op>
struct OpColCfgK<T, N, Op, 2> {
  using type = typename std::conditional<(sizeof(T) == 4 && N >= 256),
                                         typename Op::template ColGeomT<T, N>,
                                         ColCfg<T, N, OpWide<Op>::value>>::type;
};
template <typename T, int N, class Op>
struct OpColCfg {
  using type = typename OpColCfgK<T, N, Op>::type;
};

// Ops whose mid hook reads per-point inputs can have them prefetched with the
// column (Op::Aux, load_aux, mid_aux; Op::template PREFETCH<T, N> says for
// which line lengths).
template <class Op, typename T, int N, class = void>
struct OpPrefetch : std::false_type {};
template <class Op, typename T, int N>
struct OpPrefetch<Op, T, N, std::void_t<typename Op::Aux, decltype(Op::template PREFETCH<T, N>)>>
    : std::integral_constant<bool, Op::template PREFETCH<T, N>> {};
template <class Op, bool PF>
struct OpAuxOf {
  using type = int;
};
template <class Op>
struct OpAuxOf<Op, true> {
  using type = typename Op::Aux;
};

// Which elementwise hooks an op uses (ops that do not say use both).  The
// cross-row step may move past a hook that is a no-op.
template <class Op, class = void>
struct OpHooks {
  static constexpr bool pre = true, post = true;
};
template <class Op>
struct OpHooks<Op, std::void_t<decltype(Op::HAS_PRE), decltype(Op::HAS_POST)>> {
  static constexpr bool pre = Op::HAS_PRE, post = Op::HAS_POST;
};

// Ops may prepare, for a fused column pass (MODE 1, Op::template COL_CTX<T,
// N>), per-column constants (col_ctx, once per thread: uniform values the
// hook would otherwise reload per point) and a per-line table in LDS
// (Op::template LINE_TAB<T, N>: doubles per workgroup, filled by
// line_table before the first transform); their mid hook is then
// mid_line(ctx, table, line position, field, x, y, re, im, acc, active,
// aux) and must leave memory alone where `active` is false.  Such ops also
// fold the inverse transform's 1/N into the mid hook when they say so
// (Op::template SCALES_INV<T, N>).
template <class Op, typename T, int N, class = void>
struct OpColCtx : std::false_type {};
template <class Op, typename T, int N>
struct OpColCtx<Op, T, N, std::void_t<decltype(Op::template COL_CTX<T, N>)>>
    : std::integral_constant<bool, Op::template COL_CTX<T, N>> {};
template <class Op, typename T, int N, class = void>
struct OpLineTab {
  static constexpr int value = 0;
};
template <class Op, typename T, int N>
struct OpLineTab<Op, T, N, std::void_t<decltype(Op::template LINE_TAB<T, N>)>> {
  static constexpr int value = Op::template LINE_TAB<T, N>;
};
template <class Op, typename T, int N, class = void>
struct OpScalesInv : std::false_type {};
template <class Op, typename T, int N>
struct OpScalesInv<Op, T, N, std::void_t<decltype(Op::template SCALES_INV<T, N>)>>
    : std::integral_constant<bool, Op::template SCALES_INV<T, N>> {};
template <class Op, typename T, bool CC>
struct OpCtxOf {
  using type = int;
};
template <class Op, typename T>
struct OpCtxOf<Op, T, true> {
  using type = typename Op::template ColCtx<T>;
};

template <int N1>
__device__ __forceinline__ constexpr int brev_lane(int l) {
  return N1 == 4 ? ((l & 1) << 1) | ((l >> 1) & 1) : l;
}

// Lane exchange within a quad of lanes by DPP (no LDS traffic): MASK 1
// swaps lanes 0<->1 and 2<->3 (quad_perm [1,0,3,2]), MASK 2 swaps the pairs
// (quad_perm [2,3,0,1]).
template <int MASK, typename T>
__device__ __forceinline__ T shx(T v) {
  constexpr int CTRL = MASK == 1 ? 0xB1 : 0x4E;
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(
        T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
  } else {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(T, (long long)(unsigned)lo | ((long long)hi << 32));
  }
}

// One radix-2 lane butterfly: lower lane (bit clear) u+v, upper lane v-u,
// i.e. sgn*u + v with sgn = +1 / -1 (exact: one rounding, as the sum).
template <int MASK, typename T>
__device__ __forceinline__ void lane_bfly(T& ur, T& ui, T sgn) {
  const T vr = shx<MASK>(ur), vi = shx<MASK>(ui);
  ur = fma(sgn, ur, vr);
  ui = fma(sgn, ui, vi);
}

// Per-lane constants of the N1-point lane DFT (lane slot n1).
template <typename T>
struct LaneDft {
  T s1, s2;  // butterfly signs for the lane bits 1 and 2
  bool rot;  // n1 == 3: the W4 twiddle between the two butterflies
  __device__ __forceinline__ explicit LaneDft(int n1)
      : s1((n1 & 1) ? T(-1) : T(1)), s2((n1 & 2) ? T(-1) : T(1)), rot(n1 == 3) {}
};

// multiply by W4 = -i (forward) or +i (inverse) where `rot`, by selects
template <bool INV, typename T>
__device__ __forceinline__ void mul_w4(T& ur, T& ui, bool rot) {
  const T r = ur, i = ui;
  ur = rot ? (INV ? -i : i) : r;
  ui = rot ? (INV ? r : -r) : i;
}

template <typename T, int N1, bool INV, bool DIT>
__device__ __forceinline__ void lane_dft(T& ur, T& ui, const LaneDft<T>& ld) {
  if constexpr (N1 == 2) {
    lane_bfly<1>(ur, ui, ld.s1);
  } else if constexpr (N1 == 4) {
    if constexpr (!DIT) {
      lane_bfly<2>(ur, ui, ld.s2);
      mul_w4<INV>(ur, ui, ld.rot);
      lane_bfly<1>(ur, ui, ld.s1);
    } else {
      lane_bfly<1>(ur, ui, ld.s1);
      mul_w4<INV>(ur, ui, ld.rot);
      lane_bfly<2>(ur, ui, ld.s2);
    }
  }
}

// Cross-row step on the registers of a lane whose group rows sit in adjacent
// lanes (lane n1 holds row slot n1 of its group, the same positions in every
// lane).  Forward: DFT across the lanes (lane n1 then holds k1 = bitrev(n1)),
// then the twiddle W_ny^(n2*k1).  Inverse: lane n1 holds k1 = n1; conjugate
// twiddle, inverse DFT across the lanes, unscaled (the caller applies 1/N1;
// lane n1 then holds the natural row slot bitrev(n1)).
template <typename T, int P, int N1, bool INV>
__device__ __forceinline__ void cross_lanes(T (&vr)[P], T (&vi)[P], int n1, T wr, T wi) {
  const LaneDft<T> ld(n1);
#pragma unroll
  for (int i = 0; i < P; ++i) {
    if constexpr (INV) {
      // the inverse step's 1/N1 is applied by the row pass together with
      // the row transform's 1/NX (one exact power-of-two scaling)
      const T a = vr[i] * wr - vi[i] * wi;
      vi[i] = vr[i] * wi + vi[i] * wr;
      vr[i] = a;
      lane_dft<T, N1, true, false>(vr[i], vi[i], ld);
    } else {
      lane_dft<T, N1, false, false>(vr[i], vi[i], ld);
      const T a = vr[i] * wr - vi[i] * wi;
      vi[i] = vr[i] * wi + vi[i] * wr;
      vr[i] = a;
    }
  }
}

// Row-pass geometry with the cross-row step folded into the line FFT's own
// exchanges.  Stages 0 and NST-1 use the contiguous layout (the threads of
// one row are adjacent lanes, so loads and stores are 256-byte runs); stages
// 1..NST-2 use the interleaved layout (lane = t*N1 + n1), where the cross-row
// step runs on registers right after the exchange into stage 1.  After that
// step a lane holds row slot bitrev(n1) and writes it back to that line, so
// an LDS line index always equals the row slot it holds.  Row FFTs and the
// cross-row step commute (one acts along x, the other across rows with
// x-independent weights), so the order is free.  CROSS: 0 none, 1 forward,
// 2 inverse.
template <int NX, int NST, int LSTRIDE, int CROSS, typename T, int N1>
struct RowLaneGeo {
  int tc, lc, ti, li, lb, n1;
  T cwr, cwi;  // the cross-row twiddle of this lane (loaded before the transform)
  static constexpr int S_IN = 1, S_OUT = NST - 1;
  static_assert(NST >= 3, "needs two exchanges");
  static constexpr bool inter(int s) { return s >= S_IN && s < S_OUT; }
  __device__ __forceinline__ static int lidx(int line, int pos) {
    return line * LSTRIDE + pos + (pos >> 4);
  }
  // position rt + CC of `line`, rt a multiple of M: the padding splits into a
  // per-thread part and a constant when either part is a multiple of 16
  template <int CC, int M>
  __device__ __forceinline__ static int lidx_c(int line, int rt) {
    if constexpr (CC % 16 == 0 || M % 16 == 0)
      return line * LSTRIDE + (rt + (rt >> 4)) + (CC + (CC >> 4));
    else
      return lidx(line, rt + CC);
  }
  template <int S>
  __device__ __forceinline__ int t() const { return inter(S) ? ti : tc; }
  template <int S>
  __device__ __forceinline__ int ridx(int pos) const { return lidx(inter(S) ? li : lc, pos); }
  template <int S>
  __device__ __forceinline__ int widx(int pos) const {
    if constexpr (inter(S - 1)) return lidx((S - 1 == S_IN && CROSS != 0) ? lb : li, pos);
    else return lidx(lc, pos);
  }
  template <int S, int CC, int M>
  __device__ __forceinline__ int ridx_c(int rt) const {
    return lidx_c<CC, M>(inter(S) ? li : lc, rt);
  }
  template <int S, int CC, int M>
  __device__ __forceinline__ int widx_c(int rt) const {
    if constexpr (inter(S - 1)) return lidx_c<CC, M>((S - 1 == S_IN && CROSS != 0) ? lb : li, rt);
    else return lidx_c<CC, M>(lc, rt);
  }
  template <int S, typename U, int PP>
  __device__ __forceinline__ void after_read(U (&vr)[PP], U (&vi)[PP]) const {
    if constexpr (S == S_IN && CROSS == 1) cross_lanes<U, PP, N1, false>(vr, vi, n1, cwr, cwi);
    if constexpr (S == S_IN && CROSS == 2) cross_lanes<U, PP, N1, true>(vr, vi, n1, cwr, cwi);
  }
};


// Steps 1+2 of the four-step column factorisation across the N1 rows of each
// group (forward: DFT then twiddle; inverse: conjugate twiddle then inverse
// DFT scaled by 1/N1).  Registers hold positions posf(i) of line `l`.
template <typename T, int NX, int P, int N1, bool INV, int LS, class PosF>
__device__ __forceinline__ void cross_rows(T (&vr)[P], T (&vi)[P], int l, int t, const PosF& posf,
                                           T* lre, T* lim, int n2, const T* __restrict__ twyr,
                                           const T* __restrict__ twyi) {
  if constexpr (N1 > 1) {
    constexpr int T_LINE = NX / P;
    const int g = l / N1;
    const int n1 = l - g * N1;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pos = posf(i);
      const int a = l * LS + pos + (pos >> 4);
      lre[a] = vr[i];
      lim[a] = vi[i];
    }
    __syncthreads();
    // every thread of the group transforms P/N1 positions across the rows
    const int j = n1 * T_LINE + t;
    T wr[N1], wi[N1];
    wr[0] = T(1);
    wi[0] = T(0);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
      wr[k1] = twyr[n2 * k1];
      wi[k1] = INV ? -twyi[n2 * k1] : twyi[n2 * k1];
    }
#pragma unroll
    for (int m = 0; m < P / N1; ++m) {
      const int x = j + m * (N1 * T_LINE);
      const int xo = x + (x >> 4);
      T ur[N1], ui[N1];
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        ur[r] = lre[(g * N1 + r) * LS + xo];
        ui[r] = lim[(g * N1 + r) * LS + xo];
      }
      if constexpr (INV) {
#pragma unroll
        for (int r = 1; r < N1; ++r) {
          const T a = ur[r] * wr[r] - ui[r] * wi[r];
          ui[r] = ur[r] * wi[r] + ui[r] * wr[r];
          ur[r] = a;
        }
        dft<N1, true>(ur, ui);
        constexpr T s = T(1) / T(N1);
#pragma unroll
        for (int r = 0; r < N1; ++r) {
          ur[r] *= s;
          ui[r] *= s;
        }
      } else {
        dft<N1, false>(ur, ui);
#pragma unroll
        for (int r = 1; r < N1; ++r) {
          const T a = ur[r] * wr[r] - ui[r] * wi[r];
          ui[r] = ur[r] * wi[r] + ui[r] * wr[r];
          ur[r] = a;
        }
      }
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        lre[(g * N1 + r) * LS + xo] = ur[r];
        lim[(g * N1 + r) * LS + xo] = ui[r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pos = posf(i);
      const int a = l * LS + pos + (pos >> 4);
      vr[i] = lre[a];
      vi[i] = lim[a];
    }
  }
}

// ---------------------------------------------------------------------------
// Row pass.  Group gi = (field f, n2) owns rows y = n2 + N2*n1, n1 < N1.
// XIN: the input's column axis is in four-step order (apply the inverse
// cross-row step first); XOUT: leave the output in four-step order.
// ---------------------------------------------------------------------------
// The body of a row pass for logical workgroup `wb`, on the LDS buffers the
// caller provides (CF::LDS_ELEMS and CF::LDS_IM elements).  Shared by the
// one-batch kernel below and the table-driven batched kernels.
template <typename T, typename S, int NX, int N1, int MODE, bool INV1, bool INV2, bool XIN,
          bool XOUT, class Op, class CF>
__device__ __forceinline__ void row_body(const Op& op, const S* __restrict__ ire,
                                         const S* __restrict__ iim, S* __restrict__ ore,
                                         S* __restrict__ oim, int ny, long total_groups,
                                         const T* __restrict__ twr, const T* __restrict__ twi,
                                         const T* __restrict__ twyr, const T* __restrict__ twyi,
                                         int wb, T* lre, T* lim) {
  constexpr int P = CF::P;
  using PL = LinePlan<NX, P>;
  // Cross-row steps on lanes inside the line FFT's exchanges when the FFT
  // has two exchanges and the hooks they move past are no-ops; otherwise
  // through an extra LDS exchange around the FFT.
  static_assert(!(MODE == 0 && XIN && XOUT), "one cross-row step per line FFT");
  constexpr bool LANES = N1 > 1 && PL::NST >= 3 && (!XIN || !OpHooks<Op>::pre) &&
                         (!XOUT || !OpHooks<Op>::post);
  static_assert(CF::PLANES == 2 || LANES || N1 == 1, "the LDS cross-row step needs two planes");
  const int N2 = ny / N1;
  const int tid = threadIdx.x;
  const int l = tid / CF::T_LINE;  // contiguous layout: line and position index
  const int t = tid - l * CF::T_LINE;
  const int g = l / N1;
  const int n1 = l - g * N1;
  const long gi = (long)wb * CF::G + g;
  const bool active = gi < total_groups;
  const int f = active ? int(gi / N2) : 0;
  const int n2 = active ? int(gi - (long)f * N2) : 0;
  const int y = n2 + N2 * n1;
  const long base = ((long)f * ny + y) * (long)NX;
  constexpr int NA = Op::NACC > 0 ? Op::NACC : 1;
  T vr[P], vi[P];
  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pos = PL::template in_pos<false>(t, i);
      vr[i] = ld_field<T>(ire + base + pos);
      vi[i] = ld_field<T>(iim + base + pos);
    }
  } else {
#pragma unroll
    for (int i = 0; i < P; ++i) vr[i] = vi[i] = T(0);
  }
  if constexpr (XIN && !LANES)
    cross_rows<T, NX, P, N1, true, CF::LSTRIDE>(
        vr, vi, l, t, [&](int i) { return PL::template in_pos<false>(t, i); }, lre, lim, n2, twyr,
        twyi);
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) op.pre(f, PL::template in_pos<false>(t, i), y, vr[i], vi[i], acc);
  }
  // per-point inputs of the mid hook loaded with the row (see col_body)
  constexpr bool PF = MODE == 1 && OpPrefetch<Op, T, NX>::value;
  typename OpAuxOf<Op, PF>::type aux[PF ? P : 1];
  if constexpr (PF) {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) aux[i] = op.load_aux(f, PL::template out_pos<false>(t, i), y);
    }
  }
  // the two line FFTs (MODE 1) or one (MODE 0)
  constexpr int CROSS1 = !LANES ? 0 : XIN ? 2 : (MODE == 0 && XOUT) ? 1 : 0;
  constexpr int CROSS2 = !LANES ? 0 : XOUT ? 1 : 0;
  auto line_fft = [&](auto inv_c, auto rev_c, auto cross_c) {
    constexpr bool INV = decltype(inv_c)::value;
    constexpr bool REV = decltype(rev_c)::value;
    constexpr int CROSS = decltype(cross_c)::value;
    if constexpr (LANES) {
      constexpr int GT = N1 * CF::T_LINE;  // threads of one group
      const int gt = tid - g * GT;
      const int ni = gt % N1;
      // the lane's cross-row twiddle W_ny^(n2*k1): forward k1 = bitrev(slot),
      // inverse k1 = slot, conjugated
      T cwr = T(1), cwi = T(0);
      if constexpr (CROSS != 0) {
        const int k1 = CROSS == 2 ? ni : brev_lane<N1>(ni);
        cwr = twyr[n2 * k1];
        cwi = CROSS == 2 ? -twyi[n2 * k1] : twyi[n2 * k1];
      }
      RowLaneGeo<NX, PL::NST, CF::LSTRIDE, CROSS, T, N1> geo{
          t, l, gt / N1, g * N1 + ni, g * N1 + brev_lane<N1>(ni), ni, cwr, cwi};
      fft_line_g<T, NX, P, INV, REV, CF::PLANES>(vr, vi, geo, lre, lim, twr, twi);
    } else {
      const RowLIdx li{l * CF::LSTRIDE};
      fft_line<T, NX, P, INV, REV, CF::PLANES>(vr, vi, t, lre, lim, li, twr, twi);
    }
  };
  line_fft(std::bool_constant<INV1>{}, std::false_type{}, std::integral_constant<int, CROSS1>{});
  constexpr T invN = T(1) / T(NX);
  // the first transform's scale: 1/NX if inverse, times the 1/N1 of an
  // inverse cross-row step on lanes (powers of two, so one multiplication
  // gives the two scalings' result)
  constexpr bool SC1 = INV1 || CROSS1 == 2;
  constexpr T sc1 = (INV1 ? T(1) / T(NX) : T(1)) * (CROSS1 == 2 ? T(1) / T(N1) : T(1));
  constexpr bool REV_OUT = MODE == 1;
  if constexpr (MODE == 1) {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        if constexpr (SC1) {
          vr[i] *= sc1;
          vi[i] *= sc1;
        }
        if constexpr (PF)
          op.mid_aux(f, PL::template out_pos<false>(t, i), y, vr[i], vi[i], acc, aux[i]);
        else
          op.mid(f, PL::template out_pos<false>(t, i), y, vr[i], vi[i], acc);
      }
    }
    line_fft(std::bool_constant<INV2>{}, std::true_type{}, std::integral_constant<int, CROSS2>{});
  }
  // the last transform's scale (a second transform's cross-row step is
  // always the forward one)
  constexpr bool SC_LAST = MODE == 1 ? INV2 : SC1;
  constexpr T sc_last = MODE == 1 ? invN : sc1;
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      if constexpr (SC_LAST) {
        vr[i] *= sc_last;
        vi[i] *= sc_last;
      }
      op.post(f, PL::template out_pos<REV_OUT>(t, i), y, vr[i], vi[i], acc);
    }
  }
  if constexpr (XOUT && !LANES)
    cross_rows<T, NX, P, N1, false, CF::LSTRIDE>(
        vr, vi, l, t, [&](int i) { return PL::template out_pos<REV_OUT>(t, i); }, lre, lim, n2,
        twyr, twyi);
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pos = PL::template out_pos<REV_OUT>(t, i);
      st(ore + base + pos, vr[i]);
      st(oim + base + pos, vi[i]);
    }
  }
  // a workgroup's groups never span two fields (the launcher checks G | N2),
  // so slot wb belongs to field f: slots f * (N2 / G) ... (f + 1) * (N2 / G) - 1
  if constexpr (Op::NACC > 0) publish_sums<Op, CF::WG>(op, acc, tid, wb);
}

template <typename T, typename S, int NX, int N1, int MODE, bool INV1, bool INV2, bool XIN,
          bool XOUT, class Op, int RPLANES = 2>
__global__ void __launch_bounds__((RowCfg<T, NX, N1, RPLANES>::WG))
    k_row_pass(Op op, const S* __restrict__ ire, const S* __restrict__ iim, S* __restrict__ ore,
               S* __restrict__ oim, int ny, long total_groups, const T* __restrict__ twr,
               const T* __restrict__ twi, const T* __restrict__ twyr, const T* __restrict__ twyi,
               int rev) {
  using CF = RowCfg<T, NX, N1, RPLANES>;
  __shared__ T lre[CF::LDS_ELEMS];
  __shared__ T lim[CF::LDS_IM];
  // rev: walk the grid from its far end (consecutive passes alternate, so a
  // pass starts on the rows the previous one touched last, still in cache)
  const int wb = rev ? int(gridDim.x) - 1 - int(blockIdx.x) : int(blockIdx.x);
  row_body<T, S, NX, N1, MODE, INV1, INV2, XIN, XOUT, Op, CF>(op, ire, iim, ore, oim, ny,
                                                             total_groups, twr, twi, twyr, twyi,
                                                             wb, lre, lim);
}

// ---------------------------------------------------------------------------
// Column pass over N2-point lines.  A workgroup owns C adjacent columns of
// block k1 of one field.  Row of line position p:
//   forward (MODE 0, !INV1): load N2*k1 + p (four-step), store k1 + N1*p
//   inverse (MODE 0, INV1):  load k1 + N1*p (natural),   store N2*k1 + p
//   fused   (MODE 1):        load and store N2*k1 + p, natural index in mid
// Hooks see the natural row index y = k1 + N1*p.
// ---------------------------------------------------------------------------
// The body of a column pass for logical workgroup `lb` on caller-provided
// LDS (CF::LDS_ELEMS and CF::LDS_IM elements).
template <typename T, typename S, int N2, int MODE, bool INV1, bool INV2, class Op, class CF>
__device__ __forceinline__ void col_body(const Op& op, const S* __restrict__ ire,
                                         const S* __restrict__ iim, S* __restrict__ ore,
                                         S* __restrict__ oim, int nx, int tiles, int n1cnt,
                                         const T* __restrict__ twr, const T* __restrict__ twi,
                                         int lb, T* lre, T* lim, double* ltab) {
  constexpr int P = CF::P;
  constexpr int C = CF::C;
  using PL = LinePlan<N2, P>;
  const int tid = threadIdx.x;
  const int col = tid % C;
  const int t = tid / C;
  const int per_field = tiles * n1cnt;
  const int f = lb / per_field;
  const int rem = lb - f * per_field;
  const int k1 = rem / tiles;
  const int tile = rem - k1 * tiles;
  const int x = tile * C + col;
  const bool active = x < nx;
  const int ny = N2 * n1cnt;
  const long fbase = (long)f * nx * ny;
  constexpr bool NAT_IN = MODE == 0 && INV1;
  constexpr bool NAT_OUT = MODE == 0 && !INV1;
  constexpr int NA = Op::NACC > 0 ? Op::NACC : 1;
  T vr[P], vi[P];
  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;
  // Row of line position p = t + c (c a per-register constant): a workgroup-
  // uniform part (k1 and c) and a per-thread part (t), so each access is a
  // uniform base (scalar registers) plus one 32-bit per-thread offset.
  const int toff_in = (NAT_IN ? n1cnt * t : t) * nx + x;
  const int toff_out = (NAT_OUT ? n1cnt * t : t) * nx + x;
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int c = PL::template in_off<false>(i);
      const int p = t + c;
      const long urow = NAT_IN ? long(k1) + long(n1cnt) * c : long(N2) * k1 + c;
      const S* ur = ire + (fbase + urow * nx);
      const S* ui = iim + (fbase + urow * nx);
      // a row segment narrower than a 128-byte line shares its line with the
      // neighbouring tile, which finds it in L2 only after an ordinary load
      if constexpr (C * sizeof(S) < 128) {
        vr[i] = ld<T>(ur + toff_in);
        vi[i] = ld<T>(ui + toff_in);
      } else {
        vr[i] = ld_field<T>(ur + toff_in);
        vi[i] = ld_field<T>(ui + toff_in);
      }
      op.pre(f, x, k1 + n1cnt * p, vr[i], vi[i], acc);
    }
  } else {
#pragma unroll
    for (int i = 0; i < P; ++i) vr[i] = vi[i] = T(0);
  }
  // ops with per-point inputs for the mid hook (Op::Aux) may ask for them to
  // be loaded here, with the column, so their latency overlaps the field
  // loads instead of stalling the workgroup between the two transforms
  constexpr bool PF = MODE == 1 && OpPrefetch<Op, T, N2>::value;
  typename OpAuxOf<Op, PF>::type aux[PF ? P : 1];
  if constexpr (PF) {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i)
        aux[i] = op.load_aux(f, x, k1 + n1cnt * PL::template out_pos<false>(t, i));
    }
  }
  // per-line table and per-column constants of the mid hook
  constexpr bool CC = MODE == 1 && OpColCtx<Op, T, N2>::value;
  typename OpCtxOf<Op, T, CC>::type cc{};
  if constexpr (CC) {
    if constexpr (OpLineTab<Op, T, N2>::value > 0) {
      op.template line_table<T, N2>(ltab, f, k1, n1cnt, tid, CF::WG);
      if constexpr (PL::NST == 1) __syncthreads();  // no exchange barrier before the mid hook
    }
    cc = op.template col_ctx<T, N2>(f, x);
  }
  const ColLIdx<C> li{col};
  fft_line<T, N2, P, INV1, false, CF::PLANES>(vr, vi, t, lre, lim, li, twr, twi);
  constexpr T invN = T(1) / T(N2);
  constexpr bool REV_OUT = MODE == 1;
  if constexpr (MODE == 1 && CC) {
    // the hook runs for every column (columns past the grid hold zero lines)
    // and keeps its memory accesses to active ones: no branch around the
    // registers it updates
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int p = PL::template out_pos<false>(t, i);
      if constexpr (INV1) {
        vr[i] *= invN;
        vi[i] *= invN;
      }
      op.template mid_line<T, N2>(cc, ltab, p, f, x, k1 + n1cnt * p, vr[i], vi[i], acc, active,
                                  aux[PF ? i : 0]);
    }
    fft_line<T, N2, P, INV2, true, CF::PLANES>(vr, vi, t, lre, lim, li, twr, twi);
  } else if constexpr (MODE == 1) {
    if (active) {
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int p = PL::template out_pos<false>(t, i);
        if constexpr (INV1) {
          vr[i] *= invN;
          vi[i] *= invN;
        }
        if constexpr (PF)
          op.mid_aux(f, x, k1 + n1cnt * p, vr[i], vi[i], acc, aux[i]);
        else
          op.mid(f, x, k1 + n1cnt * p, vr[i], vi[i], acc);
      }
    }
    fft_line<T, N2, P, INV2, true, CF::PLANES>(vr, vi, t, lre, lim, li, twr, twi);
  }
  constexpr bool INV_LAST = MODE == 1 ? INV2 : INV1;
  // the mid hook may already carry the inverse transform's 1/N
  constexpr bool SCALE_LAST = INV_LAST && !(MODE == 1 && INV2 && OpScalesInv<Op, T, N2>::value);
  if (active) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int p = PL::template out_pos<REV_OUT>(t, i);
      if constexpr (SCALE_LAST) {
        vr[i] *= invN;
        vi[i] *= invN;
      }
      op.post(f, x, k1 + n1cnt * p, vr[i], vi[i], acc);
      const int c = PL::template out_off<REV_OUT>(i);
      const long urow = NAT_OUT ? long(k1) + long(n1cnt) * c : long(N2) * k1 + c;
      st(ore + (fbase + urow * nx) + toff_out, vr[i]);
      st(oim + (fbase + urow * nx) + toff_out, vi[i]);
    }
  }
  // slot lb = f * per_field + (k1, tile): field f's slots are contiguous
  if constexpr (Op::NACC > 0) publish_sums<Op, CF::WG>(op, acc, tid, lb);
}

template <typename T, typename S, int N2, int MODE, bool INV1, bool INV2, class Op,
          class CF = typename OpColCfg<T, N2, Op>::type>
__global__ void __launch_bounds__((CF::WG))
    k_col_pass(Op op, const S* __restrict__ ire, const S* __restrict__ iim, S* __restrict__ ore,
               S* __restrict__ oim, int nx, int tiles, int n1cnt, const T* __restrict__ twr,
               const T* __restrict__ twi, int rev) {
  __shared__ T lre[CF::LDS_ELEMS];
  __shared__ T lim[CF::LDS_IM];
  constexpr int NT = MODE == 1 ? OpLineTab<Op, T, N2>::value : 0;
  __shared__ double ltab[NT > 0 ? NT : 1];
  const int lb = xcd_remap(rev ? int(gridDim.x) - 1 - int(blockIdx.x) : int(blockIdx.x), gridDim.x);
  col_body<T, S, N2, MODE, INV1, INV2, Op, CF>(op, ire, iim, ore, oim, nx, tiles, n1cnt, twr, twi,
                                               lb, lre, lim, ltab);
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
