// One launch per pass over several batches of fields of different sides.
//
// A table of up to kMergedMax entries (one per side: the pass operation, the
// planes, twiddles, geometry, and the entry's first workgroup) travels by
// value in the kernel arguments; each workgroup finds its entry by a uniform
// scan and runs that side's pass body.  Every body is given a 512-thread
// geometry so one block size and one LDS budget serve all sides 64 ... 2048:
// row passes take as many row groups as fill 512 threads, column passes as
// many columns, with the line transforms of the per-side passes (16 points
// per thread and both LDS planes below 256-point lines, 32 points and one
// plane from 256).  A 4096-point row needs 1024 threads, so sides stop at
// 2048.  FP32 fields.
#pragma once
#include "passes.hpp"

namespace doe {

constexpr int kMergedMax = 6;  // sides 64 ... 2048
constexpr int kMWG = 512;

template <int N2>
using MColCF = typename std::conditional<
    (N2 < 256), ColGeom<float, N2, 16, 512 / (N2 / 16), 2>,
    typename std::conditional<(N2 <= 512), ColGeom<float, N2, 32, 512 / (N2 / 32), 1>,
                              ColGeom<float, N2, 32, 16, 2>>::type>::type;
// the column split of a square side (col_split<float>(n, n))
template <int NX>
constexpr int m_n1() {
  return NX == 1024 ? 2 : (NX == 2048 ? 4 : 1);
}
template <int NX>
using MRowCF = RowCfg<float, NX, m_n1<NX>(), 2, kMWG>;

template <class Op>
struct MColEntry {
  Op op;
  const float* ire;
  const float* iim;
  float* ore;
  float* oim;
  const float* twr;
  const float* twi;
  int n2, nx, tiles, n1cnt;
};
template <class Op>
struct MRowEntry {
  Op op;
  const float* ire;
  const float* iim;
  float* ore;
  float* oim;
  const float* twr;
  const float* twi;
  const float* twyr;
  const float* twyi;
  long groups;
  int nx;
};
template <class E>
struct MTable {
  int n;
  int first[kMergedMax + 1];  // first workgroup of each entry; first[n] = grid
  E e[kMergedMax];
};

__device__ __forceinline__ int m_entry(const int* first, int n, int b) {
  int c = 0;
  while (c + 1 < n && b >= first[c + 1]) ++c;
  return c;
}

template <class Op, int N2>
constexpr int mcol_lds() {
  return MColCF<N2>::LDS_ELEMS + MColCF<N2>::LDS_IM;
}
template <class Op, int N2>
constexpr int mcol_tab() {
  return OpLineTab<Op, float, N2>::value;
}
template <int NX>
constexpr int mrow_lds() {
  return MRowCF<NX>::LDS_ELEMS + MRowCF<NX>::LDS_IM;
}

template <class Op, int MODE, bool INV1, bool INV2>
__global__ void __launch_bounds__(kMWG) k_col_merged(MTable<MColEntry<Op>> tab, int rev) {
  constexpr int L = cmax(cmax(mcol_lds<Op, 64>(), mcol_lds<Op, 128>()),
                         cmax(mcol_lds<Op, 256>(), mcol_lds<Op, 512>()));
  constexpr int LT = cmax(cmax(mcol_tab<Op, 64>(), mcol_tab<Op, 128>()),
                          cmax(mcol_tab<Op, 256>(), mcol_tab<Op, 512>()));
  __shared__ float lds[L];
  __shared__ double ltab[LT > 0 ? LT : 1];
  const int b = rev ? int(gridDim.x) - 1 - int(blockIdx.x) : int(blockIdx.x);
  const int c = m_entry(tab.first, tab.n, b);
  const MColEntry<Op>& E = tab.e[c];
  const int lb = xcd_remap(b - tab.first[c], tab.first[c + 1] - tab.first[c]);
  auto body = [&](auto NC) {
    constexpr int N2 = decltype(NC)::value;
    using CF = MColCF<N2>;
    static_assert(CF::WG == kMWG, "merged column bodies are 512 threads");
    col_body<float, float, N2, MODE, INV1, INV2, Op, CF>(E.op, E.ire, E.iim, E.ore, E.oim, E.nx,
                                                         E.tiles, E.n1cnt, E.twr, E.twi, lb, lds,
                                                         lds + CF::LDS_ELEMS, ltab);
  };
  switch (E.n2) {
    case 64: body(std::integral_constant<int, 64>{}); break;
    case 128: body(std::integral_constant<int, 128>{}); break;
    case 256: body(std::integral_constant<int, 256>{}); break;
    default: body(std::integral_constant<int, 512>{}); break;
  }
}

template <class Op, int MODE, bool INV1, bool INV2, bool XIN, bool XOUT>
__global__ void __launch_bounds__(kMWG) k_row_merged(MTable<MRowEntry<Op>> tab, int rev) {
  constexpr int L = cmax(cmax(cmax(mrow_lds<64>(), mrow_lds<128>()),
                              cmax(mrow_lds<256>(), mrow_lds<512>())),
                         cmax(mrow_lds<1024>(), mrow_lds<2048>()));
  __shared__ float lds[L];
  const int b = rev ? int(gridDim.x) - 1 - int(blockIdx.x) : int(blockIdx.x);
  const int c = m_entry(tab.first, tab.n, b);
  const MRowEntry<Op>& E = tab.e[c];
  const int lb = b - tab.first[c];
  auto body = [&](auto NC) {
    constexpr int NX = decltype(NC)::value;
    using CF = MRowCF<NX>;
    static_assert(CF::WG == kMWG, "merged row bodies are 512 threads");
    row_body<float, float, NX, m_n1<NX>(), MODE, INV1, INV2, XIN, XOUT, Op, CF>(
        E.op, E.ire, E.iim, E.ore, E.oim, E.nx, E.groups, E.twr, E.twi, E.twyr, E.twyi, lb, lds,
        lds + CF::LDS_ELEMS);
  };
  switch (E.nx) {
    case 64: body(std::integral_constant<int, 64>{}); break;
    case 128: body(std::integral_constant<int, 128>{}); break;
    case 256: body(std::integral_constant<int, 256>{}); break;
    case 512: body(std::integral_constant<int, 512>{}); break;
    case 1024: body(std::integral_constant<int, 1024>{}); break;
    default: body(std::integral_constant<int, 2048>{}); break;
  }
}

// Host side: the bodies' geometry per side (0 for sides they do not serve).
inline int mcol_cols(int n2) {
  switch (n2) {
    case 64: return MColCF<64>::C;
    case 128: return MColCF<128>::C;
    case 256: return MColCF<256>::C;
    case 512: return MColCF<512>::C;
    default: return 0;
  }
}
inline int mrow_groups(int nx) {
  switch (nx) {
    case 64: return MRowCF<64>::G;
    case 128: return MRowCF<128>::G;
    case 256: return MRowCF<256>::G;
    case 512: return MRowCF<512>::G;
    case 1024: return MRowCF<1024>::G;
    case 2048: return MRowCF<2048>::G;
    default: return 0;
  }
}

}  // namespace doe
