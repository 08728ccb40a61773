// Register-resident, LDS-exchanged Stockham FFT engine for one line of a
// power-of-two 2-D grid (a row or a column).
//
// This replaces the reference's radix-2 DIT fft1d (src/fft.js:86-139) with a
// mixed-radix (16/8/4/2) Stockham auto-sort formulation: each line of N points
// is held by T = N/P threads, P points per thread.  A stage of radix R
// performs P/R radix-R butterflies per thread entirely in registers; between
// stages the line is exchanged through LDS.  Conventions match the reference:
// forward is the unscaled DFT with exp(-2*pi*i*n*k/N), the inverse conjugates
// the kernel and scales by 1/N.
//
// Register convention (all positions are indices along the line):
//   entry:  v[b*R0 + r] = x[t + b*T + r*(N/R0)]      (R0 = first radix)
//   exit:   v[b*RL + q] = X[t + b*T + q*(N/RL)]      (RL = last radix)
// Running the inverse with the stage order reversed (REV) makes its entry
// positions equal the forward's exit positions, so a forward FFT, a pointwise
// operation and an inverse FFT chain in registers without an extra exchange.
#pragma once
#include "common.hpp"

namespace doe {

// Multiply (xr + i xi) by exp(-/+ 2*pi*i*m/32): minus for the forward
// transform, plus for the inverse.  Multiples of a quarter turn are exact.
template <int M, bool INV, typename T>
__device__ __forceinline__ void mul_w32(T& xr, T& xi) {
  constexpr int m = M & 31;
  if constexpr (m == 0) {
    return;
  } else if constexpr (m == 16) {
    xr = -xr;
    xi = -xi;
  } else if constexpr (m == 8 || m == 24) {
    // exp(-i*pi/2) = -i (forward, m=8); the other three cases follow.
    constexpr bool times_minus_i = (m == 8) != INV;
    const T r = xr;
    if constexpr (times_minus_i) {
      xr = xi;
      xi = -r;
    } else {
      xr = -xi;
      xi = r;
    }
  } else {
    constexpr T c = T(c32(m));
    constexpr T s = INV ? T(s32(m)) : T(-s32(m));
    const T r = xr * c - xi * s;
    xi = xr * s + xi * c;
    xr = r;
  }
}

// In-register radix-R DFT (R <= 32), natural order in and out.
template <int R, bool INV, typename T>
__device__ __forceinline__ void dft(T* vr, T* vi) {
  if constexpr (R > 1) {
    constexpr int L = ilog2(R);
    T ar[R], ai[R];
    static_for<0, R>([&](auto I) {
      constexpr int i = decltype(I)::value;
      ar[bitrev(i, L)] = vr[i];
      ai[bitrev(i, L)] = vi[i];
    });
    static_for<1, L + 1>([&](auto S) {
      constexpr int len = 1 << decltype(S)::value;
      constexpr int half = len >> 1;
      static_for<0, R / 2>([&](auto B) {
        constexpr int b = decltype(B)::value;
        constexpr int k = b % half;
        constexpr int p = (b / half) * len + k;
        constexpr int q = p + half;
        T tr = ar[q], ti = ai[q];
        mul_w32<k * (32 / len), INV>(tr, ti);
        ar[q] = ar[p] - tr;
        ai[q] = ai[p] - ti;
        ar[p] = ar[p] + tr;
        ai[p] = ai[p] + ti;
      });
    });
    static_for<0, R>([&](auto I) {
      constexpr int i = decltype(I)::value;
      vr[i] = ar[i];
      vi[i] = ai[i];
    });
  }
}

// Stage plan for a line of N points with P points per thread: full stages of
// radix P, then one remainder stage.  REV reverses the order.
template <int N, int P>
struct LinePlan {
  static_assert((N & (N - 1)) == 0 && N >= 2, "N must be a power of two");
  static_assert(P <= N && P <= 32 && (P & (P - 1)) == 0, "bad P");
  static constexpr int T = N / P;
  static constexpr int LOGN = ilog2(N);
  static constexpr int LOGP = ilog2(P);
  static constexpr int NFULL = LOGN / LOGP;
  static constexpr int REM = 1 << (LOGN - NFULL * LOGP);
  static constexpr int NST = NFULL + (REM > 1 ? 1 : 0);
  static constexpr int radix_fwd(int s) { return s < NFULL ? P : REM; }
  template <bool REV>
  static constexpr int radix(int s) { return REV ? radix_fwd(NST - 1 - s) : radix_fwd(s); }
  template <bool REV>
  static constexpr int ns(int s) {  // product of the radices before stage s
    int p = 1;
    for (int i = 0; i < s; ++i) p *= radix<REV>(i);
    return p;
  }
  template <bool REV>
  static constexpr int first() { return radix<REV>(0); }
  template <bool REV>
  static constexpr int last() { return radix<REV>(NST - 1); }
  // Position (along the line) held by register i of thread t at entry/exit:
  // t plus a per-register constant.
  template <bool REV>
  static constexpr int in_off(int i) {
    return (i / first<REV>()) * T + (i % first<REV>()) * (N / first<REV>());
  }
  template <bool REV>
  static constexpr int out_off(int i) {
    return (i / last<REV>()) * T + (i % last<REV>()) * (N / last<REV>());
  }
  template <bool REV>
  __device__ __forceinline__ static int in_pos(int t, int i) {
    return t + in_off<REV>(i);
  }
  template <bool REV>
  __device__ __forceinline__ static int out_pos(int t, int i) {
    return t + out_off<REV>(i);
  }
};

// How a stage obtains the twiddles of its NB butterfly groups.  Group b of
// thread t multiplies input r by W_N^(r*m_b), m_b = ((t + b*TT) % NS) * (N / (NS*R)).
//   TW_SAME: TT is a multiple of NS, so m_b = (t % NS) * N/(NS*R) for every b:
//            one base twiddle and one set of powers per stage.
//   TW_EACH: each group has its own base twiddle.
// Every base twiddle is loaded from the table when the line transform starts,
// so the loads' latency overlaps the first stage.
enum { TW_NONE = 0, TW_SAME = 1, TW_EACH = 2 };
template <int TT, int NS>
constexpr int tw_mode() {
  if (NS <= 1) return TW_NONE;
  if (TT % NS == 0) return TW_SAME;
  return TW_EACH;
}

// One full line FFT with a per-stage thread geometry.  `geo` says, for each
// stage s, which line position index t the thread works on (geo.t<s>()),
// where the exchange before stage s writes stage s-1's outputs
// (geo.widx<s>(pos)) and reads stage s's inputs (geo.ridx<s>(pos)), and gets
// a hook after that read (geo.after_read<s>(vr, vi)).  A thread may change
// line or position between stages; every (line, position) must be written
// and read exactly once per exchange.  Every thread of the workgroup must
// call this (it contains barriers), active or not.
// Twiddle table: tw[m] = exp(-2*pi*i*m/N), m in [0, N), planar re/im.
// PLANES = 2: separate LDS buffers for re and im; PLANES = 1: one buffer,
// the two planes exchanged one after the other (half the LDS, two more
// barriers per exchange).
template <typename T, int N, int P, bool INV, bool REV, int PLANES = 2, class Geo>
__device__ __forceinline__ void fft_line_g(T (&vr)[P], T (&vi)[P], Geo& geo, T* lds_re, T* lds_im,
                                           const T* __restrict__ twr, const T* __restrict__ twi) {
  using PL = LinePlan<N, P>;
  constexpr int TT = PL::T;
  // base twiddles of every stage and group, loaded up front (slot s*P + b)
  T pwr[PL::NST * P], pwi[PL::NST * P];
  static_for<0, PL::NST>([&](auto SI) {
    constexpr int s = decltype(SI)::value;
    constexpr int R = PL::template radix<REV>(s);
    constexpr int NS = PL::template ns<REV>(s);
    constexpr int MODE = tw_mode<TT, NS>();
    constexpr int NL = MODE == TW_SAME ? 1 : MODE == TW_EACH ? P / R : 0;
    static_for<0, NL>([&](auto BI) {
      constexpr int b = decltype(BI)::value;
      const int m = ((geo.template t<s>() + b * TT) % NS) * (N / (NS * R));
      pwr[s * P + b] = twr[m];
      pwi[s * P + b] = INV ? -twi[m] : twi[m];
    });
  });
  static_for<0, PL::NST>([&](auto SI) {
    constexpr int s = decltype(SI)::value;
    constexpr int R = PL::template radix<REV>(s);
    constexpr int NS = PL::template ns<REV>(s);
    constexpr int NB = P / R;
    constexpr int MODE = tw_mode<TT, NS>();
    const int t = geo.template t<s>();
    if constexpr (s > 0) {
      constexpr int RP = PL::template radix<REV>(s - 1);
      constexpr int NSP = PL::template ns<REV>(s - 1);
      constexpr int NBP = P / RP;
      const int tp = geo.template t<s - 1>();
      // one plane (or both) through LDS: write stage s-1's outputs, read
      // stage s's inputs
      auto exchange = [&](auto wr_c, auto wi_c) {
        constexpr bool WRE = decltype(wr_c)::value;
        constexpr bool WIM = decltype(wi_c)::value;
        __syncthreads();
        // Positions are a per-thread part plus a compile-time part; the
        // index maps add the latter as a constant offset where their padding
        // allows (widx_c / ridx_c), so most accesses need no address math.
        // Writes: with TT a multiple of NSP, position (j/NSP)*NSP*RP + j%NSP
        // + q*NSP of j = tp + b*TT is W(tp) + b*TT*RP + q*NSP, W(tp) a
        // multiple of RP when NSP = 1.
        static_for<0, NBP>([&](auto BI) {
          constexpr int b = decltype(BI)::value;
          static_for<0, RP>([&](auto QI) {
            constexpr int q = decltype(QI)::value;
            int li;
            if constexpr (TT % NSP == 0) {
              const int w = (tp / NSP) * NSP * RP + (tp % NSP);
              li = geo.template widx_c<s, b * TT * RP + q * NSP, (NSP == 1 ? RP : 1)>(w);
            } else {
              const int j = tp + b * TT;
              li = geo.template widx<s>((j / NSP) * NSP * RP + (j % NSP) + q * NSP);
            }
            if constexpr (WRE) lds_re[li] = vr[b * RP + q];
            if constexpr (WIM) (PLANES == 1 ? lds_re : lds_im)[li] = vi[b * RP + q];
          });
        });
        __syncthreads();
        // reads: position t + b*TT + r*(N/R)
        static_for<0, NB>([&](auto BI) {
          constexpr int b = decltype(BI)::value;
          static_for<0, R>([&](auto RI) {
            constexpr int r = decltype(RI)::value;
            const int li = geo.template ridx_c<s, b * TT + r * (N / R), 1>(t);
            if constexpr (WRE) vr[b * R + r] = lds_re[li];
            if constexpr (WIM) vi[b * R + r] = (PLANES == 1 ? lds_re : lds_im)[li];
          });
        });
      };
      if constexpr (PLANES == 2) {
        exchange(std::true_type{}, std::true_type{});
      } else {
        exchange(std::true_type{}, std::false_type{});
        exchange(std::false_type{}, std::true_type{});
      }
      geo.template after_read<s>(vr, vi);
    }
    // powers W^r (r < R) of a base twiddle
    auto powers = [&](T (&wr)[R], T (&wi)[R]) {
      static_for<2, R>([&](auto RI) {
        constexpr int r = decltype(RI)::value;
        constexpr int a = (r % 2 == 0) ? r / 2 : r - 1;
        constexpr int c = (r % 2 == 0) ? r / 2 : 1;
        wr[r] = wr[a] * wr[c] - wi[a] * wi[c];
        wi[r] = wr[a] * wi[c] + wi[a] * wr[c];
      });
    };
    T swr[R], swi[R];  // TW_SAME: shared by the stage's groups
    if constexpr (MODE == TW_SAME) {
      swr[1] = pwr[s * P];
      swi[1] = pwi[s * P];
      powers(swr, swi);
    }
    static_for<0, NB>([&](auto BI) {
      constexpr int b = decltype(BI)::value;
      if constexpr (MODE != TW_NONE) {
        T wr[R], wi[R];
        if constexpr (MODE == TW_SAME) {
          static_for<1, R>([&](auto RI) {
            constexpr int r = decltype(RI)::value;
            wr[r] = swr[r];
            wi[r] = swi[r];
          });
        } else {
          wr[1] = pwr[s * P + b];
          wi[1] = pwi[s * P + b];
          powers(wr, wi);
        }
        static_for<1, R>([&](auto RI) {
          constexpr int r = decltype(RI)::value;
          const T xr = vr[b * R + r], xi = vi[b * R + r];
          vr[b * R + r] = xr * wr[r] - xi * wi[r];
          vi[b * R + r] = xr * wi[r] + xi * wr[r];
        });
      }
      dft<R, INV>(vr + b * R, vi + b * R);
    });
  });
}

// Fixed geometry: one line, one position index for every stage.
template <class LIdx>
struct FixedGeo {
  int tt;
  LIdx lidx;
  template <int S>
  __device__ __forceinline__ int t() const { return tt; }
  template <int S>
  __device__ __forceinline__ int widx(int pos) const { return lidx(pos); }
  template <int S>
  __device__ __forceinline__ int ridx(int pos) const { return lidx(pos); }
  // index of position rt + CC, rt a multiple of M
  template <int S, int CC, int M>
  __device__ __forceinline__ int widx_c(int rt) const { return lidx.template at<CC, M>(rt); }
  template <int S, int CC, int M>
  __device__ __forceinline__ int ridx_c(int rt) const { return lidx.template at<CC, M>(rt); }
  template <int S, typename T, int P>
  __device__ __forceinline__ void after_read(T (&)[P], T (&)[P]) const {}
};

template <typename T, int N, int P, bool INV, bool REV, int PLANES = 2, class LIdx>
__device__ __forceinline__ void fft_line(T (&vr)[P], T (&vi)[P], int t, T* lds_re, T* lds_im,
                                         const LIdx& lidx, const T* __restrict__ twr,
                                         const T* __restrict__ twi) {
  FixedGeo<LIdx> geo{t, lidx};
  fft_line_g<T, N, P, INV, REV, PLANES>(vr, vi, geo, lds_re, lds_im, twr, twi);
}

// ---------------------------------------------------------------------------
// Launch geometry (the row pass geometry lives with the row kernel).
// Columns: C adjacent columns per workgroup, whole line resident; the tile
// is sized so C*N complex values fit the LDS budget.
// ---------------------------------------------------------------------------
template <int N>
struct PChoice {
  static constexpr int P = N < 16 ? N : 16;
};

// Column tile geometry: P points per thread, up to CMAX adjacent columns,
// PLANES LDS buffers (1: re and im exchanged one after the other).  The tile
// holds as many columns as fit a 128 KB LDS budget and 1024 threads.
template <typename T, int N, int P_, int CMAX, int PLANES_>
struct ColGeom {
  static constexpr int P = P_ < N ? P_ : N;
  static constexpr int T_LINE = N / P;
  static constexpr int PLANES = PLANES_;
  static constexpr int BUDGET = int(131072 / sizeof(T)) / PLANES;  // values per LDS buffer
  static constexpr int C = cmax(1, cmin(CMAX, cmin(BUDGET / N, 1024 / T_LINE)));
  static constexpr int WG = C * T_LINE;
  static constexpr bool NEED_LDS = LinePlan<N, P>::NST > 1;
  static constexpr int LDS_ELEMS = !NEED_LDS ? 1 : C % 32 == 0 ? C * N : C * N + (C * N) / 32 + 1;
  static constexpr int LDS_IM = PLANES == 2 ? LDS_ELEMS : 1;
};

// Default column geometry.  FP32 lines of 512 points and more, for passes
// whose hooks are light (WIDE_OK): 32 points per thread and one LDS plane at
// a time, 32 columns (128-byte row segments) per 512-thread workgroup, two
// workgroups per CU (measured on 16 x 2048^2 at N = 512: 225 us against
// 231 us for 64-column tiles at one workgroup per CU and 240 us for 16
// points per thread with both planes).  Heavier hooks would spill at 32
// points per thread.
// FP64 lines of 512 points and more, any hooks: 16 points per thread, one
// plane, up to 16 columns (128-byte row segments; the two-plane tile held 4
// columns of a 2048-point line).
template <typename T, int N, bool WIDE_OK = false>
struct ColCfgSel {
  static constexpr bool WIDE = WIDE_OK && sizeof(T) == 4 && N >= 512;
  static constexpr bool WIDE64 = sizeof(T) == 8 && N >= 512;
  using type = typename std::conditional<
      WIDE, ColGeom<T, N, 32, 32, 1>,
      typename std::conditional<WIDE64, ColGeom<T, N, 16, 16, 1>,
                                ColGeom<T, N, PChoice<N>::P, 64, 2>>::type>::type;
};
template <typename T, int N, bool WIDE_OK = false>
using ColCfg = typename ColCfgSel<T, N, WIDE_OK>::type;

// LDS index maps.  at<CC, M>(rt) is the index of position rt + CC for a
// per-thread rt that is a multiple of M: a padded map p + (p >> k) splits
// into map(rt) + map(CC) when either part is a multiple of 2^k.
struct RowLIdx {
  int base;
  __device__ __forceinline__ int operator()(int pos) const { return base + pos + (pos >> 4); }
  template <int CC, int M>
  __device__ __forceinline__ int at(int rt) const {
    constexpr bool SPLIT = CC % 16 == 0 || M % 16 == 0;
    if constexpr (SPLIT)
      return base + (rt + (rt >> 4)) + (CC + (CC >> 4));
    else
      return (*this)(rt + CC);
  }
};
// Column tiles of 32 or more columns need no padding: a half-wave's 32
// lanes hold 32 adjacent columns of one line position, one per bank, so
// every register's LDS address is a per-thread base plus a constant.
// Narrower tiles pad one element per 32 against bank conflicts.
template <int C>
struct ColLIdx {
  int col;
  __device__ __forceinline__ int operator()(int pos) const {
    const int a = pos * C + col;
    if constexpr (C % 32 == 0)
      return a;
    else
      return a + (a >> 5);
  }
  template <int CC, int M>
  __device__ __forceinline__ int at(int rt) const {
    // a = rt*C + col + CC*C: the padding splits when CC*C is a multiple of 32
    constexpr int A = CC * C;
    if constexpr (C % 32 == 0) {
      return rt * C + col + A;
    } else if constexpr (A % 32 == 0) {
      const int a = rt * C + col;
      return a + (a >> 5) + (A + (A >> 5));
    } else {
      return (*this)(rt + CC);
    }
  }
};

// Remap a linear workgroup id so that consecutive logical tiles land on the
// same XCD (hardware deals workgroups round-robin over the 8 XCDs), which
// lets neighbouring column tiles share L2 lines.
__device__ __forceinline__ int xcd_remap(int b, int g) {
  if ((g & 7) != 0) return b;
  return (b & 7) * (g >> 3) + (b >> 3);
}

}  // namespace doe
