// Shared device/host helpers for the DOE HIP kernels (gfx950).
//
// Complex fields are planar: one array of real parts and one of imaginary
// parts, row-major with idx = x + y*nx, exactly the layout of the reference's
// ComplexField (src/field.js:22-40).  A "batch" stacks fields back to back.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <type_traits>

namespace doe {

// ---------------------------------------------------------------------------
// Storage <-> compute conversion.  Storage S is what lives in HBM (f32, f64,
// f16, bf16); compute T is f32 or f64.
// ---------------------------------------------------------------------------
template <typename T, typename S>
__device__ __forceinline__ T ld(const S* p) {
  if constexpr (std::is_same_v<S, __half>) {
    return T(__half2float(*p));
  } else if constexpr (std::is_same_v<S, __hip_bfloat16>) {
    return T(__bfloat162float(*p));
  } else {
    return T(*p);
  }
}

template <typename T, typename S>
__device__ __forceinline__ void st(S* p, T v) {
  if constexpr (std::is_same_v<S, __half>) {
    *p = __float2half(float(v));
  } else if constexpr (std::is_same_v<S, __hip_bfloat16>) {
    *p = __float2bfloat16(float(v));
  } else {
    *p = S(v);
  }
}

// ---------------------------------------------------------------------------
// Compile-time loop: static_for<B, E>(f) calls f(integral_constant<int, i>)
// for i in [B, E).  Lets the radix kernels branch on constant indices with
// `if constexpr`, so multiplications by 1 and -i vanish at compile time.
// ---------------------------------------------------------------------------
template <int I, int E>
struct StaticFor {
  template <class F>
  __device__ __forceinline__ static void run(F&& f) {
    if constexpr (I < E) {
      f(std::integral_constant<int, I>{});
      StaticFor<I + 1, E>::run(f);
    }
  }
};
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  StaticFor<B, E>::run(f);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }
constexpr int bitrev(int v, int bits) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r |= ((v >> i) & 1) << (bits - 1 - i);
  return r;
}
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// cos(2*pi*m/16), sin(2*pi*m/16) to 20 significant digits.
constexpr double c16(int m) {
  switch (m & 15) {
    case 0: return 1.0;
    case 1: return 0.92387953251128675613;
    case 2: return 0.70710678118654752440;
    case 3: return 0.38268343236508977173;
    case 4: return 0.0;
    case 5: return -0.38268343236508977173;
    case 6: return -0.70710678118654752440;
    case 7: return -0.92387953251128675613;
    case 8: return -1.0;
    case 9: return -0.92387953251128675613;
    case 10: return -0.70710678118654752440;
    case 11: return -0.38268343236508977173;
    case 12: return 0.0;
    case 13: return 0.38268343236508977173;
    case 14: return 0.70710678118654752440;
    default: return 0.92387953251128675613;
  }
}
constexpr double s16(int m) { return c16(m - 4); }  // sin(x) = cos(x - pi/2)

// cos(k*pi/16), k = 0..8
constexpr double cos_pi16(int k) {
  switch (k) {
    case 0: return 1.0;
    case 1: return 0.98078528040323044913;
    case 2: return 0.92387953251128675613;
    case 3: return 0.83146961230254523708;
    case 4: return 0.70710678118654752440;
    case 5: return 0.55557023301960222474;
    case 6: return 0.38268343236508977173;
    case 7: return 0.19509032201612826785;
    default: return 0.0;
  }
}
// cos(2*pi*m/32) and sin(2*pi*m/32) for any integer m
constexpr double c32(int m) {
  m &= 31;
  if (m <= 8) return cos_pi16(m);
  if (m <= 16) return -cos_pi16(16 - m);
  if (m <= 24) return -cos_pi16(m - 16);
  return cos_pi16(32 - m);
}
constexpr double s32(int m) { return c32(m - 8); }

}  // namespace doe
