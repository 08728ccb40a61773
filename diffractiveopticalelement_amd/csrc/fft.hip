// fft2d: one row pass then one column pass, in place (ref: src/fft.js:148-154).
#include "ctx.hpp"
#include "ops_prop.hpp"

namespace doe {

template <typename T, typename S>
int run_fft2d(doe_ctx* ctx, doe_field* f, int inverse) {
  Twiddle twx, twy;
  int rc = get_twiddle(ctx, f->nx, sizeof(T), &twx);
  if (rc) return rc;
  rc = get_twiddle(ctx, f->ny, sizeof(T), &twy);
  if (rc) return rc;
  S* re = (S*)f->re;
  S* im = (S*)f->im;
  hipError_t e = hipSuccess;
  const OpNone op{};
  with_n(f->nx, [&](auto NN) {
    constexpr int N = decltype(NN)::value;
    if (inverse)
      e = launch_row<T, S, N, 0, true, false>(ctx, "fft_row_inv", op, re, im, re, im, f->ny, f->batch, twx);
    else
      e = launch_row<T, S, N, 0, false, false>(ctx, "fft_row_fwd", op, re, im, re, im, f->ny, f->batch, twx);
  });
  DOE_HIP(e);
  with_n(f->ny, [&](auto NN) {
    constexpr int N = decltype(NN)::value;
    if (inverse)
      e = launch_col<T, S, N, 0, true, false>(ctx, "fft_col_inv", op, re, im, re, im, f->nx, f->batch, twy);
    else
      e = launch_col<T, S, N, 0, false, false>(ctx, "fft_col_fwd", op, re, im, re, im, f->nx, f->batch, twy);
  });
  DOE_HIP(e);
  return DOE_OK;
}

// Row pass only: a batch of independent 1-D transforms along x.
template <typename T, typename S>
int run_fft_rows(doe_ctx* ctx, doe_field* f, int inverse) {
  Twiddle twx;
  int rc = get_twiddle(ctx, f->nx, sizeof(T), &twx);
  if (rc) return rc;
  S* re = (S*)f->re;
  S* im = (S*)f->im;
  hipError_t e = hipSuccess;
  const OpNone op{};
  with_n(f->nx, [&](auto NN) {
    constexpr int N = decltype(NN)::value;
    if (inverse)
      e = launch_row<T, S, N, 0, true, false>(ctx, "fft_row_inv", op, re, im, re, im, f->ny, f->batch, twx);
    else
      e = launch_row<T, S, N, 0, false, false>(ctx, "fft_row_fwd", op, re, im, re, im, f->ny, f->batch, twx);
  });
  DOE_HIP(e);
  return DOE_OK;
}

// One (compute, storage) pair per object file, selected by DOE_PART.
#define DOE_FFT_INST(T, S)                                  \
  template int run_fft2d<T, S>(doe_ctx*, doe_field*, int); \
  template int run_fft_rows<T, S>(doe_ctx*, doe_field*, int);
#if DOE_PART == 1
DOE_FFT_INST(float, float)
#elif DOE_PART == 2
DOE_FFT_INST(double, double)
#elif DOE_PART == 3
DOE_FFT_INST(float, __half)
#elif DOE_PART == 4
DOE_FFT_INST(float, __hip_bfloat16)
#endif

}  // namespace doe
