// fft2d: one row pass then one column pass, in place (ref: src/fft.js:148-154).
#include "ctx.hpp"
#include "ops_prop.hpp"

#include <algorithm>
#include <string>
#include <type_traits>

namespace doe {

// The two passes over one chunk, in a row-plane / column-tile variant.
template <typename T, typename S, int RPL, int CV>
static hipError_t fft2d_chunk(doe_ctx* ctx, const Plan2& pl, S* re, S* im, S* xre, S* xim, int nb,
                              int inverse) {
  const OpNone op{};
  using NF = typename std::conditional<RPL == 2 && CV == 0, AllSizes, LargeSizes>::type;
  hipError_t e;
  if (!inverse) {
    // rows (+ the column split's first steps), then the column blocks
    e = row_pass<T, S, 0, false, false, false, true, OpNone, RPL>(ctx, "fft_row_fwd", op, re, im,
                                                                    xre, xim, nb, pl, NF{});
    if (e == hipSuccess)
      e = col_pass<T, S, 0, false, false, OpNone, CV>(ctx, "fft_col_fwd", op, xre, xim, re, im, nb,
                                                      pl, NF{});
  } else {
    e = col_pass<T, S, 0, true, false, OpNone, CV>(ctx, "fft_col_inv", op, re, im, xre, xim, nb, pl,
                                                   NF{});
    if (e == hipSuccess)
      e = row_pass<T, S, 0, true, false, true, false, OpNone, RPL>(ctx, "fft_row_inv", op, xre, xim,
                                                                     re, im, nb, pl, NF{});
  }
  return e;
}

// Row and column tile variants (six column tiles for FP32 arithmetic, two
// for FP64).
template <typename T, typename S>
static hipError_t fft2d_chunk_var(doe_ctx* ctx, const doe_tune& tu, const Plan2& pl, S* re, S* im,
                                  S* xre, S* xim, int nb, int inverse) {
  if constexpr (std::is_same<T, float>::value) {
    const int key = (tu.rowplanes == 1 ? 10 : 0) + tu.colvar;
    switch (key) {
      case 0: return fft2d_chunk<T, S, 2, 0>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 1: return fft2d_chunk<T, S, 2, 1>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 2: return fft2d_chunk<T, S, 2, 2>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 3: return fft2d_chunk<T, S, 2, 3>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 4: return fft2d_chunk<T, S, 2, 4>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 5: return fft2d_chunk<T, S, 2, 5>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 10: return fft2d_chunk<T, S, 1, 0>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 11: return fft2d_chunk<T, S, 1, 1>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 12: return fft2d_chunk<T, S, 1, 2>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 13: return fft2d_chunk<T, S, 1, 3>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 14: return fft2d_chunk<T, S, 1, 4>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 15: return fft2d_chunk<T, S, 1, 5>(ctx, pl, re, im, xre, xim, nb, inverse);
      default: return hipErrorInvalidValue;
    }
  } else {
    // FP64: the built-in tiles and the two-plane column tile, with either row
    // pass
    const int key = (tu.rowplanes == 1 ? 10 : 0) + tu.colvar;
    switch (key) {
      case 0: return fft2d_chunk<T, S, 2, 0>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 1: return fft2d_chunk<T, S, 2, 1>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 10: return fft2d_chunk<T, S, 1, 0>(ctx, pl, re, im, xre, xim, nb, inverse);
      case 11: return fft2d_chunk<T, S, 1, 1>(ctx, pl, re, im, xre, xim, nb, inverse);
      default: return hipErrorInvalidValue;
    }
  }
}

template <typename T, typename S>
int run_fft2d(doe_ctx* ctx, doe_field* f, int inverse) {
  const doe_tune tu = tune_for(ctx, f->nx, f->ny, f->batch, f->dtype);
  Plan2 pl;
  const int n1 = tu.n1;
  if (!split_fits(f->nx, f->ny, n1, int(sizeof(T)))) {
    set_error("fft2d: column split " + std::to_string(n1) + " does not fit " +
              std::to_string(f->nx) + "x" + std::to_string(f->ny));
    return DOE_ERR_UNSUPPORTED;
  }
  int rc = make_plan2(ctx, f->nx, f->ny, sizeof(T), n1, &pl);
  if (rc) return rc;
  const size_t plane = size_t(f->nx) * size_t(f->ny);
  const int cf = chunk_fields(tu.chunk >= 0 ? tu.chunk : ctx->chunk_bytes, plane * 2 * sizeof(S),
                              f->batch);
  // A split column pass reads rows in one order and writes them in another,
  // so it cannot run in place: the two passes ping-pong through work planes
  // of one chunk (same HBM traffic as in place).
  S* wre = nullptr;
  S* wim = nullptr;
  if (pl.n1 > 1) {
    rc = ensure_work(ctx, size_t(cf) * plane * sizeof(S));
    if (rc) return rc;
    wre = (S*)ctx->work_re;
    wim = (S*)ctx->work_im;
  }
  for (int f0 = 0; f0 < f->batch; f0 += cf) {
    const int nb = std::min(cf, f->batch - f0);
    S* re = (S*)f->re + size_t(f0) * plane;
    S* im = (S*)f->im + size_t(f0) * plane;
    S* xre = pl.n1 > 1 ? wre : re;
    S* xim = pl.n1 > 1 ? wim : im;
    DOE_HIP((fft2d_chunk_var<T, S>(ctx, tu, pl, re, im, xre, xim, nb, inverse)));
  }
  return DOE_OK;
}

// Row pass only: a batch of independent 1-D transforms along x.
template <typename T, typename S>
int run_fft_rows(doe_ctx* ctx, doe_field* f, int inverse) {
  Plan2 pl;
  int rc = make_plan2(ctx, f->nx, f->ny, sizeof(T), 1, &pl);
  if (rc) return rc;
  S* re = (S*)f->re;
  S* im = (S*)f->im;
  const OpNone op{};
  hipError_t e;
  if (inverse)
    e = row_pass<T, S, 0, true, false, false, false>(ctx, "fft_row_inv", op, re, im, re, im,
                                                      f->batch, pl, AllSizes{});
  else
    e = row_pass<T, S, 0, false, false, false, false>(ctx, "fft_row_fwd", op, re, im, re, im,
                                                       f->batch, pl, AllSizes{});
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
