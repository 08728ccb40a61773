// fft2d: one row pass then one column pass, in place (ref: src/fft.js:148-154).
#include "ctx.hpp"
#include "ops_prop.hpp"

#include <algorithm>

namespace doe {

template <typename T, typename S>
int run_fft2d(doe_ctx* ctx, doe_field* f, int inverse) {
  Plan2 pl;
  int rc = make_plan2(ctx, f->nx, f->ny, sizeof(T), col_split<T>(f->nx, f->ny), &pl);
  if (rc) return rc;
  const size_t plane = size_t(f->nx) * size_t(f->ny);
  const int cf = chunk_fields(ctx, plane * 2 * sizeof(S), f->batch);
  // A split column pass reads rows in one order and writes them in another,
  // so it cannot run in place: the two passes ping-pong through work planes
  // of one chunk (same HBM traffic as in place; the chunk's work planes are
  // re-read from the memory-side cache).
  S* wre = nullptr;
  S* wim = nullptr;
  if (pl.n1 > 1) {
    rc = ensure_work(ctx, size_t(cf) * plane * sizeof(S));
    if (rc) return rc;
    wre = (S*)ctx->work_re;
    wim = (S*)ctx->work_im;
  }
  const OpNone op{};
  for (int f0 = 0; f0 < f->batch; f0 += cf) {
    const int nb = std::min(cf, f->batch - f0);
    S* re = (S*)f->re + size_t(f0) * plane;
    S* im = (S*)f->im + size_t(f0) * plane;
    S* xre = pl.n1 > 1 ? wre : re;
    S* xim = pl.n1 > 1 ? wim : im;
    if (!inverse) {
      // rows (+ the column split's first steps), then the column blocks
      DOE_HIP((row_pass<T, S, 0, false, false, false, true>(ctx, "fft_row_fwd", op, re, im, xre,
                                                             xim, nb, pl, AllSizes{})));
      DOE_HIP((col_pass<T, S, 0, false, false>(ctx, "fft_col_fwd", op, xre, xim, re, im, nb, pl,
                                               AllSizes{})));
    } else {
      DOE_HIP((col_pass<T, S, 0, true, false>(ctx, "fft_col_inv", op, re, im, xre, xim, nb, pl,
                                              AllSizes{})));
      DOE_HIP((row_pass<T, S, 0, true, false, true, false>(ctx, "fft_row_inv", op, xre, xim, re,
                                                            im, nb, pl, AllSizes{})));
    }
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
