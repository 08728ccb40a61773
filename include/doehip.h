/*
 * doehip.h — C ABI of the MI355X (gfx950) FFT / free-space propagation /
 * phase-retrieval library (libdoehip.so).
 *
 * This is the native boundary behind the diffractive-optical-element toolkit's
 * JavaScript entry points.  Every function returns an int status (DOE_OK on
 * success) and never throws; doe_last_error() gives the thread's last message.
 * Device memory is owned by the library (or by the caller for wrapped fields);
 * host pointers are borrowed only for the duration of a call.  Calls are
 * synchronous with respect to the host unless stated otherwise: when they
 * return, output buffers (host or device) hold valid results.  A context owns
 * one HIP stream and is not re-entrant.
 *
 * Data layout: a field is a batch of `batch` complex nx-by-ny grids stored as
 * two planar arrays (real, imaginary), row-major with idx = x + y*nx, fields
 * back to back — the reference's ComplexField layout (src/field.js:22-40).
 * Transform grids have power-of-two sides in [2, 8192]; storage, targets,
 * metrics and images take any grid of 1..65536 rows and up to 2^32 points.
 *
 * Citations "ref:" point into the reference toolkit's sources.
 */
#ifndef DOEHIP_H
#define DOEHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define DOE_OK 0
#define DOE_ERR_INVALID 1     /* bad argument: non power-of-two, null, ... */
#define DOE_ERR_HIP 2         /* HIP runtime error */
#define DOE_ERR_OOM 3         /* device allocation failed */
#define DOE_ERR_UNSUPPORTED 4 /* size/dtype combination not built */
#define DOE_ERR_SHAPE 5       /* plane length does not match nx*ny */
#define DOE_ERR_STATE 6       /* call order (e.g. handle from another ctx) */

/* Storage precision of a field.  F32/F64 compute in their own precision;
 * F16/BF16 store half-width values and compute in FP32. */
#define DOE_F32 0
#define DOE_F64 1
#define DOE_F16 2
#define DOE_BF16 3

typedef struct doe_ctx doe_ctx;
typedef struct doe_field doe_field;

/* ---- library / context ------------------------------------------------- */
int doe_version(void);
const char* doe_last_error(void);
const char* doe_status_string(int status);
int doe_device_count(int* count);
/* stream: an existing hipStream_t to launch on, or NULL for a private one.
 * Batches (doe_gs_batch) and networks (doe_wsnet) keep their context: destroy
 * them before it. */
int doe_ctx_create(int device, void* stream, doe_ctx** out);
int doe_ctx_destroy(doe_ctx* ctx);
int doe_ctx_sync(doe_ctx* ctx);
int doe_ctx_stream(doe_ctx* ctx, void** stream);

/* Measurement plumbing (no counterpart in the reference, which times with
 * wall clocks only, tools/bench.mjs:39-47).  DOE_CTX_ASYNC: entry points on
 * device-resident fields return without synchronising the stream.
 * DOE_CTX_PROFILE: every kernel launch is timed on the context stream (the
 * row and column passes by start / stop events carried on their own
 * dispatches, other launches by events recorded around them; launches into
 * a stream under capture are not timed); doe_ctx_profile_report
 * synchronises and writes one line
 * per kernel tag, "tag count total_ms total_points" (grid points summed over
 * the launches), then clears the records.
 * doe_timer_start/stop bracket a region with two events on the stream. */
#define DOE_CTX_ASYNC 1
#define DOE_CTX_PROFILE 2
int doe_ctx_set_flags(doe_ctx* ctx, int flags);
int doe_ctx_profile_report(doe_ctx* ctx, char* buf, long cap);

/* Batch chunking (no counterpart in the reference, whose transforms take one
 * field at a time).  Batched FFT / propagation calls run chunk by chunk, a
 * chunk's planes (re + im) being at most `bytes`, so the passes of one chunk
 * can re-read it from the GPU's memory-side cache.  Default 0 = whole batch
 * in one launch per pass (environment DOE_CHUNK_MIB overrides it at context
 * creation). */
int doe_ctx_set_chunk_bytes(doe_ctx* ctx, long bytes);
int doe_ctx_get_chunk_bytes(doe_ctx* ctx, long* bytes);
int doe_timer_start(doe_ctx* ctx);
int doe_timer_stop(doe_ctx* ctx, double* ms);

/* ---- device-resident planar fields (ref: src/field.js:22-196) --------- *
 * Any grid of 1..65536 rows and at most 2^32 points can be stored,
 * uploaded, downloaded, copied, resampled, measured and imaged; transforms,
 * propagators and phase retrieval need power-of-two sides in [2, 8192]
 * (DOE_ERR_UNSUPPORTED). */
int doe_field_create(doe_ctx* ctx, int nx, int ny, int batch, int dtype, doe_field** out);
/* Non-owning view of caller device memory (re, im: batch*nx*ny elements). */
int doe_field_wrap(doe_ctx* ctx, void* re, void* im, int nx, int ny, int batch, int dtype,
                   doe_field** out);
int doe_field_free(doe_field* f);
int doe_field_info(const doe_field* f, int* nx, int* ny, int* batch, int* dtype, void** re,
                   void** im);
/* Host Float64 planes (batch*nx*ny each) <-> device storage (converted). */
int doe_field_upload(doe_ctx* ctx, doe_field* f, const double* re, const double* im);
int doe_field_download(doe_ctx* ctx, doe_field* f, double* re, double* im);
int doe_field_copy(doe_ctx* ctx, const doe_field* src, doe_field* dst);

/* ---- FFT (ref: src/fft.js:148-154 fft2d) ------------------------------- *
 * In place on every grid of the batch.  Forward is the unscaled DFT with
 * exp(-2*pi*i*n*k/N); inverse uses exp(+...) and scales by 1/(nx*ny). */
int doe_fft2d(doe_ctx* ctx, doe_field* f, int inverse);

/* fft2d tuning (no counterpart in the reference).  A context keeps one
 * configuration per (nx, ny, batch, dtype): the column split n1 (1, 2 or 4;
 * 0 = built-in), the column tile variant colvar (0 built-in; 1: 16 points x
 * 64 columns, both LDS planes; 2: 32 points x 64 columns, one plane; 3: 32
 * points x 32 columns, one plane; 4: 32 points x 16 columns, one plane; 5: 16
 * points x 32 columns, one plane), the LDS planes of the row passes (1 or 2;
 * 0 = built-in) and the chunk size in bytes (-1 = the context's).
 * doe_tune_get reports the configuration that runs, built-in choices
 * resolved.  A split the shape cannot take fails the transform with
 * DOE_ERR_UNSUPPORTED.  colvar and rowplanes
 * apply to FP32 arithmetic (FP32, FP16 and BF16 storage); FP64 fields take
 * colvar 0 or 1.  Untuned FP16 / BF16 shapes with both sides 2048 or more
 * default to n1 = 1 and one-plane rows; FP64 columns of 1024..8192 points
 * split in two (one-plane rows from 1024-point rows).  Variants are built for sides
 * 256..8192.  doe_autotune_fft2d times every configuration that exists for
 * the shape (the better of two windows of `reps` forward + inverse pairs, on
 * a scratch field of uniform noise, after 16 untimed pairs that bring the
 * clock up), keeps the fastest and writes one line per candidate
 * ("n1=.. colvar=.. rowplanes=.. chunk=.. ms=..", ms per transform) and a
 * final "best ..." line into `report`. */
int doe_tune_set(doe_ctx* ctx, int nx, int ny, int batch, int dtype, int n1, int colvar,
                 int rowplanes, long chunk_bytes);
int doe_tune_get(doe_ctx* ctx, int nx, int ny, int batch, int dtype, int* n1, int* colvar,
                 int* rowplanes, long* chunk_bytes);
int doe_autotune_fft2d(doe_ctx* ctx, int nx, int ny, int batch, int dtype, int reps, char* report,
                       long cap);

/* Host drop-in of fft2d(re, im, nx, ny, inverse) (ref: src/fft.js:148):
 * uploads the Float64 planes, transforms in `dtype`, downloads in place.
 * Returns DOE_ERR_SHAPE if len != nx*ny (the reference's length error). */
int doe_fft2d_host(doe_ctx* ctx, double* re, double* im, long len, int nx, int ny, int inverse,
                   int dtype);

/* Host drop-in of fft1d(re, im, n, stride, offset, inverse) (ref:
 * src/fft.js:86-139): one strided line of `n` points of planes of length
 * `len`; n <= 1 is a no-op, a non power of two is DOE_ERR_INVALID with the
 * reference's message. */
int doe_fft1d_host(doe_ctx* ctx, double* re, double* im, long len, long n, long stride,
                   long offset, int inverse, int dtype);

/* ---- angular spectrum (ref: src/propagate.js:57-133) ------------------- *
 * out = ASM_z(in) for each grid b with distance z[b] (in may equal out).
 * band_limit: Matsushima–Shimobaba local limit with raised-cosine edge of
 * relative width soft_edge; evanescent: 1 zeroes evanescent waves, 0 keeps
 * them as a pure phase (ref: propagate.js:88-95).  z[b] == 0 copies. */
int doe_asm(doe_ctx* ctx, const doe_field* in, doe_field* out, double dx, double dy,
            double lambda, const double* z, int band_limit, int evanescent, double soft_edge);
/* Many propagations of uneven sizes (no counterpart in the reference, which
 * propagates one ComplexField per call, propagate.js:105-133): count batches
 * in[k] (FP32 square fields of sides 64 ... 2048, each batch one side) to
 * out[k] (same shapes; may equal in[k]), batch k with pitch dx[k] x dy[k]
 * and distances z[k][0 .. batch-1], as doe_asm per batch (z = 0: H = 1).
 * Each of the three passes is one launch for up to 6 batches, driven by a
 * table of the batches' sides, planes and transfer-function parameters. */
int doe_asm_multi(doe_ctx* ctx, int count, const doe_field* const* in, doe_field* const* out,
                  const double* dx, const double* dy, double lambda, const double* const* z,
                  int band_limit, int evanescent, double soft_edge);
/* Transfer function H on the centred grid (ref: propagate.js:57-102) into a
 * batch-1 field; also returns the propagating fraction. */
int doe_asm_transfer(doe_ctx* ctx, doe_field* out, double dx, double dy, double lambda, double z,
                     int band_limit, int evanescent, double soft_edge,
                     double* propagating_fraction);

/* ---- single-FFT Fresnel (ref: src/propagate.js:155-316) ---------------- *
 * inverse = 0: forward propagation to the magnified grid; dx, dy are the
 * input pitch and out_dx/out_dy (nullable) receive lambda*|z|/(N*dx).
 * inverse = 1: exact inverse; dx, dy are the *output* (element) pitch and
 * in_dx/in_dy the magnified pitch of the input (ref: propagate.js:246-250). */
int doe_fresnel(doe_ctx* ctx, const doe_field* in, doe_field* out, double dx, double dy,
                double lambda, const double* z, int inverse, double in_dx, double in_dy,
                double* out_dx, double* out_dy);

/* ---- lens focal-plane transform (ref: src/propagate.js:342-374) -------- */
int doe_lens(doe_ctx* ctx, const doe_field* in, doe_field* out, double dx, double dy,
             double lambda, const double* f, double* out_dx, double* out_dy);

/* ---- phase retrieval (ref: src/doe.js:327-546, 639-730) ---------------- */
#define DOE_MODE_IMAGE_PLANE 0 /* angular-spectrum loop, gerchbergSaxton */
#define DOE_MODE_FAR_FIELD 1   /* unitary DFT loop, fourierCgh */
#define DOE_ALG_GS 0
#define DOE_ALG_WGS 1 /* weightedGerchbergSaxton wrapping the mode's base loop */

typedef struct doe_gs_plan {
  int mode;       /* DOE_MODE_* */
  int algorithm;  /* DOE_ALG_* */
  int iterations;
  int nx, ny;
  double dx, dy, lambda, z;
  double beta;     /* WGS exponent (ref default 0.9) */
  double feedback; /* HIO feedback, image-plane GS only (ref: doe.js:380-388) */
  int escape;      /* iterations before feedback starts */
  int dtype;       /* DOE_F32 or DOE_F64 compute */
} doe_gs_plan;

/* Host-array phase retrieval.  All planes are nx*ny Float64.  roi may be NULL
 * (whole plane); init_phase may be NULL (far-field: random from seed as the
 * reference's randomPhase; image-plane: flat).  Outputs (all nullable except
 * phase_out): phase (best phase under WGS), per-iteration error and
 * efficiency histories (iterations entries), amplitude and intensity of the
 * reported iterate, final weights, best rmse. */
int doe_gs_run(doe_ctx* ctx, const doe_gs_plan* plan, const double* illumination,
               const double* target_amp, const double* roi, const double* init_phase,
               unsigned seed, double* phase_out, double* error_out, double* efficiency_out,
               double* amplitude_out, double* intensity_out, double* weights_out,
               double* best_rmse_out);

/* Multi-plane phase retrieval (gerchbergSaxtonMultiPlane, src/doe.js:557-626):
 * one element that projects a different target at each distance.  Every
 * iteration visits the planes in order (angular spectrum to z, GS
 * projection on the plane's target inside its ROI, back to the element,
 * phase kept), bandLimit false, flat initial phase when init_phase is NULL.
 * plan: nx, ny, dx, dy, lambda, iterations and dtype (DOE_F32 / DOE_F64);
 * mode, algorithm, z, beta, feedback and escape are not used.  Outputs:
 * final phase; error[it] = sum_p weight_p * e_p / nplanes and
 * efficiency[it] = 1 - error[it] (as the reference reports them); per-plane
 * errors plane-major (nplanes x iterations).  Outputs other than phase are
 * nullable. */
typedef struct doe_mp_plane {
  double z;
  const double* target_amp; /* nx*ny */
  const double* roi;        /* nx*ny or NULL (whole plane) */
  double weight;            /* reference default 1 */
} doe_mp_plane;
int doe_gs_multiplane_run(doe_ctx* ctx, const doe_gs_plan* plan, const double* illumination,
                          int nplanes, const doe_mp_plane* planes, const double* init_phase,
                          double* phase_out, double* error_out, double* efficiency_out,
                          double* per_plane_error_out);

/* Batched far-field phase retrieval.  No single counterpart in the
 * reference, which designs one element per call: each design of a batch runs
 * fourierCgh (src/doe.js:639-730), under DOE_ALG_WGS with the weight update
 * and best-iterate tracking of weightedGerchbergSaxton (doe.js:455-546), with
 * the same results as doe_gs_run in DOE_MODE_FAR_FIELD / DOE_F32 on that
 * design alone.  A batch holds `count` square designs of power-of-two sides
 * 64 ... 4096, mixed freely, grouped into size classes.  Every iteration is
 * one column-pass, one finalize and one row-pass launch for the whole batch,
 * each workgroup finding its size class in a table of the classes (side,
 * planes, state, first workgroup) and running that side's pass body; with
 * DOE_GSB_PER_CLASS, or when a side exceeds 2048, it is one column-pass and
 * one row-pass launch per size class instead.  With DOE_GSB_GRAPH
 * the iteration loop is captured once as a hipGraph and replayed.  Inputs
 * per design (doe_gs_batch_set): illumination, target
 * amplitude, ROI (only if the batch was created with_roi; then required),
 * initial phase (NULL keeps the current one; doe_gs_batch_random_init draws
 * every design's phase on the device from a counter-based splitmix64 stream,
 * phi = (u*2 - 1)*pi in FP64).  doe_gs_batch_run restarts every design from
 * its initial phase.  Outputs per design as doe_gs_run: phase, error and
 * efficiency histories (iterations of the last run), amplitude of the
 * reported iterate, weights, best rmse (NaN for GS). */
typedef struct doe_gs_batch doe_gs_batch;
#define DOE_GSB_GRAPH 1
#define DOE_GSB_PER_CLASS 2
int doe_gs_batch_create(doe_ctx* ctx, int count, const int* sides, int algorithm, double beta,
                        int with_roi, doe_gs_batch** out);
int doe_gs_batch_destroy(doe_gs_batch* batch);
int doe_gs_batch_set(doe_gs_batch* batch, int index, const double* illumination,
                     const double* target_amp, const double* roi, const double* init_phase);
int doe_gs_batch_random_init(doe_gs_batch* batch, unsigned long long seed);
int doe_gs_batch_run(doe_gs_batch* batch, int iterations, int flags);
int doe_gs_batch_get(doe_gs_batch* batch, int index, double* phase_out, double* error_out,
                     double* efficiency_out, double* amplitude_out, double* weights_out,
                     double* best_rmse_out);
/* Per-iteration screen rmse of a design (iterations of the last run): the
 * measured far-field irradiance I of each iteration judged as
 * ProjectorSystem.screenMetrics does (projector.js:554-614): the target
 * irradiance T = t^2 power-matched to the plane's ROI power, s = sum(I*roi) /
 * sum(T), rmse = sqrt(sum_{roi>0} (I - sT)^2 / sum_{roi>0} (sT)^2). */
int doe_gs_batch_get_screen(doe_gs_batch* batch, int index, double* screen_rmse_out);

/* Learned warm start (no counterpart in the reference, whose designs start
 * from a random or flat phase: doe.js:651, 850-870).  A convolutional
 * encoder-decoder (3x3 convolutions, additive skips, ReLU; BF16 weights and
 * activations, FP32 accumulation on the matrix cores; 1.75 M parameters at
 * base width 32) maps a target irradiance pooled to res x res to a
 * target-plane phasor field (2 channels, NHWC).  All pointers below are
 * device pointers: input batch*res*res floats, output/label planes
 * batch*res*res (output interleaved re, im).  doe_wsnet_loss_grad runs the
 * forward pass, the phase-aligned, conjugate-aware weighted loss and the
 * backward pass (gradients kept for doe_wsnet_adam_step; copied to grad_host
 * if non-NULL); parameters are FP32 masters with BF16 copies. */
typedef struct doe_wsnet doe_wsnet;
int doe_wsnet_create(doe_ctx* ctx, int res, int base, int max_batch, unsigned long long seed,
                     doe_wsnet** out);
int doe_wsnet_destroy(doe_wsnet* net);
int doe_wsnet_info(doe_wsnet* net, int* res, int* max_batch);
int doe_wsnet_param_count(doe_wsnet* net, long* n);
int doe_wsnet_get_params(doe_wsnet* net, float* host);
int doe_wsnet_set_params(doe_wsnet* net, const float* host);
int doe_wsnet_forward(doe_wsnet* net, int batch, const float* input, float* output);
int doe_wsnet_loss_grad(doe_wsnet* net, int batch, const float* input, const float* label_re,
                        const float* label_im, const float* weight, double* loss,
                        float* grad_host);
int doe_wsnet_adam_step(doe_wsnet* net, float lr);

/* Hand-off between a far-field batch and the network (designs in schedule
 * order, see doe_gs_batch_design_order; every side a multiple of res):
 * network inputs from the targets; training labels (pooled target-plane
 * phasor of each design's current best phase, and the pooled target
 * amplitude as weight); and doe_warmstart_apply, which replaces every
 * design's initial phase by arg(F^-1(t * prediction)) (scratch: device
 * buffer of count*res*res*3 floats). */
int doe_gs_batch_design_order(doe_gs_batch* batch, int* order);
int doe_warmstart_inputs(doe_gs_batch* batch, int res, float* input);
int doe_warmstart_labels(doe_gs_batch* batch, int res, float* label_re, float* label_im,
                         float* weight);
int doe_warmstart_apply(doe_gs_batch* batch, doe_wsnet* net, float* scratch);
/* The same hand-off from given target-plane phasors (device, count*res*res*2
 * floats interleaved re, im, schedule order) instead of a network, e.g. a
 * quadratic target-plane phase. */
int doe_warmstart_apply_phasor(doe_gs_batch* batch, int res, const float* phasor);
/* The same hand-off from a quadratic target-plane phase computed on the
 * device: per design exp(i a |u - u0|^2) over the pooled target, u0 its
 * centroid, a = spread * pi / r with r its RMS radius in full-resolution
 * pixels (at least 4), so a target point at radius r lands near element
 * radius spread * n (scratch: count*res*res*3 floats, as doe_warmstart_apply). */
int doe_warmstart_apply_quadratic(doe_gs_batch* batch, int res, double spread, float* scratch);
/* Restart from the best iterate (no counterpart in the reference): every
 * design's initial phase becomes arg(F^-1(t * phasor(F(illumination *
 * e^{i best phase})))) — one GS projection of its best iterate — so the next
 * doe_gs_batch_run continues from it with fresh weights.  WGS on extended
 * targets lets its weights run away after its best early iterate; restarting
 * there converges where the uninterrupted run stalls. */
int doe_gs_batch_restart_best(doe_gs_batch* batch);

/* Projection-quality metrics of a screen field on the device, in FP64.
 * Replaces ProjectorSystem#screenMetrics (projector.js:554-614) and the
 * metrics.js helpers it calls (maskedSum, regionStats, rmsError,
 * correlation, ssim, spotUniformity, zeroOrderFraction, rmsRadius:
 * metrics.js:10-321) for a device-resident screen.
 *   screen   the simulated screen field (any dtype)
 *   target   FP64 field on the screen's grid: re = target irradiance,
 *            im = ROI (the mask values, as the reference's maskedSum uses)
 *   element  the element-plane field illumination * e^{i phase} (any dtype):
 *            zero-order fraction, and the incident power sum|E| * edx * edy,
 *            which equals the reference's maskedSum(illumination()) * dx * dy
 *            for that field */
typedef struct doe_screen_opts {
  double sdx, sdy;          /* screen pitch (power cell, rmsRadius) */
  double edx, edy;          /* element pitch (incident power) */
  int grid_cols, grid_rows; /* spot-array cells (target.grid); 0, 0: none */
  double grid_min_peak;     /* cells count pixels >= this * plane peak (reference default 0.05) */
  int ssim_weight_target;   /* 1: windows weighted by target energy, as screenMetrics does */
} doe_screen_opts;

typedef struct doe_metrics {
  double efficiency, flux_fraction, uniformity, spot_uniformity, flatness, contrast, rmse,
      fidelity, correlation, ssim, snr, zero_order, rms_radius, mean_intensity, peak_intensity,
      power_in_signal, power_at_screen, power_incident;
  int has_spots; /* 0: spot_uniformity is NaN (the reference's null) */
} doe_metrics;

int doe_screen_metrics(doe_ctx* ctx, const doe_field* screen, const doe_field* target,
                       const doe_field* element, const doe_screen_opts* opts, doe_metrics* out);
/* The same for every field of a batch of screens (out: screens->batch
 * entries), e.g. the planes of ProjectorSystem#defocusCurve
 * (projector.js:622-641) propagated in one batched call. */
int doe_screen_metrics_batch(doe_ctx* ctx, const doe_field* screens, const doe_field* target,
                             const doe_field* element, const doe_screen_opts* opts,
                             doe_metrics* out);
/* doe_screen_metrics without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as doe_screen_metrics; its
 * reduction partials and spot-cell bounds live in a fixed scratch array of
 * DOE_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends, set up by
 * doe_ctx_create), which bounds the screen to about 8192 x 8192 points and
 * the spot grid to 4096 cells (DOE_ERR_UNSUPPORTED beyond).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (several
 * contexts or threads at once) race; calls in sequence do not. */
#define DOE_NOALLOC_SCRATCH_BYTES (6 << 20)
int doe_screen_metrics_noalloc(doe_ctx* ctx, const doe_field* screen, const doe_field* target,
                               const doe_field* element, const doe_screen_opts* opts,
                               doe_metrics* out);

/* Same-pitch resample of ComplexField#resample (ref: src/field.js:443-465):
 * centred zero-padding or crop, dst[j][i] = src[j - oy][i - ox] with
 * ox = floor((dst.nx - src.nx) / 2).  src batch 1 is broadcast to every
 * field of dst (one element, many distances); otherwise batches match.
 * Same dtype, no overlap. */
int doe_field_resample_centred(doe_ctx* ctx, const doe_field* src, doe_field* dst);
/* ComplexField#resample (ref: src/field.js:443-495) between pitches: with
 * equal pitches the centred pad / crop above; otherwise each output sample
 * (i, j), at x = (i - dst.nx/2) dst_dx, reads the source bilinearly at the
 * fractional index x / src_dx + src.nx/2 - 1/2 (zero outside the source),
 * in FP64.  Same dtype, no overlap; src batch 1 broadcasts. */
int doe_field_resample(doe_ctx* ctx, const doe_field* src, doe_field* dst, double src_dx,
                       double src_dy, double dst_dx, double dst_dy);
/* ComplexField#sampleImage (ref: src/field.js:350-400): a host raster of
 * width x height FP64 samples (row-major) covers the central `scale` fraction
 * of the field's window; every grid sample averages the raster over the box
 * it covers (area weights) with `fill` outside it, into re; im = 0. */
int doe_field_sample_image(doe_ctx* ctx, const double* img, int width, int height, double scale,
                           double fill, doe_field* dst);
/* Copy `count` whole fields from src[src_first..] to dst[dst_first..]
 * (same grid and dtype): splits a batched result into per-plane fields. */
int doe_field_copy_slices(doe_ctx* ctx, const doe_field* src, int src_first, doe_field* dst,
                          int dst_first, int count);

/* ---- targets, regions of interest, figure images ------------------------
 * makeTarget (src/projector.js:169-255) on the device.  One spec per field
 * of `out` (FP64; the irradiance is written to re, im is left alone).
 * doe_target_spec_init fills the reference's defaults for a kind.  'text'
 * and 'image' targets are DOE_TGT_RASTER: a host raster (renderText output,
 * or an image) resampled bilinearly, 'contain' unless cover is set
 * (resampleRaster, projector.js:126-154), times gain, background outside
 * the raster; gamma != 1 gives max(0, v)^gamma ('image').  The same spec
 * with any out grid is resampleRaster.  dots/grid: missing holds n_missing
 * (col, row) cell pairs left dark.  Positive cols and rows are required. */
#define DOE_TGT_DISC 0
#define DOE_TGT_RING 1
#define DOE_TGT_DOTS 2 /* 'dots' and 'grid' */
#define DOE_TGT_CHECKER 3
#define DOE_TGT_LINES 4
#define DOE_TGT_ZONE 5
#define DOE_TGT_RASTER 6 /* 'text', 'image', resampleRaster */

typedef struct doe_target_spec {
  int kind;
  double radius;            /* disc: fraction of min(nx, ny) / 2 (0.35) */
  double inner, outer;      /* ring (0.25, 0.4) */
  double cols, rows, sigma; /* dots/grid (8, 8, 0.22 of the smaller pitch) */
  const int* missing;       /* dots/grid: dark (col, row) pairs */
  int n_missing;
  double cells;             /* checker (8) */
  double period, duty;      /* lines (12, 0.5); zone (period 16) */
  const double* raster;     /* raster_w * raster_h samples, row-major */
  int raster_w, raster_h;
  int cover;                /* 0: fit 'contain', 1: 'cover' */
  double gamma, background, gain; /* 1, 0, 1 */
} doe_target_spec;

int doe_target_spec_init(doe_target_spec* spec, int kind);
int doe_target_make(doe_ctx* ctx, const doe_target_spec* specs, doe_field* out);
/* autoRoi (projector.js:262-315) of every field of t (FP64): reads the
 * irradiance in re, writes the ROI to im.  Threshold at threshold * max
 * (reference 0.05), square dilation of half-width max(1, round(dilate)),
 * then a (2s+1)^2 box blur with s = max(1, round(soft)) unless soft <= 0.
 * Exact: equals the reference's values bit for bit for the same irradiance. */
int doe_target_auto_roi(doe_ctx* ctx, doe_field* t, double dilate, double soft, double threshold);
/* makeTarget followed by autoRoi in one call (makeTarget without a caller
 * ROI, projector.js:251): the same results as doe_target_make then
 * doe_target_auto_roi, with the irradiance maximum taken while the targets
 * are written instead of by a second read. */
int doe_target_make_roi(doe_ctx* ctx, const doe_target_spec* specs, doe_field* out, double dilate,
                        double soft, double threshold);
/* blurRaster (src/doe.js:761-794) of the re plane of every field, src to dst
 * (FP64, may be the same field): separable Gaussian, half-width
 * r = max(1, ceil(3 sigma)), clamped edges; sigma <= 0 copies.  taps: the
 * 2r+1 normalised weights, or NULL to compute them (exp in FP64). */
int doe_raster_blur(doe_ctx* ctx, const doe_field* src, doe_field* dst, double sigma, const double* taps);

/* planeToRGB / planeToGray (src/color.js:7-160) of one field's scalar
 * quantity: 8-bit RGB (a piecewise-linear colour map of 1..8 stops) or gray
 * pixels, row-major, or PNG scanlines (png_rows: each row led by filter
 * byte 0, src/png.js:96-101) ready for deflate.  The range is min / max
 * when given, else the data's (NaN skipped); with a percentile (colour
 * maps) the upper end is the sorted plane's element at
 * min(n - 1, floor(n * percentile)) and the lower its first element.
 * range_out (nullable) receives the lo, hi used. */
#define DOE_Q_RE 0
#define DOE_Q_IM 1
#define DOE_Q_INTENSITY 2 /* re^2 + im^2 */
#define DOE_Q_AMPLITUDE 3 /* hypot(re, im) */
#define DOE_Q_PHASE 4     /* atan2(im, re) */

typedef struct doe_image_opts {
  int quantity; /* DOE_Q_* */
  int gray;     /* 1: planeToGray; 0: colour map */
  int nstops;
  double stop_pos[8];
  double stop_rgb[8][3];
  int has_min, has_max;
  double min, max;
  double gamma;     /* colour map display gamma (1: none) */
  int log;          /* colour map log scaling */
  double log_floor; /* 1e-4 */
  int has_percentile;
  double percentile;
  int invert;   /* gray */
  int png_rows; /* 1: PNG scanlines */
} doe_image_opts;

int doe_image_opts_init(doe_image_opts* opts);
/* out_len >= ny * (png_rows + nx * (gray ? 1 : 3)). */
int doe_field_image(doe_ctx* ctx, const doe_field* f, int index, const doe_image_opts* opts,
                    unsigned char* out, long out_len, double* range_out);

/* ---- random phase (ref: src/doe.js:65-86 makeRng, randomPhase) --------- *
 * out re (FP64 or FP32, every point of the batch as one stream) = the
 * reference's randomPhase(nx, ny * batch, seed): xorshift32 (13, 17, 5) from
 * seed || 1, sample k = (s / 2^32 * 2 - 1) * pi with s the state after k + 1
 * steps; equal bit for bit (FP32: rounded once).  im is untouched. */
int doe_random_phase(doe_ctx* ctx, doe_field* out, unsigned seed);

/* ---- fabrication output (ref: src/doe.js:185-226, 967-976) ------------- *
 * errorDiffusionQuantize of the re plane of every field (FP64, out re gets
 * the wrapped levels; out must not be the input): Floyd–Steinberg with the
 * error taken on the circle, levels uniformly spaced, rows serpentine
 * (snake = 1, the reference's default: each image is one serial chain, the
 * images of a batch run in parallel, one workgroup each) or all left to right
 * (snake = 0: a diagonal wavefront per image, at most 8192 rows); wrap_first
 * wraps the input first.  Equal to the reference bit for bit.
 * phaseToHeight: out re = ((phase * s) % step + step) % step with
 * s = lambda / (2 pi (n - 1)), step = lambda / (n - 1). */
int doe_error_diffusion(doe_ctx* ctx, const doe_field* phase, doe_field* out, int levels, int snake,
                        int wrap_first);
int doe_phase_to_height(doe_ctx* ctx, const doe_field* phase, doe_field* out, double lambda, double n_index);

#ifdef __cplusplus
}
#endif
#endif /* DOEHIP_H */
