/* spektralab_ext.h -- entry points this engine adds to the reference ABI.
 *
 * None of these has a counterpart in the reference's include/spektralab.h; every one is an
 * addition with the same conventions (status codes, per-thread last error, borrowed inputs).
 *   - device-resident rendering for hosts and benchmarks that keep frames in HBM
 *   - the render plan and cached negative, so a CPU checker can replay a render
 *   - build-defined overrides of the benchmark workloads (explicit halation mixture, FIR/IIR
 *     crossover, spectral precision)
 *   - batched rendering of many sessions in one hipGraph launch
 *   - self-tests of the Philox generator, the Poisson sampler and the Gaussian machinery
 */
#ifndef SPEKTRALAB_EXT_H
#define SPEKTRALAB_EXT_H

#include "spektralab.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Overrides for later renders of the session (drops its cached negatives). JSON object, keys
 * optional:
 *   "halation_mixture": [[sigma_px, weight], ...]   1..8 components replacing the bounce mixture
 *                                                    (same sigma and weight on every channel)
 *   "halation_scatter": bool                         false: no scatter mixture
 *   "fir_crossover": number > 0                      FIR for 0 < sigma < crossover (3 = reference)
 *   "spectral_precision": "fp32" | "fp16"            the print's two 81-bin spectral integrals as
 *                                                    FP32 sums, or FP16 matrix-core products with
 *                                                    FP32 accumulation (keeps the negatives)
 *   "record_grain": bool                             keep each develop's grain draws (spkx_grain_record)
 * Returns SPK_OK | SPK_ERR_INVALID_ARG | SPK_ERR_USER. */
SPK_API spk_status spkx_set_overrides(spk_session* session, const char* json);

/* spk_render / spk_reprint writing RGBA16 into a caller device buffer (row pitch row_stride_px
 * pixels, >= width), enqueued on `stream` (a hipStream_t; NULL = the session's own stream).
 * The call returns after enqueueing; *out_w, *out_h receive the size.  No host copy is made. */
SPK_API spk_status spkx_render_device(spk_session* session, const char* tier, int32_t reprint, void* dev_out,
                                      uint32_t row_stride_px, void* stream, uint32_t* out_w, uint32_t* out_h);
/* spkx_render_device for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, hosts that forbid hipMalloc); no counterpart in the reference.  Every device
 * buffer of the render (planes, blur temporaries, the plan's tables when the session has none on
 * the device yet, a downscaled tier image, the auto-exposure meter image) is carved from one
 * static scratch region of 512 MiB per GPU, which the library holds from its load to the end of
 * the process; the call makes no hipMalloc/hipFree.  A render whose buffers do not fit (about
 * 4 MP for the full chain with grain and a 3-component halation mixture) is refused with
 * SPK_ERR_INVALID_ARG before anything is enqueued.  A negative developed by this call is not
 * cached; reprint != 0 reuses a negative cached by an earlier spk_render / spkx_render_device.
 * Calls on one GPU share the region: the library serialises them and each returns after `stream`
 * has drained, so they do not race; work the caller enqueues elsewhere must not read the region. */
SPK_API spk_status spkx_render_device_static(spk_session* session, const char* tier, int32_t reprint, void* dev_out,
                                             uint32_t row_stride_px, void* stream, uint32_t* out_w, uint32_t* out_h);
/* Waits for the session's stream. */
SPK_API spk_status spkx_session_sync(spk_session* session);

/* Stage timing with HIP events on the render stream (off by default).  spkx_stage_times returns
 * {"<stage>": {"ms": total, "count": renders}, ...} accumulated since enabling (or the last reset). */
SPK_API spk_status spkx_set_profiling(spk_session* session, int32_t on);
SPK_API spk_status spkx_stage_times(spk_session* session, int32_t reset, char** out_json);

/* The plan of the session's most recent render (struct spk_plan of spektralab_plan.h).
 * Returns sizeof(spk_plan); copies it when cap >= that size.  0 when nothing was rendered. */
SPK_API size_t spkx_last_plan(spk_session* session, void* out, size_t cap);
/* The cached negative of `tier` as 3 planar float32 planes (w*h*3 floats) into host memory. */
SPK_API spk_status spkx_negative(spk_session* session, const char* tier, float* out, size_t count);
/* The grain draws of the tier's last develop, kept when the override "record_grain" is true:
 * lambdas[(3 c + l) w h + i] is the Poisson mean of sub-layer l, channel c at pixel i and counts the
 * count drawn (9 planes each, count >= 9 w h floats; either output may be NULL).  For checkers of
 * the integer Poisson counts. */
SPK_API spk_status spkx_grain_record(spk_session* session, const char* tier, float* lambdas, float* counts,
                                     size_t count);

/* Host-only: the plan a session opened on `resources_dir` with `params_delta_json` (and the
 * overrides of spkx_set_overrides, NULL = none) would use for a w x h full-tier render whose film
 * seed is `film_seed`.  Needs no GPU.  Returns sizeof(spk_plan) (copied when cap allows), or 0 with
 * the reason in spk_last_error. */
SPK_API size_t spkx_plan(const char* resources_dir, const char* params_delta_json, const char* overrides_json,
                         uint32_t width, uint32_t height, uint32_t film_seed, void* out, size_t cap);

/* spk_output_transform computed on the host from a resource folder, without an engine or device
 * (for hosts that only need the matrices and constants, and for checking them without a GPU).
 * The reply and its errors are spk_output_transform's; out_cmax (may be NULL) receives the
 * 64 x 720 chroma-limit table when cap >= 46080. */
SPK_API spk_status spkx_output_transform_host(const char* resources_dir, const char* src_cs, const char* dst_cs,
                                              const char* gamut_compress_json, char** out_json, float* out_cmax,
                                              size_t cap);

/* Batched rendering.  spkx_batch_create snapshots the parameters of n open sessions of one engine
 * and captures one hipGraph that renders every session's full tier (develop + print, fresh seeds
 * from each session's own sequence at every run, as spk_render draws them) into dev_out[i]
 * (RGBA16, row pitch row_stride_px[i] >= width), on up to 4 forked streams.
 * options_json (NULL = defaults): {"autotune": bool (default true: the schedule follows the
 * sessions -- "streams" unless a session is in FP16 spectral precision, then "phased" -- and its
 * stream count, persistent-launch grid and workgroup size are chosen by timing candidate graphs;
 * "schedule" is then ignored), "schedule": "phased" (one develop and one print launch for all
 * frames between per-frame phases; the default with autotune off) | "streams" (each frame's whole
 * chain on its stream, its develop a persistent launch over that frame; no session in FP16
 * spectral precision), "streams": 1..4,
 * "record_grain": bool (keep each frame's grain draws, spkx_batch_grain_record),
 * "grid_per_cu": 1..64, "block": 64|128|256}.  Parameter changes made to the
 * sessions after creation are not seen by the batch.  Returns NULL with the reason in
 * spk_last_error on failure.
 * spkx_batch_run enqueues one batch render on `stream` (a hipStream_t; NULL = the batch's own) and
 * returns; spkx_batch_sync waits for it; spkx_batch_info returns {"frames", "pixels", "config",
 * "tuning": [{..., "ms"}]}. */
typedef struct spkx_batch spkx_batch;
SPK_API spkx_batch* spkx_batch_create(spk_engine* engine, spk_session* const* sessions, int32_t n,
                                      void* const* dev_out, const uint32_t* row_stride_px, const char* options_json);
SPK_API spk_status spkx_batch_run(spkx_batch* batch, void* stream);
SPK_API spk_status spkx_batch_sync(spkx_batch* batch);
SPK_API spk_status spkx_batch_info(spkx_batch* batch, char** out_json);
/* With the option "record_grain": true, the grain draws of frame `frame` in its last run, laid out
 * as spkx_grain_record's (9 planes of Poisson means, 9 of counts). */
SPK_API spk_status spkx_batch_grain_record(spkx_batch* batch, int32_t frame, float* lambdas, float* counts,
                                           size_t count);
SPK_API void spkx_batch_destroy(spkx_batch* batch);

/* Differentiable render (no counterpart in the reference, which renders forward only).
 * spkx_diff_create snapshots the session's full-tier parameters and draws one film and one print
 * seed from its sequence; `slots` forward passes can be held at once (checkpoint memory per slot:
 * spkx_diff_info's "checkpoint_bytes_per_slot").  The grain is frozen: each forward pass draws the
 * Poisson counts as spk_render would and the reverse pass differentiates the grain sum with those
 * counts fixed (the sample's dependence on the density through the sub-layer saturation stays).
 * Refused (SPK_ERR_USER): scatter halation, DIR couplers, grain_amount != 1, FP16 spectral
 * precision, more than 3 bounce components, the scanned-film path.
 * Quantities (float64, SPKX_DIFF_QUANTITIES): 0 exposure-compensation gain; 1-3 bounce mix a_c;
 * 4-6 mix b_c; 7-15 bounce weights [k][c]; 16-24 bounce sigmas [k][c] (px); 25-27 film curve
 * gamma_c; 28-30 film offset_c; 31 print gain; 32-34 paper curve gamma_c; 35-37 paper offset_c;
 * 38 unsharp amount; 39 reserved (0); 40 + 3 l + c enlarger weight wt_c(l), l < 81.
 * spkx_diff_defaults writes the plan's values of them.
 * spkx_diff_forward: input (device float32, interleaved h x w x 3) -> encoded output before 16-bit
 * quantisation (device float32, interleaved h x w x 3), with the quantities theta (host, NULL: the
 * defaults), into checkpoint `slot`.
 * spkx_diff_backward: the gradient of a scalar with respect to that output (device, like the
 * output) -> its gradient with respect to the input (device, like the input) and to the
 * quantities (device float64 [SPKX_DIFF_QUANTITIES], overwritten); input is the slot's forward
 * input.  Both run on `stream` (a hipStream_t; NULL = the object's own) and return when done.
 * spkx_diff_plan copies the plan of the render (spk_plan) as spkx_last_plan does. */
#define SPKX_DIFF_QUANTITIES 283
typedef struct spkx_diff spkx_diff;
SPK_API spkx_diff* spkx_diff_create(spk_session* session, int32_t slots, const char* options_json);
SPK_API spk_status spkx_diff_defaults(spkx_diff* diff, double* theta);
SPK_API spk_status spkx_diff_forward(spkx_diff* diff, int32_t slot, const double* theta, const float* input,
                                     float* output, void* stream);
SPK_API spk_status spkx_diff_backward(spkx_diff* diff, int32_t slot, const float* input, const float* grad_output,
                                      float* grad_input, double* grad_theta, void* stream);
SPK_API size_t spkx_diff_plan(spkx_diff* diff, void* out, size_t cap);
/* The enlarger's CMY filtration as quantities of the differentiable render (no counterpart in the
 * reference, whose print exposure follows c/m/y_filter_neutral and m/y_filter_shift through
 * src/print/setup.cpp:339-361).  filters = (c, m, y neutral CC, m shift, y shift), NULL: the
 * session's at spkx_diff_create.  wt (243) receives the enlarger weights theta 40..282 that
 * filtration gives -- the filtered lamp times the paper sensitivity, scaled by the exposure gain of
 * that light over the gain the render holds -- and jac (243 x 5 row-major, may be NULL) d wt / d
 * filters, so that a gradient with respect to theta 40..282 becomes one with respect to the filters
 * by jac^T.  filters_out (5, may be NULL) receives the session's filtration.  SPK_ERR_USER with a
 * preflash exposure (its offset is not a quantity). */
SPK_API spk_status spkx_diff_filter_weights(spkx_diff* diff, const double* filters, double* wt, double* jac,
                                            double* filters_out);
/* The same from a resource folder and a parameter delta, without an engine or device (the render
 * gain is the one of the delta's own filtration). */
SPK_API spk_status spkx_filter_weights(const char* resources_dir, const char* params_delta_json, const double* filters,
                                       double* wt, double* jac, double* filters_out);
SPK_API spk_status spkx_diff_info(spkx_diff* diff, char** out_json);
SPK_API void spkx_diff_destroy(spkx_diff* diff);

/* Self-tests on the engine's GPU:
 *   out[4i..4i+3] = Philox4x32-10(ctr[4i..4i+3], key[2i..2i+1])
 *   out[i] = Poisson(lam[i]) of pixel i on (stream, seed), the grain sampler itself
 *   gaussian: 3-channel interleaved host image -> per-channel Gaussian (the render's machinery) */
SPK_API spk_status spkx_philox(spk_engine* engine, const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n);
SPK_API spk_status spkx_poisson(spk_engine* engine, const float* lam, uint32_t n, uint32_t stream, uint32_t seed,
                                float* out);
SPK_API spk_status spkx_gaussian(spk_engine* engine, const float* in, float* out, uint32_t width, uint32_t height,
                                 const double* sigma3, double fir_crossover);
/* A mixture of Gaussians as the render's halation runs it: out = sum_k weight_k G_{sigma_k}(in) per
 * channel, accumulated in list order (sigma and weight: n x 3, channel fastest; 1 <= n <=
 * SPK_PLAN_MAX_MIX; the branch rules of spkx_gaussian per component). */
SPK_API spk_status spkx_mixture(spk_engine* engine, const float* in, float* out, uint32_t width, uint32_t height,
                                const double* sigma, const double* weight, uint32_t n, double fir_crossover);
/* The transpose of spkx_gaussian's operator, per channel (the reverse pass's blur of the
 * differentiable render: the FIR pass, which is symmetric, or the transposed recursion). */
SPK_API spk_status spkx_gaussian_adjoint(spk_engine* engine, const float* in, float* out, uint32_t width,
                                         uint32_t height, const double* sigma3, double fir_crossover);
/* The tier downscale (3-channel interleaved host image -> long edge `long_edge`, 0 = none): writes
 * *out_w, *out_h always and the image when out is not NULL and cap_floats allows. */
SPK_API spk_status spkx_downscale(spk_engine* engine, const float* in, uint32_t width, uint32_t height,
                                  uint32_t long_edge, float* out, size_t cap_floats, uint32_t* out_w, uint32_t* out_h);
/* Measured vector-ALU issue rates of the engine's GPU, the roofs of the compute-bound kernels (no
 * counterpart in the reference): {"compute_units", "clock_mhz", "waves", "instructions_per_wave",
 * "rates": {kind: {"ms", "wave_instructions_per_s", "per_cu_per_clock"}}} for the kinds v_fma_f32,
 * v_exp_f32, v_fma_f64 and v_mad_u64_u32 (each timed alone: 8 independent chains per thread,
 * 32 waves per compute unit, median of 5 launches). */
SPK_API spk_status spkx_valu_roofs(spk_engine* engine, char** out_json);

#ifdef __cplusplus
}
#endif

#endif /* SPEKTRALAB_EXT_H */
