/*
 * dsd_hip — MI355X (gfx950) backend for the PCM->DSD hot path of
 * foo_dsd_trellis: half-band FIR upsampling chain, Viterbi trellis SDM, and
 * Density-Aligned-Stitching (DAS) over parallel segments.
 *
 * Plain C ABI (libdsdhip.so).  Conventions follow the reference's GPU seam
 * (include/gpu_compute.h:1-9): functions return 0 on success and -1 on
 * failure, on which the caller falls back to its CPU path; host buffers are
 * caller-owned and are copied synchronously.  Functions with a _dev suffix
 * take device pointers and enqueue on the given HIP stream instead.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef DSD_HIP_H
#define DSD_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * 1. Backend seam — additions to include/gpu_compute.h
 * ---------------------------------------------------------------------- */

/* New value for gpu_backend_t (include/gpu_compute.h:24-29). */
#define GPU_BACKEND_HIP 4

/* Same layout as gpu_info_t (include/gpu_compute.h:31-36). */
typedef struct {
    int backend;
    char device_name[128];
    size_t vram_mb;
    bool available;
} gpu_hip_info_t;

/* replaces gpu_cuda_probe (include/gpu_compute.h:185) */
bool gpu_hip_probe(void);
/* replaces gpu_cuda_get_info (include/gpu_compute.h:188) */
void gpu_hip_get_info(gpu_hip_info_t *info);
/* replaces gpu_cuda_create (include/gpu_compute.h:191); the returned
 * context's first field is the backend tag GPU_BACKEND_HIP, as the
 * dispatcher expects (src/gpu_compute.c:153-157). */
void *gpu_hip_create(void);
/* replaces gpu_cuda_destroy (include/gpu_compute.h:194) */
void gpu_hip_destroy(void *ctx);
/* replaces gpu_cuda_reset_chunk (include/gpu_compute.h:195) */
void gpu_hip_reset_chunk(void *ctx);

/* replaces gpu_cuda_fir_setup (include/gpu_compute.h:240-241): every stage
 * uses `taps` (ntaps, odd).  When the taps are the reference's 63-tap
 * half-band design (src/fir.c:79-98) rounded to float, the double-precision
 * design is used, as the CPU fp64 chain does. */
int gpu_hip_fir_setup(void *ctx, const float *taps, int ntaps, int num_stages, bool upsample);
/* Exact CPU-chain plan for a rate pair (src/fir.c:523-611, incl. the 127-tap
 * first stage towards DSD512), with the engine's output gain
 * (src/engine.c:872-883).  Extension of the seam. */
int gpu_hip_fir_setup_rates(void *ctx, unsigned fs_in, unsigned fs_out, double gain);
/* replaces gpu_cuda_fir_chain (include/gpu_compute.h:242-243): one channel
 * per call, in channel order, as the plugin calls it (src/dsp_plugin.c:
 * 1490-1530); each channel slot keeps its own stage delay lines, the slot
 * advancing per call and rewound by gpu_hip_reset_chunk, so every channel
 * continues exactly like its own fir_chain_t across chunks. */
int gpu_hip_fir_chain(void *ctx, const float *in, float *out, size_t in_count, size_t *out_count);
/* replaces gpu_cuda_fir_chain_f64 (include/gpu_compute.h:244-245), per
 * channel slot like gpu_hip_fir_chain */
int gpu_hip_fir_chain_f64(void *ctx, const float *in, double *out, size_t in_count,
                          size_t *out_count);
/* replaces gpu_cuda_fir_batch (include/gpu_compute.h:248-250) */
int gpu_hip_fir_batch(void *ctx, const float *in_batch, float *out_batch, size_t samples_per_ch,
                      int num_channels, size_t *out_count_per_ch);

/* replaces gpu_cuda_trellis_setup (include/gpu_compute.h:220-223) */
int gpu_hip_trellis_setup(void *ctx, int num_cands, int order, int trellis_lat,
                          const double *ntf_a, const double *ntf_g, double state_limit,
                          int trellis_depth);
/* replaces gpu_cuda_trellis (include/gpu_compute.h:229): one channel,
 * device-resident state persisting between calls; output +-1.0 exactly as
 * sdm_process_block (src/trellis.c:1013-1040) would produce. */
int gpu_hip_trellis(void *ctx, const double *in, float *out, size_t count);
/* replaces gpu_cuda_trellis_das (include/gpu_compute.h:232-233): DAS over
 * num_channels x count samples (channel-major), same semantics as the
 * reference's CPU DAS (src/dsp_plugin.c:1314-2761).  Each output row holds
 * the produced samples (+-1.0) followed by 0.0f up to count.  With the
 * reference's seeding for the rate (seg0-first below DSD512) the produced
 * count is count, or count - trellis_lat on a fresh context, as the
 * plugin's output check expects (src/dsp_plugin.c:2733-2747). */
int gpu_hip_trellis_das(void *ctx, const double *in, float *out, size_t count, int num_channels);
/* gpu_hip_trellis_das that also returns the produced samples per channel
 * (extension of the seam; 0 on failure). */
int gpu_hip_trellis_das_ex(void *ctx, const double *in, float *out, size_t count, int num_channels,
                           size_t *out_count);
/* Produced samples per channel of the last gpu_hip_trellis_das(_ex) call. */
size_t gpu_hip_trellis_das_count(void *ctx);
/* Bit parity with the shipped Windows build: order-8 filters computed by
 * the reference's AVX2 4-candidate FMA path (src/trellis.c:348-533, chosen
 * at :904-927 for the first 4*floor(n/4) candidates of every step) instead
 * of the scalar order.  Extension of the seam; off by default. */
int gpu_hip_trellis_fma_o8(void *ctx, int on);
/* replaces gpu_cuda_fir_lowpass_setup (include/gpu_compute.h:92): taps of
 * the same-rate lowpass (1..128), widened to double */
int gpu_hip_fir_lowpass_setup(void *ctx, const float *taps, int ntaps);
/* replace gpu_cuda_fir_lowpass / _f64 (include/gpu_compute.h:97-103): one
 * channel of +-1.0 DSD per call, in channel order (the channel slot advances
 * per call and gpu_hip_reset_chunk rewinds it); y = (sum_k taps[k] s[i-k]) *
 * gain in ascending k over the channel's persistent delay line, zero before
 * the stream's start as in the engine's lowpass (src/fir.c:795-873).
 * count >= GPU_MIN_SAMPLES. */
int gpu_hip_fir_lowpass_f64(void *ctx, const float *in, double *out, size_t count, double gain);
int gpu_hip_fir_lowpass(void *ctx, const float *in, float *out, size_t count, float gain);
/* replaces gpu_cuda_boxcar_smooth (include/gpu_compute.h:112-113): the
 * engine's boxcar of taps (1..64) +-1.0 samples times gain (exact window
 * sums, src/engine.c:533-546), per channel slot like the lowpass, narrowed
 * to float */
int gpu_hip_boxcar_smooth(void *ctx, const float *in, float *out, size_t count, int taps, float gain);
/* DAS plan parameters for gpu_hip_trellis_das (segment length in samples,
 * 0 = reference auto count; seed mode 1 = seg0-first, 0 = all segments
 * seeded from the persistent state, DSDHIP_SEED_BY_RATE (the default) =
 * the reference's rule, seg0-first below DSD512 (src/dsp_plugin.c:2185);
 * reposition quiet boundaries). */
#define DSDHIP_SEED_BY_RATE (-1)
#define DSDHIP_DSD512_RATE 22579200u
int gpu_hip_das_config(void *ctx, size_t seg_len, int seg0_first, int reposition);
/* Output rate for the seed rule (also taken from gpu_hip_fir_setup_rates);
 * an unknown rate (0) seeds seg0-first. */
int gpu_hip_das_rate(void *ctx, unsigned fs_out);

/* ------------------------------------------------------------------------
 * 2. NTF tables — include/ntf.h:23-36
 * ---------------------------------------------------------------------- */
typedef struct {
    double a[10];
    double g[10];
    int order;
    unsigned freq;
    const char *name;
} dsdhip_ntf_t;

/* replaces ntf_get_filter_count / ntf_get_by_index (include/ntf.h:32-35) */
int dsdhip_ntf_count(void);
const dsdhip_ntf_t *dsdhip_ntf_get(int index);
/* replaces ntf_get_filter (include/ntf.h:23): index or -1 */
int dsdhip_ntf_lookup(int ntf_id, unsigned rate);
/* replaces ntf_auto_select (include/ntf.h:26): index or -1 */
int dsdhip_ntf_auto(unsigned rate);

/* ------------------------------------------------------------------------
 * 3. Trellis contexts on the device — include/trellis.h:61-118
 * ---------------------------------------------------------------------- */
typedef struct dsdhip_trellis dsdhip_trellis_t;

/* Pool of nctx contexts sharing one parameter set; replaces
 * sdm_context_init (include/trellis.h:62-63) for each of them. */
dsdhip_trellis_t *dsdhip_trellis_create(int ntf_index, int depth, int cands, int lat,
                                        double state_limit, int nctx);
dsdhip_trellis_t *dsdhip_trellis_create_coeffs(const double *a, const double *g, int order,
                                               int depth, int cands, int lat,
                                               double state_limit, int nctx);
void dsdhip_trellis_destroy(dsdhip_trellis_t *t);
int dsdhip_trellis_set_stream(dsdhip_trellis_t *t, void *hip_stream);
int dsdhip_trellis_num_contexts(const dsdhip_trellis_t *t);
/* replaces sdm_context_reset (include/trellis.h:86) */
int dsdhip_trellis_reset(dsdhip_trellis_t *t, int ctx);
/* replaces sdm_context_copy_state (include/trellis.h:77) */
int dsdhip_trellis_copy(dsdhip_trellis_t *t, int dst, int src);
/* DAS seeding of every survivor (src/dsp_plugin.c:2329-2338) */
int dsdhip_trellis_seed(dsdhip_trellis_t *t, int ctx, const double *state10);
/* best survivor's loop-filter state (as read by src/dsp_plugin.c:2251-2253) */
int dsdhip_trellis_best_state(dsdhip_trellis_t *t, int ctx, double *state10);
/* replaces sdm_process_block (include/trellis.h:70-71), host buffers */
int dsdhip_trellis_process(dsdhip_trellis_t *t, int ctx, const double *in, float *out,
                           size_t count, size_t *out_count);
/* replaces sdm_drain (include/trellis.h:73): up to max_out outputs from zero
 * input, flushing an unfilled latency first; host buffers */
int dsdhip_trellis_drain(dsdhip_trellis_t *t, int ctx, float *out, size_t max_out, size_t *out_count);
/* replaces sdm_state_distance (include/trellis.h:82) for contexts i, j */
int dsdhip_trellis_state_distance(dsdhip_trellis_t *t, int i, int j, double *dist);
/* replaces sdm_estimate_state_multibit (include/trellis.h:114-118); with
 * num_levels = 16 it is sdm_estimate_state (:108-110).  Host buffers; the
 * greedy recurrence (one dependent chain, latency-bound) runs on the
 * calling thread, bit-exact with the reference. */
int dsdhip_estimate_state(const double *a, const double *g, int order, const double *init_state,
                          const double *in, size_t count, double state_limit, int num_levels,
                          double *out_state);
/* replaces the estimation block of src/threadpool.c:345-366
 * (BLOCK_MODE_ESTIMATE): from init10 over x[0..count), the state at each
 * boundary offset bnd[0..nb) into out10[nb][10] (a boundary advances the
 * state only when it lies past the current position and within count).
 * Host memory, calling thread, no device use. */
int dsdhip_estimate_boundaries(const double *a, const double *g, int order, const double *init10,
                               const double *x, size_t count, const int64_t *bnd, int nb,
                               double state_limit, int levels, double *out10);
/* Validation form: index_bias shifts the fast form's index guess (results
 * must not change: wrong guesses send their block to the reference form);
 * redo_blocks receives the blocks so recomputed.  No reference counterpart. */
int dsdhip_estimate_boundaries_ex(const double *a, const double *g, int order, const double *init10,
                                  const double *x, size_t count, const int64_t *bnd, int nb,
                                  double state_limit, int levels, double *out10, double index_bias,
                                  int64_t *redo_blocks);
/* Index that the fast form's exact interval table gives v (-1: the table is
 * not used for this a0 / level count).  No reference counterpart. */
int dsdhip_est_table_index(double a0, int levels, double v);
/* The same recurrence on one device lane per channel (the DAS kernel
 * das_estimate): x is device memory, nch rows of stride doubles starting at
 * `start`; init10/out10 host arrays [nch][10] / [nch][nb][10].  For timing
 * against the host form; DAS runs the host form. */
int dsdhip_estimate_boundaries_dev(const double *a, const double *g, int order, const double *init10,
                                   const double *d_x, int64_t stride, int nch, size_t count,
                                   const int64_t *bnd, int nb, double state_limit, int levels,
                                   double *out10);
/* Batched form, device buffers (see dsdhip_trellis.cpp for field meaning):
 * one launch, one wave per job; outputs are packed bits (1 = +1.0). */
int dsdhip_trellis_run_jobs(dsdhip_trellis_t *t, int njobs, const int *ctx, const int *copy_src,
                            const double *const *d_in, const int64_t *n_in,
                            const int64_t *keep_from, uint64_t *const *d_out,
                            const double *seed, int64_t *kept, int64_t *slow_steps,
                            int *err_out);
/* dsdhip_trellis_run_jobs that allocates nothing, for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops).
 * No counterpart in the reference.  The job table lives in a static device
 * array of DSDHIP_NOALLOC_MAX_JOBS descriptors, one per GPU, held until the
 * process ends; more jobs are rejected (-1, message on stderr).  Calls that
 * overlap on one GPU share the table and race. */
#define DSDHIP_NOALLOC_MAX_JOBS 4096
int dsdhip_trellis_run_jobs_noalloc(dsdhip_trellis_t *t, int njobs, const int *ctx, const int *copy_src,
                                    const double *const *d_in, const int64_t *n_in,
                                    const int64_t *keep_from, uint64_t *const *d_out,
                                    const double *seed, int64_t *kept, int64_t *slow_steps,
                                    int *err_out);
/* Validation switch: decide every step with the literal sequential prune. */
int dsdhip_trellis_force_literal(dsdhip_trellis_t *t, int on);
/* Order 8: the AVX2 build's 4-candidate FMA filter (src/trellis.c:348-533,
 * 904-927) instead of the scalar one (see gpu_hip_trellis_fma_o8). */
int dsdhip_trellis_fma_o8(dsdhip_trellis_t *t, int on);
/* At K <= 2 (the engine's DSD-input paths, src/engine.c:43-103) jobs run one
 * per lane, 64 per wave; on = 1 runs them one per wave like K > 2 (same
 * bits; for validation and timing).  No reference counterpart. */
int dsdhip_trellis_wave_only(dsdhip_trellis_t *t, int on);
/* HIP-event timing of trellis launches on the pool's stream (no reference
 * counterpart; the reference times phases with QueryPerformanceCounter,
 * src/dsp_plugin.c:90-94). */
int dsdhip_trellis_timing(dsdhip_trellis_t *t, int on);
int dsdhip_trellis_timing_read(dsdhip_trellis_t *t, double *ms, int64_t *launches, int64_t *samples,
                               int clear);

/* ------------------------------------------------------------------------
 * 4. FIR upsampling chain — include/fir.h:22-79
 * ---------------------------------------------------------------------- */
typedef struct dsdhip_fir dsdhip_fir_t;

/* Stage plan for fs_in -> fs_out (2^n, n <= 9): per-stage tap count into
 * ntaps[]; returns the stage count or -1 (src/fir.c:523-611, :131-150). */
int dsdhip_fir_plan(unsigned fs_in, unsigned fs_out, int *ntaps);
/* The reference's half-band design (src/fir.c:79-98, beta 12). */
void dsdhip_halfband_taps(double *h, int ntaps);
/* replaces fir_chain_init_ex(..., use_fp64=true) (include/fir.h:61-62) for
 * nch channels, output multiplied by gain (src/engine.c:872-883). */
dsdhip_fir_t *dsdhip_fir_create(unsigned fs_in, unsigned fs_out, double gain, int nch);
/* uniform taps on every stage (gpu_compute.h gpu_fir_setup semantics) */
dsdhip_fir_t *dsdhip_fir_create_taps(const double *taps, int ntaps, int nstages, double gain, int nch);
/* replaces fir_chain_free (include/fir.h:79) */
void dsdhip_fir_destroy(dsdhip_fir_t *f);
int dsdhip_fir_set_stream(dsdhip_fir_t *f, void *hip_stream);
int dsdhip_fir_num_stages(const dsdhip_fir_t *f);
/* replaces fir_chain_reset (include/fir.h:76): zero delay lines */
int dsdhip_fir_reset(dsdhip_fir_t *f);
/* Adjacent 63-tap stages in one pass with the intermediate stream kept in
 * LDS (on = 1, the default) or one pass per stage (on = 0); identical
 * outputs.  No reference counterpart. */
int dsdhip_fir_set_fused(dsdhip_fir_t *f, int on);
/* replaces fir_chain_process_d (include/fir.h:70-72) for nch channels;
 * delay lines persist between calls.  Device buffers, stream-ordered. */
int dsdhip_fir_process_dev(dsdhip_fir_t *f, const float *d_in, int64_t in_stride, size_t n,
                           double *d_out, int64_t out_stride);
/* Host buffers: in nch x n floats, out nch x n*2^stages doubles. */
int dsdhip_fir_process(dsdhip_fir_t *f, const float *in, size_t n, double *out);
/* fir_chain_process_d with double input (the boxcar-smoothed signal,
 * src/engine.c:620-645); device buffers. */
int dsdhip_fir_process_f64_dev(dsdhip_fir_t *f, const double *d_in, int64_t in_stride, size_t n,
                               double *d_out, int64_t out_stride);

/* ------------------------------------------------------------------------
 * 5. DAS — src/dsp_plugin.c:1314-2761 (parallel trellis + stitching)
 * ---------------------------------------------------------------------- */
typedef struct dsdhip_das dsdhip_das_t;

/* nch persistent trellis contexts plus segment contexts grown on demand.
 * seg_len > 0: segments of seg_len samples (last takes the rest);
 * seg_len == 0: nseg equal parts shrunk until each is >= 4 overlaps
 * (:1348-1368).  Overlap = 64 * lat (:1347). */
dsdhip_das_t *dsdhip_das_create(int ntf_index, int depth, int cands, int lat, double state_limit,
                                int nch, size_t seg_len, int nseg, int seg0_first, int reposition);
dsdhip_das_t *dsdhip_das_create_coeffs(const double *a, const double *g, int order, int depth,
                                       int cands, int lat, double state_limit, int nch,
                                       size_t seg_len, int nseg, int seg0_first, int reposition);
void dsdhip_das_destroy(dsdhip_das_t *d);
int dsdhip_das_set_stream(dsdhip_das_t *d, void *hip_stream);
int dsdhip_das_force_literal(dsdhip_das_t *d, int on);
int dsdhip_das_fma_o8(dsdhip_das_t *d, int on);
int dsdhip_das_wave_only(dsdhip_das_t *d, int on);
/* Segment count (and nominal starts/sizes if given) for F samples. */
int dsdhip_das_plan(const dsdhip_das_t *d, size_t F, size_t *start, size_t *size);
/* One chunk: d_in nch x F doubles (row stride in_stride); d_out nch rows
 * of out_words packed-bit words (bit = 1 for +1.0). */
int dsdhip_das_run_dev(dsdhip_das_t *d, const double *d_in, int64_t in_stride, size_t F,
                       unsigned long long *d_out, int64_t out_words, size_t *out_count);
/* Host buffers: in nch x F doubles, out nch x F floats: *out_count
 * samples of +-1.0 per row, then 0.0f up to F. */
int dsdhip_das_run(dsdhip_das_t *d, const double *in, size_t F, float *out, size_t *out_count);
/* Fresh persistent contexts (sdm_context_reset, include/trellis.h:86). */
int dsdhip_das_reset(dsdhip_das_t *d);
/* HIP-event timing of the segment trellis launches (see dsdhip_trellis_timing). */
int dsdhip_das_timing(dsdhip_das_t *d, int on);
int dsdhip_das_timing_read(dsdhip_das_t *d, double *ms, int64_t *launches, int64_t *samples, int clear);
/* Batched form: ntracks tracks of nch channels each, every track with its
 * own persistent contexts (no reference counterpart; the reference runs one
 * DAS per plugin instance, src/dsp_plugin.c:1721-2761).  One call runs the
 * next chunk of every track with F[t] > 0: all segments of all tracks in
 * one trellis launch, all stitch boundaries in one launch. */
dsdhip_das_t *dsdhip_das_create_tracks(int ntf_index, int depth, int cands, int lat,
                                       double state_limit, int nch, int ntracks, size_t seg_len,
                                       int nseg, int seg0_first, int reposition);
int dsdhip_das_run_tracks_dev(dsdhip_das_t *d, const double *const *d_in, const int64_t *in_stride,
                              const size_t *F, unsigned long long *const *d_out,
                              const int64_t *out_words, size_t *out_count);
/* Segment count, literal-prune steps and stitch positions of the last run. */
int dsdhip_das_last_info(const dsdhip_das_t *d, int *nseg, int64_t *slow_steps, int *stitch_pos,
                         int max_pos);
/* seg0-first seeding runs its estimate (src/threadpool.c:345-366) on host
 * threads, one chain per channel, and launches each segment as soon as the
 * estimate has passed its start; on = 1 runs the estimate on one device lane
 * per channel before one launch instead (same bits; for timing).  No
 * reference counterpart. */
int dsdhip_das_estimator_device(dsdhip_das_t *d, int on);
/* seg0-first launch form: on = 0 (default) one persistent launch whose waves
 * each take the next segment and wait (bounded) until the host has
 * published its estimate, so every segment starts as soon as its seed
 * state exists; on = 1 batches of ready segments, each batch launched when
 * the previous one has finished (the form used at K <= 2 and with the FMA
 * filter).  Same bits.  No reference counterpart. */
int dsdhip_das_batch_launch(dsdhip_das_t *d, int on);
/* Longest wait, in seconds (default 20), of one wave of the host-fed launch
 * for its segment; a wave that waits longer leaves and the run returns -1.
 * No reference counterpart. */
int dsdhip_das_feed_limit(dsdhip_das_t *d, double seconds);
/* Resident waves per CU of the host-fed launch (default 6; 0: as many as
 * the kernel's registers allow, 8 up to 128-sample histories).  Fewer
 * waves per SIMD run the last
 * published segments faster; too few leave published segments waiting for
 * a slot.  No reference counterpart. */
int dsdhip_das_feed_waves(dsdhip_das_t *d, int waves_per_cu);
/* Host-side timeline of the last seg0-first run, ms from the run's start:
 * seg0 done, estimate threads started, last segment published, estimate
 * threads joined, trellis results read, run done.  No reference
 * counterpart. */
int dsdhip_das_last_timeline(const dsdhip_das_t *d, double *ms6);
/* Estimator figures of the last seg0-first run: longest chain's busy time,
 * samples estimated over all channels, blocks the fast form handed back to
 * the reference form, pipelined trellis launches. */
int dsdhip_das_last_estimate(const dsdhip_das_t *d, double *busy_ms, int64_t *samples,
                             int64_t *redo_blocks, int *batches);

/* ======================================================================
 * 6. Batched DAS: many tracks of uneven length, one trellis launch per
 *    chunk step.  No counterpart in the reference (its DAS runs one stream
 *    per plugin instance, src/dsp_plugin.c:1941-2761); a track is nch
 *    channels of float32 PCM in device memory.  Every chunk step
 *    (chunk_frames PCM frames of every unfinished track) runs one FIR chain
 *    over all streams, the segment plans with reposition, one trellis launch
 *    over all segments of all channels and tracks (with seg0-first seeding:
 *    seg0 first, then the pipelined rest), one stitch launch and the
 *    assembly; the host orchestrates the step.  Per track and chunk the
 *    result equals the single-stream DAS with the reference's seeding for
 *    fs_out (seg0-first below DSD512, src/dsp_plugin.c:2185) and the same
 *    segment length.  A device-side form of the step, made to be captured
 *    once in a hipGraph and replayed per chunk, is built but switched off:
 *    plans asking for it (use_graph = 1) run the host step and report
 *    use_graph = 0.
 * ====================================================================== */
typedef struct {
    int seg_len;            /* segment length; 0: chosen by the autotuner over the enqueued tracks */
    int waves_per_group;    /* device step's persistent launch: waves per workgroup (1, 2, 4); 0: tuned */
    int waves_per_cu;       /* device step's persistent launch: resident waves per CU; 0: tuned */
    int use_graph;          /* in: ask for the device step; out: whether it runs (0 in this build) */
    double msps_measured;   /* out: trellis Msamples/s of the tuned launch shape (trial) */
} gpu_hip_tune_t;

typedef struct gpu_hip_plan gpu_hip_plan_t;

/* Plan for tracks of nch channels, fs_in -> fs_out (2^n ratio) with the
 * chain gain, trellis (ntf_index, depth, cands, lat, state_limit);
 * tune = NULL tunes everything. */
gpu_hip_plan_t *gpu_hip_plan_create(unsigned fs_in, unsigned fs_out, double gain, int ntf_index,
                                    int depth, int cands, int lat, double state_limit, int nch,
                                    size_t chunk_frames, int reposition,
                                    const gpu_hip_tune_t *tune);
/* Use the caller's HIP stream (not the null stream) instead of the plan's own. */
int gpu_hip_plan_set_stream(gpu_hip_plan_t *p, void *stream);
/* Registers the tracks: d_pcm[t] = [nch][pcm_stride[t]] float32 device
 * memory with frames[t] frames; d_out[t] = [nch][out_stride[t]] u64 words
 * receiving the packed DSD bits (bit = 1 for +1), out_words[t] words per
 * channel.  NULL strides mean frames[t] / out_words[t]. */
int gpu_hip_plan_enqueue(gpu_hip_plan_t *p, int ntracks, const float *const *d_pcm,
                         const int64_t *pcm_stride, const int64_t *frames, uint64_t *const *d_out,
                         const int64_t *out_stride, const int64_t *out_words);
/* Runs up to max_chunks chunk steps (< 0: all that remain); the first call
 * tunes, allocates and captures.  out_bits[t] receives each track's output
 * bit count so far.  Returns the chunks run, or -1. */
int gpu_hip_plan_launch_graph(gpu_hip_plan_t *p, int max_chunks, int64_t *out_bits);
int gpu_hip_plan_get_tune(const gpu_hip_plan_t *p, gpu_hip_tune_t *out);
int gpu_hip_plan_info(const gpu_hip_plan_t *p, int64_t *chunks_total, int64_t *chunks_done,
                      int *segments_per_chunk, double *elapsed_ms);
void gpu_hip_plan_destroy(gpu_hip_plan_t *p);

/* ======================================================================
 * 8. UPOLS convolution (uniformly partitioned overlap-save), replaces
 *    conv_init / conv_load_ir / conv_process (src/convolution.c:530-1041)
 *    for a given IR per channel at the signal rate: partitions of P
 *    samples, FFT size 2P, complex FP32 spectra.  Input and output are
 *    device f64 [nch][stride].
 * ====================================================================== */
typedef struct dsdhip_conv dsdhip_conv_t;
/* ir: host f64 [nch][ir_stride], ir_len taps per channel. */
dsdhip_conv_t *dsdhip_conv_create(int nch, int P, const double *ir, int64_t ir_len, int64_t ir_stride);
void dsdhip_conv_destroy(dsdhip_conv_t *c);
/* replaces conv_reset (src/convolution.c:1036): clears the delay line. */
int dsdhip_conv_reset(dsdhip_conv_t *c);
int dsdhip_conv_set_stream(dsdhip_conv_t *c, void *hip_stream);
int dsdhip_conv_partitions(const dsdhip_conv_t *c);
/* y = x * h for the next n samples of every channel, n a multiple of P
 * (stream-ordered); -1 while the FIFO form below holds samples. */
int dsdhip_conv_process_dev(dsdhip_conv_t *c, const double *d_in, int64_t in_stride, size_t n,
                            double *d_out, int64_t out_stride);
/* Any n (replaces conv_process_direct, src/convolution.c:893-958): the
 * chunk joins the pending input, each whole block of P samples is
 * convolved and its outputs join those waiting; exactly n waiting samples
 * are written to d_out, or, when fewer are waiting (no latency is
 * inserted), all of them followed by the chunk's own input at the rest of
 * the positions, with nothing carried over.  d_out may equal d_in.  Chunks
 * that are multiples of P give dsdhip_conv_process_dev's output. */
int dsdhip_conv_process_fifo_dev(dsdhip_conv_t *c, const double *d_in, int64_t in_stride, size_t n,
                                 double *d_out, int64_t out_stride);
/* Samples the FIFO form holds: pending input, produced output waiting. */
int dsdhip_conv_fifo_state(const dsdhip_conv_t *c, int64_t *pending_in, int64_t *waiting_out);

/* ======================================================================
 * 7. Differentiable loop filter and pre-emphasis (training).
 *    Replaces the per-sample torch loop of training_pipeline/diff_sdm.py:
 *    61-100 (DiffSDM.forward) and the conv1d of model_pre_sdm.py:44-53
 *    (FixedPreEmphasis.forward).  All pointers are device memory owned by
 *    the caller; sequences are time-major [N][B] float64; the NTF
 *    coefficients a[order], g[order] are device float64.  Nothing here
 *    allocates; calls are asynchronous on `stream` (a hipStream_t, NULL =
 *    default stream).  0 = success, -1 = bad arguments or launch failure.
 * ====================================================================== */
/* Number of state checkpoints (one every ck_every samples). */
size_t dsdhip_lf_checkpoints(size_t N, int ck_every);
/* Workspace bytes of dsdhip_lf_backward processing slab_chunks checkpoint
 * intervals per slab (<= 0: all of them in one slab). */
size_t dsdhip_lf_workspace_bytes(int B, size_t N, int order, int ck_every, int slab_chunks);
/* y = relaxed loop filter of u (quantizer tanh(v * temperature)); writes
 * the state before every ck_every-th sample to ck[nck][order][B]. */
int dsdhip_lf_forward(const double *d_a, const double *d_g, int order, double temperature,
                      const double *d_u, int B, size_t N, int ck_every, double *d_y,
                      double *d_ck, void *stream);
/* Adjoint: given gy = dL/dy, writes gu = dL/du and ga, gg = dL/da, dL/dg
 * (order values each), recomputing states from the checkpoints. */
int dsdhip_lf_backward(const double *d_a, const double *d_g, int order, double temperature,
                       const double *d_u, int B, size_t N, int ck_every, const double *d_ck,
                       const double *d_gy, double *d_gu, double *d_ga, double *d_gg, void *d_ws,
                       size_t ws_bytes, int slab_chunks, void *stream);
/* dsdhip_lf_backward with trunc_span > 0: the adjoint restarts from zero at
 * the start of every trunc_span-sample interval (the gradient with the state
 * at each interval start held constant, truncated backpropagation through
 * time); the intervals then run in parallel.  trunc_span divides ck_every
 * and is a multiple of gcd(ck_every, 64); 0 is dsdhip_lf_backward. */
int dsdhip_lf_backward_ex(const double *d_a, const double *d_g, int order, double temperature,
                          const double *d_u, int B, size_t N, int ck_every, const double *d_ck,
                          const double *d_gy, double *d_gu, double *d_ga, double *d_gg, void *d_ws,
                          size_t ws_bytes, int slab_chunks, int trunc_span, void *stream);
/* Pre-emphasis: u[n][b] = sum_k w[k] p[b][n-k] (p row-major [B][N] float32,
 * zero history, 1 <= ntaps <= 128). */
size_t dsdhip_pe_workspace_bytes(int B, size_t N, int ntaps);
int dsdhip_pe_forward(const double *d_w, int ntaps, const float *d_p, int B, size_t N,
                      double *d_u, void *stream);
/* Tap gradient gw[k] = sum_{b,n} gu[n][b] p[b][n-k]. */
int dsdhip_pe_backward(const float *d_p, int B, size_t N, int ntaps, const double *d_gu,
                       double *d_gw, void *d_ws, size_t ws_bytes, void *stream);

/* ======================================================================
 * 9. Streaming chain for one multichannel stream: FIR upsampling, optional
 *    UPOLS convolution, DAS, per chunk on one HIP stream.  Replaces the
 *    per-channel engine_process_block chain (src/engine.c:478-729: FIR with
 *    gain -> conv_process -> SDM) with DAS as the SDM, as the GPU plugin
 *    path runs it (src/dsp_plugin.c:1941-2761).  The engine's soft clip
 *    (src/engine.c:704-716) is not applied.
 * ====================================================================== */
typedef struct dsdhip_stream dsdhip_stream_t;
#define DSDHIP_IN_PCM 0   /* float32 PCM, [nch][stride] */
#define DSDHIP_IN_DSD 1   /* packed DSD bits, [nch][words] */
/* Engine-level configuration (the fields of dsd_config_t and the engine
 * channel this chain needs, include/engine.h / src/engine.c:117-413). */
typedef struct {
    unsigned fs_in, fs_out;     /* fs_out == fs_in: same-rate DSD re-encode */
    int nch;
    size_t chunk_frames;        /* input samples per channel per chunk */
    int input;                  /* DSDHIP_IN_PCM or DSDHIP_IN_DSD */
    double fir_gain;            /* engine fir_gain (path table 0.708; PCM default 0.354) */
    double gain;                /* volume (cfg->gain); changes ramp over the next chunk */
    int pcm_boxcar;             /* PCM: boxcar taps before the FIR (engine_process_block: 8; 0 off) */
    int soft_clip;              /* rate conversion: the engine's tanh soft clip (1 on) */
    int dsd_boxcar;             /* DSD same-rate: boxcar taps (0: the engine's by rate);
                                   DSD rate conversion: pre-smooth taps (0 off, engine: 8) */
    int dsd_lowpass;            /* DSD same-rate: 127-tap lowpass instead of the boxcar */
    const double *ir;           /* room IR [nch][ir_stride] at fs_out, or NULL */
    int64_t ir_len, ir_stride;
    int P;                      /* convolution partition */
    int ntf_index, depth, cands, lat;
    double state_limit;
    size_t seg_len;
    int reposition;
    int fma_o8;                 /* order 8: the AVX2 build's FMA filter */
    int seeding;                /* DAS seeding: 0 the reference's rule by fs_out (seg0-first
                                   below DSD512, src/dsp_plugin.c:2185), 1 seg0-first,
                                   2 every segment from the persistent state */
} dsdhip_stream_cfg_t;
dsdhip_stream_t *dsdhip_stream_create_cfg(const dsdhip_stream_cfg_t *cfg);
/* volume: the next chunk ramps from the previous gain to this one */
int dsdhip_stream_set_gain(dsdhip_stream_t *s, double gain);
/* DSD input chunk: d_in nch rows of packed bits (words_stride words), n
 * samples per channel (<= chunk_frames); output as for PCM. */
int dsdhip_stream_process_dsd_dev(dsdhip_stream_t *s, const unsigned long long *d_in, int64_t words_stride,
                                  size_t n, unsigned long long *d_bits, int64_t out_words, size_t *out_count);
/* ir == NULL: no convolution.  Otherwise ir is host f64 [nch][ir_stride]
 * and chunk_frames * 2^stages must be a multiple of 8*P.  The DAS seeds by
 * the reference's rule for fs_out (seeding 0 of dsdhip_stream_cfg_t) with
 * segments of seg_len samples. */
dsdhip_stream_t *dsdhip_stream_create(unsigned fs_in, unsigned fs_out, double gain, int nch,
                                      size_t chunk_frames, const double *ir, int64_t ir_len,
                                      int64_t ir_stride, int P, int ntf_index, int depth, int cands,
                                      int lat, double state_limit, size_t seg_len, int reposition);
void dsdhip_stream_destroy(dsdhip_stream_t *s);
int dsdhip_stream_set_stream(dsdhip_stream_t *s, void *hip_stream);
/* replaces engine_channel_reset (src/engine.c:732) for every channel */
int dsdhip_stream_reset(dsdhip_stream_t *s);
int dsdhip_stream_ratio(const dsdhip_stream_t *s);
/* One chunk: d_pcm float32 [nch][pcm_stride], frames <= chunk_frames;
 * d_bits nch rows of out_words packed words (bit = 1 for +1.0). */
int dsdhip_stream_process_dev(dsdhip_stream_t *s, const float *d_pcm, int64_t pcm_stride,
                              size_t frames, unsigned long long *d_bits, int64_t out_words,
                              size_t *out_count);
/* SDM input of the last chunk into dst [nch][dst_stride] doubles (host or
 * device); returns the samples per channel or -1.  Synchronises. */
int64_t dsdhip_stream_copy_signal(dsdhip_stream_t *s, double *dst, int64_t dst_stride);
/* Host-side timeline of the last chunk, ms from the chunk call's start:
 * DAS run started, then the six points of dsdhip_das_last_timeline (zero
 * where the chunk was seeded all-parallel) shifted to the chunk's start.
 * No reference counterpart. */
int dsdhip_stream_last_timeline(const dsdhip_stream_t *s, double *ms7);

/* ======================================================================
 * 10. DSD wire formats: DoP and raw bit bytes (include/dop.h:12-48,
 *     src/dop.c).  Device DSD streams are packed bits: sample i of a
 *     channel is bit i % 64 of 64-bit word i / 64 of that channel's row,
 *     1 = +1.0.  Bits past a stream's count are 0.
 * ====================================================================== */
/* replaces dop_detect_interleaved (include/dop.h:21; channels == 1 is
 * dop_detect, :16): host frames, returns 1 if DoP markers are present. */
int dsdhip_dop_detect(const float *pcm24, size_t frames, int channels);
/* replaces the per-channel de-interleave + dop_unpack of the plugin's input
 * stage (src/dsp_plugin.c:1254-1276, include/dop.h:26): interleaved DoP
 * frames [frames][nch] -> packed bits [nch][words_stride], 16 per frame. */
int dsdhip_dop_unpack_dev(const float *d_pcm24, size_t frames, int nch, unsigned long long *d_bits,
                          int64_t words_stride, void *stream);
/* replaces dop_pack (include/dop.h:31) per channel, interleaved output
 * [bit_count/16][nch] floats; markers alternate from 0x05 at frame 0. */
int dsdhip_dop_pack_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t bit_count,
                        float *d_pcm24, void *stream);
/* replaces dop_pack_i24 + interleave of the plugin's output stage
 * (include/dop.h:40, src/dsp_plugin.c:2869-2945): [bit_count/16][nch][3]
 * bytes (dsd_lo, dsd_hi, marker); returns the next marker phase or -1. */
int dsdhip_dop_pack_i24_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t bit_count,
                            int marker_phase, unsigned char *d_out, void *stream);
/* replace bits_unpack / bits_pack (include/dop.h:44-48): raw MSB-first
 * bytes [nch][stride] <-> packed bits. */
int dsdhip_bytes_unpack_dev(const unsigned char *d_src, int64_t src_stride, int nch, size_t n_bits,
                            unsigned long long *d_bits, int64_t words_stride, void *stream);
int dsdhip_bytes_pack_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t n_bits,
                          unsigned char *d_dst, int64_t dst_stride, void *stream);
/* packed bits -> +-1.0f [nch][out_stride] (the plugin's ch_in form). */
int dsdhip_bits_to_f32_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t n,
                           float *d_out, int64_t out_stride, void *stream);

/* packed bits <- +-1.0f samples (x >= 0 -> 1, as src/engine.c:515 reads
 * DSD input). */
int dsdhip_f32_to_bits_dev(const float *d_in, int64_t in_stride, int nch, size_t n, unsigned long long *d_bits,
                           int64_t words_stride, void *stream);

/* ======================================================================
 * 11. Engine pre-SDM conditioning (src/engine.c:117-899): the boxcar of
 *     a same-rate DSD stream (exact window sums of +-1), the 127-tap
 *     lowpass alternative (src/fir.c:795-873), the boxcar pre-smooth of
 *     PCM before rate conversion (a running double sum, sequential per
 *     channel), the gain ramp and the tanh soft clip.  All FP64 in the
 *     reference's operation order.  Histories persist between calls like
 *     the engine channel's boxcar ring and lowpass delay line.
 * ====================================================================== */
typedef struct dsdhip_cond dsdhip_cond_t;
/* fir_lowpass_init's design for a DSD rate (127 taps); returns 127. */
int dsdhip_lowpass_taps(unsigned rate, double *h);
/* nch channels, boxcar_taps (1..64; the engine's taps, src/engine.c:131-
 * 147), lowpass_rate 0 = no lowpass, else its design rate. */
dsdhip_cond_t *dsdhip_cond_create(int nch, int boxcar_taps, unsigned lowpass_rate);
/* the same with the caller's lowpass taps (1..128) in place of the design */
dsdhip_cond_t *dsdhip_cond_create_taps(int nch, int boxcar_taps, const double *h, int ntaps);
void dsdhip_cond_destroy(dsdhip_cond_t *c);
/* replaces the boxcar / lowpass part of engine_channel_reset (:732-757) */
int dsdhip_cond_reset(dsdhip_cond_t *c);
int dsdhip_cond_set_stream(dsdhip_cond_t *c, void *hip_stream);
/* same-rate DSD branch of engine_process_fir_gain (:810-846): packed bits
 * -> f64, boxcar (or lowpass) times the ramp gs -> ge over the n samples */
int dsdhip_cond_dsd_boxcar_dev(dsdhip_cond_t *c, const unsigned long long *d_bits, int64_t words_stride,
                               size_t n, double gs, double ge, double *d_out, int64_t out_stride);
int dsdhip_cond_dsd_lowpass_dev(dsdhip_cond_t *c, const unsigned long long *d_bits, int64_t words_stride,
                                size_t n, double gs, double ge, double *d_out, int64_t out_stride);
/* boxcar pre-smooth of PCM before the FIR chain (:629-642), no gain */
int dsdhip_cond_pcm_boxcar_dev(dsdhip_cond_t *c, const float *d_in, int64_t in_stride, size_t n,
                               double *d_out, int64_t out_stride);
/* gain ramp in place: x[i] *= gs + (ge - gs) / (n - 1) * i (:646-652) */
int dsdhip_gain_ramp_dev(double *d_x, int64_t stride, int nch, size_t n, double gs, double ge, void *stream);
/* soft clip in place above +-0.95 with a tanh knee (:704-716) */
int dsdhip_soft_clip_dev(double *d_x, int64_t stride, int nch, size_t n, void *stream);

/* ======================================================================
 * 12. PreCorr SDM (src/precorr.c, include/precorr.h): greedy 1-bit
 *     quantizer through the NTF loop filter in single precision plus the
 *     8-bit-history x 8-phase prediction-table correction.  No latency.
 *     One sequential recurrence per channel on the device.
 * ====================================================================== */
/* Same layout as gpu_precorr_state_t (include/gpu_compute.h:128-136). */
typedef struct {
    float state[8];
    float prev_y;
    int history;
    int phase;
    float pad;
} dsdhip_precorr_state_t;
typedef dsdhip_precorr_state_t gpu_hip_precorr_state_t;
/* precorr_train_table (src/precorr.c:134-186) for float coefficients:
 * table[256][8] of mean corrections. */
int dsdhip_precorr_train(const float *a, const float *g, int order, float *table);
typedef struct dsdhip_precorr dsdhip_precorr_t;
/* replaces precorr_context_init (include/precorr.h:40) for nch channels;
 * state_limit as precorr_context_t.state_limit (0 = off) */
dsdhip_precorr_t *dsdhip_precorr_create(int ntf_index, int nch, float state_limit);
dsdhip_precorr_t *dsdhip_precorr_create_coeffs(const double *a, const double *g, int order, int nch,
                                               float state_limit);
void dsdhip_precorr_destroy(dsdhip_precorr_t *p);
/* replaces precorr_context_reset (include/precorr.h:49) on every channel */
int dsdhip_precorr_reset(dsdhip_precorr_t *p);
int dsdhip_precorr_set_stream(dsdhip_precorr_t *p, void *hip_stream);
int dsdhip_precorr_table(const dsdhip_precorr_t *p, float *table);
int dsdhip_precorr_get_state(dsdhip_precorr_t *p, int ch, dsdhip_precorr_state_t *st);
int dsdhip_precorr_set_state(dsdhip_precorr_t *p, int ch, const dsdhip_precorr_state_t *st);
/* replaces precorr_process_block (include/precorr.h:43) for every channel:
 * d_in [nch][in_stride] doubles, n samples each -> packed bits
 * [nch][words_stride] (n out per channel, no latency) */
int dsdhip_precorr_process_dev(dsdhip_precorr_t *p, const double *d_in, int64_t in_stride, size_t n,
                               unsigned long long *d_bits, int64_t words_stride);
/* replaces gpu_cuda_precorr_process (include/gpu_compute.h:140-144): one
 * channel, host buffers, float coefficients and table from the caller,
 * state in and out; outputs +-1.0.  No state limit (the seam has none). */
int gpu_hip_precorr_process(void *ctx, const float *in, float *out, size_t count, const float *ntf_a,
                            const float *ntf_g, int order, const float pred_table[256][8],
                            const gpu_hip_precorr_state_t *init, gpu_hip_precorr_state_t *final_state);

/* Library version string. */
const char *dsdhip_version(void);

#ifdef __cplusplus
}
#endif
#endif
