/* Internal declarations shared by the host-side translation units. */
#ifndef DSDHIP_INTERNAL_H
#define DSDHIP_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include "tr_types.h"
#include "../../include/dsd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Selects the device on first use; 0 on success. */
int dsdhip_ensure_device(void);

int dsdhip_trellis_get(dsdhip_trellis_t *t, int i, tr_ctx *out);
int dsdhip_trellis_put(dsdhip_trellis_t *t, int i, const tr_ctx *in);
int dsdhip_trellis_clone(dsdhip_trellis_t *t, int dst, int src);
int dsdhip_trellis_detach(dsdhip_trellis_t *t, int i);
int dsdhip_trellis_copy_detached(dsdhip_trellis_t *t, int dst, int src);
int dsdhip_trellis_params(const dsdhip_trellis_t *t, tr_params *out);
void *dsdhip_trellis_context_ptr(dsdhip_trellis_t *t, int i);
/* pipelined job table (dsdhip_trellis.cpp) */
int dsdhip_trellis_table_begin(dsdhip_trellis_t *t, int njobs);
int dsdhip_trellis_table_fill(dsdhip_trellis_t *t, int j, int ctx, int copy_src, const double *d_in,
                              int64_t n_in, int64_t keep_from, uint64_t *d_out, const double *seed10);
int dsdhip_trellis_table_launch(dsdhip_trellis_t *t, int first, int n, void *stream);
int dsdhip_trellis_table_end(dsdhip_trellis_t *t, void *stream, int64_t *kept, int64_t *slow_steps,
                             int *err_out);
/* host-fed launch over the table: one persistent launch whose waves wait for
 * the entries the host publishes (0: running, 1: not applicable, -1 error) */
int dsdhip_trellis_feed_begin(dsdhip_trellis_t *t, void *stream, double limit_s, int waves_per_cu);
int dsdhip_trellis_feed_publish(dsdhip_trellis_t *t, int n);
int dsdhip_trellis_feed_abort(dsdhip_trellis_t *t);
int dsdhip_trellis_timing_add(dsdhip_trellis_t *t, double ms, int64_t launches);
int dsdhip_trellis_fill_job(const dsdhip_trellis_t *t, tr_job *jb, int ctx, int src, const double *d_in,
                            int64_t n_in, int64_t keep_from, uint64_t *d_out, const double *seed10);
int dsdhip_trellis_set_snap(dsdhip_trellis_t *t, int i, int src);
/* dsdhip_trellis_run_jobs that calls queued(arg) once the launch and its
 * read-back are queued, before it waits for them (queued returns 0, or
 * nonzero to fail the call after the wait) */
int dsdhip_trellis_run_jobs_then(dsdhip_trellis_t *t, int njobs, const int *ctx, const int *copy_src,
                                 const double *const *d_in, const int64_t *n_in, const int64_t *keep_from,
                                 uint64_t *const *d_out, const double *seed, int64_t *kept,
                                 int64_t *slow_steps, int *err_out, int (*queued)(void *), void *arg);
/* 255-tap Kaiser low-pass of the stitch quality (src/dsp_plugin.c:2571-2604). */
void dsdhip_make_qfir(double *q);
/* Device scratch of the FIR chain for n input samples per channel, so later
 * dsdhip_fir_process_dev calls of at most n samples allocate nothing. */
int dsdhip_fir_reserve(dsdhip_fir_t *f, size_t n);
/* f64 -> f32 copy on the device (round to nearest). */
int dsdhip_f64_to_f32_dev(const double *d_x, float *d_y, size_t n, void *stream);

/* Host estimator of sdm_estimate_state_multibit fed block by block
 * (dsdhip_estimate.cpp); one object per channel, one thread at a time. */
typedef struct dsdhip_est dsdhip_est_t;
dsdhip_est_t *dsdhip_est_create(const double *a, const double *g, int order, double state_limit,
                                int levels, const double *init10);
void dsdhip_est_destroy(dsdhip_est_t *e);
void dsdhip_est_set_state(dsdhip_est_t *e, const double *init10);
void dsdhip_est_get_state(const dsdhip_est_t *e, double *out10);
void dsdhip_est_feed(dsdhip_est_t *e, const double *x, size_t n);
int64_t dsdhip_est_redo_blocks(const dsdhip_est_t *e);
/* the estimate over x[0..count), the state at each boundary bnd[0..nb)
 * handed to publish(ctx, b, state); wait(ctx, pos, end) (optional) returns
 * 0 once x[pos..end) may be read.  With threads = 2 the index guess runs
 * on a second thread where that applies (no limiter, fast table, FMA and
 * AVX-512 hardware). */
int dsdhip_est_chain(dsdhip_est_t *e, const double *x, size_t count, const int64_t *bnd, int nb,
                     void (*publish)(void *, int, const double *), int (*wait)(void *, size_t, size_t),
                     void *ctx, int threads);
int dsdhip_est_fast(const dsdhip_est_t *e);

#ifdef __cplusplus
}
#endif
#endif
