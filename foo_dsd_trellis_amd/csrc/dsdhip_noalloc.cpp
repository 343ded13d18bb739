/*
 * Trellis job launch that allocates nothing: dsdhip_trellis_run_jobs with
 * its job descriptor table (the per-job staging the kernel reads and whose
 * result fields it writes) in a file-scope device array instead of the
 * pool's grow-on-demand table.  For callers that cannot allocate device
 * memory around the launch (code under graph capture, tight serving
 * loops).
 *
 * The array is one per GPU, held until the process ends; its address on
 * each device comes from hipGetSymbolAddress at the first call on that
 * device (with that device current) and is what the kernels get.  Every
 * call on one GPU shares it, and the host staging of the descriptors, so
 * calls that overlap on one GPU race; each call writes every descriptor it
 * launches (result fields zeroed) before the launch.
 */
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdio.h>
#include <string.h>
#include "dsdhip_internal.h"

extern "C" int tr_launch_jobs(const tr_params *prm, tr_job *d_jobs, int njobs, tr_ctx *d_ctxs,
                              hipStream_t stream);
extern "C" hipStream_t dsdhip_trellis_stream(const dsdhip_trellis_t *t);

#define NOALLOC_MAX_DEV 64

static __device__ __align__(256) tr_job scratch[DSDHIP_NOALLOC_MAX_JOBS];
static tr_job h_stage[DSDHIP_NOALLOC_MAX_JOBS];     /* host side of the descriptors */
static tr_job *g_scratch_ptr[NOALLOC_MAX_DEV];     /* device address of `scratch` per GPU */
static std::mutex g_mu;

/* the current device's copy of `scratch` (looked up once per device) */
static tr_job *scratch_on_current_device()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV) return NULL;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = NULL;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NULL;
        g_scratch_ptr[dev] = (tr_job *)p;
    }
    return g_scratch_ptr[dev];
}

extern "C" int dsdhip_trellis_run_jobs_noalloc(dsdhip_trellis_t *t, int njobs, const int *ctx,
                                               const int *copy_src, const double *const *d_in,
                                               const int64_t *n_in, const int64_t *keep_from,
                                               uint64_t *const *d_out, const double *seed, int64_t *kept,
                                               int64_t *slow_steps, int *err_out)
{
    if (!t || njobs < 0 || (njobs > 0 && (!ctx || !d_in || !n_in || !d_out))) return -1;
    if (njobs > DSDHIP_NOALLOC_MAX_JOBS) {
        fprintf(stderr, "dsdhip_trellis_run_jobs_noalloc: %d jobs exceed the static job table of %d\n", njobs,
                DSDHIP_NOALLOC_MAX_JOBS);
        return -1;
    }
    if (njobs == 0) return 0;
    tr_job *d_jobs = scratch_on_current_device();
    if (!d_jobs) return -1;
    for (int j = 0; j < njobs; j++) {
        const int src = (copy_src && copy_src[j] >= 0) ? copy_src[j] : -1;
        if (dsdhip_trellis_fill_job(t, &h_stage[j], ctx[j], src, d_in[j], n_in[j], keep_from ? keep_from[j] : 0,
                                    d_out[j], seed ? seed + (size_t)j * 10 : NULL) != 0)
            return -1;
    }
    tr_params prm;
    if (dsdhip_trellis_params(t, &prm) != 0) return -1;
    const hipStream_t s = dsdhip_trellis_stream(t);
    tr_ctx *d_ctx = (tr_ctx *)dsdhip_trellis_context_ptr(t, 0);
    if (hipMemcpyAsync(d_jobs, h_stage, sizeof(tr_job) * (size_t)njobs, hipMemcpyHostToDevice, s) != hipSuccess ||
        tr_launch_jobs(&prm, d_jobs, njobs, d_ctx, s) != 0 ||
        hipMemcpyAsync(h_stage, d_jobs, sizeof(tr_job) * (size_t)njobs, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return -1;
    int err = 0;
    for (int j = 0; j < njobs; j++) {
        if (kept) kept[j] = h_stage[j].kept;
        if (slow_steps) slow_steps[j] = h_stage[j].slow_steps;
        if (h_stage[j].err && !err) err = h_stage[j].err;
        const int src = (copy_src && copy_src[j] >= 0) ? copy_src[j] : -1;
        if (src >= 0) dsdhip_trellis_set_snap(t, ctx[j], src);
    }
    if (err_out) *err_out = err;
    return err ? -1 : 0;
}
