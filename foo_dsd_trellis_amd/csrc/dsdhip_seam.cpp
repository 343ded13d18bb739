/*
 * The gpu_compute.h backend seam for HIP (include/gpu_compute.h:22-59,
 * 185-233): an opaque context whose first field is the backend tag, holding
 * the FIR chain, a one-context trellis and a DAS pipeline.  Host buffers in,
 * host buffers out; 0 / -1 like the reference's backends.
 */
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include "dsdhip_internal.h"

#define GPU_MIN_SAMPLES 8192   /* include/gpu_compute.h:40 */
#define SEAM_MAX_CH 8          /* channel slots of the per-call lowpass / boxcar */

struct gpu_hip_context {
    int backend;   /* GPU_BACKEND_HIP, read by the dispatcher (src/gpu_compute.c:153-157) */
    dsdhip_fir_t *fir;             /* gpu_hip_fir_batch: all channels at once */
    int fir_nch;
    /* gpu_hip_fir_chain(_f64): one channel per call, in channel order; each
     * channel slot keeps its own delay lines (the slot advances per call,
     * gpu_hip_reset_chunk rewinds it) */
    dsdhip_fir_t *fir1[SEAM_MAX_CH];
    int fir_ch;
    unsigned fs_in, fs_out;
    double fir_gain;
    /* trellis setup */
    int have_tr;
    double a[10], g[10];
    int order, cands, lat, depth;
    double state_limit;
    dsdhip_trellis_t *tr;
    dsdhip_das_t *das;
    int das_nch;
    size_t das_seg_len;
    int das_seg0_first, das_reposition;   /* das_seg0_first < 0: by output rate */
    unsigned das_rate;                    /* output rate for that rule (0: unknown) */
    size_t das_count;                     /* samples per channel of the last DAS call */
    int fma_o8;
    /* uniform-taps FIR setup (gpu_hip_fir_setup) */
    double taps[127];
    int ntaps, nstages;
    /* same-rate lowpass and boxcar of DSD input: one history per channel
     * slot, the slot advancing per call and reset by gpu_hip_reset_chunk
     * (the seam processes one channel per call, in channel order) */
    double lp_taps[128];
    int lp_ntaps, lp_ch, box_taps, box_ch;
    dsdhip_cond_t *lp[SEAM_MAX_CH], *box[SEAM_MAX_CH];
    float *d_sin;
    unsigned long long *d_sbits;
    double *d_sout;
    size_t scap;
};

extern "C" {

void dsdhip_halfband_taps(double *h, int ntaps);

void *gpu_hip_create(void)
{
    if (dsdhip_ensure_device() != 0) return NULL;
    gpu_hip_context *c = (gpu_hip_context *)calloc(1, sizeof(gpu_hip_context));
    c->backend = GPU_BACKEND_HIP;
    c->das_seg_len = 65536;
    c->das_seg0_first = -1;
    c->das_reposition = 1;
    return c;
}

void gpu_hip_destroy(void *ctx)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c) return;
    for (int i = 0; i < SEAM_MAX_CH; i++) {
        dsdhip_cond_destroy(c->lp[i]);
        dsdhip_cond_destroy(c->box[i]);
        dsdhip_fir_destroy(c->fir1[i]);
    }
    if (c->d_sin) hipFree(c->d_sin);
    if (c->d_sbits) hipFree(c->d_sbits);
    if (c->d_sout) hipFree(c->d_sout);
    dsdhip_fir_destroy(c->fir);
    dsdhip_trellis_destroy(c->tr);
    dsdhip_das_destroy(c->das);
    free(c);
}

void gpu_hip_reset_chunk(void *ctx)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (c) c->fir_ch = c->lp_ch = c->box_ch = 0;
}

/* device staging for one channel of count samples: +-1.0 floats, their
 * packed bits, the f64 result */
static int seam_stage(gpu_hip_context *c, size_t count)
{
    if (c->scap >= count) return 0;
    if (c->d_sin) hipFree(c->d_sin);
    if (c->d_sbits) hipFree(c->d_sbits);
    if (c->d_sout) hipFree(c->d_sout);
    c->d_sin = NULL;
    c->d_sbits = NULL;
    c->d_sout = NULL;
    c->scap = 0;
    if (hipMalloc(&c->d_sin, sizeof(float) * count) != hipSuccess ||
        hipMalloc(&c->d_sbits, sizeof(unsigned long long) * ((count + 63) / 64)) != hipSuccess ||
        hipMalloc(&c->d_sout, sizeof(double) * count) != hipSuccess)
        return -1;
    c->scap = count;
    return 0;
}

/* replaces gpu_cuda_fir_lowpass_setup (include/gpu_compute.h:92) */
int gpu_hip_fir_lowpass_setup(void *ctx, const float *taps, int ntaps)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !taps || ntaps < 1 || ntaps > 128) return -1;
    for (int i = 0; i < ntaps; i++) c->lp_taps[i] = (double)taps[i];
    c->lp_ntaps = ntaps;
    for (int i = 0; i < SEAM_MAX_CH; i++) {
        dsdhip_cond_destroy(c->lp[i]);
        c->lp[i] = NULL;
    }
    c->lp_ch = 0;
    return 0;
}

/* One channel of +-1.0 DSD through the lowpass (history per channel slot),
 * times gain, to f64 on the device. */
static int seam_lowpass(gpu_hip_context *c, const float *in, size_t count, double gain)
{
    if (!c || !in || c->lp_ntaps < 1 || count < GPU_MIN_SAMPLES) return -1;
    const int ch = c->lp_ch % SEAM_MAX_CH;
    if (!c->lp[ch] && !(c->lp[ch] = dsdhip_cond_create_taps(1, 1, c->lp_taps, c->lp_ntaps))) return -1;
    if (seam_stage(c, count) != 0) return -1;
    c->lp_ch++;
    const int64_t words = (int64_t)(count + 63) / 64;
    if (hipMemcpy(c->d_sin, in, sizeof(float) * count, hipMemcpyHostToDevice) != hipSuccess) return -1;
    if (dsdhip_f32_to_bits_dev(c->d_sin, (int64_t)count, 1, count, c->d_sbits, words, NULL) != 0) return -1;
    return dsdhip_cond_dsd_lowpass_dev(c->lp[ch], c->d_sbits, words, count, gain, gain, c->d_sout,
                                       (int64_t)count);
}

/* replaces gpu_cuda_fir_lowpass_f64 (include/gpu_compute.h:102-103) */
int gpu_hip_fir_lowpass_f64(void *ctx, const float *in, double *out, size_t count, double gain)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!out || seam_lowpass(c, in, count, gain) != 0) return -1;
    return hipMemcpy(out, c->d_sout, sizeof(double) * count, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

/* replaces gpu_cuda_fir_lowpass (include/gpu_compute.h:97-98): the f64
 * result narrowed to float */
int gpu_hip_fir_lowpass(void *ctx, const float *in, float *out, size_t count, float gain)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!out || seam_lowpass(c, in, count, (double)gain) != 0) return -1;
    if (dsdhip_f64_to_f32_dev(c->d_sout, c->d_sin, count, NULL) != 0) return -1;
    return hipMemcpy(out, c->d_sin, sizeof(float) * count, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

/* replaces gpu_cuda_boxcar_smooth (include/gpu_compute.h:112-113): one
 * channel of +-1.0 DSD, boxcar of taps samples (history per channel slot)
 * times gain, narrowed to float */
int gpu_hip_boxcar_smooth(void *ctx, const float *in, float *out, size_t count, int taps, float gain)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !in || !out || taps < 1 || taps > 64 || count < GPU_MIN_SAMPLES) return -1;
    if (taps != c->box_taps) {
        for (int i = 0; i < SEAM_MAX_CH; i++) {
            dsdhip_cond_destroy(c->box[i]);
            c->box[i] = NULL;
        }
        c->box_taps = taps;
        c->box_ch = 0;
    }
    const int ch = c->box_ch % SEAM_MAX_CH;
    if (!c->box[ch] && !(c->box[ch] = dsdhip_cond_create(1, taps, 0))) return -1;
    if (seam_stage(c, count) != 0) return -1;
    c->box_ch++;
    const int64_t words = (int64_t)(count + 63) / 64;
    if (hipMemcpy(c->d_sin, in, sizeof(float) * count, hipMemcpyHostToDevice) != hipSuccess) return -1;
    if (dsdhip_f32_to_bits_dev(c->d_sin, (int64_t)count, 1, count, c->d_sbits, words, NULL) != 0) return -1;
    if (dsdhip_cond_dsd_boxcar_dev(c->box[ch], c->d_sbits, words, count, (double)gain, (double)gain, c->d_sout,
                                   (int64_t)count) != 0)
        return -1;
    if (dsdhip_f64_to_f32_dev(c->d_sout, c->d_sin, count, NULL) != 0) return -1;
    return hipMemcpy(out, c->d_sin, sizeof(float) * count, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

int gpu_hip_fir_setup(void *ctx, const float *taps, int ntaps, int num_stages, bool upsample)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !taps || !upsample || (ntaps != 63 && ntaps != 127) || num_stages < 1 || num_stages > 9)
        return -1;
    double ref[127];
    dsdhip_halfband_taps(ref, ntaps);
    int same = 1;
    for (int i = 0; i < ntaps; i++) if ((float)ref[i] != taps[i]) { same = 0; break; }
    for (int i = 0; i < ntaps; i++) c->taps[i] = same ? ref[i] : (double)taps[i];
    c->ntaps = ntaps;
    c->nstages = num_stages;
    dsdhip_fir_destroy(c->fir);
    c->fir = NULL;
    for (int i = 0; i < SEAM_MAX_CH; i++) {
        dsdhip_fir_destroy(c->fir1[i]);
        c->fir1[i] = NULL;
    }
    c->fir_ch = 0;
    c->fs_in = c->fs_out = 0;
    c->fir_gain = 1.0;
    return 0;
}

int gpu_hip_fir_setup_rates(void *ctx, unsigned fs_in, unsigned fs_out, double gain)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c) return -1;
    int nt[9];
    if (dsdhip_fir_plan(fs_in, fs_out, nt) < 1) return -1;
    dsdhip_fir_destroy(c->fir);
    c->fir = NULL;
    for (int i = 0; i < SEAM_MAX_CH; i++) {
        dsdhip_fir_destroy(c->fir1[i]);
        c->fir1[i] = NULL;
    }
    c->fir_ch = 0;
    c->fs_in = fs_in;
    c->fs_out = fs_out;
    if (c->das_rate != fs_out) {
        c->das_rate = fs_out;
        dsdhip_das_destroy(c->das);
        c->das = NULL;
    }
    c->fir_gain = gain;
    c->ntaps = 0;
    return 0;
}

static dsdhip_fir_t *make_fir(gpu_hip_context *c, int nch)
{
    if (c->fs_in) return dsdhip_fir_create(c->fs_in, c->fs_out, c->fir_gain, nch);
    if (c->ntaps) return dsdhip_fir_create_taps(c->taps, c->ntaps, c->nstages, 1.0, nch);
    return NULL;
}

static int ensure_fir(gpu_hip_context *c, int nch)
{
    if (c->fir && c->fir_nch == nch) return 0;
    dsdhip_fir_destroy(c->fir);
    c->fir = make_fir(c, nch);
    c->fir_nch = nch;
    return c->fir ? 0 : -1;
}

/* the FIR chain of this call's channel slot */
static dsdhip_fir_t *channel_fir(gpu_hip_context *c)
{
    const int ch = c->fir_ch % SEAM_MAX_CH;
    if (!c->fir1[ch] && !(c->fir1[ch] = make_fir(c, 1))) return NULL;
    c->fir_ch++;
    return c->fir1[ch];
}

int gpu_hip_fir_chain_f64(void *ctx, const float *in, double *out, size_t in_count, size_t *out_count)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !in || !out || in_count < GPU_MIN_SAMPLES) return -1;
    dsdhip_fir_t *f = channel_fir(c);
    if (!f) return -1;
    if (dsdhip_fir_process(f, in, in_count, out) != 0) return -1;
    if (out_count) *out_count = in_count << dsdhip_fir_num_stages(f);
    return 0;
}

int gpu_hip_fir_chain(void *ctx, const float *in, float *out, size_t in_count, size_t *out_count)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !in || !out || in_count < GPU_MIN_SAMPLES) return -1;
    dsdhip_fir_t *f = channel_fir(c);
    if (!f) return -1;
    const size_t n = in_count << dsdhip_fir_num_stages(f);
    double *tmp = (double *)malloc(sizeof(double) * n);
    if (!tmp) return -1;
    int rc = dsdhip_fir_process(f, in, in_count, tmp);
    if (rc == 0) {
        for (size_t i = 0; i < n; i++) out[i] = (float)tmp[i];
        if (out_count) *out_count = n;
    }
    free(tmp);
    return rc;
}

int gpu_hip_fir_batch(void *ctx, const float *in_batch, float *out_batch, size_t samples_per_ch,
                      int num_channels, size_t *out_count_per_ch)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !in_batch || !out_batch || num_channels < 1 || samples_per_ch < GPU_MIN_SAMPLES) return -1;
    if (ensure_fir(c, num_channels) != 0) return -1;
    const size_t n = samples_per_ch << dsdhip_fir_num_stages(c->fir);
    double *tmp = (double *)malloc(sizeof(double) * n * (size_t)num_channels);
    if (!tmp) return -1;
    int rc = dsdhip_fir_process(c->fir, in_batch, samples_per_ch, tmp);
    if (rc == 0) {
        for (size_t i = 0; i < n * (size_t)num_channels; i++) out_batch[i] = (float)tmp[i];
        if (out_count_per_ch) *out_count_per_ch = n;
    }
    free(tmp);
    return rc;
}

int gpu_hip_trellis_setup(void *ctx, int num_cands, int order, int trellis_lat, const double *ntf_a,
                          const double *ntf_g, double state_limit, int trellis_depth)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !ntf_a || !ntf_g || order < 4 || order > 8) return -1;
    dsdhip_trellis_destroy(c->tr);
    c->tr = dsdhip_trellis_create_coeffs(ntf_a, ntf_g, order, trellis_depth, num_cands, trellis_lat,
                                         state_limit, 1);
    if (!c->tr) return -1;
    dsdhip_trellis_fma_o8(c->tr, c->fma_o8);
    memset(c->a, 0, sizeof(c->a));
    memset(c->g, 0, sizeof(c->g));
    memcpy(c->a, ntf_a, sizeof(double) * (size_t)order);
    memcpy(c->g, ntf_g, sizeof(double) * (size_t)order);
    c->order = order;
    c->cands = num_cands;
    c->lat = trellis_lat;
    c->depth = trellis_depth;
    c->state_limit = state_limit;
    c->have_tr = 1;
    dsdhip_das_destroy(c->das);
    c->das = NULL;
    return 0;
}

int gpu_hip_trellis(void *ctx, const double *in, float *out, size_t count)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !c->tr || !in || !out) return -1;
    size_t n = 0;
    return dsdhip_trellis_process(c->tr, 0, in, out, count, &n);
}

int gpu_hip_das_config(void *ctx, size_t seg_len, int seg0_first, int reposition)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c) return -1;
    c->das_seg_len = seg_len;
    c->das_seg0_first = seg0_first < 0 ? -1 : (seg0_first ? 1 : 0);
    c->das_reposition = reposition;
    dsdhip_das_destroy(c->das);
    c->das = NULL;
    return 0;
}

int gpu_hip_das_rate(void *ctx, unsigned fs_out)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c) return -1;
    if (c->das_rate != fs_out) {
        c->das_rate = fs_out;
        dsdhip_das_destroy(c->das);
        c->das = NULL;
    }
    return 0;
}

int gpu_hip_trellis_fma_o8(void *ctx, int on)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c) return -1;
    c->fma_o8 = on ? 1 : 0;
    if (c->tr) dsdhip_trellis_fma_o8(c->tr, c->fma_o8);
    if (c->das) dsdhip_das_fma_o8(c->das, c->fma_o8);
    return 0;
}

/* seed mode of the reference's DAS (src/dsp_plugin.c:2185): seg0-first
 * below DSD512, all segments from the persistent state at DSD512 and up */
static int seed_mode(const gpu_hip_context *c)
{
    if (c->das_seg0_first >= 0) return c->das_seg0_first;
    return (c->das_rate == 0 || c->das_rate < DSDHIP_DSD512_RATE) ? 1 : 0;
}

int gpu_hip_trellis_das_ex(void *ctx, const double *in, float *out, size_t count, int num_channels,
                           size_t *out_count)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    if (!c || !c->have_tr || !in || !out || num_channels < 1) return -1;
    if (!c->das || c->das_nch != num_channels) {
        dsdhip_das_destroy(c->das);
        c->das = dsdhip_das_create_coeffs(c->a, c->g, c->order, c->depth, c->cands, c->lat,
                                          c->state_limit, num_channels, c->das_seg_len,
                                          c->das_seg_len ? 0 : 2, seed_mode(c), c->das_reposition);
        c->das_nch = num_channels;
        if (!c->das) return -1;
        dsdhip_das_fma_o8(c->das, c->fma_o8);
    }
    size_t n = 0;
    c->das_count = 0;
    const int rc = dsdhip_das_run(c->das, in, count, out, &n);
    if (rc == 0) c->das_count = n;
    if (out_count) *out_count = rc == 0 ? n : 0;
    return rc;
}

int gpu_hip_trellis_das(void *ctx, const double *in, float *out, size_t count, int num_channels)
{
    return gpu_hip_trellis_das_ex(ctx, in, out, count, num_channels, NULL);
}

size_t gpu_hip_trellis_das_count(void *ctx)
{
    gpu_hip_context *c = (gpu_hip_context *)ctx;
    return c ? c->das_count : 0;
}

} /* extern "C" */
