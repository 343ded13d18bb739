// Host side of the engine's pre-SDM conditioning: per-channel boxcar and
// lowpass histories, the lowpass design, and the launches.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "dsdhip_internal.h"

extern "C" {
__global__ void cond_boxcar_bits_k(const unsigned long long *, int64_t, int, int64_t, const unsigned long long *,
                                   int64_t, int, double, double, double, double *, int64_t);
__global__ void cond_bits_hist_k(const unsigned long long *, int64_t, int, int64_t, unsigned long long *);
__global__ void cond_boxcar_pcm_k(const float *, int64_t, int, int64_t, int, double, double *, double *, int64_t);
__global__ void cond_ramp_k(double *, int64_t, int, int64_t, double, double);
__global__ void cond_clip_k(double *, int64_t, int, int64_t);
__global__ void cond_lowpass_bits_k(const unsigned long long *, int64_t, int, int64_t, const unsigned long long *,
                                    int64_t, const double *, int, double, double, double *, int64_t);
__global__ void cond_bits_hist2_k(const unsigned long long *, int64_t, int, int64_t, unsigned long long *);
__global__ void cond_narrow_k(const double *, float *, int64_t);
}

#define LP_NTAPS 127

struct dsdhip_cond {
    int nch, taps;
    int lp;                        /* lowpass designed */
    int lp_ntaps;                  /* its tap count (<= 128) */
    double *d_lp;                  /* [lp_ntaps] */
    unsigned long long *d_hist;    /* [nch] boxcar bit history */
    unsigned long long *d_hist2;   /* [nch][2] lowpass bit history */
    double *d_pcm;                 /* [nch][taps + 1] PCM boxcar state */
    int64_t done_box, done_lp;     /* samples consumed per channel */
    hipStream_t stream;
};

static unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }
static int launched(void) { return hipGetLastError() == hipSuccess ? 0 : -1; }

/* Modified Bessel I0 by its power series (src/fir.c:65-77). */
static double lp_i0(double x)
{
    double sum = 1.0, term = 1.0;
    const double q = (x / 2.0) * (x / 2.0);
    for (int k = 1; k <= 25; k++) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-20 * sum) break;
    }
    return sum;
}

extern "C" {

/* fir_lowpass_init (src/fir.c:795-861): 127-tap Kaiser (beta 10) windowed
 * sinc at 50 kHz for a DSD rate, normalised to unit sum. */
int dsdhip_lowpass_taps(unsigned rate, double *h)
{
    if (!h || rate == 0) return -1;
    const int M = LP_NTAPS - 1;
    double fc = 50000.0 / ((double)rate / 2.0);
    if (fc > 0.5) fc = 0.5;
    const double pi = 3.14159265358979323846;
    const double i0b = lp_i0(10.0);
    for (int n = 0; n <= M; n++) {
        const double x = (double)n - (double)M / 2.0;
        const double sinc = (fabs(x) < 1e-10) ? 2.0 * fc : sin(2.0 * pi * fc * x) / (pi * x);
        const double t = 2.0 * n / M - 1.0;
        const double arg = 1.0 - t * t;
        h[n] = sinc * (lp_i0(10.0 * sqrt(fabs(arg))) / i0b);
    }
    double sum = 0.0;
    for (int n = 0; n < LP_NTAPS; n++) sum += h[n];
    if (sum > 0.0)
        for (int n = 0; n < LP_NTAPS; n++) h[n] /= sum;
    return LP_NTAPS;
}

void dsdhip_cond_destroy(dsdhip_cond_t *c)
{
    if (!c) return;
    if (c->d_lp) hipFree(c->d_lp);
    if (c->d_hist) hipFree(c->d_hist);
    if (c->d_hist2) hipFree(c->d_hist2);
    if (c->d_pcm) hipFree(c->d_pcm);
    delete c;
}

dsdhip_cond_t *dsdhip_cond_create(int nch, int boxcar_taps, unsigned lowpass_rate)
{
    if (nch < 1 || boxcar_taps < 1 || boxcar_taps > 64) return nullptr;
    if (dsdhip_ensure_device() != 0) return nullptr;
    dsdhip_cond_t *c = new dsdhip_cond();
    memset((void *)c, 0, sizeof(*c));
    c->nch = nch;
    c->taps = boxcar_taps;
    bool ok = hipMalloc(&c->d_hist, sizeof(unsigned long long) * (size_t)nch) == hipSuccess &&
              hipMalloc(&c->d_hist2, sizeof(unsigned long long) * 2 * (size_t)nch) == hipSuccess &&
              hipMalloc(&c->d_pcm, sizeof(double) * (size_t)nch * (size_t)(boxcar_taps + 1)) == hipSuccess;
    if (ok && lowpass_rate) {
        double h[LP_NTAPS];
        ok = dsdhip_lowpass_taps(lowpass_rate, h) == LP_NTAPS &&
             hipMalloc(&c->d_lp, sizeof(h)) == hipSuccess &&
             hipMemcpy(c->d_lp, h, sizeof(h), hipMemcpyHostToDevice) == hipSuccess;
        c->lp = ok ? 1 : 0;
        c->lp_ntaps = LP_NTAPS;
    }
    if (!ok || dsdhip_cond_reset(c) != 0) {
        dsdhip_cond_destroy(c);
        return nullptr;
    }
    return c;
}

/* The same with the caller's lowpass taps (1..128; the delay line holds 128
 * samples), as the GPU seam's gpu_fir_lowpass_setup hands them over. */
dsdhip_cond_t *dsdhip_cond_create_taps(int nch, int boxcar_taps, const double *h, int ntaps)
{
    if (!h || ntaps < 1 || ntaps > 128) return nullptr;
    dsdhip_cond_t *c = dsdhip_cond_create(nch, boxcar_taps, 0);
    if (!c) return nullptr;
    if (hipMalloc(&c->d_lp, sizeof(double) * (size_t)ntaps) != hipSuccess ||
        hipMemcpy(c->d_lp, h, sizeof(double) * (size_t)ntaps, hipMemcpyHostToDevice) != hipSuccess) {
        dsdhip_cond_destroy(c);
        return nullptr;
    }
    c->lp = 1;
    c->lp_ntaps = ntaps;
    return c;
}

/* engine_channel_reset (src/engine.c:732-757): zero histories and sums. */
int dsdhip_cond_reset(dsdhip_cond_t *c)
{
    if (!c) return -1;
    if (hipMemsetAsync(c->d_hist, 0, sizeof(unsigned long long) * (size_t)c->nch, c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_hist2, 0, sizeof(unsigned long long) * 2 * (size_t)c->nch, c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_pcm, 0, sizeof(double) * (size_t)c->nch * (size_t)(c->taps + 1), c->stream) !=
            hipSuccess)
        return -1;
    c->done_box = c->done_lp = 0;
    return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : -1;
}

int dsdhip_cond_set_stream(dsdhip_cond_t *c, void *stream)
{
    if (!c) return -1;
    c->stream = (hipStream_t)stream;
    return 0;
}

/* Same-rate DSD: boxcar of taps samples times the ramp from gs (first
 * sample) to ge (last), src/engine.c:829-841. */
int dsdhip_cond_dsd_boxcar_dev(dsdhip_cond_t *c, const unsigned long long *d_bits, int64_t words_stride, size_t n,
                               double gs, double ge, double *d_out, int64_t out_stride)
{
    if (!c || !d_bits || !d_out) return -1;
    if (n == 0) return 0;
    if (words_stride < ((int64_t)n + 63) / 64 || out_stride < (int64_t)n) return -1;
    const double gstep = (n > 1) ? (ge - gs) / (double)(n - 1) : 0.0;
    hipLaunchKernelGGL(cond_boxcar_bits_k, dim3(nblk((int64_t)n), (unsigned)c->nch), dim3(256), 0, c->stream, d_bits,
                       words_stride, c->nch, (int64_t)n, c->d_hist, c->done_box, c->taps, 1.0 / (double)c->taps, gs,
                       gstep, d_out, out_stride);
    hipLaunchKernelGGL(cond_bits_hist_k, dim3((unsigned)((c->nch + 63) / 64)), dim3(64), 0, c->stream, d_bits, words_stride, c->nch, (int64_t)n,
                       c->d_hist);
    c->done_box += (int64_t)n;
    return launched();
}

/* Same-rate DSD with the lowpass instead (fir_mode FIR), src/engine.c:
 * 813-828: lowpass of the +-1 stream, then the ramp. */
int dsdhip_cond_dsd_lowpass_dev(dsdhip_cond_t *c, const unsigned long long *d_bits, int64_t words_stride, size_t n,
                                double gs, double ge, double *d_out, int64_t out_stride)
{
    if (!c || !c->lp || !d_bits || !d_out) return -1;
    if (n == 0) return 0;
    if (words_stride < ((int64_t)n + 63) / 64 || out_stride < (int64_t)n) return -1;
    const double gstep = (n > 1) ? (ge - gs) / (double)(n - 1) : 0.0;
    hipLaunchKernelGGL(cond_lowpass_bits_k, dim3(nblk((int64_t)n), (unsigned)c->nch), dim3(256), 0, c->stream,
                       d_bits, words_stride, c->nch, (int64_t)n, c->d_hist2, c->done_lp, c->d_lp, c->lp_ntaps, gs, gstep,
                       d_out, out_stride);
    hipLaunchKernelGGL(cond_bits_hist2_k, dim3((unsigned)((c->nch + 63) / 64)), dim3(64), 0, c->stream, d_bits, words_stride, c->nch, (int64_t)n,
                       c->d_hist2);
    c->done_lp += (int64_t)n;
    return launched();
}

/* f64 -> f32 on the device (the float outputs of the GPU seam) */
int dsdhip_f64_to_f32_dev(const double *d_x, float *d_y, size_t n, void *stream)
{
    if (!d_x || !d_y) return -1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(cond_narrow_k, dim3(nblk((int64_t)n)), dim3(256), 0, (hipStream_t)stream, d_x, d_y,
                       (int64_t)n);
    return launched();
}

/* PCM boxcar pre-smooth before rate conversion (src/engine.c:629-642):
 * float PCM -> double, no gain. */
int dsdhip_cond_pcm_boxcar_dev(dsdhip_cond_t *c, const float *d_in, int64_t in_stride, size_t n, double *d_out,
                               int64_t out_stride)
{
    if (!c || !d_in || !d_out) return -1;
    if (n == 0) return 0;
    if (in_stride < (int64_t)n || out_stride < (int64_t)n) return -1;
    hipLaunchKernelGGL(cond_boxcar_pcm_k, dim3((unsigned)c->nch), dim3(64), 0, c->stream, d_in, in_stride, c->nch,
                       (int64_t)n, c->taps, 1.0 / (double)c->taps, c->d_pcm, d_out, out_stride);
    return launched();
}

/* x[i] *= gs + (ge - gs) / (n - 1) * i (src/engine.c:646-652, 877-888). */
int dsdhip_gain_ramp_dev(double *d_x, int64_t stride, int nch, size_t n, double gs, double ge, void *stream)
{
    if (!d_x || nch < 1) return -1;
    if (n == 0) return 0;
    const double step = (n > 1) ? (ge - gs) / (double)(n - 1) : 0.0;
    hipLaunchKernelGGL(cond_ramp_k, dim3(nblk((int64_t)n), (unsigned)nch), dim3(256), 0, (hipStream_t)stream, d_x,
                       stride, nch, (int64_t)n, gs, step);
    return launched();
}

/* Soft clip above +-0.95 with a tanh knee (src/engine.c:704-716). */
int dsdhip_soft_clip_dev(double *d_x, int64_t stride, int nch, size_t n, void *stream)
{
    if (!d_x || nch < 1) return -1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(cond_clip_k, dim3(nblk((int64_t)n), (unsigned)nch), dim3(256), 0, (hipStream_t)stream, d_x,
                       stride, nch, (int64_t)n);
    return launched();
}

}  // extern "C"
