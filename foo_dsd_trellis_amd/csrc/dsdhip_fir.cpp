/*
 * Host side of the FIR upsampling chain.
 *
 * Stage plan and taps follow the reference CPU chain: half-band Kaiser
 * design (beta 12, src/fir.c:65-98), per-stage tap count (127 for the first
 * stage of a chain ending at DSD512 or DSD512/48, else 63; src/fir.c:131-150),
 * taps x2 for upsampling (src/fir.c:386-390), zero-initialised delay lines
 * that persist across calls (src/fir.c:419-426), stages chained
 * (src/fir.c:678-725), output gain multiply (src/engine.c:872-883).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "dsdhip_internal.h"

extern "C" int fir_launch_stage(const void *x, int in_f32, int64_t x_stride, const double *hist,
                                int64_t n, int nch, int ntaps, const double *d_h2, double *y,
                                int64_t y_stride, double gain, int apply_gain, hipStream_t stream);
extern "C" int fir_launch_hist_update(const void *x, int in_f32, int64_t x_stride, int64_t n,
                                      int nch, double *hist, int H, hipStream_t stream);
extern "C" int fir_launch_pair(const void *x, int in_f32, int64_t x_stride, const double *histA, int64_t n1,
                               int nch, const double *d_h2A, const double *d_h2B, double *histB,
                               double *hist_tmp, double *y, int64_t y_stride, double gain, int apply_gain,
                               hipStream_t stream);

#define FIR_MAX_STAGES 9

struct dsdhip_fir {
    int nst;
    int ntaps[FIR_MAX_STAGES];
    double *d_h2[FIR_MAX_STAGES];
    double *d_hist[FIR_MAX_STAGES];
    int nch;
    double gain;
    double *d_buf[2];
    size_t buf_cap;   /* doubles per buffer */
    hipStream_t stream;
    float *d_hin;     /* host-buffer entry: staged input / output, grow-only */
    double *d_hout;
    size_t hin_cap, hout_cap;
    int fused;        /* adjacent 63-tap stages in one pass (default) */
    double *d_htmp;   /* nch x 31: the intermediate history of a fused pair */
};

/* Modified Bessel I0 by its power series, 25 terms with early exit
 * (src/fir.c:65-77). */
static double bessel_i0(double x)
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

/* Kaiser-windowed half-band sinc, normalised to unit sum (src/fir.c:79-98). */
static void halfband_design(double *h, int ntaps, double beta)
{
    const int c = (ntaps - 1) / 2;
    const double i0b = bessel_i0(beta);
    const double pi = 3.14159265358979323846;
    double sum = 0.0;
    for (int n = 0; n < ntaps; n++) {
        const double x = (double)(n - c) / 2.0;
        const double sv = (fabs(x) < 1e-15) ? 1.0 : sin(pi * x) / (pi * x);
        const double t = (double)(n - c) / c;
        const double arg = 1.0 - t * t;
        const double w = bessel_i0(beta * sqrt(arg > 0.0 ? arg : 0.0)) / i0b;
        h[n] = sv * w;
        sum += h[n];
    }
    for (int n = 0; n < ntaps; n++) h[n] /= sum;
}

extern "C" {

void dsdhip_fir_destroy(dsdhip_fir_t *f);

void dsdhip_halfband_taps(double *h, int ntaps) { halfband_design(h, ntaps, 12.0); }

int dsdhip_fir_plan(unsigned fs_in, unsigned fs_out, int *ntaps)
{
    if (fs_in == 0 || fs_out <= fs_in || fs_out % fs_in) return -1;
    unsigned r = fs_out / fs_in;
    if (r & (r - 1)) return -1;
    if (r == 8 && (fs_out == 22579200u || fs_out == 24576000u)) return -1;   /* 4x stage: DSD input only */
    int st = 0;
    while (r > 1) { st++; r >>= 1; }
    if (st > FIR_MAX_STAGES) return -1;
    const bool to512 = fs_out == 22579200u || fs_out == 24576000u;
    for (int i = 0; i < st; i++) ntaps[i] = (st >= 2 && i == 0 && to512) ? 127 : 63;
    return st;
}

static dsdhip_fir_t *fir_alloc(int nst, const int *ntaps, const double *const *h, double gain, int nch)
{
    if (nst < 1 || nst > FIR_MAX_STAGES || nch < 1) return NULL;
    if (dsdhip_ensure_device() != 0) return NULL;
    dsdhip_fir_t *f = (dsdhip_fir_t *)calloc(1, sizeof(dsdhip_fir_t));
    f->nst = nst;
    f->nch = nch;
    f->gain = gain;
    f->fused = 1;
    if (hipMalloc(&f->d_htmp, sizeof(double) * 31 * (size_t)nch) != hipSuccess) {
        free(f);
        return NULL;
    }
    for (int s = 0; s < nst; s++) {
        const int T = ntaps[s];
        if (T != 63 && T != 127) { dsdhip_fir_destroy(f); return NULL; }
        f->ntaps[s] = T;
        double h2[127];
        for (int k = 0; k < T; k++) h2[k] = h[s][k] * 2.0;
        const int H = (T + 1) / 2 - 1;
        if (hipMalloc(&f->d_h2[s], sizeof(double) * (size_t)T) != hipSuccess ||
            hipMalloc(&f->d_hist[s], sizeof(double) * (size_t)H * (size_t)nch) != hipSuccess) {
            dsdhip_fir_destroy(f);
            return NULL;
        }
        hipMemcpy(f->d_h2[s], h2, sizeof(double) * (size_t)T, hipMemcpyHostToDevice);
        hipMemset(f->d_hist[s], 0, sizeof(double) * (size_t)H * (size_t)nch);
    }
    return f;
}

dsdhip_fir_t *dsdhip_fir_create(unsigned fs_in, unsigned fs_out, double gain, int nch)
{
    int ntaps[FIR_MAX_STAGES];
    const int st = dsdhip_fir_plan(fs_in, fs_out, ntaps);
    if (st < 1) return NULL;
    double h63[63], h127[127];
    halfband_design(h63, 63, 12.0);
    halfband_design(h127, 127, 12.0);
    const double *hp[FIR_MAX_STAGES];
    for (int s = 0; s < st; s++) hp[s] = ntaps[s] == 127 ? h127 : h63;
    return fir_alloc(st, ntaps, hp, gain, nch);
}

dsdhip_fir_t *dsdhip_fir_create_taps(const double *taps, int ntaps, int nstages, double gain, int nch)
{
    int nt[FIR_MAX_STAGES];
    const double *hp[FIR_MAX_STAGES];
    if (nstages < 1 || nstages > FIR_MAX_STAGES) return NULL;
    for (int s = 0; s < nstages; s++) { nt[s] = ntaps; hp[s] = taps; }
    return fir_alloc(nstages, nt, hp, gain, nch);
}

void dsdhip_fir_destroy(dsdhip_fir_t *f)
{
    if (!f) return;
    for (int s = 0; s < FIR_MAX_STAGES; s++) {
        if (f->d_h2[s]) hipFree(f->d_h2[s]);
        if (f->d_hist[s]) hipFree(f->d_hist[s]);
    }
    if (f->d_htmp) hipFree(f->d_htmp);
    for (int b = 0; b < 2; b++) if (f->d_buf[b]) hipFree(f->d_buf[b]);
    if (f->d_hin) hipFree(f->d_hin);
    if (f->d_hout) hipFree(f->d_hout);
    free(f);
}

int dsdhip_fir_set_stream(dsdhip_fir_t *f, void *stream)
{
    if (!f) return -1;
    f->stream = (hipStream_t)stream;
    return 0;
}

int dsdhip_fir_num_stages(const dsdhip_fir_t *f) { return f ? f->nst : 0; }
int dsdhip_fir_ratio(const dsdhip_fir_t *f) { return f ? (1 << f->nst) : 0; }

int dsdhip_fir_reset(dsdhip_fir_t *f)
{
    if (!f) return -1;
    for (int s = 0; s < f->nst; s++) {
        const int H = (f->ntaps[s] + 1) / 2 - 1;
        if (hipMemsetAsync(f->d_hist[s], 0, sizeof(double) * (size_t)H * (size_t)f->nch,
                           f->stream) != hipSuccess)
            return -1;
    }
    return 0;
}

/* d_in: nch rows of n floats (row stride in_stride); d_out: nch rows of
 * n * 2^nst doubles (row stride out_stride). */
int dsdhip_fir_reserve(dsdhip_fir_t *f, size_t n)
{
    if (!f) return -1;
    const size_t need = (f->nst > 1) ? n * ((size_t)1 << (f->nst - 1)) * (size_t)f->nch : 0;
    if (need > f->buf_cap) {
        for (int b = 0; b < 2; b++) {
            if (f->d_buf[b]) hipFree(f->d_buf[b]);
            f->d_buf[b] = NULL;
        }
        f->buf_cap = 0;
        for (int b = 0; b < 2; b++)
            if (hipMalloc(&f->d_buf[b], sizeof(double) * need) != hipSuccess) return -1;
        f->buf_cap = need;
    }
    return 0;
}

/* Stages in order; two adjacent 63-tap stages run as one pass (the
 * intermediate stream stays in LDS) unless fusion is off.  Every output is
 * the same sum in the same order either way. */
static int fir_run(dsdhip_fir_t *f, const void *d_in, int in_f32, int64_t in_stride, size_t n, double *d_out,
                   int64_t out_stride)
{
    if (!f || !d_in || !d_out) return -1;
    if (n == 0) return 0;
    if (dsdhip_fir_reserve(f, n) != 0) return -1;
    const void *src = d_in;
    int src_f32 = in_f32;
    int64_t src_stride = in_stride;
    size_t cnt = n;
    int nb = 0;   /* next ping-pong buffer */
    for (int s = 0; s < f->nst;) {
        const bool pair = f->fused && s + 1 < f->nst && f->ntaps[s] == 63 && f->ntaps[s + 1] == 63;
        const int k = pair ? 2 : 1;
        const bool last = s + k == f->nst;
        const int64_t nout = (int64_t)cnt << k;
        double *dst = last ? d_out : f->d_buf[nb];
        const int64_t dst_stride = last ? out_stride : nout;
        const int H = (f->ntaps[s] + 1) / 2 - 1;
        const int rc = pair ? fir_launch_pair(src, src_f32, src_stride, f->d_hist[s], (int64_t)cnt, f->nch, f->d_h2[s],
                                              f->d_h2[s + 1], f->d_hist[s + 1], f->d_htmp, dst, dst_stride, f->gain,
                                              (last && f->gain != 1.0) ? 1 : 0, f->stream)
                            : fir_launch_stage(src, src_f32, src_stride, f->d_hist[s], (int64_t)cnt, f->nch,
                                               f->ntaps[s], f->d_h2[s], dst, dst_stride, f->gain,
                                               (last && f->gain != 1.0) ? 1 : 0, f->stream);
        if (rc != 0) return -1;
        if (fir_launch_hist_update(src, src_f32, src_stride, (int64_t)cnt, f->nch, f->d_hist[s], H, f->stream) != 0)
            return -1;
        if (!last) nb ^= 1;
        src = dst;
        src_f32 = 0;
        src_stride = dst_stride;
        cnt = (size_t)nout;
        s += k;
    }
    return 0;
}

/* Fused passes of adjacent 63-tap stages on (default) or off (one pass per
 * stage); the outputs are identical. */
int dsdhip_fir_set_fused(dsdhip_fir_t *f, int on)
{
    if (!f) return -1;
    f->fused = on ? 1 : 0;
    return 0;
}

int dsdhip_fir_process_dev(dsdhip_fir_t *f, const float *d_in, int64_t in_stride, size_t n,
                           double *d_out, int64_t out_stride)
{
    return fir_run(f, d_in, 1, in_stride, n, d_out, out_stride);
}

/* fir_chain_process_d with double input (the engine's boxcar-smoothed
 * signal, src/engine.c:620-645). */
int dsdhip_fir_process_f64_dev(dsdhip_fir_t *f, const double *d_in, int64_t in_stride, size_t n,
                               double *d_out, int64_t out_stride)
{
    return fir_run(f, d_in, 0, in_stride, n, d_out, out_stride);
}

/* Host-buffer form: in nch x n floats, out nch x n*2^nst doubles. */
int dsdhip_fir_process(dsdhip_fir_t *f, const float *in, size_t n, double *out)
{
    if (!f) return -1;
    const size_t nout = n << f->nst;
    const size_t nin_all = n * (size_t)f->nch + 1, nout_all = nout * (size_t)f->nch + 1;
    if (f->hin_cap < nin_all) {
        if (f->d_hin) hipFree(f->d_hin);
        f->d_hin = NULL;
        f->hin_cap = 0;
        if (hipMalloc(&f->d_hin, sizeof(float) * nin_all) != hipSuccess) return -1;
        f->hin_cap = nin_all;
    }
    if (f->hout_cap < nout_all) {
        if (f->d_hout) hipFree(f->d_hout);
        f->d_hout = NULL;
        f->hout_cap = 0;
        if (hipMalloc(&f->d_hout, sizeof(double) * nout_all) != hipSuccess) return -1;
        f->hout_cap = nout_all;
    }
    if (hipMemcpy(f->d_hin, in, sizeof(float) * n * (size_t)f->nch, hipMemcpyHostToDevice) != hipSuccess ||
        dsdhip_fir_process_dev(f, f->d_hin, (int64_t)n, n, f->d_hout, (int64_t)nout) != 0 ||
        hipStreamSynchronize(f->stream) != hipSuccess ||
        hipMemcpy(out, f->d_hout, sizeof(double) * nout * (size_t)f->nch, hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    return 0;
}

} /* extern "C" */
