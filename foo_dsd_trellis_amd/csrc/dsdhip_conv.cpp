// Host side of the UPOLS convolution: IR partition spectra, FDL history,
// batched library FFTs around the hand-written spectral multiply-accumulate.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "dsdhip_internal.h"

extern "C" {
__global__ void cv_pack(const double *, int64_t, const double *, int, int, int, float *);
__global__ void cv_tail(const double *, int64_t, int, int, int, double *);
__global__ void cv_mac(const float2 *, const float2 *, const float2 *, int, int, int, int, float2 *);
__global__ void cv_hist(const float2 *, const float2 *, int, int, int, int, float2 *);
__global__ void cv_unpack(const float *, int, int, int, double *, int64_t);
__global__ void cv_cat2(const double *, int64_t, int64_t, const double *, int64_t, int64_t, int, double *,
                        int64_t);
}

#define CV_TB 8

struct dsdhip_conv {
    int nch, P, np, K;
    int plan_nb;
    hipfftHandle fwd, inv;
    float2 *d_H;          /* [nch][np][K] */
    float2 *d_hist[2];    /* [nch][np-1][K] */
    int cur;
    double *d_tail;       /* [nch][P] */
    float *d_frames;      /* [nch][nb][2P] */
    float2 *d_X, *d_Y;    /* [nch][nb][K] */
    int cap_nb;
    hipStream_t stream;
    /* FIFO form (any chunk length): pending input of the next block and
     * produced output not yet handed out, each double-buffered, plus the
     * staging of the whole blocks of a call */
    int buf_pos;          /* pending input samples per channel, < P */
    int64_t fifo_len;     /* produced samples waiting, < P */
    int fcur;
    double *d_pend[2];    /* [nch][P] */
    double *d_fifo[2];    /* [nch][P] */
    double *d_sin, *d_sout;   /* [nch][cap_st] */
    size_t cap_st;
};

static void conv_free_bufs(dsdhip_conv_t *c)
{
    if (c->d_frames) hipFree(c->d_frames);
    if (c->d_X) hipFree(c->d_X);
    if (c->d_Y) hipFree(c->d_Y);
    c->d_frames = nullptr;
    c->d_X = c->d_Y = nullptr;
    c->cap_nb = 0;
}

static int conv_plans(dsdhip_conv_t *c, int nb)
{
    if (c->plan_nb == nb) return 0;
    if (c->plan_nb) {
        hipfftDestroy(c->fwd);
        hipfftDestroy(c->inv);
        c->plan_nb = 0;
    }
    int n = 2 * c->P;
    if (hipfftPlanMany(&c->fwd, 1, &n, nullptr, 1, 2 * c->P, nullptr, 1, c->K, HIPFFT_R2C,
                       c->nch * nb) != HIPFFT_SUCCESS)
        return -1;
    if (hipfftPlanMany(&c->inv, 1, &n, nullptr, 1, c->K, nullptr, 1, 2 * c->P, HIPFFT_C2R,
                       c->nch * nb) != HIPFFT_SUCCESS) {
        hipfftDestroy(c->fwd);
        return -1;
    }
    hipfftSetStream(c->fwd, c->stream);
    hipfftSetStream(c->inv, c->stream);
    c->plan_nb = nb;
    return 0;
}

static int conv_bufs(dsdhip_conv_t *c, int nb)
{
    if (nb <= c->cap_nb) return 0;
    conv_free_bufs(c);
    const size_t fr = (size_t)c->nch * nb * 2 * c->P, sp = (size_t)c->nch * nb * c->K;
    if (hipMalloc(&c->d_frames, sizeof(float) * fr) != hipSuccess ||
        hipMalloc(&c->d_X, sizeof(float2) * sp) != hipSuccess ||
        hipMalloc(&c->d_Y, sizeof(float2) * sp) != hipSuccess)
        return -1;
    c->cap_nb = nb;
    return 0;
}

extern "C" {

void dsdhip_conv_destroy(dsdhip_conv_t *c)
{
    if (!c) return;
    if (c->plan_nb) {
        hipfftDestroy(c->fwd);
        hipfftDestroy(c->inv);
    }
    conv_free_bufs(c);
    for (int i = 0; i < 2; i++) {
        if (c->d_pend[i]) hipFree(c->d_pend[i]);
        if (c->d_fifo[i]) hipFree(c->d_fifo[i]);
    }
    if (c->d_sin) hipFree(c->d_sin);
    if (c->d_sout) hipFree(c->d_sout);
    if (c->d_H) hipFree(c->d_H);
    for (int i = 0; i < 2; i++)
        if (c->d_hist[i]) hipFree(c->d_hist[i]);
    if (c->d_tail) hipFree(c->d_tail);
    delete c;
}

dsdhip_conv_t *dsdhip_conv_create(int nch, int P, const double *ir, int64_t ir_len, int64_t ir_stride)
{
    if (nch < 1 || P < 16 || (P & (P - 1)) || !ir || ir_len < 1) return nullptr;
    if (dsdhip_ensure_device() != 0) return nullptr;
    dsdhip_conv_t *c = new dsdhip_conv();
    memset((void *)c, 0, sizeof(*c));
    c->nch = nch;
    c->P = P;
    c->K = P + 1;
    int np = (int)((ir_len + P - 1) / P);
    np = (np + CV_TB - 1) / CV_TB * CV_TB;      /* zero partitions up to a multiple of the tile */
    c->np = np;
    const size_t hsz = (size_t)nch * np * c->K, fsz = (size_t)nch * np * 2 * P;
    const size_t histsz = (size_t)nch * (np - 1) * c->K;
    float *d_fr = nullptr;
    hipfftHandle plan;
    int n = 2 * P;
    std::vector<float> fr(fsz, 0.0f);
    for (int ch = 0; ch < nch; ch++)
        for (int p = 0; p < np; p++)
            for (int j = 0; j < P; j++) {
                const int64_t t = (int64_t)p * P + j;
                if (t < ir_len) fr[((size_t)ch * np + p) * 2 * P + j] = (float)ir[(int64_t)ch * ir_stride + t];
            }
    if (hipMalloc(&c->d_H, sizeof(float2) * hsz) != hipSuccess ||
        hipMalloc(&c->d_hist[0], sizeof(float2) * (histsz ? histsz : 1)) != hipSuccess ||
        hipMalloc(&c->d_hist[1], sizeof(float2) * (histsz ? histsz : 1)) != hipSuccess ||
        hipMalloc(&c->d_tail, sizeof(double) * (size_t)nch * P) != hipSuccess ||
        hipMalloc(&c->d_pend[0], sizeof(double) * (size_t)nch * P) != hipSuccess ||
        hipMalloc(&c->d_pend[1], sizeof(double) * (size_t)nch * P) != hipSuccess ||
        hipMalloc(&c->d_fifo[0], sizeof(double) * (size_t)nch * P) != hipSuccess ||
        hipMalloc(&c->d_fifo[1], sizeof(double) * (size_t)nch * P) != hipSuccess ||
        hipMalloc(&d_fr, sizeof(float) * fsz) != hipSuccess ||
        hipMemcpy(d_fr, fr.data(), sizeof(float) * fsz, hipMemcpyHostToDevice) != hipSuccess) {
        if (d_fr) hipFree(d_fr);
        dsdhip_conv_destroy(c);
        return nullptr;
    }
    if (hipfftPlanMany(&plan, 1, &n, nullptr, 1, 2 * P, nullptr, 1, c->K, HIPFFT_R2C, nch * np) !=
        HIPFFT_SUCCESS) {
        hipFree(d_fr);
        dsdhip_conv_destroy(c);
        return nullptr;
    }
    const hipfftResult r = hipfftExecR2C(plan, d_fr, (hipfftComplex *)c->d_H);
    hipDeviceSynchronize();
    hipfftDestroy(plan);
    hipFree(d_fr);
    if (r != HIPFFT_SUCCESS || dsdhip_conv_reset(c) != 0) {
        dsdhip_conv_destroy(c);
        return nullptr;
    }
    return c;
}

int dsdhip_conv_reset(dsdhip_conv_t *c)
{
    if (!c) return -1;
    const size_t histsz = (size_t)c->nch * (c->np - 1) * c->K;
    for (int i = 0; i < 2; i++)
        if (histsz && hipMemsetAsync(c->d_hist[i], 0, sizeof(float2) * histsz, c->stream) != hipSuccess)
            return -1;
    if (hipMemsetAsync(c->d_tail, 0, sizeof(double) * (size_t)c->nch * c->P, c->stream) != hipSuccess)
        return -1;
    c->cur = 0;
    c->buf_pos = 0;
    c->fifo_len = 0;
    c->fcur = 0;
    return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : -1;
}

int dsdhip_conv_set_stream(dsdhip_conv_t *c, void *stream)
{
    if (!c) return -1;
    c->stream = (hipStream_t)stream;
    if (c->plan_nb) {
        hipfftSetStream(c->fwd, c->stream);
        hipfftSetStream(c->inv, c->stream);
    }
    return 0;
}

int dsdhip_conv_partitions(const dsdhip_conv_t *c) { return c ? c->np : 0; }

/* nb whole blocks of P samples per channel, stream-ordered, no sync. */
static int conv_blocks(dsdhip_conv_t *c, const double *d_in, int64_t in_stride, int nb, double *d_out,
                       int64_t out_stride)
{
    if (conv_bufs(c, nb) != 0 || conv_plans(c, nb) != 0) return -1;
    const int P = c->P, K = c->K, nch = c->nch, np = c->np;
    const size_t n = (size_t)nb * (size_t)P;
    hipStream_t st = c->stream;
    const unsigned gx = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(cv_pack, dim3(gx, (unsigned)nch), dim3(256), 0, st, d_in, in_stride, c->d_tail,
                       P, nb, nch, c->d_frames);
    hipLaunchKernelGGL(cv_tail, dim3((unsigned)((P + 255) / 256), (unsigned)nch), dim3(256), 0, st,
                       d_in, in_stride, P, nb, nch, c->d_tail);
    if (hipfftExecR2C(c->fwd, c->d_frames, (hipfftComplex *)c->d_X) != HIPFFT_SUCCESS) return -1;
    const float2 *hist = c->d_hist[c->cur];
    hipLaunchKernelGGL(cv_mac, dim3((unsigned)((K + 255) / 256), (unsigned)((nb + CV_TB - 1) / CV_TB), (unsigned)nch),
                       dim3(256), 0, st, hist, c->d_X, c->d_H, K, np, nb, nch, c->d_Y);
    if (np > 1)
        hipLaunchKernelGGL(cv_hist, dim3((unsigned)((K + 255) / 256), (unsigned)(np - 1), (unsigned)nch),
                           dim3(256), 0, st, hist, c->d_X, K, np, nb, nch, c->d_hist[c->cur ^ 1]);
    c->cur ^= 1;
    if (hipfftExecC2R(c->inv, (hipfftComplex *)c->d_Y, c->d_frames) != HIPFFT_SUCCESS) return -1;
    hipLaunchKernelGGL(cv_unpack, dim3(gx, (unsigned)nch), dim3(256), 0, st, c->d_frames, P, nb, nch,
                       d_out, out_stride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* n samples per channel (a multiple of P), stream-ordered, no sync. */
int dsdhip_conv_process_dev(dsdhip_conv_t *c, const double *d_in, int64_t in_stride, size_t n,
                            double *d_out, int64_t out_stride)
{
    if (!c || !d_in || !d_out) return -1;
    if (n == 0) return 0;
    if (n % (size_t)c->P) return -1;
    if (c->buf_pos != 0 || c->fifo_len != 0) return -1;   /* the FIFO form holds samples */
    return conv_blocks(c, d_in, in_stride, (int)(n / (size_t)c->P), d_out, out_stride);
}

static int cat2(dsdhip_conv_t *c, const double *a, int64_t a_stride, int64_t na, const double *b,
                int64_t b_stride, int64_t n, double *out, int64_t out_stride)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(cv_cat2, dim3((unsigned)((n + 255) / 256), (unsigned)c->nch), dim3(256), 0, c->stream,
                       a, a_stride, na, b, b_stride, n, c->nch, out, out_stride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/*
 * Any chunk length n (src/convolution.c:893-958, conv_process_direct): the
 * chunk is appended to the pending input, every whole block of P samples is
 * convolved and its P outputs join the produced samples not yet handed out;
 * then exactly n of those are written to d_out.  When fewer than n are
 * available (no latency is inserted, so this happens whenever the blocks
 * completed so far trail the input), all of them are written, the rest of
 * d_out holds the chunk's own input samples at those positions (the
 * reference convolves in place and leaves them untouched), and nothing is
 * carried over.  Chunks that are multiples of P give the block form's
 * output.  d_out may equal d_in.  Stream-ordered, no sync.
 */
int dsdhip_conv_process_fifo_dev(dsdhip_conv_t *c, const double *d_in, int64_t in_stride, size_t n,
                                 double *d_out, int64_t out_stride)
{
    if (!c || !d_in || !d_out || in_stride < (int64_t)n || out_stride < (int64_t)n) return -1;
    if (n == 0) return 0;
    const int64_t P = c->P, bp = c->buf_pos, fl = c->fifo_len;
    const int64_t total = bp + (int64_t)n, nb = total / P, nproc = nb * P;
    if (nb > INT32_MAX / 2) return -1;
    /* staging of the whole blocks (grow-only) */
    if ((size_t)nproc > c->cap_st) {
        if (c->d_sin) hipFree(c->d_sin);
        if (c->d_sout) hipFree(c->d_sout);
        c->d_sin = c->d_sout = nullptr;
        c->cap_st = 0;
        if (hipMalloc(&c->d_sin, sizeof(double) * (size_t)c->nch * (size_t)nproc) != hipSuccess ||
            hipMalloc(&c->d_sout, sizeof(double) * (size_t)c->nch * (size_t)nproc) != hipSuccess)
            return -1;
        c->cap_st = (size_t)nproc;
    }
    const int64_t cs = (int64_t)c->cap_st;
    double *pend = c->d_pend[c->fcur], *pend2 = c->d_pend[c->fcur ^ 1];
    double *fifo = c->d_fifo[c->fcur], *fifo2 = c->d_fifo[c->fcur ^ 1];
    /* the whole blocks: pending input, then the chunk's head */
    if (nb > 0) {
        if (cat2(c, pend, P, bp, d_in, in_stride, nproc, c->d_sin, cs) != 0) return -1;
        if (conv_blocks(c, c->d_sin, cs, (int)nb, c->d_sout, cs) != 0) return -1;
    }
    /* the next pending input: [pending | chunk] from nproc on (before d_out
     * is written, which may be d_in) */
    const int64_t r = total - nproc;
    if (nproc >= bp) {
        if (cat2(c, d_in + (nproc - bp), in_stride, r, d_in, in_stride, r, pend2, P) != 0) return -1;
    } else if (cat2(c, pend, P, bp, d_in, in_stride, r, pend2, P) != 0) {
        return -1;
    }
    const int64_t avail = fl + nproc;
    int64_t nfl = 0;
    if (avail >= (int64_t)n) {
        /* hand out n of [waiting | produced], keep the rest */
        nfl = avail - (int64_t)n;
        if (nfl > P) return -1;
        const int64_t a0 = fl > (int64_t)n ? fl - (int64_t)n : 0;
        const int64_t b0 = (int64_t)n > fl ? (int64_t)n - fl : 0;
        if (cat2(c, fifo + (fl > (int64_t)n ? (int64_t)n : fl), P, a0, c->d_sout + b0, cs, nfl, fifo2, P) != 0)
            return -1;
        if (cat2(c, fifo, P, fl, c->d_sout, cs, (int64_t)n, d_out, out_stride) != 0) return -1;
    } else {
        /* too few: all of them, then the chunk's own input at the rest */
        if (d_out != d_in &&
            cat2(c, d_in + avail, in_stride, (int64_t)n - avail, d_in, in_stride, (int64_t)n - avail, d_out + avail,
                 out_stride) != 0)
            return -1;
        if (cat2(c, fifo, P, fl, c->d_sout, cs, avail, d_out, out_stride) != 0) return -1;
    }
    c->fcur ^= 1;
    c->buf_pos = (int)r;
    c->fifo_len = nfl;
    return 0;
}

/* Samples held by the FIFO form: pending input, produced output waiting. */
int dsdhip_conv_fifo_state(const dsdhip_conv_t *c, int64_t *pending_in, int64_t *waiting_out)
{
    if (!c) return -1;
    if (pending_in) *pending_in = c->buf_pos;
    if (waiting_out) *waiting_out = c->fifo_len;
    return 0;
}

}  // extern "C"
