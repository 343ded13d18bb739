// Streaming PCM -> DSD chain for one multichannel stream: per chunk the FIR
// upsampling chain, the optional UPOLS room-correction convolution and the
// DAS trellis, all on one HIP stream with the signal resident in HBM.
//
// Reference order: engine_process_block (src/engine.c:478-729): FIR chain
// with gain -> conv_process (:719-721) -> SDM; the GPU form of the plugin
// runs the same chain per chunk with DAS as the SDM (src/dsp_plugin.c:
// 1941-2761).  The engine's soft clip (:704-716) is not applied here.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "dsdhip_internal.h"

struct dsdhip_stream {
    int nch;
    int R;                 /* upsampling ratio 2^stages */
    size_t chunk_frames;
    size_t cap;            /* samples per channel per chunk at the output rate */
    int conv_block;        /* 8*P when convolving, else 1 */
    dsdhip_fir_t *fir;
    dsdhip_conv_t *conv;
    dsdhip_das_t *das;
    double *d_fir;         /* [nch][cap] FIR output */
    double *d_sig;         /* [nch][cap] SDM input (conv output, or d_fir) */
    size_t last_n;
    hipStream_t stream;
};

extern "C" {

void dsdhip_stream_destroy(dsdhip_stream_t *s)
{
    if (!s) return;
    if (s->das) dsdhip_das_destroy(s->das);
    if (s->conv) dsdhip_conv_destroy(s->conv);
    if (s->fir) dsdhip_fir_destroy(s->fir);
    if (s->d_sig && s->d_sig != s->d_fir) hipFree(s->d_sig);
    if (s->d_fir) hipFree(s->d_fir);
    delete s;
}

dsdhip_stream_t *dsdhip_stream_create(unsigned fs_in, unsigned fs_out, double gain, int nch,
                                      size_t chunk_frames, const double *ir, int64_t ir_len,
                                      int64_t ir_stride, int P, int ntf_index, int depth, int cands,
                                      int lat, double state_limit, size_t seg_len, int reposition)
{
    if (nch < 1 || chunk_frames == 0) return nullptr;
    int nt[16];
    const int nst = dsdhip_fir_plan(fs_in, fs_out, nt);
    if (nst < 0) return nullptr;
    dsdhip_stream_t *s = new dsdhip_stream();
    memset((void *)s, 0, sizeof(*s));
    s->nch = nch;
    s->R = 1 << nst;
    s->chunk_frames = chunk_frames;
    s->cap = chunk_frames * (size_t)s->R;
    s->conv_block = 1;
    bool ok = (s->fir = dsdhip_fir_create(fs_in, fs_out, gain, nch)) != nullptr;
    if (ok && ir) {
        ok = (s->conv = dsdhip_conv_create(nch, P, ir, ir_len, ir_stride)) != nullptr;
        s->conv_block = 8 * P;
        ok = ok && s->cap % (size_t)s->conv_block == 0;
    }
    ok = ok && (s->das = dsdhip_das_create(ntf_index, depth, cands, lat, state_limit, nch, seg_len, 2, 0,
                                            reposition)) != nullptr;
    ok = ok && hipMalloc(&s->d_fir, sizeof(double) * s->cap * (size_t)nch) == hipSuccess;
    if (ok) {
        if (s->conv)
            ok = hipMalloc(&s->d_sig, sizeof(double) * s->cap * (size_t)nch) == hipSuccess;
        else
            s->d_sig = s->d_fir;
    }
    if (!ok) {
        dsdhip_stream_destroy(s);
        return nullptr;
    }
    return s;
}

int dsdhip_stream_set_stream(dsdhip_stream_t *s, void *hip_stream)
{
    if (!s) return -1;
    s->stream = (hipStream_t)hip_stream;
    int rc = dsdhip_fir_set_stream(s->fir, hip_stream) | dsdhip_das_set_stream(s->das, hip_stream);
    if (s->conv) rc |= dsdhip_conv_set_stream(s->conv, hip_stream);
    return rc ? -1 : 0;
}

int dsdhip_stream_reset(dsdhip_stream_t *s)
{
    if (!s) return -1;
    int rc = dsdhip_fir_reset(s->fir) | dsdhip_das_reset(s->das);
    if (s->conv) rc |= dsdhip_conv_reset(s->conv);
    s->last_n = 0;
    return rc ? -1 : 0;
}

int dsdhip_stream_ratio(const dsdhip_stream_t *s) { return s ? s->R : 0; }

/* frames PCM frames per channel (<= chunk_frames; frames * ratio a multiple
 * of 8*P when convolving).  d_bits: nch rows of out_words packed words. */
int dsdhip_stream_process_dev(dsdhip_stream_t *s, const float *d_pcm, int64_t pcm_stride, size_t frames,
                              unsigned long long *d_bits, int64_t out_words, size_t *out_count)
{
    if (!s || !d_pcm || !d_bits || frames > s->chunk_frames) return -1;
    if (out_count) *out_count = 0;
    if (frames == 0) return 0;
    const size_t n = frames * (size_t)s->R;
    if (n % (size_t)s->conv_block) return -1;
    if (out_words < (int64_t)((n + 63) / 64)) return -1;
    const int64_t st = (int64_t)s->cap;
    if (dsdhip_fir_process_dev(s->fir, d_pcm, pcm_stride, frames, s->d_fir, st) != 0) return -1;
    if (s->conv && dsdhip_conv_process_dev(s->conv, s->d_fir, st, n, s->d_sig, st) != 0) return -1;
    s->last_n = n;
    return dsdhip_das_run_dev(s->das, s->d_sig, st, n, d_bits, out_words, out_count);
}

/* SDM input of the last chunk into dst [nch][dst_stride] doubles (host or
 * device memory); returns the samples per channel, or -1. */
int64_t dsdhip_stream_copy_signal(dsdhip_stream_t *s, double *dst, int64_t dst_stride)
{
    if (!s || !dst || dst_stride < (int64_t)s->last_n) return -1;
    if (s->last_n == 0) return 0;
    if (hipMemcpy2DAsync(dst, sizeof(double) * (size_t)dst_stride, s->d_sig, sizeof(double) * s->cap,
                         sizeof(double) * s->last_n, (size_t)s->nch, hipMemcpyDefault, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return -1;
    return (int64_t)s->last_n;
}

}  // extern "C"
