// Streaming chain for one multichannel stream, per chunk on one HIP stream
// with the signal resident in HBM:
//   PCM input : [boxcar pre-smooth] -> FIR chain -> gain ramp -> [soft clip]
//               -> [UPOLS convolution] -> DAS
//   DSD input, same rate : boxcar (or lowpass) with the gain ramp
//               -> [convolution] -> DAS
//   DSD input, rate conversion : [boxcar pre-smooth] -> FIR chain -> ramp
//               -> [soft clip] -> [convolution] -> DAS
// With no pre-smooth and no clip this is engine_process_fir_gain's chain
// (src/engine.c:788-899, the plugin's DAS path); with the boxcar(8) and the
// clip it is engine_process_block's (:575-729).
//
// Reference order: engine_process_block (src/engine.c:478-729) and
// engine_process_fir_gain (:788-899); the GPU form of the plugin runs the
// same chain per chunk with DAS as the SDM (src/dsp_plugin.c:1941-2761).
// The gain ramp goes from fir_gain * previous gain to fir_gain * gain over
// each chunk (:603-606, 646-652), as the engine's volume changes do.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <chrono>

#include "dsdhip_internal.h"

struct dsdhip_stream {
    dsdhip_stream_cfg_t cfg;
    int R;                 /* upsampling ratio 2^stages (1 same-rate) */
    size_t cap;            /* samples per channel per chunk at the output rate */
    int conv_block;        /* chunk granularity the chain needs (1: any length) */
    dsdhip_fir_t *fir;     /* NULL same-rate */
    dsdhip_cond_t *cond;   /* boxcar / lowpass state, NULL when unused */
    dsdhip_conv_t *conv;
    dsdhip_das_t *das;
    double *d_pre;         /* [nch][chunk_frames] boxcar pre-smooth (rate conversion) */
    double *d_fir;         /* [nch][cap] conditioned signal */
    double *d_sig;         /* [nch][cap] SDM input (conv output, or d_fir) */
    double prev_gain, gain;
    size_t last_n;
    hipStream_t stream;
    /* host-side timeline of the last chunk (ms from the call's start): DAS
     * run started, then the DAS run's own six points */
    std::chrono::steady_clock::time_point t_call;
    double last_tl[7];
};

static double ms_since_call(const dsdhip_stream_t *s)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s->t_call).count();
}

extern "C" {

void dsdhip_stream_destroy(dsdhip_stream_t *s)
{
    if (!s) return;
    if (s->das) dsdhip_das_destroy(s->das);
    if (s->conv) dsdhip_conv_destroy(s->conv);
    if (s->fir) dsdhip_fir_destroy(s->fir);
    if (s->cond) dsdhip_cond_destroy(s->cond);
    if (s->d_sig && s->d_sig != s->d_fir) hipFree(s->d_sig);
    if (s->d_fir) hipFree(s->d_fir);
    if (s->d_pre) hipFree(s->d_pre);
    delete s;
}

/* The engine's same-rate boxcar taps by rate (src/engine.c:131-147). */
static int same_rate_taps(unsigned fs)
{
    if (fs >= 22579200u) return 16;
    if (fs >= 5644800u) return 64;
    return 32;
}

dsdhip_stream_t *dsdhip_stream_create_cfg(const dsdhip_stream_cfg_t *cfg)
{
    if (!cfg || cfg->nch < 1 || cfg->chunk_frames == 0) return nullptr;
    if (cfg->input != DSDHIP_IN_PCM && cfg->input != DSDHIP_IN_DSD) return nullptr;
    dsdhip_stream_t *s = new dsdhip_stream();
    memset((void *)s, 0, sizeof(*s));
    s->cfg = *cfg;
    const int nch = cfg->nch;
    const bool same = cfg->fs_out == cfg->fs_in;
    bool ok = true;
    if (same) {
        ok = cfg->input == DSDHIP_IN_DSD;             /* PCM -> PCM is not this path */
        s->R = 1;
    } else {
        int nt[16];
        const int nst = dsdhip_fir_plan(cfg->fs_in, cfg->fs_out, nt);
        ok = nst >= 0;
        s->R = ok ? (1 << nst) : 0;
        ok = ok && (s->fir = dsdhip_fir_create(cfg->fs_in, cfg->fs_out, 1.0, nch)) != nullptr;
    }
    s->cap = cfg->chunk_frames * (size_t)(s->R > 0 ? s->R : 1);
    s->conv_block = 1;
    /* boxcar / lowpass state: DSD input always, PCM when pre-smoothing */
    if (ok && (cfg->input == DSDHIP_IN_DSD || cfg->pcm_boxcar > 0)) {
        /* same-rate DSD: the engine's taps by rate unless given; DSD rate
         * conversion: the pre-smooth taps, 1 (the +-1 stream itself) when
         * off; PCM: the pre-smooth taps */
        const int taps = same ? (cfg->dsd_boxcar > 0 ? cfg->dsd_boxcar : same_rate_taps(cfg->fs_in))
                              : (cfg->input == DSDHIP_IN_DSD ? (cfg->dsd_boxcar > 0 ? cfg->dsd_boxcar : 1)
                                                             : cfg->pcm_boxcar);
        ok = (s->cond = dsdhip_cond_create(nch, taps, (same && cfg->dsd_lowpass) ? cfg->fs_in : 0)) != nullptr;
        if (ok && !same)
            ok = hipMalloc(&s->d_pre, sizeof(double) * cfg->chunk_frames * (size_t)nch) == hipSuccess;
    }
    if (ok && cfg->ir) {
        ok = (s->conv = dsdhip_conv_create(nch, cfg->P, cfg->ir, cfg->ir_len, cfg->ir_stride)) != nullptr;
        /* any chunk length: the convolution's FIFO form */
        s->conv_block = 1;
    }
    /* seeding: the reference's rule by output rate (src/dsp_plugin.c:2185) */
    const int seg0_first = cfg->seeding == 1 || (cfg->seeding == 0 && cfg->fs_out < DSDHIP_DSD512_RATE);
    ok = ok && cfg->seeding >= 0 && cfg->seeding <= 2;
    ok = ok && (s->das = dsdhip_das_create(cfg->ntf_index, cfg->depth, cfg->cands, cfg->lat, cfg->state_limit,
                                            nch, cfg->seg_len, 2, seg0_first, cfg->reposition)) != nullptr;
    if (ok && cfg->fma_o8) dsdhip_das_fma_o8(s->das, 1);
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
    s->gain = cfg->gain;
    s->prev_gain = 0.0;
    return s;
}

dsdhip_stream_t *dsdhip_stream_create(unsigned fs_in, unsigned fs_out, double gain, int nch,
                                      size_t chunk_frames, const double *ir, int64_t ir_len,
                                      int64_t ir_stride, int P, int ntf_index, int depth, int cands,
                                      int lat, double state_limit, size_t seg_len, int reposition)
{
    dsdhip_stream_cfg_t c;
    memset(&c, 0, sizeof(c));
    c.fs_in = fs_in;
    c.fs_out = fs_out;
    c.nch = nch;
    c.chunk_frames = chunk_frames;
    c.input = DSDHIP_IN_PCM;
    c.fir_gain = gain;
    c.gain = 1.0;
    c.ir = ir;
    c.ir_len = ir_len;
    c.ir_stride = ir_stride;
    c.P = P;
    c.ntf_index = ntf_index;
    c.depth = depth;
    c.cands = cands;
    c.lat = lat;
    c.state_limit = state_limit;
    c.seg_len = seg_len;
    c.reposition = reposition;
    return dsdhip_stream_create_cfg(&c);
}

int dsdhip_stream_set_stream(dsdhip_stream_t *s, void *hip_stream)
{
    if (!s) return -1;
    s->stream = (hipStream_t)hip_stream;
    int rc = dsdhip_das_set_stream(s->das, hip_stream);
    if (s->fir) rc |= dsdhip_fir_set_stream(s->fir, hip_stream);
    if (s->cond) rc |= dsdhip_cond_set_stream(s->cond, hip_stream);
    if (s->conv) rc |= dsdhip_conv_set_stream(s->conv, hip_stream);
    return rc ? -1 : 0;
}

int dsdhip_stream_reset(dsdhip_stream_t *s)
{
    if (!s) return -1;
    int rc = dsdhip_das_reset(s->das);
    if (s->fir) rc |= dsdhip_fir_reset(s->fir);
    if (s->cond) rc |= dsdhip_cond_reset(s->cond);
    if (s->conv) rc |= dsdhip_conv_reset(s->conv);
    s->last_n = 0;
    s->prev_gain = 0.0;
    return rc ? -1 : 0;
}

int dsdhip_stream_set_gain(dsdhip_stream_t *s, double gain)
{
    if (!s) return -1;
    s->gain = gain;
    return 0;
}

int dsdhip_stream_ratio(const dsdhip_stream_t *s) { return s ? s->R : 0; }

/* the ramp's end points for this chunk (src/engine.c:603-606) */
static void ramp_points(dsdhip_stream_t *s, double *gs, double *ge)
{
    *ge = s->cfg.fir_gain * s->gain;
    *gs = s->cfg.fir_gain * s->prev_gain;
    if (*gs == 0.0) *gs = *ge;
    s->prev_gain = s->gain;
}

/* after conditioning: clip, convolution, DAS */
static int finish(dsdhip_stream_t *s, size_t n, unsigned long long *d_bits, int64_t out_words, size_t *out_count)
{
    const int64_t st = (int64_t)s->cap;
    if (s->cfg.soft_clip && s->R > 1 &&
        dsdhip_soft_clip_dev(s->d_fir, st, s->cfg.nch, n, s->stream) != 0)
        return -1;
    if (s->conv && dsdhip_conv_process_fifo_dev(s->conv, s->d_fir, st, n, s->d_sig, st) != 0) return -1;
    s->last_n = n;
    s->last_tl[0] = ms_since_call(s);
    const int rc = dsdhip_das_run_dev(s->das, s->d_sig, st, n, d_bits, out_words, out_count);
    double tl[6];
    if (dsdhip_das_last_timeline(s->das, tl) == 0)
        for (int i = 0; i < 6; i++) s->last_tl[i + 1] = tl[i] != 0.0 ? tl[i] + s->last_tl[0] : 0.0;
    return rc;
}

/* frames PCM frames per channel (<= chunk_frames; with a room IR any length,
 * through the convolution's FIFO form).  d_bits: nch rows of out_words
 * packed words. */
int dsdhip_stream_process_dev(dsdhip_stream_t *s, const float *d_pcm, int64_t pcm_stride, size_t frames,
                              unsigned long long *d_bits, int64_t out_words, size_t *out_count)
{
    if (!s || !d_pcm || !d_bits || frames > s->cfg.chunk_frames || s->cfg.input != DSDHIP_IN_PCM) return -1;
    s->t_call = std::chrono::steady_clock::now();
    for (int i = 0; i < 7; i++) s->last_tl[i] = 0.0;
    if (out_count) *out_count = 0;
    if (frames == 0) return 0;
    const size_t n = frames * (size_t)s->R;
    if (n % (size_t)s->conv_block) return -1;
    if (out_words < (int64_t)((n + 63) / 64)) return -1;
    const int64_t st = (int64_t)s->cap;
    if (s->cfg.pcm_boxcar > 0) {
        const int64_t ps = (int64_t)s->cfg.chunk_frames;
        if (dsdhip_cond_pcm_boxcar_dev(s->cond, d_pcm, pcm_stride, frames, s->d_pre, ps) != 0 ||
            dsdhip_fir_process_f64_dev(s->fir, s->d_pre, ps, frames, s->d_fir, st) != 0)
            return -1;
    } else if (dsdhip_fir_process_dev(s->fir, d_pcm, pcm_stride, frames, s->d_fir, st) != 0) {
        return -1;
    }
    double gs, ge;
    ramp_points(s, &gs, &ge);
    if ((gs != 1.0 || ge != 1.0) && dsdhip_gain_ramp_dev(s->d_fir, st, s->cfg.nch, n, gs, ge, s->stream) != 0)
        return -1;
    return finish(s, n, d_bits, out_words, out_count);
}

/* DSD input: d_in nch rows of packed bits (words_stride words each), n
 * samples per channel at fs_in (<= chunk_frames). */
int dsdhip_stream_process_dsd_dev(dsdhip_stream_t *s, const unsigned long long *d_in, int64_t words_stride,
                                  size_t n, unsigned long long *d_bits, int64_t out_words, size_t *out_count)
{
    if (!s || !d_in || !d_bits || n > s->cfg.chunk_frames || s->cfg.input != DSDHIP_IN_DSD) return -1;
    s->t_call = std::chrono::steady_clock::now();
    for (int i = 0; i < 7; i++) s->last_tl[i] = 0.0;
    if (out_count) *out_count = 0;
    if (n == 0) return 0;
    const size_t m = n * (size_t)s->R;
    if (m % (size_t)s->conv_block) return -1;
    if (out_words < (int64_t)((m + 63) / 64)) return -1;
    const int64_t st = (int64_t)s->cap;
    double gs, ge;
    ramp_points(s, &gs, &ge);
    if (s->R == 1) {
        const int rc = s->cfg.dsd_lowpass
                           ? dsdhip_cond_dsd_lowpass_dev(s->cond, d_in, words_stride, n, gs, ge, s->d_fir, st)
                           : dsdhip_cond_dsd_boxcar_dev(s->cond, d_in, words_stride, n, gs, ge, s->d_fir, st);
        if (rc != 0) return -1;
    } else {
        /* boxcar pre-smooth of the +-1 stream (exact window sums; 1 tap is
         * the stream widened to double), FIR, ramp */
        const int64_t ps = (int64_t)s->cfg.chunk_frames;
        if (dsdhip_cond_dsd_boxcar_dev(s->cond, d_in, words_stride, n, 1.0, 1.0, s->d_pre, ps) != 0 ||
            dsdhip_fir_process_f64_dev(s->fir, s->d_pre, ps, n, s->d_fir, st) != 0)
            return -1;
        if ((gs != 1.0 || ge != 1.0) && dsdhip_gain_ramp_dev(s->d_fir, st, s->cfg.nch, m, gs, ge, s->stream) != 0)
            return -1;
    }
    return finish(s, m, d_bits, out_words, out_count);
}

/* SDM input of the last chunk into dst [nch][dst_stride] doubles (host or
 * device memory); returns the samples per channel, or -1. */
/* Host-side timeline of the last chunk, ms from the call's start: DAS run
 * started, then the DAS run's six points (seg0 done, estimate started, last
 * segment published, estimate joined, trellis read, run done; zero where
 * the DAS did not pass them). */
int dsdhip_stream_last_timeline(const dsdhip_stream_t *s, double *ms7)
{
    if (!s || !ms7) return -1;
    for (int i = 0; i < 7; i++) ms7[i] = s->last_tl[i];
    return 0;
}

int64_t dsdhip_stream_copy_signal(dsdhip_stream_t *s, double *dst, int64_t dst_stride)
{
    if (!s || !dst || dst_stride < (int64_t)s->last_n) return -1;
    if (s->last_n == 0) return 0;
    if (hipMemcpy2DAsync(dst, sizeof(double) * (size_t)dst_stride, s->d_sig, sizeof(double) * s->cap,
                         sizeof(double) * s->last_n, (size_t)s->cfg.nch, hipMemcpyDefault, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return -1;
    return (int64_t)s->last_n;
}

}  // extern "C"
