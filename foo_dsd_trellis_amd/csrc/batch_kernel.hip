/*
 * Device-side chunk pipeline of the batched DAS (many tracks of uneven
 * length, each with nch channels, in one persistent trellis launch).
 *
 * Every kernel reads the chunk index from device memory, so one chunk step
 * (gather -> FIR -> plan -> trellis -> stitch -> assemble -> copy-back ->
 * advance) can be captured once in a hipGraph and replayed for every
 * chunk.  The semantics per track and chunk are those of the single-stream
 * DAS driver (src/dsp_plugin.c:1941-2761 with segments of a given length,
 * all segments seeded from the persistent context), decided on the device:
 *
 *   bt_gather     PCM frames of the chunk -> staging (zeros past a track's end)
 *   bt_plan       one wave per track: segment plan, quiet-boundary reposition
 *                 (:1963-2034), job slots for every channel, snapshot of the
 *                 persistent contexts and the seed (best survivor state)
 *   bt_post       one wave per track: stitch boundaries from the kept counts
 *   das_stitch    (das_kernel.hip) candidates and decode quality per boundary
 *   bt_select     one wave per track: candidate selection (0.8 / 1.2 rule,
 *                 :2652-2681) and the output spans of every channel
 *   bt_assemble   spans -> packed output, appended at the track's bit count
 *   bt_finish     copy-back of the last segment's context (links detached)
 *   bt_advance    bit counts, chunk index, queue counter
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "tr_types.h"
#include "bt_types.h"

typedef unsigned long long u64;

__device__ __forceinline__ int64_t bt_imin(int64_t a, int64_t b) { return a < b ? a : b; }

/* ---- gather ------------------------------------------------------------ */
extern "C" __global__ void bt_gather(bt_cfg cfg, const bt_track *__restrict__ tr,
                                     const bt_state *__restrict__ st, float *__restrict__ stage)
{
    const int s = blockIdx.y;
    const int t = s / cfg.nch, ch = s - t * cfg.nch;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cfg.CF) return;
    const int64_t f = st->chunk * cfg.CF + i;
    const bt_track &k = tr[t];
    stage[(int64_t)s * cfg.CF + i] = (f < k.frames) ? k.pcm[(int64_t)ch * k.pcm_stride + f] : 0.0f;
}

/* ---- plan -------------------------------------------------------------- */
/* sum of squares of the 2048 samples around c, ascending (as a scalar loop) */
__device__ double bt_sumsq(const double *x, int64_t c)
{
    double s = 0.0;
    for (int64_t k = c - 1024; k < c + 1024; k++) {
        const double v = x[k];
        s = s + v * v;
    }
    return s;
}

extern "C" __global__ __launch_bounds__(64) void bt_plan(
    bt_cfg cfg, const bt_track *__restrict__ tr, const bt_state *__restrict__ stt,
    const double *__restrict__ fout, int64_t *__restrict__ seg_st, int64_t *__restrict__ seg_sz,
    int *__restrict__ nseg, int64_t *__restrict__ fcount, tr_job *__restrict__ jobs,
    tr_ctx *__restrict__ ctxs, u64 *__restrict__ segbits, bt_state *__restrict__ stw)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    const int64_t c = stt->chunk;
    const bt_track &k = tr[t];
    int64_t v = k.frames - c * cfg.CF;
    v = v < 0 ? 0 : (v > cfg.CF ? cfg.CF : v);
    const int64_t F = v * cfg.R;
    int S = F > 0 ? (int)(F / cfg.seg_len) : 0;
    if (F > 0 && S < 1) S = 1;
    if (S > cfg.smax) {
        if (lane == 0) atomicOr(&stw->err, 1);
        S = 0;
    }
    int64_t *st = seg_st + (int64_t)t * cfg.smax;
    int64_t *sz = seg_sz + (int64_t)t * cfg.smax;
    /* sizes by the reference's rule for S segments (src/dsp_plugin.c:1944-1953) */
    {
        const int64_t base = S > 0 ? F / S : 0, rem = S > 0 ? F % S : 0;
        for (int s = lane; s < S; s += 64) {
            st[s] = (int64_t)s * base + (s < rem ? s : rem);
            sz[s] = (s == S - 1) ? F - st[s] : base + (s < rem ? 1 : 0);
        }
    }
    __syncthreads();
    /* quiet-boundary reposition on channel 0 (src/dsp_plugin.c:1963-2034) */
    if (cfg.reposition && S > 2) {
        const double *x0 = fout + (int64_t)(t * cfg.nch) * cfg.fstride;
        const int64_t w = 2048, step = 1024;
        const double loud = 0.005;
        const int64_t base = F / S;
        int64_t max_shift = base * 2 / 5;
        if (max_shift > 200000) max_shift = 200000;
        const int64_t nc = 2 * (max_shift / step);     /* candidates of one boundary, visiting order */
        __shared__ double csum[64];
        for (int seg = 2; seg < S; seg++) {
            const int64_t b = st[seg];
            if (b < w || b + w >= F) continue;
            const double cur = sqrt(bt_sumsq(x0, b) / (double)w);   /* same value on every lane */
            if (cur > loud) continue;
            int64_t best = b;
            double best_rms = cur;
            bool stop = false;
            for (int64_t c0 = 0; c0 < nc && !stop; c0 += 64) {
                /* candidate i: offset (i/2+1)*step, direction - for even i, + for odd */
                const int64_t i = c0 + lane;
                double r = 0.0;
                if (i < nc) {
                    const int64_t off = (i / 2 + 1) * step;
                    const int64_t cand = (i & 1) ? b + off : (b > off ? b - off : 0);
                    if (!(cand < w / 2 || cand + w / 2 >= F)) r = bt_sumsq(x0, cand);
                }
                csum[lane] = r;
                __syncthreads();
                if (lane == 0) {
                    for (int q = 0; q < 64 && c0 + q < nc; q++) {
                        const int64_t ii = c0 + q;
                        const int64_t off = (ii / 2 + 1) * step;
                        const int64_t cand = (ii & 1) ? b + off : (b > off ? b - off : 0);
                        if (!(cand < w / 2 || cand + w / 2 >= F) &&
                            !(cand <= st[seg - 1] + cfg.O * 2) &&
                            !(seg < S - 1 && cand >= st[seg + 1] - cfg.O * 2)) {
                            const double rr = sqrt(csum[q] / (double)w);
                            if (rr > best_rms) {
                                best_rms = rr;
                                best = cand;
                            }
                        }
                        if ((ii & 1) && best_rms > loud) {   /* both directions of this offset done */
                            stop = true;
                            break;
                        }
                    }
                    csum[0] = stop ? 1.0 : 0.0;
                }
                __syncthreads();
                stop = csum[0] != 0.0;
                __syncthreads();
            }
            if (lane == 0 && best != b) {
                st[seg] = best;
                sz[seg - 1] = best - st[seg - 1];
                sz[seg] = (seg < S - 1) ? st[seg + 1] - best : F - best;
            }
            __syncthreads();
        }
    }
    if (lane == 0) {
        nseg[t] = S;
        fcount[t] = F;
    }
    /* job slots of every channel (full-parallel seeding, :2268-2363) */
    const int O = (int)cfg.O;
    for (int ch = 0; ch < cfg.nch; ch++) {
        const int s = t * cfg.nch + ch;
        const int persist = s, snap = cfg.nstreams + s;
        for (int seg = lane; seg < cfg.smax; seg += 64) {
            tr_job &jb = jobs[(int64_t)s * cfg.smax + seg];
            jb.kept = 0;
            jb.slow_steps = 0;
            jb.err = 0;
            if (seg >= S) {
                jb.skip = 1;
                continue;
            }
            int64_t is, cnt, keep;
            const int64_t ns = st[seg], nz = sz[seg];
            if (seg == 0) {
                cnt = nz;
                if (S > 1 && nz + O <= F) cnt += O;
                is = 0;
                keep = 0;
            } else {
                const bool ext = seg < S - 1;
                is = ns >= O ? ns - O : 0;
                cnt = ns + nz - is;
                if (ext && ns + nz + O <= F) cnt += O;
                const int64_t disc = ns >= O ? O - cfg.L : 0;
                keep = disc ? (int64_t)cfg.L + disc : 0;
            }
            if ((cnt + 63) / 64 + 1 > cfg.seg_words) {   /* would overrun its output region */
                atomicOr(&stw->err, 1);
                jb.skip = 1;
                continue;
            }
            jb.in = fout + (int64_t)s * cfg.fstride + is;
            jb.out = (uint64_t *)(segbits + ((int64_t)s * cfg.smax + seg) * cfg.seg_words);
            jb.n_in = cnt;
            jb.keep_from = keep;
            jb.skip = 0;
            if (seg == 0) {
                jb.ctx_in = persist;
                jb.ctx_out = persist;
                jb.snap = -1;
                jb.copy = 0;
                jb.seed = 0;
            } else {
                jb.ctx_in = snap;
                jb.ctx_out = 2 * cfg.nstreams + s * cfg.smax + seg;
                jb.snap = snap;
                jb.copy = 1;
                jb.seed = 1;
                for (int q = 0; q < TR_MAXP; q++) jb.seed_st[q] = q < cfg.P ? ctxs[persist].st[0][q] : 0.0;
            }
        }
        /* snapshot of the persistent context, read by the copied segments */
        if (S > 1) {
            const u64 *src = (const u64 *)&ctxs[persist];
            u64 *dst = (u64 *)&ctxs[snap];
            for (int i = lane; i < (int)(sizeof(tr_ctx) / 8); i += 64) dst[i] = src[i];
        }
    }
}

/* ---- stitch inputs ----------------------------------------------------- */
extern "C" __global__ __launch_bounds__(64) void bt_post(
    bt_cfg cfg, const int *__restrict__ nseg, const int64_t *__restrict__ seg_sz,
    const tr_job *__restrict__ jobs, int64_t *__restrict__ pbit, int64_t *__restrict__ cbit,
    int *__restrict__ ovl, bt_state *__restrict__ stw)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    const int S = nseg[t];
    const int s0 = t * cfg.nch;                     /* channel 0 of the track */
    const tr_job *J = jobs + (int64_t)s0 * cfg.smax;
    const int64_t O = cfg.O;
    for (int k = lane; k < cfg.smax; k += 64) {
        const int64_t b = (int64_t)t * cfg.smax + k;
        ovl[b] = 0;
        if (k == 0 || k >= S) continue;
        int64_t prev_end = J[k - 1].kept;
        if (k - 1 == 0 && prev_end > seg_sz[(int64_t)t * cfg.smax]) prev_end = seg_sz[(int64_t)t * cfg.smax];
        const int64_t kk = J[k].kept;
        if (prev_end < 2 * O || kk == 0) {
            atomicOr(&stw->err, 4);
            continue;
        }
        pbit[b] = ((int64_t)s0 * cfg.smax + (k - 1)) * cfg.seg_words * 64 + prev_end - O;
        cbit[b] = ((int64_t)s0 * cfg.smax + k) * cfg.seg_words * 64;
        ovl[b] = (int)(O < kk ? O : kk);
    }
    /* trellis errors of every channel */
    for (int i = lane; i < cfg.nch * cfg.smax; i += 64) {
        const tr_job &jb = jobs[(int64_t)s0 * cfg.smax + i];
        if (!jb.skip && jb.err) atomicOr(&stw->err, 2);
    }
}

/* ---- selection and spans ---------------------------------------------- */
extern "C" __global__ __launch_bounds__(64) void bt_select(
    bt_cfg cfg, const int *__restrict__ nseg, const int64_t *__restrict__ seg_sz,
    const tr_job *__restrict__ jobs, const int *__restrict__ ncand, const int *__restrict__ peak,
    const int *__restrict__ cpos, const int *__restrict__ cden, const double *__restrict__ cmax,
    int *__restrict__ bpos, int64_t *__restrict__ sp_src, int64_t *__restrict__ sp_dst,
    int64_t *__restrict__ sp_len, int64_t *__restrict__ total)
{
    const int t = blockIdx.x, lane = threadIdx.x;
    const int S = nseg[t];
    int *bp = bpos + (int64_t)t * cfg.smax;
    if (lane == 0) {
        for (int k = 0; k < cfg.smax; k++) bp[k] = 0;
        for (int k = 1; k < S; k++) {
            const int64_t b = (int64_t)t * cfg.smax + k;
            const int n = ncand[b];
            if (n <= 0) continue;
            const int pk = peak[b];
            const int *pos = cpos + b * (BT_MAXQ + 1);
            const int *den = cden + b * (BT_MAXQ + 1);
            const double *mx = cmax + b * (BT_MAXQ + 1);
            double qual[BT_MAXQ + 1];
            for (int q = 0; q < n; q++) {
                if (n == 1) { qual[0] = 0.0; break; }
                double dr = (double)den[q] / (double)(pk > 0 ? pk : 1);
                if (dr < 0.5) dr = 0.5;
                qual[q] = mx[q] / dr;
            }
            int best = 0;
            for (int q = 1; q < n; q++) {
                const double rq = qual[q] / (qual[best] + 1e-15);
                if (rq < 0.8) best = q;
                else if (rq < 1.2 && den[q] > den[best]) best = q;
            }
            bp[k] = pos[best];
        }
    }
    __syncthreads();
    /* spans of each channel: same positions on every channel (:2683-2713) */
    const int64_t O = cfg.O;
    for (int ch = lane; ch < cfg.nch; ch += 64) {
        const int s = t * cfg.nch + ch;
        const tr_job *J = jobs + (int64_t)s * cfg.smax;
        int64_t *src = sp_src + (int64_t)s * cfg.smax;
        int64_t *dst = sp_dst + (int64_t)s * cfg.smax;
        int64_t *len = sp_len + (int64_t)s * cfg.smax;
        int64_t wp = 0;
        if (S > 0) {
            int64_t c0 = J[0].kept;
            if (c0 > seg_sz[(int64_t)t * cfg.smax]) c0 = seg_sz[(int64_t)t * cfg.smax];
            src[0] = ((int64_t)s * cfg.smax) * cfg.seg_words * 64;
            dst[0] = 0;
            len[0] = c0;
            wp = c0;
        }
        for (int k = 1; k < S; k++) {
            const int64_t kk = J[k].kept;
            if (kk == 0) { dst[k] = wp; src[k] = 0; len[k] = 0; continue; }
            const int64_t pst = wp >= O ? wp - O : 0;
            const int64_t b = bp[k];
            dst[k] = pst + b;
            src[k] = ((int64_t)s * cfg.smax + k) * cfg.seg_words * 64 + b;
            len[k] = kk - b;
            wp = dst[k] + len[k];
        }
        /* effective span: up to where the next non-empty span starts */
        for (int k = 0; k < S; k++) {
            int64_t e = dst[k] + len[k];
            for (int k2 = k + 1; k2 < S; k2++)
                if (len[k2] > 0) { if (dst[k2] < e) e = dst[k2]; break; }
            len[k] = e > dst[k] ? e - dst[k] : 0;
        }
        for (int k = S; k < cfg.smax; k++) { src[k] = 0; dst[k] = wp; len[k] = 0; }
        if (ch == 0) total[t] = wp;
    }
}

/* ---- assemble ---------------------------------------------------------- */
/* One lane per output word touched by this chunk; bits below the track's
 * current bit count keep their value (the first word is merged). */
extern "C" __global__ void bt_assemble(bt_cfg cfg, const bt_track *__restrict__ tr,
                                       const int *__restrict__ nseg, const u64 *__restrict__ segbits,
                                       const int64_t *__restrict__ sp_src,
                                       const int64_t *__restrict__ sp_dst,
                                       const int64_t *__restrict__ sp_len,
                                       const int64_t *__restrict__ total,
                                       bt_state *__restrict__ stw)
{
    const int s = blockIdx.y;
    const int t = s / cfg.nch, ch = s - t * cfg.nch;
    const int S = nseg[t];
    if (S == 0) return;
    const bt_track &k = tr[t];
    const int64_t tot = total[t], pos0 = k.out_pos;
    const int64_t wl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = pos0 / 64 + wl;
    if (w * 64 >= pos0 + tot) return;
    if (w >= k.out_words) {
        atomicOr(&stw->err, 8);
        return;
    }
    const int64_t *dst = sp_dst + (int64_t)s * cfg.smax;
    const int64_t *src = sp_src + (int64_t)s * cfg.smax;
    const int64_t *len = sp_len + (int64_t)s * cfg.smax;
    u64 *out = k.out + (int64_t)ch * k.out_stride;
    u64 word = (w * 64 < pos0) ? out[w] : 0ull;
    /* chunk-relative position of bit 0 of this word */
    const int64_t r0 = w * 64 - pos0;
    int lo = 0, hi = S - 1, sp = 0;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (dst[mid] <= (r0 < 0 ? 0 : r0)) { sp = mid; lo = mid + 1; } else hi = mid - 1;
    }
    for (int i = 0; i < 64; i++) {
        const int64_t r = r0 + i;
        if (r < 0) continue;
        if (r >= tot) {
            word &= ~(1ull << i);
            continue;
        }
        while (sp + 1 < S && dst[sp + 1] <= r) sp++;
        const int64_t d0 = dst[sp];
        u64 bit = 0;
        if (r >= d0 && r < d0 + len[sp]) {
            const int64_t sb = src[sp] + (r - d0);
            bit = (segbits[sb >> 6] >> (sb & 63)) & 1ull;
        }
        word = (word & ~(1ull << i)) | (bit << i);
    }
    out[w] = word;
}

/* ---- copy-back --------------------------------------------------------- */
extern "C" __global__ __launch_bounds__(64) void bt_finish(bt_cfg cfg, const int *__restrict__ nseg,
                                                           tr_ctx *__restrict__ ctxs)
{
    const int s = blockIdx.x, lane = threadIdx.x;
    const int t = s / cfg.nch;
    const int S = nseg[t];
    if (S <= 1) return;
    const tr_ctx *src = &ctxs[2 * cfg.nstreams + s * cfg.smax + (S - 1)];
    tr_ctx *dst = &ctxs[s];
    const u64 *a = (const u64 *)src;
    u64 *b = (u64 *)dst;
    for (int i = lane; i < (int)(sizeof(tr_ctx) / 8); i += 64) b[i] = a[i];
    __syncthreads();
    /* detach: the segment contexts' links led into the snapshot */
    int8_t *lk = &dst->link[0][0];
    for (int i = lane; i < 2 * TR_SLOTS; i += 64) lk[i] = TR_LINK_NONE;
    if (lane == 0) dst->has_foreign = 0;
}

extern "C" __global__ void bt_advance(bt_cfg cfg, bt_track *__restrict__ tr,
                                      const int64_t *__restrict__ total, bt_state *__restrict__ stw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cfg.ntracks) tr[t].out_pos += total[t];
    if (t == 0) {
        stw->chunk += 1;
        stw->queue = 0;
    }
}

/* zeroes per-chunk totals (a fill kernel, so a captured graph holds no memset) */
extern "C" __global__ void bt_zero_i64(int64_t *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
