/*
 * Trellis sigma-delta modulator for at most two survivors per step (K <= 2,
 * the engine's candidate count on every DSD-input path, src/engine.c:43-103):
 * one LANE per trellis context, so a wave runs 64 segment jobs at once.
 *
 * With K <= 2 a step has at most four children, and the reference's
 * candidate sort (src/trellis.c:621-731) reduces to a handful of compares:
 *   - the path-hash chains only ever hold this step's registrants (the heads
 *     of the children's buckets are cleared first, src/trellis.c:629-633, and
 *     the write past act[31] that can re-link a bucket needs 32 listed
 *     survivors), so a lookup is "the earlier registrant with the same path";
 *   - the active list holds at most two entries, kept here as two child
 *     indices with their metrics.
 * Each lane runs the sequential sort literally over its own children, in the
 * reference's order and with its int64 metric compares; the loop filter is
 * the scalar operation order (FP64, no contraction, src/trellis.c:37-191).
 * The order-8 FMA filter of the AVX2 build applies to groups of four
 * candidates (src/trellis.c:904-927), which never form at K <= 2.
 *
 * Jobs are the same tr_job/tr_ctx records as the one-wave-per-job kernel;
 * the final context is stored in the same layout.  The per-slot hash links
 * it stores are all "none": at K <= 2 no chain outlives its step.
 */
#include <hip/hip_runtime.h>
#include "tr_types.h"

typedef unsigned long long u64;
#define GPTR(T) T __attribute__((address_space(1))) *

namespace {

__device__ __forceinline__ long long lbits(double d) { return __double_as_longlong(d); }

/* CRFB loop filter of one parent (y = 0 here; the +-1 of the branch is added
 * to d[0] by the caller), as the wave kernel computes it. */
template <int P>
__device__ __forceinline__ double lane_filt(const double (&s)[P], double x, const tr_params &prm, double (&d)[P])
{
    d[0] = s[0] - prm.g[0] * s[1] + x - 0.0;
    double v = x + prm.a[0] * d[0];
#pragma unroll
    for (int q = 1; q < P - 1; q++) {
        d[q] = s[q] + s[q - 1] - prm.g[q] * s[q + 1];
        v = v + prm.a[q] * d[q];
    }
    d[P - 1] = s[P - 1] + s[P - 2];
    v = v + prm.a[P - 1] * d[P - 1];
    return v;
}

__device__ __forceinline__ double clampl(double v, double lim)
{
    return (v > lim) ? lim : ((v < -lim) ? -lim : v);
}

template <int HW>
__device__ __forceinline__ int hist_bit(const u64 (&h)[HW], int i)
{
    u64 w = h[0];
#pragma unroll
    for (int k = 1; k < HW; k++)
        if (k == (i >> 6)) w = h[k];
    return (int)((w >> (i & 63)) & 1ull);
}

template <int P, int HW>
__device__ void tr_lane_body(const tr_params &prm, tr_job *jb, tr_ctx *ctxs)
{
    const tr_ctx *c0 = &ctxs[jb->ctx_in];
    const int K = prm.K, L = prm.L;
    const uint32_t mask = prm.mask;
    const double a0 = prm.a[0];
    const bool limit = prm.state_limit > 0.0;
    const double lim = prm.state_limit;

    int n = c0->num_cands, pos = c0->pos, pending = c0->pending, idx = c0->idx;
    const int draining = c0->draining;
    int err = 0;

    /* the (at most two) survivors: loop-filter state, metric, path,
     * traceback bit, history */
    double st0[P], st1[P];
    double cs0, cs1;
    uint32_t pa0, pa1;
    int nx0, nx1;
    int sl0, sl1;               /* slots of the survivors in the active bank */
    u64 h0[HW], h1[HW];
    {
        const bool sd = jb->seed != 0;
        const bool two = n > 1;
#pragma unroll
        for (int q = 0; q < P; q++) {
            st0[q] = sd ? jb->seed_st[q] : c0->st[0][q];
            st1[q] = two ? (sd ? jb->seed_st[q] : c0->st[1][q]) : 0.0;
        }
        cs0 = sd ? 0.0 : c0->cost[0];
        cs1 = two ? (sd ? 0.0 : c0->cost[1]) : 0.0;
        pa0 = c0->path[0];
        pa1 = two ? c0->path[1] : 0u;
        nx0 = c0->next[0];
        nx1 = two ? c0->next[1] : 0;
        sl0 = c0->slot[0];
        sl1 = two ? c0->slot[1] : 0;
#pragma unroll
        for (int w = 0; w < HW; w++) {
            h0[w] = c0->hist[0][w];
            h1[w] = two ? c0->hist[1][w] : 0ull;
        }
    }

    const GPTR(double) in = (const GPTR(double))jb->in;
    GPTR(u64) out = (GPTR(u64))jb->out;
    const int64_t n_in = jb->n_in, keep_from = jb->keep_from;
    u64 acc = 0;
    int nacc = 0;
    int64_t kept = 0;

    /* input in blocks of 4 samples; the next block is loaded while this
     * one runs (past the end: the last sample again, never used) */
    double xr[4], xn[4];
#pragma unroll
    for (int u = 0; u < 4; u++) xr[u] = (n_in > 0) ? in[u < n_in ? u : n_in - 1] : 0.0;

    for (int64_t t0 = 0; t0 < n_in; t0 += 4) {
        /* this block's output bits, stored once per 64 (keeps the
         * stores out of the steps, so the input loads' waits stay exact) */
        unsigned ob = 0;
        int ne = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t tn = t0 + 4 + u;
            xn[u] = in[tn < n_in ? tn : n_in - 1];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t t = t0 + u;
            if (t >= n_in) break;
            const double x = xr[u];
            int next_pos = pos + 1;
            if (next_pos == L) next_pos = 0;
            const bool two = n > 1;
            const int m = 2 * n;

            /* ---- expand (src/trellis.c:904-936) ---- */
            double d0[P], d1[P];
            const double v0 = lane_filt<P>(st0, x, prm, d0);
            const double v1 = lane_filt<P>(st1, x, prm, d1);
            const double e00 = v0 + a0, e01 = v0 - a0, e10 = v1 + a0, e11 = v1 - a0;
            const double c[4] = {cs0 + e00 * e00, cs0 + e01 * e01, cs1 + e10 * e10, cs1 + e11 * e11};
            const long long cb[4] = {lbits(c[0]), lbits(c[1]), lbits(c[2]), lbits(c[3])};
            const uint32_t p[4] = {(pa0 << 1) & mask, ((pa0 << 1) | 1u) & mask,
                                   (pa1 << 1) & mask, ((pa1 << 1) | 1u) & mask};
            const int nn0 = hist_bit<HW>(h0, next_pos);
            const int nn1 = two ? hist_bit<HW>(h1, next_pos) : 0;
            if (!__builtin_isfinite(c[0]) || !__builtin_isfinite(c[1]) ||
                (two && (!__builtin_isfinite(c[2]) || !__builtin_isfinite(c[3]))))
                err = 1;

            /* ---- prune (src/trellis.c:621-731), as selects ---- */
            /* first strict minimum: per parent, then across the two */
            const int lo0 = (cb[1] < cb[0]) ? 1 : 0;
            const int lo1 = (cb[3] < cb[2]) ? 3 : 2;
            const long long lb0 = (lo0 == 1) ? cb[1] : cb[0];
            const long long lb1 = (lo1 == 3) ? cb[3] : cb[2];
            int mn = (two && lb1 < lb0) ? lo1 : lo0;
            /* majority of the children's traceback bits (a parent's two
             * children share its bit; tie -> 0): with two parents it is
             * nx0 & nx1, with one it is that parent's bit.  When the
             * minimum's bit is not the majority, the cheapest majority
             * child is the other parent's cheaper child. */
            {
                const int pm = mn >> 1;
                const bool ovr = two && nx0 != nx1 && ((pm ? nx1 : nx0) == 1);
                const int bm = pm ? lo0 : lo1;
                const double cbm = (bm == 0) ? c[0] : (bm == 1) ? c[1] : (bm == 2) ? c[2] : c[3];
                const double cmn = (mn == 0) ? c[0] : (mn == 1) ? c[1] : (mn == 2) ? c[2] : c[3];
                mn = (ovr && cbm < cmn * 1.1) ? bm : mn;
            }
            const int M = (mn >> 1) ? nx1 : nx0;
            const bool cand0 = nx0 == M, cand1 = two && nx1 == M;

            /* scan in index order into the active list (a0s, a1s) of at
             * most K entries with metrics (ab0, ab1); reg: registered
             * children */
            int a0s = 0, a1s = 0, no = 0;
            long long ab0 = 0, ab1 = 0;
            unsigned reg = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const bool cand = (i >> 1) ? cand1 : cand0;
                const long long ci = cb[i];
                const long long lastb = (no == 2) ? ab1 : ab0;
                const bool go = cand && !(no == K && lastb <= ci);
                /* the earlier registrant with the same path (at most one) */
                bool found = false;
                int tg = 0;
                long long tb = 0;
#pragma unroll
                for (int j = 0; j < i; j++) {
                    const bool hit = ((reg >> j) & 1u) && p[j] == p[i];
                    found = found || hit;
                    tg = hit ? j : tg;
                    tb = hit ? cb[j] : tb;
                }
                const bool proceed = go && !(found && tb <= ci);
                reg |= (go && !found) ? (1u << i) : 0u;
                /* a cheaper duplicate replaces its registrant if listed */
                const bool rm0 = proceed && found && no >= 1 && a0s == tg;
                const bool rm1 = proceed && found && !rm0 && no == 2 && a1s == tg;
                a0s = rm0 ? a1s : a0s;
                ab0 = rm0 ? ab1 : ab0;
                no -= (rm0 || rm1) ? 1 : 0;
                /* stable insertion behind equal metrics; an entry pushed
                 * past K drops out */
                const bool p0 = proceed && (no == 0 || ab0 > ci);
                const bool p1 = proceed && !p0;
                const bool sh = p0 && no >= 1;
                a1s = sh ? a0s : (p1 ? i : a1s);
                ab1 = sh ? ab0 : (p1 ? ci : ab1);
                a0s = p0 ? i : a0s;
                ab0 = p0 ? ci : ab0;
                no += (proceed && no < K) ? 1 : 0;
            }

            /* ---- commit (src/trellis.c:938-972) ---- */
            const int outbit = (a0s >> 1) ? nx1 : nx0;
            const double minc = __longlong_as_double(ab0);
            /* survivors from their parents' filter outputs */
            const bool q0 = (a0s >> 1) != 0, q1 = (a1s >> 1) != 0;
            const double s00 = (q0 ? d1[0] : d0[0]) + ((a0s & 1) ? -1.0 : 1.0);
            const double s10 = (q1 ? d1[0] : d0[0]) + ((a1s & 1) ? -1.0 : 1.0);
            double ns0[P], ns1[P];
            ns0[0] = s00;
            ns1[0] = s10;
#pragma unroll
            for (int q = 1; q < P; q++) {
                ns0[q] = q0 ? d1[q] : d0[q];
                ns1[q] = q1 ? d1[q] : d0[q];
            }
            if (limit) {
#pragma unroll
                for (int q = 0; q < P; q++) {
                    ns0[q] = clampl(ns0[q], lim);
                    ns1[q] = clampl(ns1[q], lim);
                }
            }
            const uint32_t np0 = (((q0 ? pa1 : pa0) << 1) | (uint32_t)(a0s & 1)) & mask;
            const uint32_t np1 = (((q1 ? pa1 : pa0) << 1) | (uint32_t)(a1s & 1)) & mask;
            const int pos_w = pos >> 6;
            const u64 pbit = 1ull << (pos & 63);
            u64 nh0[HW], nh1[HW];
#pragma unroll
            for (int w = 0; w < HW; w++) {
                u64 a = q0 ? h1[w] : h0[w], b = q1 ? h1[w] : h0[w];
                if (w == pos_w) {
                    a = (np0 & 1u) ? (a | pbit) : (a & ~pbit);
                    b = (np1 & 1u) ? (b | pbit) : (b & ~pbit);
                }
                nh0[w] = a;
                nh1[w] = b;
            }
            pa0 = np0;
            pa1 = np1;
            cs0 = __longlong_as_double(ab0) - minc;
            cs1 = __longlong_as_double(ab1) - minc;
            nx0 = q0 ? nn1 : nn0;
            nx1 = q1 ? nn1 : nn0;
            sl0 = a0s;                /* child i of a step is slot i: 2*parent + branch */
            sl1 = a1s;
#pragma unroll
            for (int q = 0; q < P; q++) {
                st0[q] = ns0[q];
                st1[q] = ns1[q];
            }
#pragma unroll
            for (int w = 0; w < HW; w++) {
                h0[w] = nh0[w];
                h1[w] = nh1[w];
            }
            n = no;
            pos = next_pos;
            idx ^= 1;

            /* output bookkeeping (src/trellis.c:1013-1040) */
            const bool emit = pending >= L && t >= keep_from;
            pending += (pending < L) ? 1 : 0;
            ob |= emit ? ((unsigned)outbit << ne) : 0u;
            ne += emit ? 1 : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) xr[u] = xn[u];
        const u64 lo = acc | ((u64)ob << nacc);
        const int tot = nacc + ne;
        if (tot >= 64) {                      /* nacc >= 60 here */
            out[kept >> 6] = lo;
            kept += 64;
            acc = (u64)ob >> (64 - nacc);
            nacc = tot - 64;
        } else {
            acc = lo;
            nacc = tot;
        }
    }
    if (nacc) {
        out[kept >> 6] = acc;
        kept += nacc;
    }

    if (jb->ctx_out >= 0) {
        tr_ctx *co = &ctxs[jb->ctx_out];
#pragma unroll
        for (int q = 0; q < P; q++) co->st[0][q] = st0[q];
        co->cost[0] = cs0;
        co->path[0] = pa0;
        co->next[0] = (uint8_t)nx0;
        co->slot[0] = (uint8_t)sl0;
#pragma unroll
        for (int w = 0; w < HW; w++) co->hist[0][w] = h0[w];
        if (n > 1) {
#pragma unroll
            for (int q = 0; q < P; q++) co->st[1][q] = st1[q];
            co->cost[1] = cs1;
            co->path[1] = pa1;
            co->next[1] = (uint8_t)nx1;
            co->slot[1] = (uint8_t)sl1;
#pragma unroll
            for (int w = 0; w < HW; w++) co->hist[1][w] = h1[w];
        }
        for (int b = 0; b < 2; b++)
            for (int i = 0; i < TR_SLOTS; i++) co->link[b][i] = TR_LINK_NONE;
        co->num_cands = n;
        co->pos = pos;
        co->pending = pending;
        co->idx = idx;
        co->draining = draining;
        co->has_foreign = 0;
        co->err = err;
    }
    jb->kept = kept;
    jb->slow_steps = 0;
    jb->err = err;
}

} // namespace

/* Lane j of workgroup b runs job 64 b + j; unused slots are skipped. */
template <int P, int HW>
__global__ __launch_bounds__(64) void tr_lane_jobs(tr_params prm, tr_job *jobs, int njobs, tr_ctx *ctxs)
{
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= njobs) return;
    tr_job *jb = &jobs[j];
    if (jb->skip) return;
    tr_lane_body<P, HW>(prm, jb, ctxs);
}

/* Whether tr_lane_launch takes prm: K <= 2, histories up to 128 samples,
 * not forced onto the one-wave-per-job kernel. */
extern "C" int tr_lane_eligible(const tr_params *prm)
{
    return prm->K >= 1 && prm->K <= 2 && prm->L >= 1 && prm->L <= 128 && prm->P >= 4 && prm->P <= 8 &&
           !prm->force_slow && !prm->wave_only;
}

extern "C" int tr_lane_launch(const tr_params *prm, tr_job *d_jobs, int njobs, tr_ctx *d_ctxs, hipStream_t stream)
{
    if (njobs <= 0) return 0;
    if (!tr_lane_eligible(prm)) return -1;
    const int P = prm->P, HW = prm->L > 64 ? 2 : 1;
    dim3 grid((unsigned)((njobs + 63) / 64)), block(64);
#define TL_CASE(PP, HH) \
    if (P == PP && HW == HH) { \
        hipLaunchKernelGGL((tr_lane_jobs<PP, HH>), grid, block, 0, stream, *prm, d_jobs, njobs, d_ctxs); \
        return hipGetLastError() == hipSuccess ? 0 : -1; }
    TL_CASE(4, 1) TL_CASE(5, 1) TL_CASE(6, 1) TL_CASE(7, 1) TL_CASE(8, 1)
    TL_CASE(4, 2) TL_CASE(5, 2) TL_CASE(6, 2) TL_CASE(7, 2) TL_CASE(8, 2)
    return -1;
}
