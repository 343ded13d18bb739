/*
 * DAS state estimator on host threads: the 16-level greedy loop-filter
 * recurrence of sdm_estimate_state_multibit (src/trellis.c:1118-1187), run
 * from seg1's start across every later segment boundary as the estimation
 * block of src/threadpool.c:345-366 does.
 *
 * Why the host: the recurrence is one dependent chain per channel (the
 * state after sample n feeds sample n+1 through the quantizer), so its
 * speed is the latency of that chain.  A GPU lane issues a dependent FP64
 * operation every few cycles at 2.4 GHz and the reference's form divides
 * twice per sample; one host core runs the same chain several times faster.
 *
 * Bit-exact with the reference, faster per sample:
 *   - the quantizer idx = clamp((int)((v / a0 + 1) / step + 0.5)) is a
 *     non-decreasing step function of v (for a0 > 0) up to the point where
 *     the int conversion overflows to INT_MIN (idx 0).  Its exact
 *     transition points are found once by bisection over the ordered
 *     doubles, evaluating the reference formula itself, so idx k holds
 *     exactly for v in [lo[k], hi[k]);
 *   - on the per-sample chain, idx is guessed by a truncation of a scaled,
 *     reassociated v: the terms that do not depend on the previous sample's
 *     quantizer output are summed first, so only one product and one
 *     addition separate that output from the next guess; off the chain,
 *     v in the reference's order is formed and checked against
 *     [lo[idx], hi[idx]).  A block whose check fails anywhere is
 *     recomputed from its start state with the reference's formula, so
 *     every state equals the reference's.
 * The state update itself (loop filter, y subtraction, clamp) is the
 * reference's operation order; build with -ffp-contract=off.
 */
#include <emmintrin.h>
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <thread>
#include <vector>

#include "dsdhip_internal.h"

namespace {

inline int x86_trunc(double w) { return _mm_cvttsd_si32(_mm_set_sd(w)); }

/* the reference's quantizer index (src/trellis.c:1162-1168) */
inline int ref_idx(double v, double a0, double step, int levels)
{
    const double yi = v / a0;
    int idx = x86_trunc((yi + 1.0) / step + 0.5);
    if (idx < 0) idx = 0;
    if (idx > levels - 1) idx = levels - 1;
    return idx;
}

inline uint64_t dkey(double d)
{
    uint64_t u;
    memcpy(&u, &d, 8);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}

inline double dunkey(uint64_t k)
{
    const uint64_t u = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    double d;
    memcpy(&d, &u, 8);
    return d;
}

struct Quant {
    int ok;              /* thresholds valid: the fast form may be used */
    int levels;
    double a0, step;
    double lo[32], hi[32], y[32];
    double over;         /* v >= over: conversion overflow, idx 0 */
    double c1, c2;       /* idx ~ trunc(v * c1 + c2) */
    /* one-step lookahead of the scaled quantizer input (no limiter):
     * c1 v_{t+1} + c2 ~ c2 + cx x_{t+1} + h1 (x_t - y_t) + r1 . s_{t-1} */
    double cx, h1, r1[TR_MAXP], hy[32];
};

void quant_init(Quant &q, double a0, int levels)
{
    memset(&q, 0, sizeof(q));
    q.levels = levels;
    q.a0 = a0;
    q.step = 2.0 / (double)(levels - 1);
    for (int k = 0; k < levels && k < 32; k++) q.y[k] = -1.0 + (double)k * q.step;
    if (!(a0 > 0.0) || levels < 3 || levels > 32 || !isfinite(a0)) return;
    /* overflow point: first v above a0 * 2 (idx = levels - 1) where idx is 0 */
    uint64_t l = dkey(2.0 * a0), h = dkey(INFINITY);
    if (ref_idx(dunkey(l), a0, q.step, levels) != levels - 1) return;
    while (h - l > 1) {
        const uint64_t m = l + (h - l) / 2;
        if (ref_idx(dunkey(m), a0, q.step, levels) == 0) h = m;
        else l = m;
    }
    q.over = dunkey(h);
    const uint64_t top = h - 1;   /* last v before the overflow */
    double t[33];
    t[0] = -INFINITY;
    for (int k = 1; k < levels; k++) {
        uint64_t lo = dkey(-INFINITY), hi = top;
        while (hi - lo > 1) {
            const uint64_t m = lo + (hi - lo) / 2;
            if (ref_idx(dunkey(m), a0, q.step, levels) >= k) hi = m;
            else lo = m;
        }
        t[k] = dunkey(hi);
        /* a step function: exactly k at the point, k-1 just below */
        if (ref_idx(t[k], a0, q.step, levels) != k || ref_idx(dunkey(hi - 1), a0, q.step, levels) != k - 1)
            return;
    }
    t[levels] = q.over;
    for (int k = 0; k < levels; k++) {
        q.lo[k] = t[k];
        q.hi[k] = t[k + 1];
    }
    q.c1 = 1.0 / (a0 * q.step);
    q.c2 = 1.0 / q.step + 0.5;
    q.ok = 1;
}

/* One sample in the reference's form (src/trellis.c:1140-1186). */
template <int P>
inline void ref_step(const double *a, const double *g, double *s, double x, const Quant &q, double lim)
{
    double ns[P];
    ns[0] = s[0] - g[0] * s[1] + x;
    double v = x + a[0] * ns[0];
    for (int i = 1; i < P - 1; i++) {
        ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        v = v + a[i] * ns[i];
    }
    ns[P - 1] = s[P - 1] + s[P - 2];
    v = v + a[P - 1] * ns[P - 1];
    double y;
    if (q.levels <= 2) y = (v * a[0] > 0.0) ? 1.0 : -1.0;
    else y = -1.0 + (double)ref_idx(v, a[0], q.step, q.levels) * q.step;
    ns[0] = ns[0] - y;
    if (lim > 0.0)
        for (int i = 0; i < P; i++) {
            if (ns[i] > lim) ns[i] = lim;
            else if (ns[i] < -lim) ns[i] = -lim;
        }
    for (int i = 0; i < P; i++) s[i] = ns[i];
}

/* n samples with the fast index; returns 0 if every index was confirmed
 * against the reference-order sum (s then holds the exact state), 1 if not
 * (s is then unspecified). */
template <int P, bool LIM>
int fast_block(const double *a, const double *g, double *sv, const double *x, size_t n, const Quant &q,
               double lim)
{
    double s[P];
    for (int i = 0; i < P; i++) s[i] = sv[i];
    const int top = q.levels - 1;
    const double c1 = q.c1, c2 = q.c2, over = q.over;
    const double e0 = a[0] * c1, e1 = a[1] * c1, es = e0 + e1;
    int bad = 0;
    for (size_t t = 0; t < n; t++) {
        const double xv = x[t];
        double ns[P], p[P];
        ns[0] = s[0] - g[0] * s[1] + xv;
        for (int i = 1; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        for (int i = 0; i < P; i++) p[i] = a[i] * ns[i];
        /* reference-order sum (checked, off the chain) */
        double ve = xv + p[0];
        for (int i = 1; i < P; i++) ve = ve + p[i];
        /* scaled, reassociated sum: ns0 and ns1 are s0 plus terms known
         * before the previous output, so only e * s0 sits on the chain */
        double r = xv + p[2];
        for (int i = 3; i < P; i++) r = r + p[i];
        const double base = ((r * c1 + c2) + e0 * (xv - g[0] * s[1])) + e1 * (s[1] - g[1] * s[2]);
        const double w = base + es * s[0];
        int k = x86_trunc(w);
        k = k < 0 ? 0 : (k > top ? top : k);
        const bool in = (ve >= q.lo[k]) & (ve < q.hi[k]);
        const bool zero_ok = (k == 0) & ((ve >= over) | (ve != ve));
        bad |= !(in | zero_ok);
        ns[0] = ns[0] - q.y[k];
        if (LIM)
            for (int i = 0; i < P; i++) {
                if (ns[i] > lim) ns[i] = lim;
                else if (ns[i] < -lim) ns[i] = -lim;
            }
        for (int i = 0; i < P; i++) s[i] = ns[i];
    }
    for (int i = 0; i < P; i++) sv[i] = s[i];
    return bad;
}

/* The same without a limiter, with the index of sample t+1 guessed one
 * step ahead: in exact arithmetic the loop filter is linear, so
 * c1 v_{t+1} + c2 = c2 + cx x_{t+1} + h1 (x_t - y_t) + r1 . s_{t-1}, and
 * the chain from one guess to the next is a table read, a subtraction and
 * the truncation.  The state still advances in the reference's operations
 * and every guess is checked against the reference-order sum. */
template <int P>
int fast_block_la(const double *a, const double *g, double *sv, const double *x, size_t n, const Quant &q)
{
    double s[P];
    for (int i = 0; i < P; i++) s[i] = sv[i];
    const int top = q.levels - 1;
    const double c1 = q.c1, c2 = q.c2, over = q.over, cx = q.cx, h1 = q.h1;
    const double e0 = a[0] * c1, e1 = a[1] * c1, es = e0 + e1;
    int bad = 0;
    /* first guess from the start state (the single-step form) */
    int k;
    {
        double ns[P];
        for (int i = 2; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        double r = x[0] + a[2] * ns[2];
        for (int i = 3; i < P; i++) r = r + a[i] * ns[i];
        const double w = (((r * c1 + c2) + e0 * (x[0] - g[0] * s[1])) + e1 * (s[1] - g[1] * s[2])) + es * s[0];
        k = x86_trunc(w);
        k = k < 0 ? 0 : (k > top ? top : k);
    }
    for (size_t t = 0; t < n; t++) {
        const double xv = x[t];
        const double xn = (t + 1 < n) ? x[t + 1] : 0.0;
        double ns[P], p[P];
        ns[0] = s[0] - g[0] * s[1] + xv;
        for (int i = 1; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        for (int i = 0; i < P; i++) p[i] = a[i] * ns[i];
        double ve = xv + p[0];
        for (int i = 1; i < P; i++) ve = ve + p[i];
        const bool in = (ve >= q.lo[k]) & (ve < q.hi[k]);
        const bool zero_ok = (k == 0) & ((ve >= over) | (ve != ve));
        bad |= !(in | zero_ok);
        /* next guess: terms of s_{t-1} other than s0 first, s0's term last */
        double qn = (c2 + cx * xn) + h1 * xv;
        for (int i = P - 1; i >= 1; i--) qn = qn + q.r1[i] * s[i];
        qn = qn + q.r1[0] * s[0];
        int kn = x86_trunc(qn - q.hy[k]);
        kn = kn < 0 ? 0 : (kn > top ? top : kn);
        ns[0] = ns[0] - q.y[k];
        for (int i = 0; i < P; i++) s[i] = ns[i];
        k = kn;
    }
    for (int i = 0; i < P; i++) sv[i] = s[i];
    return bad;
}

/* The same on FMA hardware, with the chain from one index guess to the
 * next shortened: the index is kept as a double (k = 0..levels-1 exactly),
 * y_t = -1 + k step is formed from it as the reference forms it, and the
 * next guess is w = (qn + h1) - (h1 step) k, truncated and clamped in
 * doubles: one multiply-add, a round toward zero, a max and a min, where
 * the integer form needs a table read, a subtraction and a conversion.
 * The index as an integer (for the interval check) is off the chain. */
/* trunc(w) clamped to [0, top] in doubles; NaN gives 0 */
__attribute__((target("sse4.1"))) static inline double clampk(double top, double w)
{
    const __m128d z = _mm_setzero_pd();
    const __m128d k = _mm_round_sd(z, _mm_set_sd(w), _MM_FROUND_TO_ZERO | _MM_FROUND_NO_EXC);
    return _mm_cvtsd_f64(_mm_min_sd(_mm_max_sd(k, z), _mm_set_sd(top)));
}

/* m1 * m2 + ad, rounded once (used only for the index guess) */
__attribute__((target("fma"))) static inline double fmad(double m1, double m2, double ad)
{
    return _mm_cvtsd_f64(_mm_fmadd_sd(_mm_set_sd(m1), _mm_set_sd(m2), _mm_set_sd(ad)));
}

template <int P>
__attribute__((target("avx512f,fma,sse4.1"))) int fast_block_f(const double *a, const double *g, double *sv,
                                                       const double *x, size_t n, const Quant &q)
{
    double s[P];
    for (int i = 0; i < P; i++) s[i] = sv[i];
    const double step = q.step, over = q.over, cx = q.cx, h1 = q.h1;
    const double ch = q.c2 + h1, bh = h1 * step;
    const double top = (double)(q.levels - 1);
    const double c1 = q.c1, c2 = q.c2;
    const double e0 = a[0] * c1, e1 = a[1] * c1, es = e0 + e1;
    double kd;
    {
        double ns[P];
        for (int i = 2; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        double r = x[0] + a[2] * ns[2];
        for (int i = 3; i < P; i++) r = r + a[i] * ns[i];
        kd = clampk(top, (((r * c1 + c2) + e0 * (x[0] - g[0] * s[1])) + e1 * (s[1] - g[1] * s[2])) + es * s[0]);
    }
    int bad = 0;
    for (size_t t = 0; t < n; t++) {
        const double xv = x[t];
        const double xn = (t + 1 < n) ? x[t + 1] : 0.0;
        double ns[P], p[P];
        ns[0] = s[0] - g[0] * s[1] + xv;
        for (int i = 1; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        for (int i = 0; i < P; i++) p[i] = a[i] * ns[i];
        double ve = xv + p[0];
        for (int i = 1; i < P; i++) ve = ve + p[i];
        const int k = (int)kd;
        const bool in = (ve >= q.lo[k]) & (ve < q.hi[k]);
        const bool zero_ok = (k == 0) & ((ve >= over) | (ve != ve));
        bad |= !(in | zero_ok);
        /* next guess: the terms of s_{t-1} other than s0 (they arrive a
         * step before s0) in four short multiply-add chains, which keeps the
         * adders (the reference-order work needs them) free; s0's term last */
        double c[4] = {fmad(cx, xn, ch), h1 * xv, 0.0, 0.0};
        for (int i = 1; i < P; i++) {
            const int j = (i - 1) & 3;
            c[j] = (i <= 2 || i >= 5) ? fmad(q.r1[i], s[i], c[j]) : q.r1[i] * s[i];
        }
        const double qn = fmad(q.r1[0], s[0], (c[0] + c[1]) + (c[2] + c[3]));
        const double kn = clampk(top, fmad(-bh, kd, qn));
        const double y = -1.0 + kd * step;
        ns[0] = ns[0] - y;
        for (int i = 0; i < P; i++) s[i] = ns[i];
        kd = kn;
    }
    for (int i = 0; i < P; i++) sv[i] = s[i];
    return bad;
}

/* The chain alone: the exact state advanced with guessed indices (no
 * check), each index written to kb.  Exact wherever every guess is right,
 * which a separate pass over the indices establishes (check_run). */
template <int P>
__attribute__((target("avx512f,fma,sse4.1"))) void guess_run(const double *a, const double *g, double *sv,
                                                             const double *x, size_t n, const Quant &q,
                                                             uint8_t *kb)
{
    double s[P];
    for (int i = 0; i < P; i++) s[i] = sv[i];
    const double step = q.step, cx = q.cx, h1 = q.h1;
    const double ch = q.c2 + h1, bh = h1 * step;
    const double top = (double)(q.levels - 1);
    const double c1 = q.c1, c2 = q.c2;
    const double e0 = a[0] * c1, e1 = a[1] * c1, es = e0 + e1;
    double kd;
    {
        double ns[P];
        for (int i = 2; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        double r = x[0] + a[2] * ns[2];
        for (int i = 3; i < P; i++) r = r + a[i] * ns[i];
        kd = clampk(top, (((r * c1 + c2) + e0 * (x[0] - g[0] * s[1])) + e1 * (s[1] - g[1] * s[2])) + es * s[0]);
    }
    for (size_t t = 0; t < n; t++) {
        const double xv = x[t];
        const double xn = (t + 1 < n) ? x[t + 1] : 0.0;
        double ns[P];
        ns[0] = s[0] - g[0] * s[1] + xv;
        for (int i = 1; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        double c[4] = {fmad(cx, xn, ch), h1 * xv, 0.0, 0.0};
        for (int i = 1; i < P; i++) {
            const int j = (i - 1) & 3;
            c[j] = (i <= 2 || i >= 5) ? fmad(q.r1[i], s[i], c[j]) : q.r1[i] * s[i];
        }
        const double qn = fmad(q.r1[0], s[0], (c[0] + c[1]) + (c[2] + c[3]));
        const double kn = clampk(top, fmad(-bh, kd, qn));
        kb[t] = (uint8_t)(int)kd;
        const double y = -1.0 + kd * step;
        ns[0] = ns[0] - y;
        for (int i = 0; i < P; i++) s[i] = ns[i];
        kd = kn;
    }
    for (int i = 0; i < P; i++) sv[i] = s[i];
}

/* The check of a guessed run: from the exact state sv, every index kb[t]
 * against the reference-order v; returns n if all hold (sv then holds the
 * exact state after the run), else the first sample whose index is wrong
 * (sv holds the exact state before it). */
template <int P>
__attribute__((target("avx512f,fma,sse4.1"))) size_t check_run(const double *a, const double *g, double *sv,
                                                               const double *x, size_t n, const Quant &q,
                                                               const uint8_t *kb)
{
    double s[P];
    for (int i = 0; i < P; i++) s[i] = sv[i];
    const double step = q.step, over = q.over;
    size_t t = 0;
    for (; t < n; t++) {
        const double xv = x[t];
        double ns[P], p[P];
        ns[0] = s[0] - g[0] * s[1] + xv;
        for (int i = 1; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        ns[P - 1] = s[P - 1] + s[P - 2];
        for (int i = 0; i < P; i++) p[i] = a[i] * ns[i];
        double ve = xv + p[0];
        for (int i = 1; i < P; i++) ve = ve + p[i];
        const int k = kb[t];
        const bool in = (ve >= q.lo[k]) & (ve < q.hi[k]);
        const bool zero_ok = (k == 0) & ((ve >= over) | (ve != ve));
        if (!(in | zero_ok)) break;
        const double y = -1.0 + (double)k * step;
        ns[0] = ns[0] - y;
        for (int i = 0; i < P; i++) s[i] = ns[i];
    }
    for (int i = 0; i < P; i++) sv[i] = s[i];
    return t;
}

static bool have_fma()
{
    static const int ok = __builtin_cpu_supports("sse4.1") && __builtin_cpu_supports("fma") ? 1 : 0;
    return ok != 0;
}

template <int P>
void advance(const double *a, const double *g, double *s, const double *x, size_t n, const Quant &q,
             double lim, int64_t *redo)
{
    const size_t BLK = 1024;
    const bool vec = have_fma();
    for (size_t off = 0; off < n; off += BLK) {
        const size_t m = (n - off < BLK) ? n - off : BLK;
        if (q.ok) {
            double keep[P];
            for (int i = 0; i < P; i++) keep[i] = s[i];
            const int bad = (lim > 0.0) ? fast_block<P, true>(a, g, s, x + off, m, q, lim)
                            : vec       ? fast_block_f<P>(a, g, s, x + off, m, q)
                                        : fast_block_la<P>(a, g, s, x + off, m, q);
            if (!bad) continue;
            for (int i = 0; i < P; i++) s[i] = keep[i];
            if (redo) (*redo)++;
        }
        for (size_t t = 0; t < m; t++) ref_step<P>(a, g, s, x[off + t], q, lim);
    }
}

} /* namespace */

struct dsdhip_est {
    Quant q;
    double a[TR_MAXP], g[TR_MAXP];
    int P;
    double lim;
    double s[10];
    int64_t redo;
};

/* ---- the chain on two threads ------------------------------------------
 * The guessing loop alone (exact state, guessed indices) and the check of
 * its indices (reference-order v against the index's interval, exact
 * state) each cost less than the two together, so a chain runs them on
 * two threads: the guesser writes indices into a ring, the checker (the
 * calling thread) follows it and owns the exact state.  A wrong index
 * stops the guesser; the checker takes that sample in the reference's
 * form and the next block in the single-thread form, then starts the
 * guesser again from its exact state.  Every published state is the
 * checker's, so the results are those of the single-thread form. */
#define EST_RING ((size_t)1 << 24)      /* indices in flight */
#define EST_CHUNK ((size_t)1 << 16)     /* guesser step, divides EST_RING */

struct EstSplit {
    const dsdhip_est_t *proto;
    const double *x;
    size_t count;
    int (*wait)(void *, size_t, size_t);
    void *ctx;
    std::vector<uint8_t> ring;
    std::atomic<size_t> a_pos{0}, c_pos{0};
    std::atomic<int> stop{0}, a_fail{0};
};

struct EstStart {
    double s[10];
};

template <int P>
static void guess_thread(EstSplit *sp, EstStart st, size_t pos)
{
    const dsdhip_est_t *e = sp->proto;
    double s[P];
    for (int i = 0; i < P; i++) s[i] = st.s[i];
    while (pos < sp->count && !sp->stop.load(std::memory_order_relaxed)) {
        const size_t end = std::min(sp->count, (pos / EST_CHUNK + 1) * EST_CHUNK);
        while (end - sp->c_pos.load(std::memory_order_acquire) > EST_RING) {
            if (sp->stop.load(std::memory_order_relaxed)) return;
            std::this_thread::yield();
        }
        if (sp->wait && sp->wait(sp->ctx, pos, end) != 0) {
            sp->a_fail.store(1);
            return;
        }
        guess_run<P>(e->a, e->g, s, sp->x + pos, end - pos, e->q, &sp->ring[pos % EST_RING]);
        pos = end;
        sp->a_pos.store(pos, std::memory_order_release);
    }
}

static bool split_usable(const dsdhip_est_t *e, size_t count)
{
    return e->q.ok && !(e->lim > 0.0) && have_fma() && __builtin_cpu_supports("avx512f") &&
           count >= 4 * EST_CHUNK && std::thread::hardware_concurrency() > 1;
}

template <int P>
static int chain_split(dsdhip_est_t *e, const double *x, size_t count, const int64_t *bnd, int nb,
                       void (*publish)(void *, int, const double *), int (*wait)(void *, size_t, size_t),
                       void *ctx)
{
    EstSplit sp;
    sp.proto = e;
    sp.x = x;
    sp.count = count;
    sp.wait = wait;
    sp.ctx = ctx;
    sp.ring.resize(EST_RING);
    auto start = [&]() {
        EstStart st;
        memcpy(st.s, e->s, sizeof(st.s));
        return st;
    };
    std::thread guesser(guess_thread<P>, &sp, start(), (size_t)0);
    int fails = 0, rc = 0;
    size_t pos = 0;
    for (int b = 0; b < nb && rc == 0; b++) {
        const int64_t bd = bnd[b];
        if (bd > (int64_t)pos && (size_t)bd <= count) {
            const size_t target = (size_t)bd;
            while (pos < target && rc == 0) {
                size_t avail;
                while ((avail = sp.a_pos.load(std::memory_order_acquire)) <= pos) {
                    if (sp.a_fail.load()) {
                        rc = -1;
                        break;
                    }
                    std::this_thread::yield();
                }
                if (rc) break;
                size_t end = std::min(target, avail);
                end = std::min(end, (pos / EST_RING + 1) * EST_RING);    /* no ring wrap inside a check */
                const size_t ok = check_run<P>(e->a, e->g, e->s, x + pos, end - pos, e->q, &sp.ring[pos % EST_RING]);
                if (ok == end - pos) {
                    pos = end;
                    sp.c_pos.store(pos, std::memory_order_release);
                    continue;
                }
                /* a wrong index at pos + ok: the guesser's later indices are void */
                sp.stop.store(1);
                guesser.join();
                pos += ok;
                e->redo++;
                ref_step<P>(e->a, e->g, e->s, x[pos], e->q, e->lim);
                pos++;
                const size_t blk = std::min(target - pos, (size_t)1024);
                if (blk && wait && wait(ctx, pos, pos + blk) != 0) return -1;
                if (blk) {
                    advance<P>(e->a, e->g, e->s, x + pos, blk, e->q, e->lim, &e->redo);
                    pos += blk;
                }
                sp.c_pos.store(pos, std::memory_order_release);
                sp.a_pos.store(pos, std::memory_order_release);
                if (++fails > 64) {
                    /* guesses keep failing: finish in the single-thread form */
                    for (int b2 = b; b2 < nb; b2++) {
                        const int64_t bd2 = bnd[b2];
                        if (bd2 > (int64_t)pos && (size_t)bd2 <= count) {
                            if (wait && wait(ctx, pos, (size_t)bd2) != 0) return -1;
                            advance<P>(e->a, e->g, e->s, x + pos, (size_t)bd2 - pos, e->q, e->lim, &e->redo);
                            pos = (size_t)bd2;
                        }
                        publish(ctx, b2, e->s);
                    }
                    return 0;
                }
                sp.stop.store(0);
                guesser = std::thread(guess_thread<P>, &sp, start(), pos);
            }
        }
        if (rc == 0) publish(ctx, b, e->s);
    }
    sp.stop.store(1);
    if (guesser.joinable()) guesser.join();
    return rc;
}

extern "C" {

/* Estimator over a stream fed block by block (state carried between calls). */
dsdhip_est_t *dsdhip_est_create(const double *a, const double *g, int order, double state_limit,
                                int levels, const double *init10)
{
    if (!a || !g || order < 4 || order > TR_MAXP || levels < 2 || levels > 32) return NULL;
    dsdhip_est_t *e = new dsdhip_est_t();
    memset(e->a, 0, sizeof(e->a));
    memset(e->g, 0, sizeof(e->g));
    memcpy(e->a, a, sizeof(double) * (size_t)order);
    memcpy(e->g, g, sizeof(double) * (size_t)order);
    e->P = order;
    e->lim = state_limit;
    quant_init(e->q, a[0], levels);
    if (e->q.ok) {
        /* loop filter without quantizer: ns = M s + e0 x (src/trellis.c:1140-1150) */
        const int P = order;
        double M[TR_MAXP][TR_MAXP] = {{0}};
        M[0][0] = 1.0;
        M[0][1] = -g[0];
        for (int i = 1; i < P - 1; i++) {
            M[i][i - 1] = 1.0;
            M[i][i] = 1.0;
            M[i][i + 1] = -g[i];
        }
        M[P - 1][P - 2] = 1.0;
        M[P - 1][P - 1] = 1.0;
        double aM[TR_MAXP] = {0}, aM2[TR_MAXP] = {0};
        for (int j = 0; j < P; j++)
            for (int i = 0; i < P; i++) aM[j] += a[i] * M[i][j];
        for (int j = 0; j < P; j++)
            for (int i = 0; i < P; i++) aM2[j] += aM[i] * M[i][j];
        e->q.cx = e->q.c1 * (1.0 + a[0]);
        e->q.h1 = e->q.c1 * aM[0];
        for (int j = 0; j < TR_MAXP; j++) e->q.r1[j] = j < P ? e->q.c1 * aM2[j] : 0.0;
        for (int k = 0; k < levels; k++) e->q.hy[k] = e->q.h1 * e->q.y[k];
    }
    for (int i = 0; i < 10; i++) e->s[i] = (init10 && i < order) ? init10[i] : 0.0;
    e->redo = 0;
    return e;
}

void dsdhip_est_destroy(dsdhip_est_t *e) { delete e; }

void dsdhip_est_set_state(dsdhip_est_t *e, const double *init10)
{
    for (int i = 0; i < 10; i++) e->s[i] = (i < e->P) ? init10[i] : 0.0;
}

void dsdhip_est_get_state(const dsdhip_est_t *e, double *out10)
{
    for (int i = 0; i < 10; i++) out10[i] = (i < e->P) ? e->s[i] : 0.0;
}

void dsdhip_est_feed(dsdhip_est_t *e, const double *x, size_t n)
{
    switch (e->P) {
    case 4: advance<4>(e->a, e->g, e->s, x, n, e->q, e->lim, &e->redo); break;
    case 5: advance<5>(e->a, e->g, e->s, x, n, e->q, e->lim, &e->redo); break;
    case 6: advance<6>(e->a, e->g, e->s, x, n, e->q, e->lim, &e->redo); break;
    case 7: advance<7>(e->a, e->g, e->s, x, n, e->q, e->lim, &e->redo); break;
    default: advance<8>(e->a, e->g, e->s, x, n, e->q, e->lim, &e->redo); break;
    }
}

/* The estimate over x[0..count) with the state published at each boundary
 * bnd[0..nb) (advanced only to boundaries past the current position and
 * within count, src/threadpool.c:345-366).  wait(ctx, pos, end), if given,
 * returns once x[pos..end) may be read (0) or on failure (nonzero).
 * threads = 2 lets the index guess run on a thread of its own. */
int dsdhip_est_chain(dsdhip_est_t *e, const double *x, size_t count, const int64_t *bnd, int nb,
                     void (*publish)(void *, int, const double *), int (*wait)(void *, size_t, size_t),
                     void *ctx, int threads)
{
    if (!e || !publish || (nb > 0 && !bnd)) return -1;
    if (threads > 1 && split_usable(e, count)) {
        switch (e->P) {
        case 4: return chain_split<4>(e, x, count, bnd, nb, publish, wait, ctx);
        case 5: return chain_split<5>(e, x, count, bnd, nb, publish, wait, ctx);
        case 6: return chain_split<6>(e, x, count, bnd, nb, publish, wait, ctx);
        case 7: return chain_split<7>(e, x, count, bnd, nb, publish, wait, ctx);
        default: return chain_split<8>(e, x, count, bnd, nb, publish, wait, ctx);
        }
    }
    size_t pos = 0;
    for (int b = 0; b < nb; b++) {
        const int64_t bd = bnd[b];
        if (bd > (int64_t)pos && (size_t)bd <= count) {
            if (wait && wait(ctx, pos, (size_t)bd) != 0) return -1;
            dsdhip_est_feed(e, x + pos, (size_t)bd - pos);
            pos = (size_t)bd;
        }
        publish(ctx, b, e->s);
    }
    return 0;
}

int64_t dsdhip_est_redo_blocks(const dsdhip_est_t *e) { return e->redo; }
int dsdhip_est_fast(const dsdhip_est_t *e) { return e->q.ok; }

/* The estimation block of src/threadpool.c:345-366 on host memory: from
 * init over x[0..count), boundaries bnd[0..nb) (offsets into x, the state
 * being advanced to a boundary only when it lies after the current position
 * and within count); out[b*10..] = the state at boundary b.  Pure host
 * code.  index_bias shifts the fast index guess (validation: the results
 * must not depend on it); redo_blocks receives the blocks recomputed in the
 * reference form. */
int dsdhip_estimate_boundaries_ex(const double *a, const double *g, int order, const double *init10,
                                  const double *x, size_t count, const int64_t *bnd, int nb,
                                  double state_limit, int levels, double *out10, double index_bias,
                                  int64_t *redo_blocks)
{
    if (!a || !g || !init10 || (!x && count) || (nb > 0 && (!bnd || !out10)) || order < 4 ||
        order > TR_MAXP || levels < 2 || levels > 32)
        return -1;
    dsdhip_est_t *e = dsdhip_est_create(a, g, order, state_limit, levels, init10);
    if (!e) return -1;
    e->q.c2 += index_bias;
    struct Out {
        double *o;
        int P;
    } out = {out10, order};
    const int rc = dsdhip_est_chain(e, x, count, bnd, nb,
                                    [](void *c, int b, const double *st) {
                                        Out *u = (Out *)c;
                                        for (int i = 0; i < 10; i++) u->o[(size_t)b * 10 + i] = i < u->P ? st[i] : 0.0;
                                    },
                                    nullptr, &out, 2);
    if (redo_blocks) *redo_blocks = e->redo;
    dsdhip_est_destroy(e);
    return rc;
}

int dsdhip_estimate_boundaries(const double *a, const double *g, int order, const double *init10,
                               const double *x, size_t count, const int64_t *bnd, int nb,
                               double state_limit, int levels, double *out10)
{
    return dsdhip_estimate_boundaries_ex(a, g, order, init10, x, count, bnd, nb, state_limit, levels, out10,
                                         0.0, NULL);
}

/* The quantizer index the fast form's exact interval table assigns to v
 * (levels 3..32, a0 > 0), or -1 when the table is not used for (a0,
 * levels).  For checking the table against the reference formula. */
int dsdhip_est_table_index(double a0, int levels, double v)
{
    Quant q;
    quant_init(q, a0, levels);
    if (!q.ok) return -1;
    if (v >= q.over || v != v) return 0;
    for (int k = levels - 1; k >= 0; k--)
        if (v >= q.lo[k]) return k;
    return 0;
}

} /* extern "C" */
