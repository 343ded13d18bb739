/*
 * Host orchestration of Density-Aligned Stitching on the device.
 *
 * Reference: src/dsp_plugin.c:1314-2761 (one chunk), generalised from its
 * per-chunk limit of 8 segments to any segment count.  Per run:
 *   plan       nominal split (:1941-1957) by segment length or count;
 *              boundary reposition on channel 0 (:1963-2034): window sums
 *              on the device, sqrt and the decisions on the host;
 *   seeding    seg0-first (:2185-2266): seg0 on the persistent context,
 *              seg1 an exact continuation, 16-level estimate for seg >= 2;
 *              or all segments seeded from the persistent state (:2268-2283);
 *   segments   every segment of every channel in one trellis launch
 *              (:2288-2363, src/engine.c:915-938);
 *   stitch     density/candidate/decode search on the device per boundary
 *              (channel 0), ratio + selection on the host (:2411-2681);
 *   assemble   the same positions applied to all channels (:2683-2713);
 *   copy-back  persistent = last segment (:2715-2724).
 * In full-parallel mode the copied segments read a duplicate of the
 * persistent context taken before seg0 runs (the reference reads the
 * persistent context while seg0 runs on it); after the copy-back the
 * persistent contexts' hash links are dropped, as the oracle does.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "dsdhip_internal.h"

extern "C" __global__ void das_sumsq(const double *, const int64_t *, int, double *);
extern "C" __global__ void das_estimate(const double *, int64_t, int64_t, int64_t, const int64_t *,
                                        int, const double *, const double *, const double *, int,
                                        double, int, int, double *);
extern "C" __global__ void das_stitch(const unsigned long long *, const int64_t *, const int64_t *,
                                      const int *, int, const double *, int *, int *, int *, int *,
                                      double *);
extern "C" __global__ void das_assemble(const unsigned long long *, const int64_t *,
                                        const int64_t *, const int64_t *, int,
                                        unsigned long long *, int64_t, int64_t);
extern "C" __global__ void bits_to_f32(const unsigned long long *, int64_t, int64_t, float *, int64_t);

#define QN 33

static double nan_val() { return __builtin_nan(""); }

struct dsdhip_das {
    double a[10], g[10];
    int order, depth, K, L, nch;
    double state_limit;
    size_t O;
    size_t seg_len;
    int nseg_req, seg0_first, reposition;
    dsdhip_trellis_t *tr;
    int cap_seg;                 /* segments per channel the pool holds */
    int timing;
    hipStream_t stream;
    double qfir[255];
    double *d_qfir;
    /* last run */
    int last_S;
    std::vector<int> last_stitch;
    int64_t last_slow;
};

/* 255-tap Kaiser (beta 8) low-pass used to score stitch candidates
 * (src/dsp_plugin.c:2571-2604). */
extern "C" void dsdhip_make_qfir(double *q)
{
    const int N = 255;
    const double fc = 20000.0 / (5644800.0 / 2.0);
    double sum = 0;
    for (int i = 0; i < N; i++) {
        const double n = i - 127.0;
        if (n > -0.001 && n < 0.001) q[i] = 2.0 * fc;
        else q[i] = sin(2.0 * 3.14159265358979 * fc * n) / (3.14159265358979 * n);
        const double x = 2.0 * i / (N - 1) - 1.0;
        const double kb = 8.0;
        const double arg = kb * sqrt(1.0 - x * x > 0 ? 1.0 - x * x : 0);
        double i0a = 1.0, term = 1.0;
        for (int m = 1; m <= 20; m++) {
            term *= (arg / 2.0 / m) * (arg / 2.0 / m);
            i0a += term;
        }
        double i0k = 1.0;
        term = 1.0;
        for (int m = 1; m <= 20; m++) {
            term *= (kb / 2.0 / m) * (kb / 2.0 / m);
            i0k += term;
        }
        q[i] *= i0a / i0k;
        sum += q[i];
    }
    for (int i = 0; i < N; i++) q[i] /= sum;
}

/* context indices in the pool */
static inline int ix_persist(const dsdhip_das_t *, int ch) { return ch; }
static inline int ix_snap(const dsdhip_das_t *d, int ch) { return d->nch + ch; }
static inline int ix_seg(const dsdhip_das_t *d, int ch, int seg) { return 2 * d->nch + ch * d->cap_seg + seg; }

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    int ensure(size_t m)
    {
        if (m <= n) return 0;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc(&p, sizeof(T) * (m ? m : 1)) != hipSuccess) return -1;
        n = m;
        return 0;
    }
    ~DevBuf() { if (p) hipFree(p); }
};

extern "C" {

dsdhip_das_t *dsdhip_das_create_coeffs(const double *a, const double *g, int order, int depth,
                                       int cands, int lat, double state_limit, int nch,
                                       size_t seg_len, int nseg, int seg0_first, int reposition)
{
    if (!a || !g || order < 4 || order > TR_MAXP) return NULL;
    if (nch < 1 || lat < 1 || (seg_len == 0 && nseg < 1)) return NULL;
    if ((size_t)64 * (size_t)lat > 8192) return NULL;   /* overlap held in LDS by the stitch kernel */
    if (seg_len > 0 && seg_len < (size_t)128 * (size_t)lat) return NULL;   /* segments >= 2 overlaps */
    dsdhip_das_t *d = new dsdhip_das();
    memset(d->a, 0, sizeof(d->a));
    memset(d->g, 0, sizeof(d->g));
    memcpy(d->a, a, sizeof(double) * (size_t)order);
    memcpy(d->g, g, sizeof(double) * (size_t)order);
    d->order = order;
    d->depth = depth;
    d->K = cands;
    d->L = lat;
    d->nch = nch;
    d->state_limit = state_limit;
    d->O = (size_t)64 * (size_t)lat;
    d->seg_len = seg_len;
    d->nseg_req = nseg;
    d->seg0_first = seg0_first;
    d->reposition = reposition;
    d->cap_seg = 0;
    d->timing = 0;
    d->tr = NULL;
    d->stream = 0;
    d->d_qfir = NULL;
    d->last_S = 0;
    d->last_slow = 0;
    dsdhip_make_qfir(d->qfir);
    if (dsdhip_ensure_device() != 0 ||
        hipMalloc(&d->d_qfir, sizeof(d->qfir)) != hipSuccess ||
        hipMemcpy(d->d_qfir, d->qfir, sizeof(d->qfir), hipMemcpyHostToDevice) != hipSuccess) {
        delete d;
        return NULL;
    }
    /* persistent contexts exist from the start (fresh) */
    d->tr = dsdhip_trellis_create_coeffs(d->a, d->g, order, depth, cands, lat, state_limit, 2 * nch);
    if (!d->tr) {
        hipFree(d->d_qfir);
        delete d;
        return NULL;
    }
    return d;
}

dsdhip_das_t *dsdhip_das_create(int ntf_index, int depth, int cands, int lat, double state_limit,
                                int nch, size_t seg_len, int nseg, int seg0_first, int reposition)
{
    const dsdhip_ntf_t *f = dsdhip_ntf_get(ntf_index);
    if (!f) return NULL;
    return dsdhip_das_create_coeffs(f->a, f->g, f->order, depth, cands, lat, state_limit, nch,
                                    seg_len, nseg, seg0_first, reposition);
}

void dsdhip_das_destroy(dsdhip_das_t *d)
{
    if (!d) return;
    dsdhip_trellis_destroy(d->tr);
    if (d->d_qfir) hipFree(d->d_qfir);
    delete d;
}

int dsdhip_das_set_stream(dsdhip_das_t *d, void *stream)
{
    if (!d) return -1;
    d->stream = (hipStream_t)stream;
    return dsdhip_trellis_set_stream(d->tr, stream);
}

int dsdhip_das_force_literal(dsdhip_das_t *d, int on) { return d ? dsdhip_trellis_force_literal(d->tr, on) : -1; }

/* Plan for F samples: returns the segment count; fills start/size if given. */
int dsdhip_das_plan(const dsdhip_das_t *d, size_t F, size_t *start, size_t *size)
{
    int S;
    if (d->seg_len > 0) {
        S = (int)(F / d->seg_len);
        if (S < 1) S = 1;
    } else {
        S = d->nseg_req;
        size_t min_seg = d->O * 4;
        if (min_seg < 256) min_seg = 256;
        while (S > 1 && F / (size_t)S < min_seg) S--;
    }
    if (start && size) {
        size_t pos = 0;
        if (d->seg_len > 0) {
            for (int s = 0; s < S; s++) {
                start[s] = pos;
                size[s] = (s == S - 1) ? F - pos : d->seg_len;
                pos += size[s];
            }
        } else {
            const size_t base = F / (size_t)S, rem = F % (size_t)S;
            for (int s = 0; s < S; s++) {
                start[s] = pos;
                size[s] = (s == S - 1) ? F - pos : base + ((size_t)s < rem ? 1 : 0);
                pos += size[s];
            }
        }
    }
    return S;
}

/* grow the context pool to hold S segments per channel, keeping the
 * persistent contexts */
static int ensure_pool(dsdhip_das_t *d, int S)
{
    if (S <= d->cap_seg) return 0;
    int cap = S;
    dsdhip_trellis_t *nt = dsdhip_trellis_create_coeffs(d->a, d->g, d->order, d->depth, d->K, d->L,
                                                        d->state_limit, 2 * d->nch + d->nch * cap);
    if (!nt) return -1;
    dsdhip_trellis_set_stream(nt, d->stream);
    for (int ch = 0; ch < d->nch; ch++) {
        tr_ctx c;
        if (dsdhip_trellis_get(d->tr, ix_persist(d, ch), &c) != 0) return -1;
        if (dsdhip_trellis_put(nt, ch, &c) != 0) return -1;
    }
    if (d->timing) {
        double ms;
        int64_t nl, ns;
        dsdhip_trellis_timing_read(d->tr, &ms, &nl, &ns, 0);
        dsdhip_trellis_timing(nt, 1);
    }
    dsdhip_trellis_destroy(d->tr);
    d->tr = nt;
    d->cap_seg = cap;
    return 0;
}

/* Fresh persistent contexts (sdm_context_reset on every channel). */
int dsdhip_das_reset(dsdhip_das_t *d)
{
    if (!d) return -1;
    for (int ch = 0; ch < d->nch; ch++)
        if (dsdhip_trellis_reset(d->tr, ix_persist(d, ch)) != 0) return -1;
    return 0;
}

int dsdhip_das_timing(dsdhip_das_t *d, int on)
{
    if (!d) return -1;
    d->timing = on ? 1 : 0;
    return dsdhip_trellis_timing(d->tr, on);
}

int dsdhip_das_timing_read(dsdhip_das_t *d, double *ms, int64_t *launches, int64_t *samples, int clear)
{
    if (!d) return -1;
    return dsdhip_trellis_timing_read(d->tr, ms, launches, samples, clear);
}

static int reposition(dsdhip_das_t *d, const double *d_ch0, size_t F, std::vector<size_t> &st,
                      std::vector<size_t> &sz)
{
    const int S = (int)st.size();
    if (S <= 2) return 0;
    const size_t w = 2048;
    const double loud = 0.005;
    const size_t base = F / (size_t)S;
    size_t max_shift = base * 2 / 5;
    if (max_shift > 200000) max_shift = 200000;
    const size_t step = w / 2;
    const size_t O = d->O;
    /* pass 1: current boundaries */
    std::vector<int64_t> centers;
    std::vector<int> segs;
    for (int seg = 2; seg < S; seg++) {
        const size_t b = st[(size_t)seg];
        if (b < w || b + w >= F) continue;
        centers.push_back((int64_t)b);
        segs.push_back(seg);
    }
    if (centers.empty()) return 0;
    DevBuf<int64_t> dc;
    DevBuf<double> ds;
    std::vector<double> sums(centers.size());
    auto run = [&](const std::vector<int64_t> &cv, std::vector<double> &out) -> int {
        out.resize(cv.size());
        if (cv.empty()) return 0;
        if (dc.ensure(cv.size()) || ds.ensure(cv.size())) return -1;
        if (hipMemcpyAsync(dc.p, cv.data(), sizeof(int64_t) * cv.size(), hipMemcpyHostToDevice,
                           d->stream) != hipSuccess)
            return -1;
        const int n = (int)cv.size();
        hipLaunchKernelGGL(das_sumsq, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream,
                           d_ch0, dc.p, n, ds.p);
        if (hipMemcpyAsync(out.data(), ds.p, sizeof(double) * cv.size(), hipMemcpyDeviceToHost,
                           d->stream) != hipSuccess)
            return -1;
        return hipStreamSynchronize(d->stream) == hipSuccess ? 0 : -1;
    };
    if (run(centers, sums) != 0) return -1;
    /* pass 2: candidates around quiet boundaries (all offsets, decided below) */
    std::vector<int64_t> cc;
    std::vector<size_t> first(centers.size(), 0), count(centers.size(), 0);
    for (size_t i = 0; i < centers.size(); i++) {
        const double cur = sqrt(sums[i] / (double)w);
        if (cur > loud) continue;
        first[i] = cc.size();
        const size_t b = (size_t)centers[i];
        for (size_t off = step; off <= max_shift; off += step)
            for (int dir = -1; dir <= 1; dir += 2) {
                const size_t cand = (dir < 0) ? (b > off ? b - off : 0) : b + off;
                const bool inr = !(cand < w / 2 || cand + w / 2 >= F);
                cc.push_back(inr ? (int64_t)cand : (int64_t)w);   /* placeholder when out of range */
                count[i]++;
            }
    }
    std::vector<double> csum;
    if (run(cc, csum) != 0) return -1;
    for (size_t i = 0; i < centers.size(); i++) {
        const int seg = segs[i];
        const double cur = sqrt(sums[i] / (double)w);
        if (cur > loud) continue;
        const size_t b = st[(size_t)seg];
        size_t best = b;
        double best_rms = cur;
        size_t k = first[i];
        for (size_t off = step; off <= max_shift; off += step) {
            for (int dir = -1; dir <= 1; dir += 2, k++) {
                const size_t cand = (dir < 0) ? (b > off ? b - off : 0) : b + off;
                if (cand < w / 2 || cand + w / 2 >= F) continue;
                if (cand <= st[(size_t)seg - 1] + O * 2) continue;
                if (seg < S - 1 && cand >= st[(size_t)seg + 1] - O * 2) continue;
                const double r = sqrt(csum[k] / (double)w);
                if (r > best_rms) {
                    best_rms = r;
                    best = cand;
                }
            }
            if (best_rms > loud) break;
        }
        if (best != b) {
            st[(size_t)seg] = best;
            sz[(size_t)seg - 1] = best - st[(size_t)seg - 1];
            sz[(size_t)seg] = (seg < S - 1) ? st[(size_t)seg + 1] - best : F - best;
        }
    }
    return 0;
}

/*
 * Run DAS over one chunk.  d_in: nch rows of F doubles (row stride
 * in_stride); d_out: nch rows of out_words packed-bit words (1 = +1.0).
 * *out_count receives the assembled length (channel 0's, as the reference).
 */
int dsdhip_das_run_dev(dsdhip_das_t *d, const double *d_in, int64_t in_stride, size_t F,
                       unsigned long long *d_out, int64_t out_words, size_t *out_count)
{
    if (!d || !d_in || !d_out || F == 0) return -1;
    const int nch = d->nch;
    const size_t O = d->O;
    const int L = d->L;
    const int S = dsdhip_das_plan(d, F, NULL, NULL);
    std::vector<size_t> st((size_t)S), sz((size_t)S);
    dsdhip_das_plan(d, F, st.data(), sz.data());
    if (d->reposition && reposition(d, d_in, F, st, sz) != 0) return -1;
    if (ensure_pool(d, S) != 0) return -1;
    dsdhip_trellis_t *tr = d->tr;
    const int seg0_first = d->seg0_first;

    /* segment inputs / keep rules */
    std::vector<int64_t> in_start((size_t)S), in_cnt((size_t)S), keep((size_t)S);
    for (int seg = 0; seg < S; seg++) {
        const size_t ns = st[(size_t)seg], nz = sz[(size_t)seg];
        if (seg == 0) {
            size_t cnt = nz;
            if (!seg0_first && S > 1 && nz + O <= F) cnt += O;
            in_start[0] = 0;
            in_cnt[0] = (int64_t)cnt;
            keep[0] = 0;
        } else if (seg == 1 && seg0_first) {
            size_t cnt = nz;
            if (S > 2 && ns + nz + O <= F) cnt += O;
            in_start[1] = (int64_t)ns;
            in_cnt[1] = (int64_t)cnt;
            keep[1] = 0;
        } else {
            const bool ext = seg < S - 1;
            const size_t is = ns >= O ? ns - O : 0;
            size_t cnt = ns + nz - is;
            if (ext && ns + nz + O <= F) cnt += O;
            const size_t disc = ns >= O ? O - (size_t)L : 0;
            in_start[(size_t)seg] = (int64_t)is;
            in_cnt[(size_t)seg] = (int64_t)cnt;
            keep[(size_t)seg] = disc ? (int64_t)L + (int64_t)disc : 0;
        }
    }
    /* packed outputs: one region per (ch, seg) */
    std::vector<int64_t> woff((size_t)(nch * S));
    int64_t wtot = 0;
    for (int ch = 0; ch < nch; ch++)
        for (int seg = 0; seg < S; seg++) {
            woff[(size_t)(ch * S + seg)] = wtot;
            wtot += (in_cnt[(size_t)seg] + 63) / 64 + 1;
        }
    DevBuf<unsigned long long> segbits;
    if (segbits.ensure((size_t)wtot)) return -1;
    std::vector<int64_t> kept((size_t)(nch * S), 0), slow((size_t)(nch * S), 0);

    auto run = [&](const std::vector<int> &segs, const std::vector<int> &srcs,
                   const std::vector<double> &seeds) -> int {
        const int nj = (int)segs.size();
        if (nj == 0) return 0;
        std::vector<int> ctx((size_t)nj), cs((size_t)nj);
        std::vector<const double *> pin((size_t)nj);
        std::vector<uint64_t *> pout((size_t)nj);
        std::vector<int64_t> ni((size_t)nj), kf((size_t)nj), kp((size_t)nj), sl((size_t)nj);
        for (int j = 0; j < nj; j++) {
            const int ch = segs[(size_t)j] / S, seg = segs[(size_t)j] % S;
            ctx[(size_t)j] = (seg == 0) ? ix_persist(d, ch) : ix_seg(d, ch, seg);
            cs[(size_t)j] = srcs[(size_t)j];
            pin[(size_t)j] = d_in + (int64_t)ch * in_stride + in_start[(size_t)seg];
            pout[(size_t)j] = (uint64_t *)(segbits.p + woff[(size_t)(ch * S + seg)]);
            ni[(size_t)j] = in_cnt[(size_t)seg];
            kf[(size_t)j] = keep[(size_t)seg];
        }
        int err = 0;
        int rc = dsdhip_trellis_run_jobs(tr, nj, ctx.data(), cs.data(), pin.data(), ni.data(),
                                         kf.data(), pout.data(), seeds.data(), kp.data(),
                                         sl.data(), &err);
        for (int j = 0; j < nj; j++) {
            kept[(size_t)segs[(size_t)j]] = kp[(size_t)j];
            slow[(size_t)segs[(size_t)j]] = sl[(size_t)j];
        }
        return rc;
    };
    const double nan = nan_val();
    std::vector<int> segs, srcs;
    std::vector<double> seeds;
    auto add = [&](int ch, int seg, int src, const double *seed) {
        segs.push_back(ch * S + seg);
        srcs.push_back(src);
        for (int q = 0; q < 10; q++) seeds.push_back(seed ? seed[q] : nan);
    };
    if (seg0_first) {
        for (int ch = 0; ch < nch; ch++) add(ch, 0, -1, NULL);
        if (run(segs, srcs, seeds) != 0) return -1;
        segs.clear(); srcs.clear(); seeds.clear();
        std::vector<double> est;
        if (S > 2) {
            /* estimate from seg1's start over the rest (src/threadpool.c:345-366) */
            std::vector<double> init((size_t)nch * 10);
            for (int ch = 0; ch < nch; ch++)
                if (dsdhip_trellis_best_state(tr, ix_persist(d, ch), &init[(size_t)ch * 10]) != 0) return -1;
            const int64_t es = (int64_t)st[1];
            std::vector<int64_t> bnd;
            for (int seg = 1; seg < S; seg++) bnd.push_back(seg >= 2 ? (int64_t)st[(size_t)seg] - es : 0);
            tr_params prm;
            dsdhip_trellis_params(tr, &prm);
            DevBuf<int64_t> db;
            DevBuf<double> dinit, dest, dab;
            if (db.ensure(bnd.size()) || dinit.ensure(init.size()) ||
                dest.ensure((size_t)nch * bnd.size() * 10) || dab.ensure(2 * TR_MAXP))
                return -1;
            double ab[2 * TR_MAXP];
            memcpy(ab, prm.a, sizeof(prm.a));
            memcpy(ab + TR_MAXP, prm.g, sizeof(prm.g));
            hipMemcpyAsync(db.p, bnd.data(), sizeof(int64_t) * bnd.size(), hipMemcpyHostToDevice, d->stream);
            hipMemcpyAsync(dinit.p, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice, d->stream);
            hipMemcpyAsync(dab.p, ab, sizeof(ab), hipMemcpyHostToDevice, d->stream);
            hipLaunchKernelGGL(das_estimate, dim3((unsigned)((nch + 63) / 64)), dim3(64), 0, d->stream, d_in, in_stride, es,
                               (int64_t)F - es, db.p, (int)bnd.size(), dinit.p, dab.p,
                               dab.p + TR_MAXP, prm.P, prm.state_limit, 16, nch, dest.p);
            est.resize((size_t)nch * bnd.size() * 10);
            hipMemcpyAsync(est.data(), dest.p, sizeof(double) * est.size(), hipMemcpyDeviceToHost, d->stream);
            if (hipStreamSynchronize(d->stream) != hipSuccess) return -1;
        }
        const int nb = S - 1;
        for (int ch = 0; ch < nch; ch++) {
            if (S > 1) add(ch, 1, ix_persist(d, ch), NULL);
            for (int seg = 2; seg < S; seg++)
                add(ch, seg, ix_persist(d, ch), &est[((size_t)ch * nb + (size_t)(seg - 1)) * 10]);
        }
        if (run(segs, srcs, seeds) != 0) return -1;
    } else {
        std::vector<double> pst((size_t)nch * 10);
        for (int ch = 0; ch < nch; ch++) {
            if (dsdhip_trellis_best_state(tr, ix_persist(d, ch), &pst[(size_t)ch * 10]) != 0) return -1;
            if (dsdhip_trellis_clone(tr, ix_snap(d, ch), ix_persist(d, ch)) != 0) return -1;
        }
        for (int ch = 0; ch < nch; ch++) {
            add(ch, 0, -1, NULL);
            for (int seg = 1; seg < S; seg++) add(ch, seg, ix_snap(d, ch), &pst[(size_t)ch * 10]);
        }
        if (run(segs, srcs, seeds) != 0) return -1;
    }
    int64_t tslow = 0;
    for (size_t i = 0; i < slow.size(); i++) tslow += slow[i];
    d->last_slow = tslow;

    /* ---- stitch on channel 0 ---- */
    const int first = seg0_first ? 2 : 1;
    std::vector<int> bp((size_t)S, 0);
    int64_t s0 = kept[0];
    if (!seg0_first && s0 > (int64_t)sz[0]) s0 = (int64_t)sz[0];
    const int nbd = S - first > 0 ? S - first : 0;
    std::vector<int64_t> pbit((size_t)nbd > 0 ? (size_t)nbd : 1), cbit((size_t)nbd > 0 ? (size_t)nbd : 1);
    std::vector<int> ovl((size_t)nbd > 0 ? (size_t)nbd : 1);
    for (int k = first; k < S; k++) {
        const int64_t prev_end = (k - 1 == 0) ? s0 : kept[(size_t)k - 1];
        if (prev_end < 2 * (int64_t)O || kept[(size_t)k] == 0) return -1;   /* plan too fine for span-local stitching */
        pbit[(size_t)(k - first)] = woff[(size_t)(k - 1)] * 64 + prev_end - (int64_t)O;
        cbit[(size_t)(k - first)] = woff[(size_t)k] * 64;
        ovl[(size_t)(k - first)] = (int)((int64_t)O < kept[(size_t)k] ? (int64_t)O : kept[(size_t)k]);
    }
    if (nbd > 0) {
        DevBuf<int64_t> dpb, dcb;
        DevBuf<int> dov, dnc, dpk, dpos, dden;
        DevBuf<double> dmx;
        if (dpb.ensure((size_t)nbd) || dcb.ensure((size_t)nbd) || dov.ensure((size_t)nbd) ||
            dnc.ensure((size_t)nbd) || dpk.ensure((size_t)nbd) || dpos.ensure((size_t)nbd * QN) ||
            dden.ensure((size_t)nbd * QN) || dmx.ensure((size_t)nbd * QN))
            return -1;
        hipMemcpyAsync(dpb.p, pbit.data(), sizeof(int64_t) * (size_t)nbd, hipMemcpyHostToDevice, d->stream);
        hipMemcpyAsync(dcb.p, cbit.data(), sizeof(int64_t) * (size_t)nbd, hipMemcpyHostToDevice, d->stream);
        hipMemcpyAsync(dov.p, ovl.data(), sizeof(int) * (size_t)nbd, hipMemcpyHostToDevice, d->stream);
        hipLaunchKernelGGL(das_stitch, dim3((unsigned)nbd), dim3(256), 0, d->stream,
                           segbits.p, dpb.p, dcb.p, dov.p, L, d->d_qfir, dnc.p, dpk.p, dpos.p,
                           dden.p, dmx.p);
        std::vector<int> nc((size_t)nbd), pk((size_t)nbd), pos((size_t)nbd * QN), den((size_t)nbd * QN);
        std::vector<double> mx((size_t)nbd * QN);
        hipMemcpyAsync(nc.data(), dnc.p, sizeof(int) * nc.size(), hipMemcpyDeviceToHost, d->stream);
        hipMemcpyAsync(pk.data(), dpk.p, sizeof(int) * pk.size(), hipMemcpyDeviceToHost, d->stream);
        hipMemcpyAsync(pos.data(), dpos.p, sizeof(int) * pos.size(), hipMemcpyDeviceToHost, d->stream);
        hipMemcpyAsync(den.data(), dden.p, sizeof(int) * den.size(), hipMemcpyDeviceToHost, d->stream);
        hipMemcpyAsync(mx.data(), dmx.p, sizeof(double) * mx.size(), hipMemcpyDeviceToHost, d->stream);
        if (hipStreamSynchronize(d->stream) != hipSuccess) return -1;
        for (int i = 0; i < nbd; i++) {
            const int n = nc[(size_t)i];
            const int peak = pk[(size_t)i];
            std::vector<double> qual((size_t)n);
            for (int c = 0; c < n; c++) {
                if (n == 1) { qual[0] = 0.0; break; }
                double dr = (double)den[(size_t)i * QN + c] / (double)(peak > 0 ? peak : 1);
                if (dr < 0.5) dr = 0.5;
                qual[(size_t)c] = mx[(size_t)i * QN + c] / dr;
            }
            int best = 0;
            for (int c = 1; c < n; c++) {
                const double rq = qual[(size_t)c] / (qual[(size_t)best] + 1e-15);
                if (rq < 0.8) best = c;
                else if (rq < 1.2 && den[(size_t)i * QN + c] > den[(size_t)i * QN + best]) best = c;
            }
            bp[(size_t)(i + first)] = pos[(size_t)i * QN + best];
        }
    }

    /* ---- assemble (same positions on every channel) ---- */
    std::vector<int64_t> sbit, dbit, len;
    int64_t total = 0;
    for (int ch = 0; ch < nch; ch++) {
        std::vector<int64_t> at((size_t)S), src((size_t)S), cc((size_t)S);
        int64_t c0 = kept[(size_t)(ch * S)];
        if (!seg0_first && c0 > (int64_t)sz[0]) c0 = (int64_t)sz[0];
        at[0] = 0;
        src[0] = woff[(size_t)(ch * S)] * 64;
        cc[0] = c0;
        int64_t wp = c0;
        for (int k = 1; k < S; k++) {
            const int64_t kk = kept[(size_t)(ch * S + k)];
            if (kk == 0) { at[(size_t)k] = wp; src[(size_t)k] = 0; cc[(size_t)k] = 0; continue; }
            int64_t a, b;
            if (k == 1 && seg0_first) { a = wp; b = 0; }
            else {
                const int64_t pst = wp >= (int64_t)O ? wp - (int64_t)O : 0;
                b = bp[(size_t)k];
                a = pst + b;
            }
            at[(size_t)k] = a;
            src[(size_t)k] = woff[(size_t)(ch * S + k)] * 64 + b;
            cc[(size_t)k] = kk - b;
            wp = a + (kk - b);
        }
        if (ch == 0) total = wp;
        for (int k = 0; k < S; k++) {
            /* effective span: up to where the next non-empty span starts */
            int64_t e = at[(size_t)k] + cc[(size_t)k];
            for (int k2 = k + 1; k2 < S; k2++)
                if (cc[(size_t)k2] > 0) { if (at[(size_t)k2] < e) e = at[(size_t)k2]; break; }
            sbit.push_back(src[(size_t)k]);
            dbit.push_back(at[(size_t)k]);
            len.push_back(e > at[(size_t)k] ? e - at[(size_t)k] : 0);
        }
    }
    if ((total + 63) / 64 > out_words) return -1;
    {
        DevBuf<int64_t> ds, dd, dl;
        if (ds.ensure(sbit.size()) || dd.ensure(dbit.size()) || dl.ensure(len.size())) return -1;
        hipMemcpyAsync(ds.p, sbit.data(), sizeof(int64_t) * sbit.size(), hipMemcpyHostToDevice, d->stream);
        hipMemcpyAsync(dd.p, dbit.data(), sizeof(int64_t) * dbit.size(), hipMemcpyHostToDevice, d->stream);
        hipMemcpyAsync(dl.p, len.data(), sizeof(int64_t) * len.size(), hipMemcpyHostToDevice, d->stream);
        const int64_t words = out_words;
        hipLaunchKernelGGL(das_assemble, dim3((unsigned)((words + 255) / 256), (unsigned)nch), dim3(256), 0,
                           d->stream, segbits.p, ds.p, dd.p, dl.p, S, d_out, words, total);
        if (hipStreamSynchronize(d->stream) != hipSuccess) return -1;
    }
    /* ---- copy-back ---- */
    if (S > 1)
        for (int ch = 0; ch < nch; ch++)
            if (dsdhip_trellis_copy_detached(tr, ix_persist(d, ch), ix_seg(d, ch, S - 1)) != 0) return -1;
    d->last_S = S;
    d->last_stitch = bp;
    if (out_count) *out_count = (size_t)total;
    return 0;
}

int dsdhip_das_last_info(const dsdhip_das_t *d, int *nseg, int64_t *slow_steps, int *stitch_pos,
                         int max_pos)
{
    if (!d) return -1;
    if (nseg) *nseg = d->last_S;
    if (slow_steps) *slow_steps = d->last_slow;
    if (stitch_pos)
        for (int i = 0; i < max_pos && i < (int)d->last_stitch.size(); i++) stitch_pos[i] = d->last_stitch[(size_t)i];
    return 0;
}

/* Host-buffer form: in nch x F doubles, out nch x F floats (+-1.0). */
int dsdhip_das_run(dsdhip_das_t *d, const double *in, size_t F, float *out, size_t *out_count)
{
    if (!d) return -1;
    const int nch = d->nch;
    const int64_t words = (int64_t)(F + 63) / 64 + 1;
    double *din = NULL;
    unsigned long long *dbits = NULL;
    float *dout = NULL;
    int rc = -1;
    size_t cnt = 0;
    if (hipMalloc(&din, sizeof(double) * F * (size_t)nch) == hipSuccess &&
        hipMalloc(&dbits, sizeof(unsigned long long) * (size_t)words * (size_t)nch) == hipSuccess &&
        hipMalloc(&dout, sizeof(float) * F * (size_t)nch) == hipSuccess &&
        hipMemcpy(din, in, sizeof(double) * F * (size_t)nch, hipMemcpyHostToDevice) == hipSuccess &&
        dsdhip_das_run_dev(d, din, (int64_t)F, F, dbits, words, &cnt) == 0) {
        hipLaunchKernelGGL(bits_to_f32, dim3((unsigned)((cnt + 255) / 256 + 1), (unsigned)nch), dim3(256),
                           0, d->stream, dbits, words, (int64_t)cnt, dout, (int64_t)F);
        if (hipStreamSynchronize(d->stream) == hipSuccess &&
            hipMemcpy(out, dout, sizeof(float) * F * (size_t)nch, hipMemcpyDeviceToHost) == hipSuccess) {
            rc = 0;
            if (out_count) *out_count = cnt;
        }
    }
    if (din) hipFree(din);
    if (dbits) hipFree(dbits);
    if (dout) hipFree(dout);
    return rc;
}

} /* extern "C" */
