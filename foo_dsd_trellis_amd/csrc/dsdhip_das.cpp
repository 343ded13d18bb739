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
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
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

/* device buffer that only grows; freed with its owner */
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
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

/* per-object device scratch of a run, kept between calls (grow-only); like
 * the trellis contexts, one object is used from one thread at a time */
/* seg0-first estimate input staged in pinned host memory, copied in blocks
 * of EST_BLOCK samples with an event after each */
#define EST_BLOCK ((size_t)1 << 22)
struct DasScratch {
    DevBuf<int64_t> dc, db, dpb, dcb;
    DevBuf<double> ds, dinit, dest, dab, dmx;
    DevBuf<int> dov, dnc, dpk, dpos, dden;
    DevBuf<unsigned long long> segbits;
    std::vector<DevBuf<int64_t>> dsb, ddb, dlb;   /* per track */
    DevBuf<double> hin;                           /* host-buffer entry: staged input */
    DevBuf<unsigned long long> hbits;             /* ... packed output */
    DevBuf<float> hout;                           /* ... +-1.0 output */
    /* seg0-first pipeline: the batch streams, batch completion events, the
     * estimate input staged in pinned host memory and its block events */
    static const int NBS = 2;
    hipStream_t bs[NBS] = {0, 0};
    std::vector<hipEvent_t> bev, blk;
    hipEvent_t ev_go = 0, ev_t0 = 0, ev_t1 = 0;
    double *h_est = nullptr;
    size_t h_est_cap = 0;
    /* the staging copies run on their own stream, behind ev_in (the input
     * is ready) only, so they overlap segment 0; per estimate chain: its
     * offset in h_est and its first block event (npos: not staged) */
    hipStream_t cs = 0;
    hipEvent_t ev_in = 0;
    int staged = 0;
    std::vector<size_t> st_base, st_ev;
    int pipe_ok = 0;
    int init_pipe()
    {
        if (pipe_ok) return 0;
        for (int i = 0; i < NBS; i++)
            if (hipStreamCreateWithFlags(&bs[i], hipStreamNonBlocking) != hipSuccess) return -1;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return -1;
        if (hipEventCreateWithFlags(&ev_go, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_in, hipEventDisableTiming) != hipSuccess ||
            hipEventCreate(&ev_t0) != hipSuccess || hipEventCreate(&ev_t1) != hipSuccess)
            return -1;
        pipe_ok = 1;
        return 0;
    }
    static int grow_events(std::vector<hipEvent_t> &v, size_t n)
    {
        while (v.size() < n) {
            hipEvent_t e;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return -1;
            v.push_back(e);
        }
        return 0;
    }
    int ensure_host(size_t n)
    {
        if (n <= h_est_cap) return 0;
        if (h_est) hipHostFree(h_est);
        h_est = nullptr;
        h_est_cap = 0;
        if (hipHostMalloc((void **)&h_est, sizeof(double) * n, 0) != hipSuccess) return -1;
        h_est_cap = n;
        return 0;
    }
    ~DasScratch()
    {
        for (hipEvent_t e : bev) hipEventDestroy(e);
        for (hipEvent_t e : blk) hipEventDestroy(e);
        for (hipStream_t x : bs)
            if (x) hipStreamDestroy(x);
        if (cs) {
            hipStreamSynchronize(cs);     /* no staging copy may outlive h_est */
            hipStreamDestroy(cs);
        }
        if (ev_in) hipEventDestroy(ev_in);
        if (ev_go) hipEventDestroy(ev_go);
        if (ev_t0) hipEventDestroy(ev_t0);
        if (ev_t1) hipEventDestroy(ev_t1);
        if (h_est) hipHostFree(h_est);
    }
};

struct dsdhip_das {
    double a[10], g[10];
    int order, depth, K, L, nch;
    int ntracks, ns;             /* tracks of nch channels; ns = ntracks * nch streams */
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
    int est_device;              /* seg0-first estimate on one device lane per channel (timing) */
    int batch_launch;            /* seg0-first: batches of ready segments instead of the host-fed launch */
    double feed_limit_s;         /* host-fed launch: longest wait of one wave for its entry */
    int feed_waves_cu;           /* host-fed launch: resident waves per CU (0: the kernel's own) */
    /* last run */
    int last_S;
    std::vector<int> last_stitch;
    int64_t last_slow;
    double last_est_ms;          /* estimator threads: longest busy time */
    int64_t last_est_samples;    /* samples estimated (all channels) */
    int64_t last_est_redo;       /* estimator blocks recomputed in the reference form */
    int last_batches;            /* pipelined trellis launches */
    /* host-side timeline of the last seg0-first run (ms from its start): seg0
     * done, estimate started, last segment published, estimate threads
     * joined, trellis results read, run done */
    std::chrono::steady_clock::time_point t_run0;
    double last_tl[6];
    DasScratch sc;
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

/* context indices in the pool; stream s = track * nch + channel */
static inline int ix_persist(const dsdhip_das_t *, int s) { return s; }
static inline int ix_snap(const dsdhip_das_t *d, int s) { return d->ns + s; }
static inline int ix_seg(const dsdhip_das_t *d, int s, int seg) { return 2 * d->ns + s * d->cap_seg + seg; }


extern "C" {

static dsdhip_das_t *das_create(const double *a, const double *g, int order, int depth, int cands,
                                int lat, double state_limit, int nch, int ntracks, size_t seg_len,
                                int nseg, int seg0_first, int reposition)
{
    if (!a || !g || order < 4 || order > TR_MAXP) return NULL;
    if (nch < 1 || ntracks < 1 || lat < 1 || (seg_len == 0 && nseg < 1)) return NULL;
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
    d->ntracks = ntracks;
    d->ns = nch * ntracks;
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
    d->est_device = 0;
    d->batch_launch = 0;
    d->feed_limit_s = 20.0;
    d->feed_waves_cu = 6;         /* 30 s stereo to DSD256, 65,536-sample segments: time from the last
                                     estimate to the end, 4 waves per CU 203 ms, 5: 128, 6: 138, 8: 145 */
    d->last_est_ms = 0.0;
    d->last_est_samples = 0;
    d->last_est_redo = 0;
    d->last_batches = 0;
    dsdhip_make_qfir(d->qfir);
    if (dsdhip_ensure_device() != 0 ||
        hipMalloc(&d->d_qfir, sizeof(d->qfir)) != hipSuccess ||
        hipMemcpy(d->d_qfir, d->qfir, sizeof(d->qfir), hipMemcpyHostToDevice) != hipSuccess) {
        delete d;
        return NULL;
    }
    /* persistent contexts exist from the start (fresh) */
    d->tr = dsdhip_trellis_create_coeffs(d->a, d->g, order, depth, cands, lat, state_limit, 2 * d->ns);
    if (!d->tr) {
        hipFree(d->d_qfir);
        delete d;
        return NULL;
    }
    return d;
}

dsdhip_das_t *dsdhip_das_create_coeffs(const double *a, const double *g, int order, int depth,
                                       int cands, int lat, double state_limit, int nch,
                                       size_t seg_len, int nseg, int seg0_first, int reposition)
{
    return das_create(a, g, order, depth, cands, lat, state_limit, nch, 1, seg_len, nseg,
                      seg0_first, reposition);
}

dsdhip_das_t *dsdhip_das_create(int ntf_index, int depth, int cands, int lat, double state_limit,
                                int nch, size_t seg_len, int nseg, int seg0_first, int reposition)
{
    const dsdhip_ntf_t *f = dsdhip_ntf_get(ntf_index);
    if (!f) return NULL;
    return das_create(f->a, f->g, f->order, depth, cands, lat, state_limit, nch, 1, seg_len, nseg,
                      seg0_first, reposition);
}

dsdhip_das_t *dsdhip_das_create_tracks(int ntf_index, int depth, int cands, int lat,
                                       double state_limit, int nch, int ntracks, size_t seg_len,
                                       int nseg, int seg0_first, int reposition)
{
    const dsdhip_ntf_t *f = dsdhip_ntf_get(ntf_index);
    if (!f) return NULL;
    return das_create(f->a, f->g, f->order, depth, cands, lat, state_limit, nch, ntracks, seg_len,
                      nseg, seg0_first, reposition);
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
int dsdhip_das_fma_o8(dsdhip_das_t *d, int on) { return d ? dsdhip_trellis_fma_o8(d->tr, on) : -1; }
int dsdhip_das_wave_only(dsdhip_das_t *d, int on) { return d ? dsdhip_trellis_wave_only(d->tr, on) : -1; }

/* sdm_estimate_state_multibit (src/trellis.c:1118-1187): the greedy
 * num_levels quantizer run through the loop filter from init_state over
 * count host samples, on the calling thread (dsdhip_estimate.cpp). */
int dsdhip_estimate_state(const double *a, const double *g, int order, const double *init_state,
                          const double *in, size_t count, double state_limit, int num_levels,
                          double *out_state)
{
    if (!a || !g || !init_state || !out_state || (!in && count) || order < 4 || order > TR_MAXP ||
        num_levels < 2)
        return -1;
    double init[10] = {0}, est[10];
    memcpy(init, init_state, sizeof(double) * (size_t)order);
    const int64_t bnd = (int64_t)count;
    if (dsdhip_estimate_boundaries(a, g, order, init, in, count, &bnd, 1, state_limit, num_levels, est) != 0)
        return -1;
    memcpy(out_state, est, sizeof(double) * (size_t)order);
    return 0;
}

/* The estimation block on one device lane per channel (das_estimate), for
 * timing against the host form.  d_x: nch rows of stride doubles. */
int dsdhip_estimate_boundaries_dev(const double *a, const double *g, int order, const double *init10,
                                   const double *d_x, int64_t stride, int nch, size_t count,
                                   const int64_t *bnd, int nb, double state_limit, int levels,
                                   double *out10)
{
    if (!a || !g || !init10 || !d_x || nch < 1 || nb < 1 || !bnd || !out10 || order < 4 ||
        order > TR_MAXP || levels < 2)
        return -1;
    if (dsdhip_ensure_device() != 0) return -1;
    double ab[2 * TR_MAXP] = {0};
    memcpy(ab, a, sizeof(double) * (size_t)order);
    memcpy(ab + TR_MAXP, g, sizeof(double) * (size_t)order);
    DevBuf<double> dinit, dab, dest;
    DevBuf<int64_t> db;
    if (dinit.ensure((size_t)nch * 10) || dab.ensure(2 * TR_MAXP) || dest.ensure((size_t)nch * nb * 10) ||
        db.ensure((size_t)nb) ||
        hipMemcpy(dinit.p, init10, sizeof(double) * (size_t)nch * 10, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dab.p, ab, sizeof(ab), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db.p, bnd, sizeof(int64_t) * (size_t)nb, hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(das_estimate, dim3((unsigned)((nch + 63) / 64)), dim3(64), 0, 0, d_x, stride,
                       (int64_t)0, (int64_t)count, db.p, nb, dinit.p, dab.p, dab.p + TR_MAXP, order,
                       state_limit, levels, nch, dest.p);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out10, dest.p, sizeof(double) * (size_t)nch * nb * 10, hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    return 0;
}

/* Plan for F samples: returns the segment count; fills start/size if given.
 * With a segment length the count is floor(F / seg_len) (at least 1), else
 * the requested count reduced while segments would be shorter than 4
 * overlaps (src/dsp_plugin.c:1362-1366); either way the sizes follow the
 * reference's rule for that count, base + (s < rem) with the last taking
 * the rest (:1944-1953), so no segment is longer than the others by more
 * than one sample. */
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
        const size_t base = F / (size_t)S, rem = F % (size_t)S;
        for (int s = 0; s < S; s++) {
            start[s] = pos;
            size[s] = (s == S - 1) ? F - pos : base + ((size_t)s < rem ? 1 : 0);
            pos += size[s];
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
                                                        d->state_limit, 2 * d->ns + d->ns * cap);
    if (!nt) return -1;
    dsdhip_trellis_set_stream(nt, d->stream);
    {   /* the pool's mode switches carry over */
        tr_params op;
        dsdhip_trellis_params(d->tr, &op);
        dsdhip_trellis_force_literal(nt, op.force_slow);
        dsdhip_trellis_fma_o8(nt, op.fma_o8);
        dsdhip_trellis_wave_only(nt, op.wave_only);
    }
    for (int s = 0; s < d->ns; s++) {
        tr_ctx c;
        if (dsdhip_trellis_get(d->tr, ix_persist(d, s), &c) != 0) return -1;
        if (dsdhip_trellis_put(nt, s, &c) != 0) return -1;
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
    for (int s = 0; s < d->ns; s++)
        if (dsdhip_trellis_reset(d->tr, ix_persist(d, s)) != 0) return -1;
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
    DevBuf<int64_t> &dc = d->sc.dc;
    DevBuf<double> &ds = d->sc.ds;
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
/* One track of a run: its plan, segment geometry and results. */
struct TrackRun {
    const double *in = nullptr;
    int64_t in_stride = 0;
    size_t F = 0;
    unsigned long long *out = nullptr;
    int64_t out_words = 0;
    int S = 0;
    std::vector<size_t> st, sz;
    std::vector<int64_t> in_start, in_cnt, keep;   /* per segment */
    std::vector<int64_t> woff;                     /* per (ch, seg): word offset in segbits */
    std::vector<int64_t> kept, slow;               /* per (ch, seg) */
    std::vector<int> bp;                           /* stitch position per segment */
    int64_t total = 0;
    const double *hin = nullptr;                   /* host copy of the input rows, or null */
};

/* One channel's estimator chain: from seg1's start over the rest of the
 * chunk, the state at the start of every segment >= 2. */
struct EstChain {
    int t = 0, ch = 0;
    const double *dx = nullptr;    /* device input at seg1's start */
    const double *hx = nullptr;    /* host input at seg1's start (caller's or staged) */
    const hipEvent_t *blk = nullptr;   /* staged: event of each EST_BLOCK block, or null */
    size_t count = 0;
    std::vector<int64_t> bnd;      /* offsets of segments 2.. from seg1's start */
    double init[10] = {};
    std::vector<double> est;       /* [nb][10] */
    std::atomic<int> ready{0};     /* boundaries whose state is in est */
    double busy_ms = 0.0;
    int64_t redo = 0;
};

/* the staged blocks of chain c covering [pos, end) have arrived */
static int est_wait(void *cv, size_t pos, size_t end)
{
    const EstChain *c = (const EstChain *)cv;
    if (!c->blk || end <= pos) return 0;
    for (size_t b = pos / EST_BLOCK; b <= (end - 1) / EST_BLOCK; b++)
        if (hipEventSynchronize(c->blk[b]) != hipSuccess) return -1;
    return 0;
}

static void est_publish(void *cv, int b, const double *st)
{
    EstChain *c = (EstChain *)cv;
    for (int i = 0; i < 10; i++) c->est[(size_t)b * 10 + i] = st[i];
    c->ready.store(b + 1, std::memory_order_release);
}

static void est_worker(const dsdhip_das_t *d, std::vector<EstChain *> chains, std::atomic<int> *fail,
                       int threads)
{
    for (EstChain *c : chains) {
        const int nb = (int)c->bnd.size();
        const auto t0 = std::chrono::steady_clock::now();
        dsdhip_est_t *e = dsdhip_est_create(d->a, d->g, d->order, d->state_limit, 16, c->init);
        bool ok = e != nullptr &&
                  dsdhip_est_chain(e, c->hx, c->count, c->bnd.data(), nb, est_publish, est_wait, c, threads) == 0;
        if (!ok) {
            for (int b = c->ready.load(); b < nb; b++) c->ready.store(b + 1, std::memory_order_release);
            fail->store(1);
        }
        if (e) {
            c->redo = dsdhip_est_redo_blocks(e);
            dsdhip_est_destroy(e);
        }
        c->busy_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

/*
 * seg0-first seeding after seg0 has run (src/dsp_plugin.c:2185-2266, 2288-
 * 2363): seg1 continues from the persistent context at once; the estimate
 * from seg1's start runs on host threads (one chain per channel, reading
 * the input staged to pinned host memory block by block), and every
 * segment >= 2 starts as soon as the estimate has passed its start: one
 * host-fed persistent launch whose waves each take the next table entry
 * and wait until this thread has published it (seg1 first, then the
 * segments in the order their estimates arrive).  Where that launch does
 * not apply (the one-lane kernel at K <= 2, the FMA filter), or when
 * batch_launch is set, the entries are launched in batches instead: every
 * entry ready when the previous batch has finished (at least a minimum
 * size), on one stream, and the last batch at once when the estimate ends,
 * on a second stream.
 * Fills kept/slow of every segment >= 1.
 */
static double ms_since(const dsdhip_das_t *d)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - d->t_run0).count();
}

/* Queues the copies of the estimate input (each channel from seg1's start
 * to the end, for tracks of more than two segments) into pinned host memory
 * on the copy stream, behind ev_in only: block i of every chain before
 * block i + 1 of any, each followed by its event.  Called once segment 0's
 * launch is queued, so the copies run beside segment 0 and the estimator
 * chains, which start when segment 0 ends, find their input across. */
static int estimate_stage(dsdhip_das_t *d, const std::vector<TrackRun> &R)
{
    DasScratch &sc = d->sc;
    const int nch = d->nch, T = d->ntracks;
    sc.staged = 0;
    sc.st_base.clear();
    sc.st_ev.clear();
    std::vector<const double *> dx;
    std::vector<size_t> cnt;
    size_t staged = 0, nblk = 0;
    for (int t = 0; t < T; t++) {
        const TrackRun &r = R[(size_t)t];
        if (r.S <= 2) continue;
        const int64_t es = (int64_t)r.st[1];
        for (int ch = 0; ch < nch; ch++) {
            if (r.hin) {
                sc.st_base.push_back((size_t)-1);
                sc.st_ev.push_back(0);
                dx.push_back(nullptr);
                cnt.push_back(0);
                continue;
            }
            const size_t n = r.F - (size_t)es;
            sc.st_base.push_back(staged);
            sc.st_ev.push_back(nblk);
            dx.push_back(r.in + (int64_t)ch * r.in_stride + es);
            cnt.push_back(n);
            staged += n;
            nblk += (n + EST_BLOCK - 1) / EST_BLOCK;
        }
    }
    if (staged == 0) return 0;
    /* copies of an earlier call that failed may still be in flight into h_est */
    if (hipStreamSynchronize(sc.cs) != hipSuccess || sc.ensure_host(staged) != 0 ||
        DasScratch::grow_events(sc.blk, nblk) != 0 || hipStreamWaitEvent(sc.cs, sc.ev_in, 0) != hipSuccess)
        return -1;
    for (size_t b = 0;; b++) {
        bool any = false;
        for (size_t i = 0; i < dx.size(); i++) {
            if (!dx[i] || b * EST_BLOCK >= cnt[i]) continue;
            any = true;
            const size_t n = std::min(EST_BLOCK, cnt[i] - b * EST_BLOCK);
            if (hipMemcpyAsync(sc.h_est + sc.st_base[i] + b * EST_BLOCK, dx[i] + b * EST_BLOCK,
                               sizeof(double) * n, hipMemcpyDeviceToHost, sc.cs) != hipSuccess ||
                hipEventRecord(sc.blk[sc.st_ev[i] + b], sc.cs) != hipSuccess)
                return -1;
        }
        if (!any) break;
    }
    sc.staged = 1;
    return 0;
}

static int seg0_first_rest(dsdhip_das_t *d, std::vector<TrackRun> &R, unsigned long long *segbits)
{
    const int nch = d->nch, T = d->ntracks;
    dsdhip_trellis_t *tr = d->tr;
    DasScratch &sc = d->sc;
    if (sc.init_pipe() != 0) return -1;
    struct Ent { int t, ch, seg; };
    std::vector<Ent> ent;
    size_t nchains = 0;
    for (int t = 0; t < T; t++)
        if (R[(size_t)t].S > 2) nchains += (size_t)nch;
    std::vector<EstChain> chains(nchains);
    int total = 0;
    size_t ci = 0, staged = 0, nblk = 0;
    for (int t = 0; t < T; t++) {
        TrackRun &r = R[(size_t)t];
        if (r.S >= 2) total += nch * (r.S - 1);
        if (r.S <= 2) continue;
        const int64_t es = (int64_t)r.st[1];
        for (int ch = 0; ch < nch; ch++, ci++) {
            EstChain &c = chains[ci];
            c.t = t;
            c.ch = ch;
            c.dx = r.in + (int64_t)ch * r.in_stride + es;
            c.hx = r.hin ? r.hin + (size_t)ch * r.F + (size_t)es : nullptr;
            c.count = r.F - (size_t)es;
            if (!c.hx) {
                staged += c.count;
                nblk += (c.count + EST_BLOCK - 1) / EST_BLOCK;
            }
            for (int seg = 2; seg < r.S; seg++) c.bnd.push_back((int64_t)r.st[(size_t)seg] - es);
            c.est.assign(c.bnd.size() * 10, 0.0);
            if (dsdhip_trellis_best_state(tr, ix_persist(d, t * nch + ch), c.init) != 0) return -1;
        }
    }
    /* the staging copies, queued beside segment 0 (estimate_stage) */
    const bool prestaged = sc.staged && sc.st_base.size() == nchains;
    sc.staged = 0;
    d->last_batches = 0;
    d->last_est_ms = 0.0;
    d->last_est_samples = 0;
    d->last_est_redo = 0;
    d->last_tl[0] = ms_since(d);      /* seg0 done: its best states were read */
    if (total == 0) return 0;
    if (dsdhip_trellis_table_begin(tr, total) != 0) return -1;
    if (!prestaged && staged &&
        (hipStreamSynchronize(sc.cs) != hipSuccess || sc.ensure_host(staged) != 0 ||
         DasScratch::grow_events(sc.blk, nblk) != 0))
        return -1;
    if (DasScratch::grow_events(sc.bev, (size_t)total + 1) != 0) return -1;
    ent.reserve((size_t)total);
    auto fill = [&](int t, int ch, int seg, const double *seed) -> int {
        TrackRun &r = R[(size_t)t];
        const int sidx = t * nch + ch;
        const int j = (int)ent.size();
        ent.push_back({t, ch, seg});
        return dsdhip_trellis_table_fill(tr, j, ix_seg(d, sidx, seg), ix_persist(d, sidx),
                                         r.in + (int64_t)ch * r.in_stride + r.in_start[(size_t)seg],
                                         r.in_cnt[(size_t)seg], r.keep[(size_t)seg],
                                         (uint64_t *)(segbits + r.woff[(size_t)(ch * r.S + seg)]), seed);
    };
    /* the trellis streams start after seg0 and all earlier work */
    const bool timing = d->timing != 0;
    if (hipEventRecord(sc.ev_go, d->stream) != hipSuccess) return -1;
    if (timing) hipEventRecord(sc.ev_t0, d->stream);
    for (int i = 0; i < DasScratch::NBS; i++)
        if (hipStreamWaitEvent(sc.bs[i], sc.ev_go, 0) != hipSuccess) return -1;
    /* stage the estimate input if segment 0's call did not: every copy is
     * queued now, block i of every chain before block i + 1 of any, each
     * followed by its event */
    int rc = 0;
    if (prestaged) {
        for (size_t i = 0; i < nchains; i++) {
            if (chains[i].hx || sc.st_base[i] == (size_t)-1) continue;
            chains[i].hx = sc.h_est + sc.st_base[i];
            chains[i].blk = &sc.blk[sc.st_ev[i]];
        }
    } else {
        size_t off = 0, ei = 0;
        std::vector<size_t> base(nchains, 0), first_ev(nchains, 0);
        for (size_t i = 0; i < nchains; i++) {
            if (chains[i].hx) continue;
            base[i] = off;
            off += chains[i].count;
            first_ev[i] = ei;
            ei += (chains[i].count + EST_BLOCK - 1) / EST_BLOCK;
            chains[i].hx = sc.h_est + base[i];
            chains[i].blk = &sc.blk[first_ev[i]];
        }
        for (size_t b = 0; rc == 0; b++) {
            bool any = false;
            for (size_t i = 0; i < nchains && rc == 0; i++) {
                EstChain &c = chains[i];
                if (!c.blk || b * EST_BLOCK >= c.count) continue;
                any = true;
                const size_t n = std::min(EST_BLOCK, c.count - b * EST_BLOCK);
                if (hipMemcpyAsync(sc.h_est + base[i] + b * EST_BLOCK, c.dx + b * EST_BLOCK, sizeof(double) * n,
                                   hipMemcpyDeviceToHost, d->stream) != hipSuccess ||
                    hipEventRecord(sc.blk[first_ev[i] + b], d->stream) != hipSuccess)
                    rc = -1;
            }
            if (!any) break;
        }
    }
    /* host-fed launch: queued after every staging copy (on the DAS stream
     * behind them, or on the copy stream before it), since its waves wait
     * for estimates that need those copies (a launch queued ahead of them,
     * on any stream sharing their hardware queue, would never see its
     * entries published) */
    int fed = 1;
    if (rc == 0 && !d->batch_launch) {
        fed = dsdhip_trellis_feed_begin(tr, d->stream, d->feed_limit_s, d->feed_waves_cu);
        if (fed < 0) rc = -1;
    }
    const bool feed = fed == 0;
    int launched = 0, nbatch = feed ? 1 : 0;
    const int NBS = DasScratch::NBS;
    int last[DasScratch::NBS];       /* last batch on each stream */
    for (int i = 0; i < NBS; i++) last[i] = -1;
    auto launch = [&](int si) -> int {
        const int n = (int)ent.size() - launched;
        if (n <= 0) return 0;
        int rc2 = dsdhip_trellis_table_launch(tr, launched, n, sc.bs[si]);
        if (rc2 == 0 && hipEventRecord(sc.bev[(size_t)nbatch], sc.bs[si]) != hipSuccess) rc2 = -1;
        last[si] = nbatch;
        launched += n;
        nbatch++;
        return rc2;
    };
    auto idle = [&](int si) -> bool {
        return last[si] < 0 || hipEventQuery(sc.bev[(size_t)last[si]]) == hipSuccess;
    };
    /* batch form: batches before the last run on stream 0, one after the
     * other, each once it has a useful size (a batch holds its stream for
     * about one segment's latency); stream 1 is kept for the last batch,
     * which starts as soon as the estimate ends (streams that share a
     * hardware queue would serialise, so there are only two) */
    auto pick = [&](bool final_batch) -> int {
        if (final_batch) return idle(0) ? 0 : 1;
        return idle(0) ? 0 : -1;
    };
    auto publish = [&]() -> int {
        const int n = (int)ent.size();
        if (dsdhip_trellis_feed_publish(tr, n) != 0) return -1;
        launched = n;
        return 0;
    };
    const int min_batch = std::max(32, total / 32);
    /* seg1: exact continuation, no estimate needed; the host-fed launch
     * starts it at once, the batch form with the first estimated batch */
    for (int t = 0; t < T && rc == 0; t++)
        for (int ch = 0; ch < nch && R[(size_t)t].S >= 2 && rc == 0; ch++) rc = fill(t, ch, 1, NULL);
    if (feed && rc == 0) rc = publish();
    if (!feed && nchains == 0 && rc == 0) rc = launch(0);
    if (nchains > 0 && rc == 0) {
        /* estimator threads: chains dealt round-robin over at most 16
         * threads, one core left to this one */
        const unsigned hw = std::thread::hardware_concurrency();
        const int nw = (int)std::min<size_t>(nchains, std::max(1u, std::min(hw > 1 ? hw - 1 : 1, 16u)));
        std::vector<std::vector<EstChain *>> part((size_t)nw);
        for (size_t i = 0; i < nchains; i++) part[i % (size_t)nw].push_back(&chains[i]);
        std::atomic<int> fail{0};
        std::vector<std::thread> th;
        /* a chain's index guess gets a thread of its own where the cores allow */
        const int per = (2 * (unsigned)nw + 1 <= hw) ? 2 : 1;
        d->last_tl[1] = ms_since(d);
        for (int i = 0; i < nw; i++) th.emplace_back(est_worker, d, part[(size_t)i], &fail, per);
        std::vector<int> taken(nchains, 0);
        for (;;) {
            bool all = true;
            for (size_t i = 0; i < nchains && rc == 0; i++) {
                EstChain &c = chains[i];
                const int nb = (int)c.bnd.size();
                const int rd = c.ready.load(std::memory_order_acquire);
                for (int b = taken[i]; b < rd && rc == 0; b++)
                    rc = fill(c.t, c.ch, b + 2, &c.est[(size_t)b * 10]);
                taken[i] = rd;
                if (rd < nb) all = false;
            }
            if (rc != 0 || fail.load()) break;
            const int pending = (int)ent.size() - launched;
            if (feed) {
                if (pending > 0) rc = publish();
            } else {
                const int si = (pending > 0 && (all || pending >= min_batch)) ? pick(all) : -1;
                if (si >= 0) rc = launch(si);
            }
            if (rc != 0) break;
            if (all) break;
            std::this_thread::sleep_for(std::chrono::microseconds(feed ? 50 : 200));
        }
        d->last_tl[2] = ms_since(d);
        for (auto &x : th) x.join();
        d->last_tl[3] = ms_since(d);
        if (fail.load()) rc = -1;
        for (size_t i = 0; i < nchains; i++) {
            d->last_est_ms = std::max(d->last_est_ms, chains[i].busy_ms);
            d->last_est_samples += (int64_t)chains[i].count;
            d->last_est_redo += chains[i].redo;
        }
    }
    /* a host-fed launch must not wait for entries that will never come; it
     * runs on the DAS stream itself, so it needs no join */
    if (feed && (rc != 0 || launched != total)) dsdhip_trellis_feed_abort(tr);
    /* join the batch streams back into the DAS stream, then read results */
    for (int si = 0; si < NBS; si++)
        if (last[si] >= 0 && hipStreamWaitEvent(d->stream, sc.bev[(size_t)last[si]], 0) != hipSuccess) rc = -1;
    if (rc != 0) {
        for (int si = 0; si < NBS; si++) hipStreamSynchronize(sc.bs[si]);
        hipStreamSynchronize(d->stream);
        if (feed) dsdhip_trellis_table_end(tr, d->stream, nullptr, nullptr, nullptr);
        return -1;
    }
    if (launched != total) return -1;
    if (timing) hipEventRecord(sc.ev_t1, d->stream);
    std::vector<int64_t> kp((size_t)total), sl((size_t)total);
    int err = 0;
    if (dsdhip_trellis_table_end(tr, d->stream, kp.data(), sl.data(), &err) != 0) return -1;
    d->last_tl[4] = ms_since(d);
    if (timing) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, sc.ev_t0, sc.ev_t1) == hipSuccess) dsdhip_trellis_timing_add(tr, ms, nbatch);
    }
    for (int j = 0; j < total; j++) {
        const Ent &q = ent[(size_t)j];
        TrackRun &r = R[(size_t)q.t];
        r.kept[(size_t)(q.ch * r.S + q.seg)] = kp[(size_t)j];
        r.slow[(size_t)(q.ch * r.S + q.seg)] = sl[(size_t)j];
    }
    d->last_batches = nbatch;
    return 0;
}

/* DAS over every track (stream s = t * nch + ch) with all segments of all
 * tracks in one trellis launch (two with seg0-first seeding), all
 * boundaries in one stitch launch.  Tracks with F[t] == 0 are left alone. */
static int run_tracks(dsdhip_das_t *d, const double *const *d_in, const int64_t *in_stride,
                      const size_t *F, unsigned long long *const *d_out, const int64_t *out_words,
                      size_t *out_count, const double *const *h_in)
{
    const int nch = d->nch, T = d->ntracks;
    const size_t O = d->O;
    const int L = d->L;
    const int seg0_first = d->seg0_first;
    d->t_run0 = std::chrono::steady_clock::now();
    for (int i = 0; i < 6; i++) d->last_tl[i] = 0.0;
    std::vector<TrackRun> R((size_t)T);
    int Smax = 0;
    for (int t = 0; t < T; t++) {
        TrackRun &r = R[(size_t)t];
        if (F[t] == 0) continue;
        if (!d_in[t] || !d_out[t]) return -1;
        r.in = d_in[t];
        r.hin = h_in ? h_in[t] : nullptr;
        r.in_stride = in_stride[t];
        r.F = F[t];
        r.out = d_out[t];
        r.out_words = out_words[t];
        r.S = dsdhip_das_plan(d, r.F, NULL, NULL);
        r.st.resize((size_t)r.S);
        r.sz.resize((size_t)r.S);
        dsdhip_das_plan(d, r.F, r.st.data(), r.sz.data());
        if (d->reposition && reposition(d, r.in, r.F, r.st, r.sz) != 0) return -1;
        if (r.S > Smax) Smax = r.S;
    }
    if (Smax == 0) {
        for (int t = 0; t < T; t++) if (out_count) out_count[t] = 0;
        return 0;
    }
    if (ensure_pool(d, Smax) != 0) return -1;
    dsdhip_trellis_t *tr = d->tr;

    /* segment inputs / keep rules, packed output regions */
    int64_t wtot = 0;
    for (int t = 0; t < T; t++) {
        TrackRun &r = R[(size_t)t];
        const int S = r.S;
        if (S == 0) continue;
        r.in_start.assign((size_t)S, 0);
        r.in_cnt.assign((size_t)S, 0);
        r.keep.assign((size_t)S, 0);
        for (int seg = 0; seg < S; seg++) {
            const size_t ns = r.st[(size_t)seg], nz = r.sz[(size_t)seg];
            if (seg == 0) {
                size_t cnt = nz;
                if (!seg0_first && S > 1 && nz + O <= r.F) cnt += O;
                r.in_cnt[0] = (int64_t)cnt;
            } else if (seg == 1 && seg0_first) {
                size_t cnt = nz;
                if (S > 2 && ns + nz + O <= r.F) cnt += O;
                r.in_start[1] = (int64_t)ns;
                r.in_cnt[1] = (int64_t)cnt;
            } else {
                const bool ext = seg < S - 1;
                const size_t is = ns >= O ? ns - O : 0;
                size_t cnt = ns + nz - is;
                if (ext && ns + nz + O <= r.F) cnt += O;
                const size_t disc = ns >= O ? O - (size_t)L : 0;
                r.in_start[(size_t)seg] = (int64_t)is;
                r.in_cnt[(size_t)seg] = (int64_t)cnt;
                r.keep[(size_t)seg] = disc ? (int64_t)L + (int64_t)disc : 0;
            }
        }
        r.woff.assign((size_t)(nch * S), 0);
        for (int ch = 0; ch < nch; ch++)
            for (int seg = 0; seg < S; seg++) {
                r.woff[(size_t)(ch * S + seg)] = wtot;
                wtot += (r.in_cnt[(size_t)seg] + 63) / 64 + 1;
            }
        r.kept.assign((size_t)(nch * S), 0);
        r.slow.assign((size_t)(nch * S), 0);
    }
    DevBuf<unsigned long long> &segbits = d->sc.segbits;
    if (segbits.ensure((size_t)wtot)) return -1;

    struct JobRef { int t, ch, seg, src; };
    std::vector<JobRef> jl;
    std::vector<double> seeds;
    auto add = [&](int t, int ch, int seg, int src, const double *seed) {
        jl.push_back({t, ch, seg, src});
        const double nan = nan_val();
        for (int q = 0; q < 10; q++) seeds.push_back(seed ? seed[q] : nan);
    };
    auto run = [&](int (*queued)(void *) = nullptr, void *qarg = nullptr) -> int {
        const int nj = (int)jl.size();
        if (nj == 0) return 0;
        std::vector<int> ctx((size_t)nj), cs((size_t)nj);
        std::vector<const double *> pin((size_t)nj);
        std::vector<uint64_t *> pout((size_t)nj);
        std::vector<int64_t> ni((size_t)nj), kf((size_t)nj), kp((size_t)nj), sl((size_t)nj);
        for (int j = 0; j < nj; j++) {
            const JobRef &q = jl[(size_t)j];
            TrackRun &r = R[(size_t)q.t];
            const int sidx = q.t * nch + q.ch;
            ctx[(size_t)j] = (q.seg == 0) ? ix_persist(d, sidx) : ix_seg(d, sidx, q.seg);
            cs[(size_t)j] = q.src;
            pin[(size_t)j] = r.in + (int64_t)q.ch * r.in_stride + r.in_start[(size_t)q.seg];
            pout[(size_t)j] = (uint64_t *)(segbits.p + r.woff[(size_t)(q.ch * r.S + q.seg)]);
            ni[(size_t)j] = r.in_cnt[(size_t)q.seg];
            kf[(size_t)j] = r.keep[(size_t)q.seg];
        }
        int err = 0;
        const int rc = dsdhip_trellis_run_jobs_then(tr, nj, ctx.data(), cs.data(), pin.data(), ni.data(),
                                                    kf.data(), pout.data(), seeds.data(), kp.data(),
                                                    sl.data(), &err, queued, qarg);
        for (int j = 0; j < nj; j++) {
            const JobRef &q = jl[(size_t)j];
            TrackRun &r = R[(size_t)q.t];
            r.kept[(size_t)(q.ch * r.S + q.seg)] = kp[(size_t)j];
            r.slow[(size_t)(q.ch * r.S + q.seg)] = sl[(size_t)j];
        }
        jl.clear();
        seeds.clear();
        return rc;
    };

    if (seg0_first) {
        for (int t = 0; t < T; t++)
            for (int ch = 0; ch < nch && R[(size_t)t].S > 0; ch++) add(t, ch, 0, -1, NULL);
        d->sc.staged = 0;
        if (!d->est_device) {
            /* the estimate input's copies run beside segment 0 */
            struct StageArg { dsdhip_das_t *d; const std::vector<TrackRun> *R; } sa{d, &R};
            if (d->sc.init_pipe() != 0 || hipEventRecord(d->sc.ev_in, d->stream) != hipSuccess) return -1;
            if (run([](void *p) -> int {
                    const StageArg *a = (const StageArg *)p;
                    return estimate_stage(a->d, *a->R);
                }, &sa) != 0)
                return -1;
        } else if (run() != 0) {
            return -1;
        }
        if (!d->est_device) {
            if (seg0_first_rest(d, R, segbits.p) != 0) return -1;
        } else {
        /* the same estimate on one device lane per channel (for timing) */
        std::vector<std::vector<double>> est((size_t)T);
        tr_params prm;
        dsdhip_trellis_params(tr, &prm);
        for (int t = 0; t < T; t++) {
            TrackRun &r = R[(size_t)t];
            const int S = r.S;
            if (S <= 2) continue;
            /* estimate from seg1's start over the rest (src/threadpool.c:345-366) */
            std::vector<double> init((size_t)nch * 10);
            for (int ch = 0; ch < nch; ch++)
                if (dsdhip_trellis_best_state(tr, ix_persist(d, t * nch + ch), &init[(size_t)ch * 10]) != 0)
                    return -1;
            const int64_t es = (int64_t)r.st[1];
            std::vector<int64_t> bnd;
            for (int seg = 1; seg < S; seg++) bnd.push_back(seg >= 2 ? (int64_t)r.st[(size_t)seg] - es : 0);
            DevBuf<int64_t> &db = d->sc.db;
            DevBuf<double> &dinit = d->sc.dinit, &dest = d->sc.dest, &dab = d->sc.dab;
            if (db.ensure(bnd.size()) || dinit.ensure(init.size()) ||
                dest.ensure((size_t)nch * bnd.size() * 10) || dab.ensure(2 * TR_MAXP))
                return -1;
            double ab[2 * TR_MAXP];
            memcpy(ab, prm.a, sizeof(prm.a));
            memcpy(ab + TR_MAXP, prm.g, sizeof(prm.g));
            bool ok =
                hipMemcpyAsync(db.p, bnd.data(), sizeof(int64_t) * bnd.size(), hipMemcpyHostToDevice,
                               d->stream) == hipSuccess &&
                hipMemcpyAsync(dinit.p, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice,
                               d->stream) == hipSuccess &&
                hipMemcpyAsync(dab.p, ab, sizeof(ab), hipMemcpyHostToDevice, d->stream) == hipSuccess;
            if (!ok) return -1;
            hipLaunchKernelGGL(das_estimate, dim3((unsigned)((nch + 63) / 64)), dim3(64), 0, d->stream, r.in,
                               r.in_stride, es, (int64_t)r.F - es, db.p, (int)bnd.size(), dinit.p,
                               dab.p, dab.p + TR_MAXP, prm.P, prm.state_limit, 16, nch, dest.p);
            if (hipGetLastError() != hipSuccess) return -1;
            est[(size_t)t].resize((size_t)nch * bnd.size() * 10);
            if (hipMemcpyAsync(est[(size_t)t].data(), dest.p, sizeof(double) * est[(size_t)t].size(),
                               hipMemcpyDeviceToHost, d->stream) != hipSuccess ||
                hipStreamSynchronize(d->stream) != hipSuccess)
                return -1;
        }
        for (int t = 0; t < T; t++) {
            TrackRun &r = R[(size_t)t];
            const int S = r.S;
            const int nb = S - 1;
            for (int ch = 0; ch < nch && S > 0; ch++) {
                const int sidx = t * nch + ch;
                if (S > 1) add(t, ch, 1, ix_persist(d, sidx), NULL);
                for (int seg = 2; seg < S; seg++)
                    add(t, ch, seg, ix_persist(d, sidx),
                        &est[(size_t)t][((size_t)ch * nb + (size_t)(seg - 1)) * 10]);
            }
        }
        if (run() != 0) return -1;
        }
    } else {
        std::vector<double> pst((size_t)d->ns * 10);
        for (int t = 0; t < T; t++)
            for (int ch = 0; ch < nch && R[(size_t)t].S > 0; ch++) {
                const int sidx = t * nch + ch;
                if (dsdhip_trellis_best_state(tr, ix_persist(d, sidx), &pst[(size_t)sidx * 10]) != 0) return -1;
                if (dsdhip_trellis_clone(tr, ix_snap(d, sidx), ix_persist(d, sidx)) != 0) return -1;
            }
        for (int t = 0; t < T; t++)
            for (int ch = 0; ch < nch && R[(size_t)t].S > 0; ch++) {
                const int sidx = t * nch + ch;
                add(t, ch, 0, -1, NULL);
                for (int seg = 1; seg < R[(size_t)t].S; seg++)
                    add(t, ch, seg, ix_snap(d, sidx), &pst[(size_t)sidx * 10]);
            }
        if (run() != 0) return -1;
    }
    int64_t tslow = 0;
    for (int t = 0; t < T; t++)
        for (size_t i = 0; i < R[(size_t)t].slow.size(); i++) tslow += R[(size_t)t].slow[i];
    d->last_slow = tslow;

    /* ---- stitch on channel 0 of every track, all boundaries in one launch ---- */
    const int first = seg0_first ? 2 : 1;
    std::vector<int64_t> pbit, cbit;
    std::vector<int> ovl, bt, bk;           /* boundary -> (track, segment) */
    for (int t = 0; t < T; t++) {
        TrackRun &r = R[(size_t)t];
        const int S = r.S;
        r.bp.assign((size_t)(S > 0 ? S : 1), 0);
        if (S == 0) continue;
        int64_t s0 = r.kept[0];
        if (!seg0_first && s0 > (int64_t)r.sz[0]) s0 = (int64_t)r.sz[0];
        for (int k = first; k < S; k++) {
            const int64_t prev_end = (k - 1 == 0) ? s0 : r.kept[(size_t)k - 1];
            if (prev_end < 2 * (int64_t)O || r.kept[(size_t)k] == 0) return -1;   /* plan too fine for span-local stitching */
            pbit.push_back(r.woff[(size_t)(k - 1)] * 64 + prev_end - (int64_t)O);
            cbit.push_back(r.woff[(size_t)k] * 64);
            ovl.push_back((int)((int64_t)O < r.kept[(size_t)k] ? (int64_t)O : r.kept[(size_t)k]));
            bt.push_back(t);
            bk.push_back(k);
        }
    }
    const int nbd = (int)pbit.size();
    if (nbd > 0) {
        DevBuf<int64_t> &dpb = d->sc.dpb, &dcb = d->sc.dcb;
        DevBuf<int> &dov = d->sc.dov, &dnc = d->sc.dnc, &dpk = d->sc.dpk, &dpos = d->sc.dpos,
                    &dden = d->sc.dden;
        DevBuf<double> &dmx = d->sc.dmx;
        if (dpb.ensure((size_t)nbd) || dcb.ensure((size_t)nbd) || dov.ensure((size_t)nbd) ||
            dnc.ensure((size_t)nbd) || dpk.ensure((size_t)nbd) || dpos.ensure((size_t)nbd * QN) ||
            dden.ensure((size_t)nbd * QN) || dmx.ensure((size_t)nbd * QN))
            return -1;
        bool ok =
            hipMemcpyAsync(dpb.p, pbit.data(), sizeof(int64_t) * (size_t)nbd, hipMemcpyHostToDevice,
                           d->stream) == hipSuccess &&
            hipMemcpyAsync(dcb.p, cbit.data(), sizeof(int64_t) * (size_t)nbd, hipMemcpyHostToDevice,
                           d->stream) == hipSuccess &&
            hipMemcpyAsync(dov.p, ovl.data(), sizeof(int) * (size_t)nbd, hipMemcpyHostToDevice,
                           d->stream) == hipSuccess;
        if (!ok) return -1;
        hipLaunchKernelGGL(das_stitch, dim3((unsigned)nbd), dim3(256), 0, d->stream,
                           segbits.p, dpb.p, dcb.p, dov.p, L, d->d_qfir, dnc.p, dpk.p, dpos.p,
                           dden.p, dmx.p);
        if (hipGetLastError() != hipSuccess) return -1;
        std::vector<int> nc((size_t)nbd), pk((size_t)nbd), pos((size_t)nbd * QN), den((size_t)nbd * QN);
        std::vector<double> mx((size_t)nbd * QN);
        ok = hipMemcpyAsync(nc.data(), dnc.p, sizeof(int) * nc.size(), hipMemcpyDeviceToHost,
                            d->stream) == hipSuccess &&
             hipMemcpyAsync(pk.data(), dpk.p, sizeof(int) * pk.size(), hipMemcpyDeviceToHost,
                            d->stream) == hipSuccess &&
             hipMemcpyAsync(pos.data(), dpos.p, sizeof(int) * pos.size(), hipMemcpyDeviceToHost,
                            d->stream) == hipSuccess &&
             hipMemcpyAsync(den.data(), dden.p, sizeof(int) * den.size(), hipMemcpyDeviceToHost,
                            d->stream) == hipSuccess &&
             hipMemcpyAsync(mx.data(), dmx.p, sizeof(double) * mx.size(), hipMemcpyDeviceToHost,
                            d->stream) == hipSuccess;
        if (hipStreamSynchronize(d->stream) != hipSuccess || !ok) return -1;
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
            R[(size_t)bt[(size_t)i]].bp[(size_t)bk[(size_t)i]] = pos[(size_t)i * QN + best];
        }
    }

    /* ---- assemble each track (same positions on every channel) ---- */
    if (d->sc.dsb.size() < (size_t)T) {
        d->sc.dsb.resize((size_t)T);
        d->sc.ddb.resize((size_t)T);
        d->sc.dlb.resize((size_t)T);
    }
    std::vector<DevBuf<int64_t>> &dsb = d->sc.dsb, &ddb = d->sc.ddb, &dlb = d->sc.dlb;
    for (int t = 0; t < T; t++) {
        TrackRun &r = R[(size_t)t];
        const int S = r.S;
        if (S == 0) continue;
        std::vector<int64_t> sbit, dbit, len;
        int64_t total = 0;
        for (int ch = 0; ch < nch; ch++) {
            std::vector<int64_t> at((size_t)S), src((size_t)S), cc((size_t)S);
            int64_t c0 = r.kept[(size_t)(ch * S)];
            if (!seg0_first && c0 > (int64_t)r.sz[0]) c0 = (int64_t)r.sz[0];
            at[0] = 0;
            src[0] = r.woff[(size_t)(ch * S)] * 64;
            cc[0] = c0;
            int64_t wp = c0;
            for (int k = 1; k < S; k++) {
                const int64_t kk = r.kept[(size_t)(ch * S + k)];
                if (kk == 0) { at[(size_t)k] = wp; src[(size_t)k] = 0; cc[(size_t)k] = 0; continue; }
                int64_t a, b;
                if (k == 1 && seg0_first) { a = wp; b = 0; }
                else {
                    const int64_t pst = wp >= (int64_t)O ? wp - (int64_t)O : 0;
                    b = r.bp[(size_t)k];
                    a = pst + b;
                }
                at[(size_t)k] = a;
                src[(size_t)k] = r.woff[(size_t)(ch * S + k)] * 64 + b;
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
        if ((total + 63) / 64 > r.out_words) return -1;
        r.total = total;
        if (dsb[(size_t)t].ensure(sbit.size()) || ddb[(size_t)t].ensure(dbit.size()) ||
            dlb[(size_t)t].ensure(len.size()))
            return -1;
        hipMemcpyAsync(dsb[(size_t)t].p, sbit.data(), sizeof(int64_t) * sbit.size(), hipMemcpyHostToDevice, d->stream);
        hipMemcpyAsync(ddb[(size_t)t].p, dbit.data(), sizeof(int64_t) * dbit.size(), hipMemcpyHostToDevice, d->stream);
        hipMemcpyAsync(dlb[(size_t)t].p, len.data(), sizeof(int64_t) * len.size(), hipMemcpyHostToDevice, d->stream);
        const int64_t words = r.out_words;
        hipLaunchKernelGGL(das_assemble, dim3((unsigned)((words + 255) / 256), (unsigned)nch), dim3(256), 0,
                           d->stream, segbits.p, dsb[(size_t)t].p, ddb[(size_t)t].p, dlb[(size_t)t].p, S,
                           r.out, words, total);
    }
    if (hipStreamSynchronize(d->stream) != hipSuccess) return -1;
    /* ---- copy-back ---- */
    for (int t = 0; t < T; t++) {
        const int S = R[(size_t)t].S;
        if (S > 1)
            for (int ch = 0; ch < nch; ch++) {
                const int sidx = t * nch + ch;
                if (dsdhip_trellis_copy_detached(tr, ix_persist(d, sidx), ix_seg(d, sidx, S - 1)) != 0)
                    return -1;
            }
        if (out_count) out_count[t] = (size_t)R[(size_t)t].total;
    }
    d->last_S = R[0].S;
    d->last_stitch = R[0].bp;
    d->last_tl[5] = ms_since(d);
    return 0;
}

int dsdhip_das_run_dev(dsdhip_das_t *d, const double *d_in, int64_t in_stride, size_t F,
                       unsigned long long *d_out, int64_t out_words, size_t *out_count)
{
    if (!d || !d_in || !d_out || F == 0 || d->ntracks != 1) return -1;
    return run_tracks(d, &d_in, &in_stride, &F, &d_out, &out_words, out_count, NULL);
}

/* Every track of a multi-track object: d_in[t] = [nch][in_stride[t]] doubles
 * with F[t] samples (0: the track is not run), d_out[t] = [nch][out_words[t]]
 * packed bits; out_count[t] receives the bits produced per channel. */
int dsdhip_das_run_tracks_dev(dsdhip_das_t *d, const double *const *d_in, const int64_t *in_stride,
                              const size_t *F, unsigned long long *const *d_out,
                              const int64_t *out_words, size_t *out_count)
{
    if (!d || !d_in || !in_stride || !F || !d_out || !out_words) return -1;
    return run_tracks(d, d_in, in_stride, F, d_out, out_words, out_count, NULL);
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

/* Host-buffer form: in nch x F doubles, out nch x F floats.  Row c gets
 * the *out_count produced samples (+-1.0) followed by 0.0f up to F (a value
 * no DSD sample takes), so no earlier call's output shows through. */
int dsdhip_das_run(dsdhip_das_t *d, const double *in, size_t F, float *out, size_t *out_count)
{
    if (!d) return -1;
    if (F == 0 || !in || !out) return -1;
    const int nch = d->nch;
    const int64_t words = (int64_t)(F + 63) / 64 + 1;
    DasScratch &sc = d->sc;
    size_t cnt = 0;
    if (sc.hin.ensure(F * (size_t)nch) || sc.hbits.ensure((size_t)words * (size_t)nch) ||
        sc.hout.ensure(F * (size_t)nch) ||
        hipMemcpy(sc.hin.p, in, sizeof(double) * F * (size_t)nch, hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    /* the estimator threads read the caller's host rows directly */
    const double *din = sc.hin.p, *hin = in;
    int64_t stride = (int64_t)F;
    if (run_tracks(d, &din, &stride, &F, &sc.hbits.p, &words, &cnt, &hin) != 0) return -1;
    if (cnt > F) return -1;
    if (cnt > 0) {
        hipLaunchKernelGGL(bits_to_f32, dim3((unsigned)((cnt + 255) / 256 + 1), (unsigned)nch), dim3(256), 0,
                           d->stream, sc.hbits.p, words, (int64_t)cnt, sc.hout.p, (int64_t)F);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(d->stream) != hipSuccess ||
            hipMemcpy2D(out, sizeof(float) * F, sc.hout.p, sizeof(float) * F, sizeof(float) * cnt, (size_t)nch,
                        hipMemcpyDeviceToHost) != hipSuccess)
            return -1;
    }
    for (int c = 0; c < nch; c++)
        for (size_t i = cnt; i < F; i++) out[(size_t)c * F + i] = 0.0f;
    if (out_count) *out_count = cnt;
    return 0;
}

/* Seeding estimate on one device lane per channel instead of host threads
 * (same states; for timing the two against each other). */
int dsdhip_das_estimator_device(dsdhip_das_t *d, int on)
{
    if (!d) return -1;
    d->est_device = on ? 1 : 0;
    return 0;
}

/* seg0-first launch form: 0 (default) one host-fed persistent launch whose
 * waves start each segment as soon as its estimate is published; 1 batches
 * of ready segments (the form used where the host-fed launch does not
 * apply). */
int dsdhip_das_batch_launch(dsdhip_das_t *d, int on)
{
    if (!d) return -1;
    d->batch_launch = on ? 1 : 0;
    return 0;
}

/* Longest wait of one wave of the host-fed launch for its entry, in
 * seconds (default 20); a wave that waits longer leaves and the run fails. */
int dsdhip_das_feed_limit(dsdhip_das_t *d, double seconds)
{
    if (!d || !(seconds >= 0.0) || seconds > 3600.0) return -1;
    d->feed_limit_s = seconds;
    return 0;
}

/* Host-side timeline of the last seg0-first run, in ms from the run's
 * start: seg0 done, estimate threads started, last segment published,
 * estimate threads joined, trellis results read, run done. */
int dsdhip_das_last_timeline(const dsdhip_das_t *d, double *ms6)
{
    if (!d || !ms6) return -1;
    for (int i = 0; i < 6; i++) ms6[i] = d->last_tl[i];
    return 0;
}

/* Resident waves per CU of the host-fed launch (default 6, one or two per
 * SIMD; 0: as many as the kernel's registers allow, 8 up to 128-sample
 * histories).  The segments arrive
 * at the estimator's pace, which a few waves per SIMD keep up with; fewer
 * resident waves leave the last segments fewer neighbours on their SIMD, so
 * the step ends sooner after the last estimate, until published segments
 * start to wait for a slot. */
int dsdhip_das_feed_waves(dsdhip_das_t *d, int waves_per_cu)
{
    if (!d || waves_per_cu < 0 || waves_per_cu > 64) return -1;
    d->feed_waves_cu = waves_per_cu;
    return 0;
}

/* Estimator figures of the last seg0-first run: longest chain busy time,
 * samples estimated over all channels, blocks recomputed in the reference
 * form, pipelined trellis launches. */
int dsdhip_das_last_estimate(const dsdhip_das_t *d, double *busy_ms, int64_t *samples, int64_t *redo_blocks,
                             int *batches)
{
    if (!d) return -1;
    if (busy_ms) *busy_ms = d->last_est_ms;
    if (samples) *samples = d->last_est_samples;
    if (redo_blocks) *redo_blocks = d->last_est_redo;
    if (batches) *batches = d->last_batches;
    return 0;
}

} /* extern "C" */
