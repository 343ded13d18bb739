/*
 * Host side of the trellis backend: device context pool, the reference's
 * sdm_* operations expressed on it, and batched job launch.
 *
 * Reference operations and where they live here:
 *   sdm_context_init / reset      (src/trellis.c:977-1011, 1189-1224)  -> dsdhip_trellis_reset
 *   sdm_process_block             (src/trellis.c:1013-1040)            -> dsdhip_trellis_run
 *   sdm_context_copy_state        (src/trellis.c:1064-1092)            -> dsdhip_trellis_copy
 *   DAS candidate seeding         (src/dsp_plugin.c:2329-2338)         -> dsdhip_trellis_seed
 */
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "dsdhip_internal.h"

extern "C" int tr_launch_jobs(const tr_params *prm, tr_job *d_jobs, int njobs,
                              tr_ctx *d_ctxs, hipStream_t stream);
extern "C" int tr_feed_eligible(const tr_params *prm);
extern "C" int tr_launch_feed(const tr_params *prm, const tr_job *h_jobs, tr_job *d_jobs, int njobs,
                              const tr_feed_host *hc, tr_feed_dev *dc, tr_ctx *d_ctxs, uint64_t poll,
                              uint64_t limit, uint64_t max_iter, size_t lds_pad, hipStream_t stream);
extern "C" int tr_feed_static_lds(const tr_params *prm);

struct dsdhip_trellis {
    tr_params prm;
    int nctx;
    tr_ctx *d_ctx;             /* device context pool */
    std::vector<int> snap_of;  /* context each one was copied from (-1 none) */
    tr_job *d_jobs;
    int jobs_cap;
    hipStream_t stream;
    /* optional HIP-event timing of the trellis launches */
    int timing;
    hipEvent_t ev0, ev1;
    double kernel_ms;        /* accumulated */
    int64_t kernel_launches;
    int64_t samples;         /* inputs consumed by timed launches */
    double *d_pin;           /* host-buffer entry: staged input / packed output, grow-only */
    uint64_t *d_pout;
    size_t pin_cap, pout_cap;
    tr_job *table;               /* pipelined job table, coherent pinned host memory (grow-only) */
    size_t table_cap, table_n;
    std::vector<int> table_src;  /* copy source of each entry */
    /* host-fed launch over the table: control words and whether it runs */
    tr_feed_host *h_feed;
    tr_feed_dev *d_feed;
    int feed_on;
};

static void ctx_fresh(tr_ctx *c)
{
    memset(c, 0, sizeof(*c));
    c->num_cands = 1;
    memset(c->link, 0xff, sizeof(c->link));
}

extern "C" {

dsdhip_trellis_t *dsdhip_trellis_create_coeffs(const double *a, const double *g, int order,
                                               int depth, int cands, int lat,
                                               double state_limit, int nctx)
{
    if (!a || !g || nctx < 1) return NULL;
    if (depth < 1 || depth > 32 || cands < 1 || cands > TR_MAXK || lat < 1 || lat > TR_MAXL)
        return NULL;
    if (order < 4 || order > TR_MAXP) return NULL;
    if (dsdhip_ensure_device() != 0) return NULL;
    dsdhip_trellis_t *t = new dsdhip_trellis();
    memset(&t->prm, 0, sizeof(t->prm));
    for (int i = 0; i < order; i++) {
        t->prm.a[i] = a[i];
        t->prm.g[i] = g[i];
    }
    t->prm.P = order;
    t->prm.K = cands;
    t->prm.L = lat;
    t->prm.B = depth;
    t->prm.mask = (uint32_t)(((uint64_t)1 << depth) - 1);
    t->prm.state_limit = state_limit;
    t->nctx = nctx;
    t->snap_of.assign(nctx, -1);
    t->d_jobs = NULL;
    t->jobs_cap = 0;
    t->d_pin = NULL;
    t->d_pout = NULL;
    t->pin_cap = t->pout_cap = 0;
    t->table = NULL;
    t->table_cap = t->table_n = 0;
    t->h_feed = NULL;
    t->d_feed = NULL;
    t->feed_on = 0;
    t->stream = 0;
    t->timing = 0;
    t->kernel_ms = 0.0;
    t->kernel_launches = 0;
    t->samples = 0;
    if (hipMalloc(&t->d_ctx, sizeof(tr_ctx) * (size_t)nctx) != hipSuccess) {
        delete t;
        return NULL;
    }
    /* every context starts fresh: one host copy, then doubling
     * device-to-device copies (a DAS pool holds thousands of contexts) */
    tr_ctx c;
    ctx_fresh(&c);
    bool ok = hipMemcpy(&t->d_ctx[0], &c, sizeof(c), hipMemcpyHostToDevice) == hipSuccess;
    for (int done = 1; ok && done < nctx; done *= 2) {
        const int n = std::min(done, nctx - done);
        ok = hipMemcpy(&t->d_ctx[done], &t->d_ctx[0], sizeof(tr_ctx) * (size_t)n,
                       hipMemcpyDeviceToDevice) == hipSuccess;
    }
    if (!ok) {
        hipFree(t->d_ctx);
        delete t;
        return NULL;
    }
    return t;
}

dsdhip_trellis_t *dsdhip_trellis_create(int ntf_index, int depth, int cands, int lat,
                                        double state_limit, int nctx)
{
    const dsdhip_ntf_t *f = dsdhip_ntf_get(ntf_index);
    if (!f) return NULL;
    return dsdhip_trellis_create_coeffs(f->a, f->g, f->order, depth, cands, lat, state_limit, nctx);
}

/* Enables HIP-event timing of the trellis launches on the pool's stream;
 * the accumulated kernel time, launch count and consumed input samples are
 * read (and cleared) with dsdhip_trellis_timing_read. */
int dsdhip_trellis_timing(dsdhip_trellis_t *t, int on)
{
    if (!t) return -1;
    if (on && !t->timing) {
        if (hipEventCreate(&t->ev0) != hipSuccess || hipEventCreate(&t->ev1) != hipSuccess) return -1;
    }
    if (!on && t->timing) {
        hipEventDestroy(t->ev0);
        hipEventDestroy(t->ev1);
    }
    t->timing = on ? 1 : 0;
    t->kernel_ms = 0.0;
    t->kernel_launches = 0;
    t->samples = 0;
    return 0;
}

int dsdhip_trellis_timing_read(dsdhip_trellis_t *t, double *ms, int64_t *launches, int64_t *samples,
                               int clear)
{
    if (!t) return -1;
    if (ms) *ms = t->kernel_ms;
    if (launches) *launches = t->kernel_launches;
    if (samples) *samples = t->samples;
    if (clear) {
        t->kernel_ms = 0.0;
        t->kernel_launches = 0;
        t->samples = 0;
    }
    return 0;
}

void dsdhip_trellis_destroy(dsdhip_trellis_t *t)
{
    if (!t) return;
    if (t->timing) {
        hipEventDestroy(t->ev0);
        hipEventDestroy(t->ev1);
    }
    hipFree(t->d_ctx);
    if (t->d_jobs) hipFree(t->d_jobs);
    if (t->d_pin) hipFree(t->d_pin);
    if (t->d_pout) hipFree(t->d_pout);
    if (t->table) hipHostFree(t->table);
    if (t->h_feed) hipHostFree(t->h_feed);
    if (t->d_feed) hipFree(t->d_feed);
    delete t;
}

int dsdhip_trellis_set_stream(dsdhip_trellis_t *t, void *stream)
{
    if (!t) return -1;
    t->stream = (hipStream_t)stream;
    return 0;
}

int dsdhip_trellis_force_literal(dsdhip_trellis_t *t, int on)
{
    if (!t) return -1;
    t->prm.force_slow = on ? 1 : 0;
    return 0;
}

int dsdhip_trellis_fma_o8(dsdhip_trellis_t *t, int on)
{
    if (!t) return -1;
    t->prm.fma_o8 = on ? 1 : 0;
    return 0;
}

int dsdhip_trellis_wave_only(dsdhip_trellis_t *t, int on)
{
    if (!t) return -1;
    t->prm.wave_only = on ? 1 : 0;
    return 0;
}

int dsdhip_trellis_num_contexts(const dsdhip_trellis_t *t) { return t ? t->nctx : 0; }

void *dsdhip_trellis_context_ptr(dsdhip_trellis_t *t, int i)
{
    if (!t || i < 0 || i >= t->nctx) return NULL;
    return &t->d_ctx[i];
}

int dsdhip_trellis_reset(dsdhip_trellis_t *t, int i)
{
    if (!t || i < 0 || i >= t->nctx) return -1;
    tr_ctx c;
    ctx_fresh(&c);
    t->snap_of[i] = -1;
    return hipMemcpy(&t->d_ctx[i], &c, sizeof(c), hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

int dsdhip_trellis_get(dsdhip_trellis_t *t, int i, tr_ctx *out)
{
    if (!t || i < 0 || i >= t->nctx || !out) return -1;
    if (hipStreamSynchronize(t->stream) != hipSuccess) return -1;
    return hipMemcpy(out, &t->d_ctx[i], sizeof(*out), hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

int dsdhip_trellis_put(dsdhip_trellis_t *t, int i, const tr_ctx *in)
{
    if (!t || i < 0 || i >= t->nctx || !in) return -1;
    return hipMemcpy(&t->d_ctx[i], in, sizeof(*in), hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

/* sdm_context_copy_state: dst becomes a copy of src; its hash links keep
 * pointing into src, so later lookups through them read src's slots. */
int dsdhip_trellis_copy(dsdhip_trellis_t *t, int dst, int src)
{
    if (!t || dst < 0 || src < 0 || dst >= t->nctx || src >= t->nctx || dst == src) return -1;
    tr_ctx c;
    if (dsdhip_trellis_get(t, src, &c) != 0) return -1;
    int bad = 0;
    for (int b = 0; b < 2; b++)
        for (int i = 0; i < TR_SLOTS; i++) {
            int8_t l = c.link[b][i];
            if (l >= TR_LINK_FOREIGN) bad = 1;
            else if (l >= 0) c.link[b][i] = (int8_t)(l + TR_LINK_FOREIGN);
        }
    c.has_foreign = 1;
    if (bad) return -1;   /* links two contexts deep are not represented */
    t->snap_of[dst] = src;
    return dsdhip_trellis_put(t, dst, &c);
}

/* Byte-exact duplicate of a context (no copy_state semantics). */
int dsdhip_trellis_clone(dsdhip_trellis_t *t, int dst, int src)
{
    if (!t || dst < 0 || src < 0 || dst >= t->nctx || src >= t->nctx) return -1;
    if (dst == src) return 0;
    t->snap_of[dst] = t->snap_of[src];
    return hipMemcpyAsync(&t->d_ctx[dst], &t->d_ctx[src], sizeof(tr_ctx), hipMemcpyDeviceToDevice,
                          t->stream) == hipSuccess ? 0 : -1;
}

/* Drops every path-hash link of a context (they are only followed by
 * lookups that walk past a bucket overwritten at K=32). */
int dsdhip_trellis_detach(dsdhip_trellis_t *t, int i)
{
    tr_ctx c;
    if (dsdhip_trellis_get(t, i, &c) != 0) return -1;
    memset(c.link, 0xff, sizeof(c.link));
    c.has_foreign = 0;
    t->snap_of[i] = -1;
    return dsdhip_trellis_put(t, i, &c);
}

/* copy_state followed by detach, in one round trip. */
int dsdhip_trellis_copy_detached(dsdhip_trellis_t *t, int dst, int src)
{
    if (!t || dst < 0 || src < 0 || dst >= t->nctx || src >= t->nctx) return -1;
    tr_ctx c;
    if (dsdhip_trellis_get(t, src, &c) != 0) return -1;
    memset(c.link, 0xff, sizeof(c.link));
    c.has_foreign = 0;
    t->snap_of[dst] = -1;
    return dsdhip_trellis_put(t, dst, &c);
}

int dsdhip_trellis_params(const dsdhip_trellis_t *t, tr_params *out)
{
    if (!t || !out) return -1;
    *out = t->prm;
    return 0;
}

hipStream_t dsdhip_trellis_stream(const dsdhip_trellis_t *t) { return t ? t->stream : 0; }

int dsdhip_trellis_seed(dsdhip_trellis_t *t, int i, const double *st)
{
    tr_ctx c;
    if (dsdhip_trellis_get(t, i, &c) != 0) return -1;
    for (int k = 0; k < c.num_cands; k++) {
        for (int q = 0; q < t->prm.P; q++) c.st[k][q] = st[q];
        c.cost[k] = 0.0;
    }
    return dsdhip_trellis_put(t, i, &c);
}

int dsdhip_trellis_best_state(dsdhip_trellis_t *t, int i, double *out)
{
    tr_ctx c;
    if (dsdhip_trellis_get(t, i, &c) != 0) return -1;
    for (int q = 0; q < 10; q++) out[q] = q < t->prm.P ? c.st[0][q] : 0.0;
    return 0;
}

/*
 * Run a batch of jobs in one launch.  d_in[j] / d_out[j] are device
 * pointers; job j runs n_in[j] samples on context ctx[j] (optionally
 * starting from a copy of copy_src[j] >= 0 seeded with seed[j*10..]),
 * keeps outputs produced from input index keep_from[j] on, and stores the
 * final state back into ctx[j].  kept[j] receives the kept output count.
 * Returns 0, or -1 on launch failure / error flag (err_out gets the flag).
 */
/* one job descriptor; -1 on a context index outside the pool */
static int fill_job(const dsdhip_trellis_t *t, tr_job &jb, int ctx, int src, const double *d_in,
                    int64_t n_in, int64_t keep_from, uint64_t *d_out, const double *seed10)
{
    memset(&jb, 0, sizeof(jb));
    if (ctx < 0 || ctx >= t->nctx || src >= t->nctx) return -1;
    jb.in = d_in;
    jb.out = d_out;
    jb.n_in = n_in;
    jb.keep_from = keep_from;
    jb.ctx_out = ctx;
    if (src >= 0) {
        jb.ctx_in = src;
        jb.copy = 1;
        jb.snap = src;
    } else {
        jb.ctx_in = ctx;
        jb.copy = 0;
        jb.snap = t->snap_of[ctx];
    }
    if (seed10 && !(seed10[0] != seed10[0])) {   /* NaN in slot 0 = no seed */
        jb.seed = 1;
        for (int q = 0; q < TR_MAXP; q++) jb.seed_st[q] = seed10[q];
    }
    return 0;
}

static int ensure_jobs(dsdhip_trellis_t *t, int njobs)
{
    if (t->jobs_cap >= njobs) return 0;
    if (t->d_jobs) hipFree(t->d_jobs);
    t->d_jobs = NULL;
    t->jobs_cap = 0;
    if (hipMalloc(&t->d_jobs, sizeof(tr_job) * (size_t)njobs) != hipSuccess) return -1;
    t->jobs_cap = njobs;
    return 0;
}

/*
 * Pipelined launches over one job table: entries are filled on the host and
 * launched in batches as their inputs become known (each batch on the
 * stream the caller names, so batches overlap), and the results of every
 * entry are read back once at the end.  The pool's device job table is
 * grown to njobs first; nothing else of the pool may launch until
 * dsdhip_trellis_table_end.
 */
int dsdhip_trellis_table_begin(dsdhip_trellis_t *t, int njobs)
{
    if (!t || njobs < 0) return -1;
    if (ensure_jobs(t, njobs > 0 ? njobs : 1) != 0) return -1;
    if (t->table_cap < (size_t)njobs) {
        if (t->table) hipHostFree(t->table);
        t->table = NULL;
        t->table_cap = 0;
        /* coherent: a host-fed launch reads entries the host writes while it runs */
        if (hipHostMalloc((void **)&t->table, sizeof(tr_job) * (size_t)njobs, hipHostMallocCoherent) !=
            hipSuccess)
            return -1;
        t->table_cap = (size_t)njobs;
    }
    t->table_n = (size_t)njobs;
    t->table_src.assign((size_t)njobs, -1);
    t->feed_on = 0;
    return 0;
}

/*
 * Host-fed launch over the whole table (after table_begin): one launch on
 * `stream` with one wave per entry, each waiting until the host has
 * published its entry with dsdhip_trellis_feed_publish (entries 0..n-1
 * complete, published in index order).  Every entry must be published, or the launch
 * aborted with dsdhip_trellis_feed_abort, before the stream is joined; a
 * wave that waits longer than limit_s seconds gives up (table_end then
 * reports an error).  The launch waits for the host, so nothing the host
 * needs before it can publish (input copies it reads, events it waits on)
 * may be queued after it: streams can share a hardware queue, so queue
 * such work before this call, on `stream`.  Returns 0, 1 when the host-fed
 * form does not apply
 * to these parameters (launch the entries with table_launch instead), or
 * -1 on failure.
 */
int dsdhip_trellis_feed_begin(dsdhip_trellis_t *t, void *stream, double limit_s, int waves_per_cu)
{
    hipStream_t s = (hipStream_t)stream;
    if (!t) return -1;
    if (!tr_feed_eligible(&t->prm)) return 1;
    const int n = (int)t->table_n;
    if (n == 0) return 1;
    if (!t->h_feed &&
        hipHostMalloc((void **)&t->h_feed, sizeof(tr_feed_host), hipHostMallocCoherent) != hipSuccess) {
        t->h_feed = NULL;
        return -1;
    }
    if (!t->d_feed && hipMalloc(&t->d_feed, sizeof(tr_feed_dev)) != hipSuccess) {
        t->d_feed = NULL;
        return -1;
    }
    int dev = 0, khz = 0, lds_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0 ||
        hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
        return -1;
    /* at most waves_per_cu resident workgroups per CU: pad each one's LDS to
     * just over lds_cu / (waves_per_cu + 1) */
    size_t pad = 0;
    if (waves_per_cu > 0) {
        const int st = tr_feed_static_lds(&t->prm);
        if (st < 0) return -1;
        const size_t want = (size_t)lds_cu / (size_t)(waves_per_cu + 1) + 256;
        if (want > (size_t)st) pad = want - (size_t)st;
        if ((size_t)st + pad > 65536) pad = 65536 > (size_t)st ? 65536 - (size_t)st : 0;
    }
    tr_job *hd = NULL;
    tr_feed_host *hc = NULL;
    if (hipHostGetDevicePointer((void **)&hd, t->table, 0) != hipSuccess ||
        hipHostGetDevicePointer((void **)&hc, t->h_feed, 0) != hipSuccess)
        return -1;
    __atomic_store_n(&t->h_feed->ready, 0, __ATOMIC_RELEASE);
    __atomic_store_n(&t->h_feed->abort, 0, __ATOMIC_RELEASE);
    const uint64_t poll = (uint64_t)khz * 2 / 1000;           /* 2 us */
    const uint64_t limit = (uint64_t)((double)khz * 1000.0 * limit_s);
    /* a second bound in wait rounds: a round sleeps about 1000 clocks (at
     * least 0.3 us), so this is at most a few times limit_s */
    const uint64_t max_iter = (uint64_t)(limit_s * 3.0e6) + 1;
    if (tr_launch_feed(&t->prm, hd, t->d_jobs, n, hc, t->d_feed, t->d_ctx, poll ? poll : 1, limit, max_iter, pad,
                       s) != 0)
        return -1;
    t->feed_on = 1;
    return 0;
}

/* entries 0..n-1 of the table are complete (host-fed launch) */
int dsdhip_trellis_feed_publish(dsdhip_trellis_t *t, int n)
{
    if (!t || !t->feed_on || n < 0 || (size_t)n > t->table_n) return -1;
    __atomic_store_n(&t->h_feed->ready, n, __ATOMIC_RELEASE);
    return 0;
}

/* the host-fed launch leaves without taking further entries */
int dsdhip_trellis_feed_abort(dsdhip_trellis_t *t)
{
    if (!t || !t->feed_on) return -1;
    __atomic_store_n(&t->h_feed->abort, 1, __ATOMIC_RELEASE);
    return 0;
}

int dsdhip_trellis_table_fill(dsdhip_trellis_t *t, int j, int ctx, int copy_src, const double *d_in,
                              int64_t n_in, int64_t keep_from, uint64_t *d_out, const double *seed10)
{
    if (!t || j < 0 || (size_t)j >= t->table_n) return -1;
    t->table_src[(size_t)j] = copy_src;
    return fill_job(t, t->table[j], ctx, copy_src, d_in, n_in, keep_from, d_out, seed10);
}

int dsdhip_trellis_table_launch(dsdhip_trellis_t *t, int first, int n, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!t || first < 0 || n < 0 || (size_t)first + (size_t)n > t->table_n) return -1;
    if (n == 0) return 0;
    if (hipMemcpyAsync(t->d_jobs + first, t->table + first, sizeof(tr_job) * (size_t)n,
                       hipMemcpyHostToDevice, s) != hipSuccess)
        return -1;
    return tr_launch_jobs(&t->prm, t->d_jobs + first, n, t->d_ctx, s);
}

/* After every batch has been joined into stream s: results of all entries. */
int dsdhip_trellis_table_end(dsdhip_trellis_t *t, void *stream, int64_t *kept, int64_t *slow_steps,
                             int *err_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!t) return -1;
    const size_t n = t->table_n;
    tr_feed_dev fd;
    memset(&fd, 0, sizeof(fd));
    if (t->feed_on &&
        hipMemcpyAsync(&fd, t->d_feed, sizeof(fd), hipMemcpyDeviceToHost, s) != hipSuccess)
        return -1;
    if (n && hipMemcpyAsync(t->table, t->d_jobs, sizeof(tr_job) * n, hipMemcpyDeviceToHost, s) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    if (t->feed_on) {
        t->feed_on = 0;
        /* a wave gave up waiting, or the launch was aborted: entries may not have run */
        if (fd.err || fd.abort || fd.done != (int)n) {
            if (err_out) *err_out = 3;
            return -1;
        }
    }
    int err = 0;
    for (size_t j = 0; j < n; j++) {
        if (kept) kept[j] = t->table[j].kept;
        if (slow_steps) slow_steps[j] = t->table[j].slow_steps;
        if (t->table[j].err && !err) err = t->table[j].err;
        if (t->table_src[j] >= 0) t->snap_of[t->table[j].ctx_out] = t->table_src[j];
    }
    if (t->timing)
        for (size_t j = 0; j < n; j++) t->samples += t->table[j].n_in;
    if (err_out) *err_out = err;
    return err ? -1 : 0;
}

/* fill_job for other translation units */
int dsdhip_trellis_fill_job(const dsdhip_trellis_t *t, tr_job *jb, int ctx, int src, const double *d_in,
                            int64_t n_in, int64_t keep_from, uint64_t *d_out, const double *seed10)
{
    if (!t || !jb) return -1;
    return fill_job(t, *jb, ctx, src, d_in, n_in, keep_from, d_out, seed10);
}

/* records that context i now follows links into context src */
int dsdhip_trellis_set_snap(dsdhip_trellis_t *t, int i, int src)
{
    if (!t || i < 0 || i >= t->nctx) return -1;
    t->snap_of[i] = src;
    return 0;
}

/* Adds a measured span to the launch timing (pipelined runs time their
 * trellis work as one span). */
int dsdhip_trellis_timing_add(dsdhip_trellis_t *t, double ms, int64_t launches)
{
    if (!t) return -1;
    if (t->timing) {
        t->kernel_ms += ms;
        t->kernel_launches += launches;
    }
    return 0;
}

int dsdhip_trellis_run_jobs(dsdhip_trellis_t *t, int njobs, const int *ctx, const int *copy_src,
                            const double *const *d_in, const int64_t *n_in,
                            const int64_t *keep_from, uint64_t *const *d_out,
                            const double *seed, int64_t *kept, int64_t *slow_steps,
                            int *err_out)
{
    return dsdhip_trellis_run_jobs_then(t, njobs, ctx, copy_src, d_in, n_in, keep_from, d_out, seed, kept,
                                        slow_steps, err_out, nullptr, nullptr);
}

int dsdhip_trellis_run_jobs_then(dsdhip_trellis_t *t, int njobs, const int *ctx, const int *copy_src,
                                 const double *const *d_in, const int64_t *n_in, const int64_t *keep_from,
                                 uint64_t *const *d_out, const double *seed, int64_t *kept,
                                 int64_t *slow_steps, int *err_out, int (*queued)(void *), void *arg)
{
    if (!t || njobs < 0) return -1;
    if (njobs == 0) return 0;
    std::vector<tr_job> jobs((size_t)njobs);
    for (int j = 0; j < njobs; j++) {
        const int src = (copy_src && copy_src[j] >= 0) ? copy_src[j] : -1;
        if (fill_job(t, jobs[(size_t)j], ctx[j], src, d_in[j], n_in[j], keep_from ? keep_from[j] : 0,
                     d_out[j], seed ? seed + (size_t)j * 10 : NULL) != 0)
            return -1;
    }
    if (ensure_jobs(t, njobs) != 0) return -1;
    if (hipMemcpyAsync(t->d_jobs, jobs.data(), sizeof(tr_job) * (size_t)njobs,
                       hipMemcpyHostToDevice, t->stream) != hipSuccess)
        return -1;
    if (t->timing) hipEventRecord(t->ev0, t->stream);
    if (tr_launch_jobs(&t->prm, t->d_jobs, njobs, t->d_ctx, t->stream) != 0) return -1;
    if (t->timing) hipEventRecord(t->ev1, t->stream);
    if (hipMemcpyAsync(jobs.data(), t->d_jobs, sizeof(tr_job) * (size_t)njobs,
                       hipMemcpyDeviceToHost, t->stream) != hipSuccess)
        return -1;
    const int qrc = queued ? queued(arg) : 0;
    if (hipStreamSynchronize(t->stream) != hipSuccess || qrc != 0) return -1;
    if (t->timing) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t->ev0, t->ev1) == hipSuccess) t->kernel_ms += ms;
        t->kernel_launches++;
        for (int j = 0; j < njobs; j++) t->samples += n_in[j];
    }
    int err = 0;
    for (int j = 0; j < njobs; j++) {
        if (kept) kept[j] = jobs[(size_t)j].kept;
        if (slow_steps) slow_steps[j] = jobs[(size_t)j].slow_steps;
        if (jobs[(size_t)j].err && !err) err = jobs[(size_t)j].err;
        int src = (copy_src && copy_src[j] >= 0) ? copy_src[j] : -1;
        if (src >= 0) t->snap_of[ctx[j]] = src;
    }
    if (err_out) *err_out = err;
    return err ? -1 : 0;
}

/* sdm_process_block on one context with host buffers: float +-1.0 out. */
int dsdhip_trellis_process(dsdhip_trellis_t *t, int i, const double *in, float *out,
                           size_t count, size_t *out_count)
{
    if (!t || i < 0 || i >= t->nctx) return -1;
    size_t nw = (count + 63) / 64 + 1;
    const size_t nin = count ? count : 1;
    if (t->pin_cap < nin) {
        if (t->d_pin) hipFree(t->d_pin);
        t->d_pin = NULL;
        t->pin_cap = 0;
        if (hipMalloc(&t->d_pin, sizeof(double) * nin) != hipSuccess) return -1;
        t->pin_cap = nin;
    }
    if (t->pout_cap < nw) {
        if (t->d_pout) hipFree(t->d_pout);
        t->d_pout = NULL;
        t->pout_cap = 0;
        if (hipMalloc(&t->d_pout, sizeof(uint64_t) * nw) != hipSuccess) return -1;
        t->pout_cap = nw;
    }
    double *d_in = t->d_pin;
    uint64_t *d_out = t->d_pout;
    int rc = -1;
    if (hipMemcpy(d_in, in, sizeof(double) * count, hipMemcpyHostToDevice) == hipSuccess) {
        const double *pin = d_in;
        uint64_t *pout = d_out;
        int64_t n = (int64_t)count, kf = 0, kept = 0;
        int err = 0;
        rc = dsdhip_trellis_run_jobs(t, 1, &i, NULL, &pin, &n, &kf, &pout, NULL, &kept, NULL, &err);
        if (rc == 0) {
            std::vector<uint64_t> bitsv(nw);
            if (hipMemcpy(bitsv.data(), d_out, sizeof(uint64_t) * nw, hipMemcpyDeviceToHost) != hipSuccess)
                rc = -1;
            else {
                for (int64_t k = 0; out && k < kept; k++)
                    out[k] = ((bitsv[(size_t)(k >> 6)] >> (k & 63)) & 1) ? 1.0f : -1.0f;
                if (out_count) *out_count = (size_t)kept;
            }
        }
    }
    return rc;
}

/* sdm_state_distance (src/trellis.c:1094-1109): squared distance of the
 * best survivors' loop-filter states of two contexts (1e30 if either has
 * none). */
int dsdhip_trellis_state_distance(dsdhip_trellis_t *t, int i, int j, double *dist)
{
    if (!t || !dist || i < 0 || j < 0 || i >= t->nctx || j >= t->nctx) return -1;
    tr_ctx a, b;
    if (dsdhip_trellis_get(t, i, &a) != 0 || dsdhip_trellis_get(t, j, &b) != 0) return -1;
    if (a.num_cands < 1 || b.num_cands < 1) {
        *dist = 1e30;
        return 0;
    }
    double d = 0.0;
    for (int k = 0; k < t->prm.P; k++) {
        const double e = a.st[0][k] - b.st[0][k];
        d += e * e;
    }
    *dist = d;
    return 0;
}

/* sdm_drain (src/trellis.c:1042-1062): up to max_out outputs from zero
 * input.  A context that has not filled its latency and is not draining
 * first runs the missing fill steps on zeros without output; then `pending`
 * zero steps (at most max_out) each give one output, and the context is
 * left draining with pending reduced by that count. */
int dsdhip_trellis_drain(dsdhip_trellis_t *t, int i, float *out, size_t max_out, size_t *out_count)
{
    if (!t || i < 0 || i >= t->nctx || (!out && max_out)) return -1;
    tr_ctx c;
    if (dsdhip_trellis_get(t, i, &c) != 0) return -1;
    const int L = t->prm.L;
    const size_t p0 = (size_t)c.pending;
    const size_t len = p0 < max_out ? p0 : max_out;
    if (!c.draining && c.pending < L) {
        std::vector<double> z((size_t)(L - c.pending), 0.0);
        size_t n = 0;
        /* the fill steps of process produce no output */
        if (dsdhip_trellis_process(t, i, z.data(), NULL, z.size(), &n) != 0 || n != 0) return -1;
        if (dsdhip_trellis_get(t, i, &c) != 0) return -1;
    }
    size_t n = 0;
    if (len) {
        c.pending = L;                 /* every zero step gives an output */
        if (dsdhip_trellis_put(t, i, &c) != 0) return -1;
        std::vector<double> z(len, 0.0);
        if (dsdhip_trellis_process(t, i, z.data(), out, len, &n) != 0 || n != len) return -1;
        if (dsdhip_trellis_get(t, i, &c) != 0) return -1;
    }
    c.pending = (int32_t)(p0 - len);
    c.draining = 1;
    if (dsdhip_trellis_put(t, i, &c) != 0) return -1;
    if (out_count) *out_count = n;
    return 0;
}

} /* extern "C" */
