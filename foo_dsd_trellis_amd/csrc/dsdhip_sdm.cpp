/*
 * The reference's sdm_* / ntf_* C surface (include/trellis.h:61-118,
 * include/ntf.h:23-36) on caller-owned contexts in the reference's layout.
 *
 * sdm_process_block / sdm_drain convert the context to the device form
 * (tr_ctx), run one trellis job on a per-thread device pool of two
 * contexts, and write the result back into the caller's struct:
 *
 *   survivors      act[c] of the active bank (idx) -> st/cost/path/next, the
 *                  node's slot, and the bits of its history buffer;
 *   nodes          path, metric and path_list of the 64 nodes of both banks,
 *                  the only node fields the sort reads again once the node
 *                  has stopped being a survivor (src/trellis.c:599-619,
 *                  621-731: stale chains compare path and metric);
 *   chain links    path_list as a node index of the same bank, or, for a
 *                  link into another context (sdm_context_copy_state copies
 *                  chain pointers unchanged, src/trellis.c:1064-1092), as a
 *                  node of the job's second context, which holds every
 *                  outside node reachable from this context's chains with
 *                  its path, metric and onward link;
 *   bookkeeping    num_cands, pos, pending, draining, idx.
 *
 * Outside nodes are read only inside contexts the library knows to be
 * sources of this one (registered by sdm_context_copy_state), so no
 * pointer is followed that the reference would not be able to follow.
 * The path-hash heads are not carried: every step clears the heads of all
 * its children's buckets before its first lookup (src/trellis.c:629-633),
 * so a head left by an earlier step is never read.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <unordered_map>
#include <vector>

#include "dsdhip_internal.h"
#include "../../include/dsd_sdm.h"

static_assert(sizeof(ntf_filter_t) == sizeof(dsdhip_ntf_t), "ntf_filter_t layout");
static_assert(sizeof(sdm_state_t) == 112, "sdm_state_t layout");
static_assert(sizeof(sdm_context_t) == 46752, "sdm_context_t layout");

namespace {

/* ---- copy-source registry: dst -> the context it was copied from ---- */
std::mutex g_reg_mu;
std::unordered_map<const sdm_context_t *, const sdm_context_t *> g_src;

void reg_set(const sdm_context_t *dst, const sdm_context_t *src)
{
    std::lock_guard<std::mutex> lk(g_reg_mu);
    if (src) g_src[dst] = src;
    else g_src.erase(dst);
}

/* the chain of copy sources of c (c itself excluded), nearest first */
std::vector<const sdm_context_t *> sources_of(const sdm_context_t *c)
{
    std::vector<const sdm_context_t *> out;
    std::lock_guard<std::mutex> lk(g_reg_mu);
    const sdm_context_t *k = c;
    for (int depth = 0; depth < 64; depth++) {
        auto it = g_src.find(k);
        if (it == g_src.end() || it->second == c) break;
        bool seen = false;
        for (const sdm_context_t *o : out) seen = seen || o == it->second;
        if (seen) break;
        out.push_back(it->second);
        k = it->second;
    }
    return out;
}

bool node_of(const sdm_context_t *c, int b, const sdm_state_t *p, int *slot)
{
    const sdm_state_t *base = &c->trellis[b].sdm[0];
    if (p < base || p >= base + TR_SLOTS) return false;
    *slot = (int)(p - base);
    return true;
}

bool within(const sdm_context_t *c, const void *p)
{
    const char *q = (const char *)p, *b = (const char *)c;
    return q >= b && q < b + sizeof(*c);
}

/* a node in the banks of one of the known sources */
bool source_node(const std::vector<const sdm_context_t *> &srcs, const sdm_state_t *p)
{
    for (const sdm_context_t *s : srcs)
        for (int b = 0; b < 2; b++) {
            const sdm_state_t *base = &s->trellis[b].sdm[0];
            if (p >= base && p < base + SDM_MB_LEVELS * SDM_TRELLIS_MAX_NUM &&
                ((const char *)p - (const char *)base) % sizeof(sdm_state_t) == 0)
                return true;
        }
    return false;
}

/* outside nodes reached by a context's chains, per bank: index = slot of
 * the job's second context */
struct Outside {
    const sdm_state_t *node[2][TR_SLOTS];
    int n[2];
};

int outside_index(Outside &o, int b, const sdm_state_t *p)
{
    for (int k = 0; k < o.n[b]; k++)
        if (o.node[b][k] == p) return k;
    if (o.n[b] == TR_SLOTS) return -1;
    o.node[b][o.n[b]] = p;
    return o.n[b]++;
}

void ctx_clear(tr_ctx *t)
{
    memset(t, 0, sizeof(*t));
    memset(t->link, 0xff, sizeof(t->link));
}

/* reference layout -> device form (t), with the outside nodes in snap */
const char *to_device(const sdm_context_t *c, int P, tr_ctx *t, tr_ctx *snap, Outside &o)
{
    ctx_clear(t);
    ctx_clear(snap);
    o.n[0] = o.n[1] = 0;
    const int K = (int)c->trellis_num, L = (int)c->trellis_lat;
    const int n = (int)c->num_cands, bank = (int)(c->idx & 1);
    if (n < 1 || n > K) return "num_cands out of range";
    if ((int)c->pos >= L || (int)c->pending > L) return "pos / pending out of range";
    t->num_cands = n;
    t->pos = (int)c->pos;
    t->pending = (int)c->pending;
    t->draining = c->draining ? 1 : 0;
    t->idx = bank;
    const size_t hb = (size_t)(L + 7) / 8;
    for (int i = 0; i < n; i++) {
        const sdm_state_t *s = c->trellis[bank].act[i];
        int slot = 0;
        if (!s || !node_of(c, bank, s, &slot)) return "survivor outside its bank";
        if (s->hist >= 2 * K) return "survivor history index out of range";
        t->slot[i] = (uint8_t)slot;
        for (int q = 0; q < P; q++) t->st[i][q] = s->state[q];
        t->cost[i] = s->cost;
        t->path[i] = s->path;
        t->next[i] = s->next;
        memcpy(&t->hist[i][0], c->hist[s->hist], hb);
    }
    std::vector<const sdm_context_t *> srcs;
    bool have_srcs = false;
    for (int b = 0; b < 2; b++)
        for (int j = 0; j < TR_SLOTS; j++) {
            const sdm_state_t *nd = &c->trellis[b].sdm[j];
            t->node_path[b][j] = nd->path;
            t->node_cost[b][j] = nd->cost;
            const sdm_state_t *p = nd->path_list;
            int l = TR_LINK_NONE, slot = 0;
            if (!p) {
                l = TR_LINK_NONE;
            } else if (node_of(c, b, p, &slot)) {
                l = slot;
            } else if (within(c, p)) {
                return "chain link into the other bank";
            } else {
                if (!have_srcs) {
                    srcs = sources_of(c);
                    have_srcs = true;
                }
                if (!source_node(srcs, p)) return "chain link into memory that is not a copy source of this context";
                const int k = outside_index(o, b, p);
                if (k < 0) return "more than 64 outside nodes in one bank";
                l = TR_LINK_FOREIGN + k;
            }
            t->link[b][j] = (int8_t)l;
        }
    /* outside nodes, with their onward links (the list grows while walked) */
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < o.n[b]; k++) {
            const sdm_state_t *p = o.node[b][k];
            snap->node_path[b][k] = p->path;
            snap->node_cost[b][k] = p->cost;
            const sdm_state_t *q = p->path_list;
            if (!q) continue;
            if (within(c, q)) return "outside chain leads back into this context";
            if (!source_node(srcs, q)) return "outside chain leaves the copy sources";
            const int k2 = outside_index(o, b, q);
            if (k2 < 0) return "more than 64 outside nodes in one bank";
            snap->link[b][k] = (int8_t)k2;
        }
    t->has_foreign = (o.n[0] + o.n[1]) > 0 ? 1 : 0;
    return NULL;
}

/* device form -> reference layout, in place */
void from_device(sdm_context_t *c, int P, const tr_ctx *t, const Outside &o)
{
    const int K = (int)c->trellis_num, L = (int)c->trellis_lat;
    const int n = t->num_cands, bank = t->idx & 1;
    for (int b = 0; b < 2; b++)
        for (int j = 0; j < TR_SLOTS; j++) {
            sdm_state_t *nd = &c->trellis[b].sdm[j];
            nd->path = t->node_path[b][j];
            nd->cost = t->node_cost[b][j];
            const int l = t->link[b][j];
            nd->path_list = l < 0 ? NULL
                            : l < TR_LINK_FOREIGN ? &c->trellis[b].sdm[l]
                                                  : (sdm_state_t *)o.node[b][l - TR_LINK_FOREIGN];
        }
    /* survivor i owns history buffer i; the others are free */
    c->hist_fnum = 0;
    for (int h = 2 * K - 1; h >= n; h--) c->hist_free[c->hist_fnum++] = (uint8_t)h;
    const size_t hb = (size_t)(L + 7) / 8;
    for (int i = 0; i < n; i++) {
        sdm_state_t *s = &c->trellis[bank].sdm[t->slot[i]];
        for (int q = 0; q < P; q++) s->state[q] = t->st[i][q];
        s->cost = t->cost[i];
        s->path = t->path[i];
        s->next = t->next[i];
        s->hist = (uint8_t)i;
        s->hist_used = 0;
        memcpy(c->hist[i], &t->hist[i][0], hb);
        c->trellis[bank].act[i] = s;
    }
    c->num_cands = (unsigned)n;
    c->pos = (unsigned)t->pos;
    c->pending = (unsigned)t->pending;
    c->draining = (unsigned)t->draining;
    c->idx = (unsigned)bank;
}

/* parameters of a context the device can run, or the reason it cannot */
const char *check_params(const sdm_context_t *c)
{
    if (!c) return "null context";
    if (!c->filter) return "context has no filter";
    if (c->filter->order < 4 || c->filter->order > TR_MAXP) return "loop-filter order outside 4..8";
    if (c->mb_levels != 2) return "only 1-bit branching (mb_levels = 2) runs on the GPU";
    if (c->trellis_num < 1 || c->trellis_num > TR_MAXK) return "trellis_num outside 1..32";
    if (c->trellis_lat < 1 || c->trellis_lat > TR_MAXL) return "trellis_lat outside 1..2048";
    if (c->trellis_mask == 0 || (c->trellis_mask & (c->trellis_mask + 1u)) != 0)
        return "trellis_mask is not 2^depth - 1";
    return NULL;
}

/* ---- device pools: two contexts each, keyed by the parameter set ---- */
struct Engine {
    double a[TR_MAXP], g[TR_MAXP];
    int P, K, L;
    uint32_t mask;
    double limit;
    dsdhip_trellis_t *pool;
    hipStream_t stream;
    bool busy;
};

std::mutex g_eng_mu;
std::vector<Engine *> g_engines;     /* kept until the process ends */

/* depth B of a mask 2^B - 1 */
int depth_of(uint32_t mask)
{
    int d = 0;
    while (d < 32 && ((mask >> d) & 1u)) d++;
    return d;
}

Engine *acquire(const sdm_context_t *c)
{
    const ntf_filter_t *f = c->filter;
    const int P = f->order;
    std::lock_guard<std::mutex> lk(g_eng_mu);
    for (Engine *e : g_engines) {
        if (e->busy || e->P != P || e->K != (int)c->trellis_num || e->L != (int)c->trellis_lat ||
            e->mask != c->trellis_mask || memcmp(&e->limit, &c->state_limit, sizeof(double)) != 0 ||
            memcmp(e->a, f->a, sizeof(double) * (size_t)P) != 0 ||
            memcmp(e->g, f->g, sizeof(double) * (size_t)P) != 0)
            continue;
        e->busy = true;
        return e;
    }
    Engine *e = new Engine();
    memcpy(e->a, f->a, sizeof(double) * (size_t)P);
    memcpy(e->g, f->g, sizeof(double) * (size_t)P);
    e->P = P;
    e->K = (int)c->trellis_num;
    e->L = (int)c->trellis_lat;
    e->mask = c->trellis_mask;
    e->limit = c->state_limit;
    const int depth = depth_of(c->trellis_mask);
    e->pool = dsdhip_trellis_create_coeffs(f->a, f->g, P, depth, e->K, e->L, c->state_limit, 2);
    if (!e->pool || hipStreamCreate(&e->stream) != hipSuccess) {
        if (e->pool) dsdhip_trellis_destroy(e->pool);
        delete e;
        return nullptr;
    }
    dsdhip_trellis_set_stream(e->pool, e->stream);
    /* one job per call: the one-wave kernel at every K, which carries every
     * node's path, metric and link (the lane-per-job form at K <= 2 keeps
     * only the survivors) */
    dsdhip_trellis_wave_only(e->pool, 1);
    e->busy = true;
    g_engines.push_back(e);
    return e;
}

void release(Engine *e)
{
    std::lock_guard<std::mutex> lk(g_eng_mu);
    e->busy = false;
}

void say(const char *fn, const char *why)
{
    fprintf(stderr, "libdsdhip %s: %s; context left unchanged\n", fn, why);
}

/* one GPU call on ctx: process (drain = false) or drain */
size_t run(sdm_context_t *ctx, const double *in, float *out, size_t count, bool drain, const char *fn)
{
    const char *why = check_params(ctx);
    if (why) {
        say(fn, why);
        return 0;
    }
    if (!drain && count == 0) return 0;
    if (!drain && !in) {
        say(fn, "null input");
        return 0;
    }
    if (!out && (drain ? count > 0 : ctx->pending + count > ctx->trellis_lat)) {
        say(fn, "null output");
        return 0;
    }
    const int P = ctx->filter->order;
    tr_ctx t, snap;
    Outside o;
    why = to_device(ctx, P, &t, &snap, o);
    if (why) {
        say(fn, why);
        return 0;
    }
    Engine *e = acquire(ctx);
    if (!e) {
        say(fn, "no GPU device pool (HIP device unavailable)");
        return 0;
    }
    size_t n = 0;
    int rc = dsdhip_trellis_put(e->pool, 0, &t);
    if (rc == 0) rc = dsdhip_trellis_put(e->pool, 1, &snap);
    if (rc == 0) rc = dsdhip_trellis_set_snap(e->pool, 0, t.has_foreign ? 1 : -1);
    if (rc == 0) rc = drain ? dsdhip_trellis_drain(e->pool, 0, out, count, &n)
                            : dsdhip_trellis_process(e->pool, 0, in, out, count, &n);
    tr_ctx r;
    if (rc == 0) rc = dsdhip_trellis_get(e->pool, 0, &r);
    release(e);
    if (rc != 0) {
        say(fn, "trellis job failed");
        return 0;
    }
    if (r.err) {
        say(fn, r.err == 1 ? "non-finite path metric" : "hash chain the device form cannot follow");
        return 0;
    }
    from_device(ctx, P, &r, o);
    return n;
}

void init_candidate(sdm_context_t *ctx)
{
    ctx->hist_fnum = 0;
    for (unsigned i = 0; i < 2u * ctx->trellis_num; i++) ctx->hist_free[ctx->hist_fnum++] = (uint8_t)i;
    sdm_state_t *s0 = &ctx->trellis[0].sdm[0];
    s0->hist = ctx->hist_free[--ctx->hist_fnum];
    s0->path = 0;
    ctx->trellis[0].act[0] = s0;
}

}  // namespace

extern "C" {

int sdm_context_init(sdm_context_t *ctx, const ntf_filter_t *filter, int trellis_depth, int trellis_cands,
                     int trellis_lat)
{
    if (!ctx) return -1;
    memset(ctx, 0, sizeof(*ctx));
    reg_set(ctx, NULL);
    if (!filter) return -1;
    if (trellis_depth < 1 || trellis_depth > SDM_TRELLIS_MAX_ORDER || trellis_cands < 1 ||
        trellis_cands > SDM_TRELLIS_MAX_NUM || trellis_lat < 1 || trellis_lat > SDM_TRELLIS_MAX_LAT)
        return -1;
    ctx->filter = filter;
    ctx->trellis_num = (uint32_t)trellis_cands;
    ctx->trellis_lat = (uint32_t)trellis_lat;
    ctx->trellis_mask = (uint32_t)(((uint64_t)1 << trellis_depth) - 1);
    ctx->num_cands = 1;
    ctx->mb_levels = 2;
    ctx->state_limit = 0.0;
    init_candidate(ctx);
    return 0;
}

size_t sdm_process_block(sdm_context_t *ctx, const double *in, float *out, size_t count)
{
    return run(ctx, in, out, count, false, "sdm_process_block");
}

size_t sdm_drain(sdm_context_t *ctx, float *out, size_t max_out)
{
    return run(ctx, NULL, out, max_out, true, "sdm_drain");
}

void sdm_context_copy_state(sdm_context_t *dst, const sdm_context_t *src)
{
    if (!dst || !src || dst == src) return;
    memcpy(dst, src, sizeof(*dst));
    const unsigned nc = dst->num_cands < SDM_TRELLIS_MAX_NUM ? dst->num_cands : SDM_TRELLIS_MAX_NUM;
    for (int b = 0; b < 2; b++)
        for (unsigned i = 0; i < nc; i++)
            if (src->trellis[b].act[i])
                dst->trellis[b].act[i] = dst->trellis[b].sdm + (src->trellis[b].act[i] - src->trellis[b].sdm);
    memset(dst->path_hash, 0, sizeof(dst->path_hash));
    dst->conv_fail = 0;
    dst->cands_collapse = 0;
    dst->next_filter_drops = 0;
    dst->total_children = 0;
    reg_set(dst, src);
}

double sdm_state_distance(const sdm_context_t *a, const sdm_context_t *b)
{
    if (!a || !b || !a->filter) return 1e30;
    const sdm_state_t *sa = a->trellis[a->idx & 1].act[0];
    const sdm_state_t *sb = b->trellis[b->idx & 1].act[0];
    if (!sa || !sb) return 1e30;
    double d = 0.0;
    for (int i = 0; i < a->filter->order; i++) {
        const double e = sa->state[i] - sb->state[i];
        d += e * e;
    }
    return d;
}

void sdm_context_reset(sdm_context_t *ctx)
{
    if (!ctx || !ctx->filter) return;
    memset(ctx->trellis, 0, sizeof(ctx->trellis));
    memset(ctx->path_hash, 0, sizeof(ctx->path_hash));
    memset(ctx->hist, 0, sizeof(ctx->hist));
    ctx->num_cands = 1;
    ctx->pos = 0;
    ctx->pending = 0;
    ctx->draining = 0;
    ctx->idx = 0;
    ctx->prev_y = 0.0;
    ctx->conv_fail = 0;
    init_candidate(ctx);
    reg_set(ctx, NULL);
}

void sdm_context_free(sdm_context_t *ctx)
{
    if (!ctx) return;
    memset(ctx, 0, sizeof(*ctx));
    reg_set(ctx, NULL);
}

int sdm_context_init_fast(sdm_context_t *ctx, const ntf_filter_t *filter, int trellis_depth, int num_cands,
                          int trellis_lat)
{
    return sdm_context_init(ctx, filter, trellis_depth, num_cands, trellis_lat);
}
size_t sdm_process_block_fast(sdm_context_t *ctx, const double *in, float *out, size_t count)
{
    return run(ctx, in, out, count, false, "sdm_process_block_fast");
}
size_t sdm_drain_fast(sdm_context_t *ctx, float *out, size_t max_out)
{
    return run(ctx, NULL, out, max_out, true, "sdm_drain_fast");
}
void sdm_context_copy_state_fast(sdm_context_t *dst, const sdm_context_t *src) { sdm_context_copy_state(dst, src); }
double sdm_state_distance_fast(const sdm_context_t *a, const sdm_context_t *b) { return sdm_state_distance(a, b); }
void sdm_context_reset_fast(sdm_context_t *ctx) { sdm_context_reset(ctx); }
void sdm_context_free_fast(sdm_context_t *ctx) { sdm_context_free(ctx); }

void sdm_estimate_state_multibit(const ntf_filter_t *filter, const double *init_state, const double *in,
                                 size_t count, double state_limit, int num_levels, double *out_state)
{
    if (!filter || !init_state || !out_state) return;
    if (dsdhip_estimate_state(filter->a, filter->g, filter->order, init_state, in, count, state_limit, num_levels,
                              out_state) != 0) {
        fprintf(stderr, "libdsdhip sdm_estimate_state_multibit: unsupported order %d or levels %d; "
                        "state returned unchanged\n", filter->order, num_levels);
        for (int i = 0; i < filter->order && i < MAX_NTF_ORDER; i++) out_state[i] = init_state[i];
    }
}

void sdm_estimate_state(const ntf_filter_t *filter, const double *init_state, const double *in, size_t count,
                        double state_limit, double *out_state)
{
    sdm_estimate_state_multibit(filter, init_state, in, count, state_limit, 16, out_state);
}

/* ---- NTF table ---- */
const ntf_filter_t *ntf_get_by_index(int index)
{
    return (const ntf_filter_t *)dsdhip_ntf_get(index);
}

int ntf_get_filter_count(void) { return dsdhip_ntf_count(); }

const ntf_filter_t *ntf_get_filter(ntf_filter_id_t id, unsigned dsd_rate)
{
    return ntf_get_by_index(dsdhip_ntf_lookup((int)id, dsd_rate));
}

const ntf_filter_t *ntf_auto_select(unsigned dsd_rate)
{
    return ntf_get_by_index(dsdhip_ntf_auto(dsd_rate));
}

/* src/ntf.c:857-873: clans-7 from DSD128 up, clans-6 below */
const ntf_filter_t *ntf_auto_select_precorr(unsigned dsd_rate)
{
    return ntf_get_filter(dsd_rate >= 5644800u ? NTF_CLANS_7 : NTF_CLANS_6, dsd_rate);
}

int dsdhip_sdm_roundtrip(sdm_context_t *ctx)
{
    const char *why = check_params(ctx);
    if (why) return -1;
    const int P = ctx->filter->order;
    tr_ctx t, snap;
    Outside o;
    if (to_device(ctx, P, &t, &snap, o)) return -1;
    /* wipe what the device form carries, then write it back */
    for (int b = 0; b < 2; b++) {
        for (int j = 0; j < TR_SLOTS; j++) {
            sdm_state_t *nd = &ctx->trellis[b].sdm[j];
            memset(nd->state, 0, sizeof(nd->state));
            nd->cost = 0.0;
            nd->path = 0;
            nd->next = 0;
            nd->hist = 0;
            nd->path_list = NULL;
        }
        memset(ctx->trellis[b].act, 0, sizeof(ctx->trellis[b].act));
    }
    memset(ctx->hist, 0, sizeof(ctx->hist));
    memset(ctx->hist_free, 0, sizeof(ctx->hist_free));
    ctx->hist_fnum = 0;
    from_device(ctx, P, &t, o);
    return 0;
}

}  // extern "C"
