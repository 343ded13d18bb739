/*
 * Oracle restatement of the trellis SDM core.  TEST INFRASTRUCTURE ONLY
 * (see dsd_oracle.h).  Must be compiled with -ffp-contract=off.
 *
 * Follows src/trellis.c of the reference:
 *   loop filter     src/trellis.c:37-151 (scalar order), children :153-191
 *   step            src/trellis.c:735-746
 *   sort/prune      src/trellis.c:582-731
 *   per sample      src/trellis.c:868-973 (scalar branch, 1-bit)
 *   init/process    src/trellis.c:977-1062, copy :1064-1092, reset :1189-1224
 *   estimator       src/trellis.c:1118-1187
 * History buffers are carried by value per survivor: in the reference every
 * survivor either keeps its parent's buffer or receives a full copy of it
 * before writing its own bit at `pos`, so its contents are always
 * "parent history with bit pos replaced".
 */
#include "dsd_oracle.h"
#include <string.h>
#include <math.h>

static const or_ntf_t g_ntf[] = {
#include "../foo_dsd_trellis_amd/csrc/ntf_table.inc"
};

int or_ntf_count(void) { return (int)(sizeof(g_ntf) / sizeof(g_ntf[0])); }

const or_ntf_t *or_ntf_get(int index)
{
    if (index < 0 || index >= or_ntf_count()) return NULL;
    return &g_ntf[index];
}

/* NTF id -> name, ids as include/dsd_types.h:57-68. */
static const char *const k_ntf_names[10] = {
    "clans-4", "sdm-4", "clans-5", "sdm-5", "clans-6",
    "sdm-6", "clans-7", "sdm-7", "clans-8", "sdm-8"};

int or_ntf_lookup(int id, unsigned rate)
{
    if (id < 0 || id >= 10) return -1;
    for (int i = 0; i < or_ntf_count(); i++)
        if (strcmp(g_ntf[i].name, k_ntf_names[id]) == 0 && g_ntf[i].freq <= rate)
            return i;
    return -1;
}

int or_ntf_auto(unsigned rate)
{
    int id;
    if (rate >= 11289600u) id = 8;        /* DSD256 and up: clans-8 */
    else if (rate >= 5644800u) id = 4;    /* DSD128: clans-6 */
    else id = 2;                          /* DSD64: clans-5 */
    return or_ntf_lookup(id, rate);
}

/* ---- cost ordering: signed 64-bit compare of the IEEE bits (:582-597) ---- */
static inline int64_t bits64(double d) { int64_t i; memcpy(&i, &d, 8); return i; }
static inline int cost_lt(double a, double b) { return bits64(a) < bits64(b); }
static inline int cost_le(double a, double b) { return bits64(a) <= bits64(b); }

/* Loop filter with y = 0: writes d[] and returns v (src/trellis.c:37-151).
 * Addition order is strictly left to right. */
static double filt(const double *s, double *d, const or_ntf_t *f, double x)
{
    const double *a = f->a, *g = f->g;
    int P = f->order, i;
    double v;
    d[0] = s[0] - g[0] * s[1] + x - 0.0;
    v = x + a[0] * d[0];
    for (i = 1; i < P - 1; i++) {
        d[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        v = v + a[i] * d[i];
    }
    d[i] = s[i] + s[i - 1];
    v = v + a[i] * d[i];
    return v;
}

/* The order-8 filter of the reference's 4-candidate AVX2 step
 * (src/trellis.c:395-421): each d[i] fuses its g term, v accumulates by
 * FMA in ascending i. */
static double filt_fma8(const double *s, double *d, const or_ntf_t *f, double x)
{
    const double *a = f->a, *g = f->g;
    int i;
    d[0] = fma(-g[0], s[1], s[0] + x);
    for (i = 1; i < 7; i++) d[i] = fma(-g[i], s[i + 1], s[i] + s[i - 1]);
    d[7] = s[7] + s[6];
    double v = fma(a[0], d[0], x);
    for (i = 1; i < 8; i++) v = fma(a[i], d[i], v);
    return v;
}

int or_fma_o8_default = 0;

static void clamp_state(double *s, int P, double lim)
{
    for (int i = 0; i < P; i++) {
        if (s[i] > lim) s[i] = lim;
        else if (s[i] < -lim) s[i] = -lim;
    }
}

int or_sdm_init(or_ctx *c, const or_ntf_t *f, int depth, int cands, int lat)
{
    memset(c, 0, sizeof(*c));
    if (!f) return -1;
    if (depth > 32 || cands > OR_MAXK || lat > OR_MAXL) return -1;
    if (depth < 1 || cands < 1 || lat < 1) return -1;
    c->f = f;
    c->fma_o8 = or_fma_o8_default;
    c->K = (unsigned)cands;
    c->L = (unsigned)lat;
    c->mask = (uint32_t)(((uint64_t)1 << depth) - 1);
    c->num_cands = 1;
    c->act[0][0] = &c->node[0][0];
    return 0;
}

void or_sdm_reset(or_ctx *c)
{
    if (!c->f) return;
    memset(c->node, 0, sizeof(c->node));
    memset(c->act, 0, sizeof(c->act));
    memset(c->head, 0, sizeof(c->head));
    memset(c->hist, 0, sizeof(c->hist));
    c->num_cands = 1;
    c->pos = c->pending = c->draining = c->idx = 0;
    c->conv_fail = 0;
    c->act[0][0] = &c->node[0][0];
}

/* Optional observer of every prune step (tests use it to compare a
 * parallel formulation with this sequential one); NULL in normal use. */
or_prune_observer_fn or_prune_observer = NULL;
static uint64_t g_reg_mask;

/* Path dedup lookup (src/trellis.c:599-619): first chain entry with an equal
 * path, else register s at the head of its bucket. */
static or_node *lookup(or_ctx *c, or_node *s)
{
    unsigned b = s->path & (OR_HASH - 1);
    for (or_node *t = c->head[b]; t; t = t->plist)
        if (t->path == s->path) return t;
    s->plist = c->head[b];
    c->head[b] = s;
    if (or_prune_observer) {
        int bank = (s >= &c->node[1][0] && s < &c->node[1][OR_NODES]) ? 1 : 0;
        g_reg_mask |= 1ull << (s - &c->node[bank][0]);
    }
    return NULL;
}

/* Write act slot j of `bank`.  The reference's act[] holds 32 entries; its
 * un-guarded insertion writes act[32] when the list is full, which lands on
 * the first path-hash bucket for bank 1 (the next struct member) and on a
 * bank-0 node state that is rewritten before it is read for bank 0. */
static inline void act_write(or_ctx *c, int bank, unsigned j, or_node *v)
{
    if (j == OR_MAXK) {
        if (bank == 1) c->head[0] = v;
        return;
    }
    c->act[bank][j] = v;
}

/* Candidate sort/prune (src/trellis.c:621-731). */
static unsigned sort_cands(or_ctx *c, int bank)
{
    or_node *nd = c->node[bank];
    or_node **act = c->act[bank];
    unsigned m = 2 * c->num_cands, i, j, n;
    unsigned K = c->K;
    or_node *mn = NULL, *s, *t, *r;

    for (i = 0; i < m; i++) {
        s = &nd[i];
        c->head[s->path & (OR_HASH - 1)] = NULL;
        if (!mn || cost_lt(s->cost, mn->cost)) mn = s;
    }
    c->total_children += m;

    unsigned votes1 = 0;
    for (i = 0; i < m; i++) votes1 += nd[i].next & 1;
    unsigned maj = (votes1 > m - votes1) ? 1u : 0u;
    if (mn->next != maj) {
        or_node *bm = NULL;
        for (i = 0; i < m; i++) {
            s = &nd[i];
            if (s->next == maj && (!bm || cost_lt(s->cost, bm->cost))) bm = s;
        }
        if (bm && bm->cost < mn->cost * 1.1) mn = bm;
    }

    for (i = 0, n = 0; i < m; i++) {
        s = &nd[i];
        if (s->next != mn->next) { c->next_filter_drops++; continue; }
        if (n == K && cost_le(act[n - 1]->cost, s->cost)) continue;

        t = lookup(c, s);
        if (!t) {
            for (j = n; j > 0; j--) {
                r = act[j - 1];
                if (cost_le(r->cost, s->cost)) break;
                act_write(c, bank, j, r);
            }
            if (j < K) act[j] = s;
            if (n < K) n++;
            continue;
        }
        if (cost_le(t->cost, s->cost)) continue;

        unsigned tp;
        for (tp = 0; tp < n; tp++)
            if (act[tp] == t) break;
        if (tp == n) {
            /* t is no longer listed: insert s as a new entry (guarded). */
            for (j = n; j > 0; j--) {
                r = act[j - 1];
                if (cost_le(r->cost, s->cost)) break;
                if (j < K) act[j] = r;
            }
            if (j < K) act[j] = s;
            if (n < K) n++;
            continue;
        }
        for (j = tp; j + 1 < n; j++) act[j] = act[j + 1];
        n--;
        for (j = n; j > 0; j--) {
            r = act[j - 1];
            if (cost_le(r->cost, s->cost)) break;
            if (j < K) act[j] = r;
        }
        if (j < K) act[j] = s;
        n++;
    }
    return n;
}

static inline unsigned hist_get(const uint64_t *h, unsigned i)
{
    return (unsigned)(h[i >> 6] >> (i & 63)) & 1u;
}

static inline void hist_put(uint64_t *h, unsigned i, unsigned v)
{
    h[i >> 6] = (h[i >> 6] & ~(1ull << (i & 63))) | ((uint64_t)v << (i & 63));
}

int or_sdm_sample(or_ctx *c, double x)
{
    int cur = (int)c->idx, nxt = cur ^ 1;
    const or_ntf_t *f = c->f;
    int P = f->order;
    unsigned next_pos = c->pos + 1;
    if (next_pos == c->L) next_pos = 0;
    unsigned hw = (c->L + 63) / 64;

    /* the AVX2 build takes candidates four at a time while four remain
     * (src/trellis.c:904-927) */
    const unsigned nfma = (c->fma_o8 && P == 8) ? (c->num_cands & ~3u) : 0;
    for (unsigned i = 0; i < c->num_cands; i++) {
        or_node *p = c->act[cur][i];
        or_node *ch = &c->node[nxt][2 * i];
        double v = i < nfma ? filt_fma8(p->st, ch[0].st, f, x) : filt(p->st, ch[0].st, f, x);
        for (int k = 0; k < P; k++) ch[1].st[k] = ch[0].st[k];
        ch[0].st[0] += 1.0;
        ch[1].st[0] -= 1.0;
        if (c->state_limit > 0.0) {
            clamp_state(ch[0].st, P, c->state_limit);
            clamp_state(ch[1].st, P, c->state_limit);
        }
        double e0 = v + f->a[0], e1 = v - f->a[0];
        ch[0].cost = p->cost + e0 * e0;
        ch[1].cost = p->cost + e1 * e1;
        for (int b = 0; b < 2; b++) {
            ch[b].path = ((p->path << 1) | (uint32_t)b) & c->mask;
            ch[b].next = p->next;
            ch[b].parent = (uint8_t)i;
        }
        c->next_new[i] = (uint8_t)hist_get(c->hist[cur][i], next_pos);
        p->next = c->next_new[i];
    }

    double obs_cost[64];
    if (or_prune_observer) {
        for (unsigned i = 0; i < 2 * c->num_cands; i++) obs_cost[i] = c->node[nxt][i].cost;
        g_reg_mask = 0;
    }
    unsigned nn = sort_cands(c, nxt);
    if (or_prune_observer) {
        or_prune_obs o;
        o.m = (int)(2 * c->num_cands);
        o.K = (int)c->K;
        o.bank = nxt;
        o.n = (int)nn;
        o.reg = g_reg_mask;
        for (int i = 0; i < o.m; i++) {
            or_node *q = &c->node[nxt][i];
            o.cost[i] = obs_cost[i];
            o.path[i] = q->path;
            o.next[i] = q->next;
            o.plist[i] = -2;
            if ((g_reg_mask >> i) & 1) {
                if (!q->plist) o.plist[i] = -1;
                else if (q->plist >= &c->node[nxt][0] && q->plist < &c->node[nxt][OR_NODES])
                    o.plist[i] = (int)(q->plist - &c->node[nxt][0]);
                else o.plist[i] = -3;
            }
        }
        for (unsigned k = 0; k < nn; k++) o.act[k] = (int)(c->act[nxt][k] - &c->node[nxt][0]);
        or_prune_observer(&o);
    }
    double min_cost = c->act[nxt][0]->cost;
    int out = c->act[nxt][0]->next;

    for (unsigned k = 0; k < nn; k++) {
        or_node *s = c->act[nxt][k];
        uint64_t *dst = c->hist[nxt][k];
        memcpy(dst, c->hist[cur][s->parent], hw * sizeof(uint64_t));
        s->cost -= min_cost;
        s->next = c->next_new[s->parent];
        hist_put(dst, c->pos, s->path & 1);
    }
    if (nn < c->num_cands) { c->conv_fail++; c->cands_collapse++; }
    c->num_cands = nn;
    c->pos = next_pos;
    c->idx ^= 1;
    return out;
}

size_t or_sdm_process(or_ctx *c, const double *in, float *out, size_t count)
{
    size_t len = count, o = 0;
    if (c->pending < c->L) {
        size_t pre = c->L - c->pending;
        if (pre > len) pre = len;
        c->pending += (unsigned)pre;
        len -= pre;
        while (pre--) or_sdm_sample(c, *in++);
    }
    while (len--) {
        int b = or_sdm_sample(c, *in++);
        if (out) out[o] = b ? 1.0f : -1.0f;
        o++;
    }
    return o;
}

size_t or_sdm_drain(or_ctx *c, float *out, size_t max_out)
{
    size_t len = c->pending;
    if (len > max_out) len = max_out;
    if (!c->draining && c->pending < c->L) {
        unsigned flush = c->L - c->pending;
        while (flush--) or_sdm_sample(c, 0.0);
    }
    c->draining = 1;
    c->pending -= (unsigned)len;
    for (size_t i = 0; i < len; i++) out[i] = or_sdm_sample(c, 0.0) ? 1.0f : -1.0f;
    return len;
}

/* Deep copy (src/trellis.c:1064-1092): act lists are re-based onto dst's own
 * nodes; node hash links keep pointing where they pointed (into src), the
 * hash heads are cleared and the diagnostic counters zeroed. */
void or_sdm_copy_state(or_ctx *dst, const or_ctx *src)
{
    memcpy(dst, src, sizeof(*dst));
    for (int bank = 0; bank < 2; bank++) {
        unsigned nc = dst->num_cands;
        if (nc > OR_MAXK) nc = OR_MAXK;
        for (unsigned i = 0; i < nc; i++) {
            if (src->act[bank][i])
                dst->act[bank][i] = &dst->node[bank][0] + (src->act[bank][i] - &src->node[bank][0]);
        }
    }
    memset(dst->head, 0, sizeof(dst->head));
    dst->conv_fail = dst->cands_collapse = dst->next_filter_drops = dst->total_children = 0;
}

void or_sdm_best_state(const or_ctx *c, double *out)
{
    const or_node *s = c->act[c->idx & 1][0];
    for (int i = 0; i < OR_MAX_ORDER; i++) out[i] = s ? s->st[i] : 0.0;
}

/* Segment wrapper (src/engine.c:915-938): the first trellis_lat + discard
 * inputs are run with their outputs thrown away. */
size_t or_sdm_segment(or_ctx *c, const double *in, float *out, size_t count,
                      size_t discard)
{
    if (discard == 0) return or_sdm_process(c, in, out, count);
    size_t warm = (size_t)c->L + discard;
    if (warm > count) {
        or_sdm_process(c, in, NULL, count);
        return 0;
    }
    or_sdm_process(c, in, NULL, warm);
    return or_sdm_process(c, in + warm, out, count - warm);
}

/* Greedy multi-level estimator (src/trellis.c:1118-1187). */
void or_estimate_state(const or_ntf_t *f, const double *init, const double *in,
                       size_t count, double state_limit, int levels,
                       double *out)
{
    const int P = f->order;
    const double *a = f->a, *g = f->g;
    double s[OR_MAX_ORDER], ns[OR_MAX_ORDER];
    const double step = 2.0 / (double)(levels - 1);
    for (int i = 0; i < P; i++) s[i] = init[i];
    for (size_t n = 0; n < count; n++) {
        double x = in[n];
        ns[0] = s[0] - g[0] * s[1] + x;
        double v = x + a[0] * ns[0];
        for (int i = 1; i < P - 1; i++) {
            ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
            v += a[i] * ns[i];
        }
        ns[P - 1] = s[P - 1] + s[P - 2];
        v += a[P - 1] * ns[P - 1];
        double y;
        if (levels <= 2) {
            y = (v * a[0] > 0.0) ? 1.0 : -1.0;
        } else {
            double yi = v / a[0];
            double q = (yi + 1.0) / step + 0.5;
            /* C (int) conversion as executed on x86-64 (cvttsd2si): values
             * outside the int range and NaN yield INT_MIN. */
            int idx;
            if (q != q || q >= 2147483648.0 || q <= -2147483649.0) idx = (int)0x80000000u;
            else idx = (int)q;
            if (idx < 0) idx = 0;
            if (idx > levels - 1) idx = levels - 1;
            y = -1.0 + (double)idx * step;
        }
        ns[0] -= y;
        if (state_limit > 0.0) clamp_state(ns, P, state_limit);
        for (int i = 0; i < P; i++) s[i] = ns[i];
    }
    for (int i = 0; i < P; i++) out[i] = s[i];
}
