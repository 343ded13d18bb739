/*
 * Thin driver around the reference's own trellis.c / ntf.c, compiled from
 * their sources under /root/reference into oracle/_ref/libdsdref.so by
 * oracle/Makefile.  TEST INFRASTRUCTURE ONLY: used to pin the oracle and to
 * generate the committed fixtures in tests/golden/.
 * It calls only the reference's public API (include/trellis.h:61-118,
 * include/ntf.h:23-36).
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include "trellis.h"
#include "ntf.h"

int ref_ntf_count(void) { return ntf_get_filter_count(); }

int ref_ntf_get(int i, double *a, double *g, int *order, unsigned *freq,
                char *name, int name_len)
{
    const ntf_filter_t *f = ntf_get_by_index(i);
    if (!f) return -1;
    memcpy(a, f->a, sizeof(f->a));
    memcpy(g, f->g, sizeof(f->g));
    *order = f->order;
    *freq = f->freq;
    strncpy(name, f->name, (size_t)name_len - 1);
    name[name_len - 1] = 0;
    return 0;
}

static int index_of(const ntf_filter_t *f)
{
    for (int i = 0; i < ntf_get_filter_count(); i++)
        if (ntf_get_by_index(i) == f) return i;
    return -1;
}

int ref_ntf_lookup(int id, unsigned rate) { return index_of(ntf_get_filter((ntf_filter_id_t)id, rate)); }
int ref_ntf_auto(unsigned rate) { return index_of(ntf_auto_select(rate)); }

void *ref_ctx_new(int filter_index, int depth, int cands, int lat, double state_limit)
{
    sdm_context_t *c = (sdm_context_t *)calloc(1, sizeof(sdm_context_t));
    if (!c) return NULL;
    if (sdm_context_init(c, ntf_get_by_index(filter_index), depth, cands, lat) != 0) {
        free(c);
        return NULL;
    }
    c->state_limit = state_limit;
    return c;
}

void ref_ctx_free(void *c) { sdm_context_free((sdm_context_t *)c); free(c); }
void ref_ctx_reset(void *c) { sdm_context_reset((sdm_context_t *)c); }

size_t ref_process(void *c, const double *in, float *out, size_t n)
{
    return sdm_process_block((sdm_context_t *)c, in, out, n);
}

size_t ref_drain(void *c, float *out, size_t max_out)
{
    return sdm_drain((sdm_context_t *)c, out, max_out);
}

void ref_copy_state(void *dst, const void *src)
{
    sdm_context_copy_state((sdm_context_t *)dst, (const sdm_context_t *)src);
}

void ref_best_state(const void *cv, double *out)
{
    const sdm_context_t *c = (const sdm_context_t *)cv;
    const sdm_state_t *s = c->trellis[c->idx & 1].act[0];
    for (int i = 0; i < MAX_NTF_ORDER; i++) out[i] = s ? s->state[i] : 0.0;
}

/* Seed every active candidate of the current bank with a state vector and
 * zero cost, as the DAS dispatch does (src/dsp_plugin.c:2329-2338). */
void ref_seed(void *cv, const double *st)
{
    sdm_context_t *c = (sdm_context_t *)cv;
    int bank = c->idx & 1;
    int order = c->filter->order;
    for (unsigned i = 0; i < c->num_cands; i++) {
        sdm_state_t *s = c->trellis[bank].act[i];
        if (s) {
            for (int k = 0; k < order; k++) s->state[k] = st[k];
            s->cost = 0.0;
        }
    }
}

void ref_counters(const void *cv, unsigned long long *out)
{
    const sdm_context_t *c = (const sdm_context_t *)cv;
    out[0] = c->num_cands;
    out[1] = c->pos;
    out[2] = c->pending;
    out[3] = c->idx;
    out[4] = c->conv_fail;
    out[5] = c->cands_collapse;
    out[6] = c->next_filter_drops;
    out[7] = c->total_children;
}

void ref_estimate(int filter_index, const double *init, const double *in,
                  size_t count, double state_limit, int levels, double *out)
{
    const ntf_filter_t *f = ntf_get_by_index(filter_index);
    if (levels == 16)
        sdm_estimate_state(f, init, in, count, state_limit, out);
    else
        sdm_estimate_state_multibit(f, init, in, count, state_limit, levels, out);
}

size_t ref_ctx_size(void) { return sizeof(sdm_context_t); }

/* Byte offset of the node the first path-hash bucket points at, -1 if empty
 * (diagnostic used to observe the bucket state after a step). */
long ref_hash0_offset(const void *cv)
{
    const sdm_context_t *c = (const sdm_context_t *)cv;
    if (!c->path_hash[0]) return -1;
    return (long)((const char *)c->path_hash[0] - (const char *)c);
}

/* Layout probe: sizes and field offsets of the context types, for checking
 * a restatement of the layout. out: sizeof state/trellis/context, then the
 * offsets of act, path_hash, hist_free, hist_fnum, trellis_mask, num_cands,
 * idx, filter, prev_y, conv_fail, total_children, state_limit, hist, and of
 * sdm_state_t's cost, path, next, hist, parent, path_list. */
void ref_layout(long *out)
{
    int k = 0;
    out[k++] = (long)sizeof(sdm_state_t);
    out[k++] = (long)sizeof(sdm_trellis_t);
    out[k++] = (long)sizeof(sdm_context_t);
    out[k++] = (long)offsetof(sdm_trellis_t, act);
    out[k++] = (long)offsetof(sdm_context_t, path_hash);
    out[k++] = (long)offsetof(sdm_context_t, hist_free);
    out[k++] = (long)offsetof(sdm_context_t, hist_fnum);
    out[k++] = (long)offsetof(sdm_context_t, trellis_mask);
    out[k++] = (long)offsetof(sdm_context_t, num_cands);
    out[k++] = (long)offsetof(sdm_context_t, idx);
    out[k++] = (long)offsetof(sdm_context_t, filter);
    out[k++] = (long)offsetof(sdm_context_t, prev_y);
    out[k++] = (long)offsetof(sdm_context_t, conv_fail);
    out[k++] = (long)offsetof(sdm_context_t, total_children);
    out[k++] = (long)offsetof(sdm_context_t, state_limit);
    out[k++] = (long)offsetof(sdm_context_t, hist);
    out[k++] = (long)offsetof(sdm_state_t, cost);
    out[k++] = (long)offsetof(sdm_state_t, path);
    out[k++] = (long)offsetof(sdm_state_t, next);
    out[k++] = (long)offsetof(sdm_state_t, hist);
    out[k++] = (long)offsetof(sdm_state_t, parent);
    out[k++] = (long)offsetof(sdm_state_t, path_list);
}

/* What a context carries into its next steps, flattened:
 *   meta[8]   num_cands, pos, pending, draining, idx & 1, trellis_num,
 *             trellis_lat, hist_fnum
 *   per survivor c < num_cands (act[c] of the active bank):
 *             st[c*10 + k], cost[c], path[c], next[c], slot[c] (node index),
 *             hist bits hist[c*256 + byte] (trellis_lat bits)
 *   per node j < 64 of bank b (index b*64 + j): path, cost, and link =
 *             -1 none, 0..63 node of the same bank, 1000 + 128*b' + j' node
 *             j' of bank b' of `other`, -2 anything else. */
void ref_export(const void *cv, const void *other, int *meta, double *st, double *cost, unsigned *path,
                unsigned char *next, int *slot, unsigned char *hist, unsigned *npath, double *ncost, int *nlink)
{
    const sdm_context_t *c = (const sdm_context_t *)cv;
    const sdm_context_t *o = (const sdm_context_t *)other;
    const int bank = c->idx & 1;
    meta[0] = (int)c->num_cands;
    meta[1] = (int)c->pos;
    meta[2] = (int)c->pending;
    meta[3] = (int)c->draining;
    meta[4] = bank;
    meta[5] = (int)c->trellis_num;
    meta[6] = (int)c->trellis_lat;
    meta[7] = (int)c->hist_fnum;
    for (unsigned i = 0; i < c->num_cands && i < SDM_TRELLIS_MAX_NUM; i++) {
        const sdm_state_t *s = c->trellis[bank].act[i];
        for (int k = 0; k < MAX_NTF_ORDER; k++) st[i * 10 + k] = s->state[k];
        cost[i] = s->cost;
        path[i] = s->path;
        next[i] = s->next;
        slot[i] = (int)(s - &c->trellis[bank].sdm[0]);
        memcpy(hist + i * 256, c->hist[s->hist], (c->trellis_lat + 7) / 8);
    }
    for (int b = 0; b < 2; b++)
        for (int j = 0; j < 64; j++) {
            const sdm_state_t *s = &c->trellis[b].sdm[j];
            const sdm_state_t *p = s->path_list;
            int l = -2;
            if (!p) l = -1;
            else if (p >= &c->trellis[b].sdm[0] && p < &c->trellis[b].sdm[64]) l = (int)(p - &c->trellis[b].sdm[0]);
            else if (o) {
                for (int b2 = 0; b2 < 2; b2++)
                    if (p >= &o->trellis[b2].sdm[0] && p < &o->trellis[b2].sdm[128])
                        l = 1000 + 128 * b2 + (int)(p - &o->trellis[b2].sdm[0]);
            }
            npath[b * 64 + j] = s->path;
            ncost[b * 64 + j] = s->cost;
            nlink[b * 64 + j] = l;
        }
}

/* Raw bytes of a context (size ref_ctx_size()), to load it into another
 * implementation of the same layout, and back. */
void ref_ctx_read(const void *cv, void *dst) { memcpy(dst, cv, sizeof(sdm_context_t)); }
void ref_ctx_write(void *cv, const void *src) { memcpy(cv, src, sizeof(sdm_context_t)); }
