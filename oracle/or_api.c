/*
 * Flat entry points of the oracle for ctypes (same shapes as ref_driver.c),
 * so tests drive the oracle and the compiled reference identically.
 * TEST INFRASTRUCTURE ONLY.
 */
#include "dsd_oracle.h"
#include <stdlib.h>
#include <string.h>

int ora_ntf_count(void) { return or_ntf_count(); }

int ora_ntf_get(int i, double *a, double *g, int *order, unsigned *freq,
                char *name, int name_len)
{
    const or_ntf_t *f = or_ntf_get(i);
    if (!f) return -1;
    memcpy(a, f->a, sizeof(f->a));
    memcpy(g, f->g, sizeof(f->g));
    *order = f->order;
    *freq = f->freq;
    strncpy(name, f->name, (size_t)name_len - 1);
    name[name_len - 1] = 0;
    return 0;
}

int ora_ntf_lookup(int id, unsigned rate) { return or_ntf_lookup(id, rate); }
int ora_ntf_auto(unsigned rate) { return or_ntf_auto(rate); }

void *ora_ctx_new(int filter_index, int depth, int cands, int lat, double state_limit)
{
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    if (!c) return NULL;
    if (or_sdm_init(c, or_ntf_get(filter_index), depth, cands, lat) != 0) {
        free(c);
        return NULL;
    }
    c->state_limit = state_limit;
    return c;
}

void ora_ctx_free(void *c) { free(c); }
void ora_ctx_reset(void *c) { or_sdm_reset((or_ctx *)c); }

size_t ora_process(void *c, const double *in, float *out, size_t n)
{
    return or_sdm_process((or_ctx *)c, in, out, n);
}

size_t ora_drain(void *c, float *out, size_t max_out)
{
    return or_sdm_drain((or_ctx *)c, out, max_out);
}

void ora_copy_state(void *dst, const void *src)
{
    or_sdm_copy_state((or_ctx *)dst, (const or_ctx *)src);
}

void ora_best_state(const void *c, double *out) { or_sdm_best_state((const or_ctx *)c, out); }

void ora_seed(void *cv, const double *st)
{
    or_ctx *c = (or_ctx *)cv;
    int bank = c->idx & 1;
    for (unsigned i = 0; i < c->num_cands; i++) {
        or_node *s = c->act[bank][i];
        if (s) {
            for (int k = 0; k < c->f->order; k++) s->st[k] = st[k];
            s->cost = 0.0;
        }
    }
}

void ora_counters(const void *cv, unsigned long long *out)
{
    const or_ctx *c = (const or_ctx *)cv;
    out[0] = c->num_cands;
    out[1] = c->pos;
    out[2] = c->pending;
    out[3] = c->idx;
    out[4] = c->conv_fail;
    out[5] = c->cands_collapse;
    out[6] = c->next_filter_drops;
    out[7] = c->total_children;
}

void ora_estimate(int filter_index, const double *init, const double *in,
                  size_t count, double state_limit, int levels, double *out)
{
    or_estimate_state(or_ntf_get(filter_index), init, in, count, state_limit, levels, out);
}

size_t ora_segment(void *c, const double *in, float *out, size_t count, size_t discard)
{
    return or_sdm_segment((or_ctx *)c, in, out, count, discard);
}

size_t ora_ctx_size(void) { return sizeof(or_ctx); }

/* (bank * 256 + slot) of the node the first hash bucket points at, -1 if
 * empty; matches ref_hash0_offset decoded with the reference's layout. */
long ora_hash0_slot(const void *cv)
{
    const or_ctx *c = (const or_ctx *)cv;
    const or_node *t = c->head[0];
    if (!t) return -1;
    for (int b = 0; b < 2; b++)
        if (t >= &c->node[b][0] && t < &c->node[b][OR_NODES])
            return b * 256 + (long)(t - &c->node[b][0]);
    return -2;   /* points into another context */
}

/* DAS over nch channels of F samples (channel-major) on fresh contexts of
 * the given parameters; returns assembled length, fills out (nch*F floats)
 * and stitch positions (nseg ints).  plan_out receives nom_start/nom_size
 * (2*nseg) if non-NULL.  seg_len 0 => nseg equal parts. */
long ora_das(int filter_index, int depth, int cands, int lat, const double *in, int nch,
             size_t F, size_t seg_len, int nseg, int seg0_first, int reposition,
             float *out, int *stitch_pos, long *plan_out, int max_seg)
{
    or_das_plan_t p;
    int S = or_das_plan_init(&p, F, seg_len, nseg, (size_t)64 * (size_t)lat, seg0_first, reposition);
    if (S > max_seg) { or_das_plan_free(&p); return -2; }
    if (reposition) or_das_reposition(&p, in, F);
    or_ctx *pc = (or_ctx *)calloc((size_t)nch, sizeof(or_ctx));
    for (int ch = 0; ch < nch; ch++)
        or_sdm_init(&pc[ch], or_ntf_get(filter_index), depth, cands, lat);
    size_t n = or_das_run(&p, pc, nch, in, F, out, stitch_pos);
    if (plan_out)
        for (int s = 0; s < S; s++) {
            plan_out[2 * s] = (long)p.nom_start[s];
            plan_out[2 * s + 1] = (long)p.nom_size[s];
        }
    free(pc);
    or_das_plan_free(&p);
    return (long)n;
}

void ora_das_set_threads(int n) { or_das_set_threads(n); }

int ora_das_nseg(size_t F, size_t seg_len, int nseg, int lat)
{
    or_das_plan_t p;
    int S = or_das_plan_init(&p, F, seg_len, nseg, (size_t)64 * (size_t)lat, 1, 0);
    or_das_plan_free(&p);
    return S;
}

/* intermediates of one stitch search: ints[0..5] = half_w, peak,
 * peak_pos, scanned, ncand, best; cand_i[3][33] = pos, run, density;
 * cand_d[2][33] = quality, largest decoded difference; dm[ovl] */
int ora_stitch_detail(const float *prev, const float *cur, int ovl, int lat, int *ints, int *cand_i,
                      double *cand_d, int *dm)
{
    or_stitch_detail_t det;
    memset(&det, 0, sizeof(det));
    det.dm = dm;
    const int pos = or_das_stitch_ex(prev, cur, ovl, lat, &det);
    ints[0] = det.half_w;
    ints[1] = det.peak;
    ints[2] = det.peak_pos;
    ints[3] = det.scanned;
    ints[4] = det.ncand;
    ints[5] = det.best;
    for (int i = 0; i < det.ncand; i++) {
        cand_i[i] = det.pos[i];
        cand_i[33 + i] = det.run[i];
        cand_i[66 + i] = det.density[i];
        cand_d[i] = det.quality[i];
        cand_d[33 + i] = det.max_decoded[i];
    }
    return pos;
}

int ora_pick(int n, const double *quality, const int *density) { return or_das_pick(n, quality, density); }

/* reposition of a plan over channel-0 data: starts/sizes in/out */
void ora_reposition(size_t F, size_t seg_len, int lat, const double *ch0, long *start, long *size, int *nseg)
{
    or_das_plan_t p;
    int S = or_das_plan_init(&p, F, seg_len, 0, (size_t)64 * (size_t)lat, 1, 1);
    or_das_reposition(&p, ch0, F);
    for (int s = 0; s < S; s++) {
        start[s] = (long)p.nom_start[s];
        size[s] = (long)p.nom_size[s];
    }
    *nseg = S;
    or_das_plan_free(&p);
}

int ora_stitch(const float *prev, const float *cur, int ovl, int lat, double *q, int *nc)
{
    return or_das_stitch(prev, cur, ovl, lat, q, nc);
}

size_t ora_fir_chain(unsigned fs_in, unsigned fs_out, const float *in, size_t n, double gain,
                     double *out)
{
    return or_fir_chain(fs_in, fs_out, in, n, gain, out);
}

size_t ora_fir_chain_ld(unsigned fs_in, unsigned fs_out, const float *in, size_t n, double gain,
                        double *out)
{
    size_t cnt = n;
    int nt[OR_FIR_MAX_STAGES];
    int st = or_fir_plan(fs_in, fs_out, nt);
    if (st < 0) return 0;
    cnt = n << st;
    long double *tmp = (long double *)malloc(cnt * sizeof(long double));
    size_t r = or_fir_chain_ld(fs_in, fs_out, in, n, gain, tmp);
    for (size_t i = 0; i < r; i++) out[i] = (double)tmp[i];
    free(tmp);
    return r;
}

void ora_halfband(double *h, int ntaps) { or_design_halfband(h, ntaps, 12.0); }

/* DAS over a stream of chunks with persistent contexts carried between
 * calls (the plugin's per-chunk DAS, src/dsp_plugin.c:1721-2761). */
typedef struct {
    or_ctx *pc;
    int nch, lat;
} ora_das_stream_t;

void *ora_das_stream_new(int filter_index, int depth, int cands, int lat, int nch)
{
    ora_das_stream_t *h = (ora_das_stream_t *)calloc(1, sizeof(*h));
    h->pc = (or_ctx *)calloc((size_t)nch, sizeof(or_ctx));
    h->nch = nch;
    h->lat = lat;
    for (int ch = 0; ch < nch; ch++)
        if (or_sdm_init(&h->pc[ch], or_ntf_get(filter_index), depth, cands, lat) != 0) {
            free(h->pc);
            free(h);
            return NULL;
        }
    return h;
}

long ora_das_stream_chunk_n(void *hv, const double *in, size_t F, size_t seg_len, int nseg, int seg0_first,
                            int reposition, float *out)
{
    ora_das_stream_t *h = (ora_das_stream_t *)hv;
    or_das_plan_t p;
    or_das_plan_init(&p, F, seg_len, nseg, (size_t)64 * (size_t)h->lat, seg0_first, reposition);
    if (reposition) or_das_reposition(&p, in, F);
    size_t n = or_das_run(&p, h->pc, h->nch, in, F, out, NULL);
    or_das_plan_free(&p);
    return (long)n;
}

long ora_das_stream_chunk(void *hv, const double *in, size_t F, size_t seg_len, int seg0_first,
                          int reposition, float *out)
{
    ora_das_stream_t *h = (ora_das_stream_t *)hv;
    or_das_plan_t p;
    or_das_plan_init(&p, F, seg_len, 0, (size_t)64 * (size_t)h->lat, seg0_first, reposition);
    if (reposition) or_das_reposition(&p, in, F);
    size_t n = or_das_run(&p, h->pc, h->nch, in, F, out, NULL);
    or_das_plan_free(&p);
    return (long)n;
}

void ora_das_stream_free(void *hv)
{
    ora_das_stream_t *h = (ora_das_stream_t *)hv;
    if (!h) return;
    free(h->pc);
    free(h);
}

/* Select the order-8 FMA filter for contexts created from now on. */
void ora_set_fma_o8(int on) { or_fma_o8_default = on ? 1 : 0; }
