/*
 * Thin driver around the reference's own precorr.c, compiled from its
 * sources under /root/reference into oracle/_ref/libdsdref.so by
 * oracle/Makefile.  TEST INFRASTRUCTURE ONLY: pins the oracle's PreCorr.
 * It calls only the reference's public API (include/precorr.h:40-52).
 */
#include <stdlib.h>
#include <string.h>
#include "precorr.h"

void *ref_pc_new(int filter_index, float state_limit)
{
    precorr_context_t *c = (precorr_context_t *)calloc(1, sizeof(*c));
    if (!c || precorr_context_init(c, ntf_get_by_index(filter_index)) != 0) {
        free(c);
        return NULL;
    }
    c->state_limit = state_limit;
    return c;
}
void ref_pc_free(void *c) { free(c); }
void ref_pc_reset(void *c) { precorr_context_reset((precorr_context_t *)c); }
size_t ref_pc_process(void *c, const double *in, float *out, size_t n)
{
    return precorr_process_block((precorr_context_t *)c, in, out, n);
}
void ref_pc_table(const void *c, float *t) { memcpy(t, ((const precorr_context_t *)c)->pred_table, 256 * 8 * sizeof(float)); }
void ref_pc_state(const void *cv, float *st, int *hist, int *phase, float *prev_y)
{
    const precorr_context_t *c = (const precorr_context_t *)cv;
    memcpy(st, c->state, 8 * sizeof(float));
    *hist = c->history;
    *phase = c->phase;
    *prev_y = c->prev_y;
}
