/*
 * Oracle restatement of the PreCorr SDM.  TEST INFRASTRUCTURE ONLY.
 * Follows src/precorr.c: float loop filter :20-120 (left-to-right sums,
 * no contraction), LCG :124-135, table training :139-186, init :190-208,
 * per-sample quantize + correction :212-262, reset :269-278.  Pinned
 * against the reference's own precorr.c compiled into oracle/_ref.
 */
#include <string.h>

#include "dsd_oracle.h"

static float pc_filt(const float *s, float *d, const float *a, const float *g, int order, float x, float y)
{
    d[0] = s[0] - g[0] * s[1] + x - y;
    float v = x + a[0] * d[0];
    int i;
    for (i = 1; i < order - 1; i++) {
        d[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        v += a[i] * d[i];
    }
    d[i] = s[i] + s[i - 1];
    v += a[i] * d[i];
    return v;
}

int or_pc_init(or_pc *c, int filter_index, float state_limit)
{
    const or_ntf_t *f = or_ntf_get(filter_index);
    if (!f) return -1;
    memset(c, 0, sizeof(*c));
    c->order = f->order;
    for (int i = 0; i < f->order; i++) {
        c->a[i] = (float)f->a[i];
        c->g[i] = (float)f->g[i];
    }
    c->state_limit = state_limit;
    /* training */
    static float accum[256][8];
    static int hits[256][8];
    memset(accum, 0, sizeof(accum));
    memset(hits, 0, sizeof(hits));
    float st[OR_MAX_ORDER] = {0}, d[OR_MAX_ORDER];
    unsigned hist = 0x69;
    int ph = 0;
    float prev = 0.0f;
    uint32_t rng = 0xDEADBEEFu;
    for (int i = 0; i < 65536; i++) {
        rng = rng * 1664525u + 1013904223u;
        float x = 0.1f * ((float)(int32_t)rng / 2147483648.0f);
        float v = pc_filt(st, d, c->a, c->g, c->order, x, prev);
        float y = (v >= 0.0f) ? 1.0f : -1.0f;
        memcpy(st, d, (size_t)c->order * sizeof(float));
        accum[hist][ph] += y - v;
        hits[hist][ph]++;
        hist = ((hist << 1) | (y > 0.0f ? 1u : 0u)) & 255u;
        ph = (ph + 1) & 7;
        prev = y;
    }
    for (int h = 0; h < 256; h++)
        for (int p = 0; p < 8; p++)
            c->table[h][p] = hits[h][p] >= 4 ? accum[h][p] / (float)hits[h][p] : 0.0f;
    or_pc_reset(c);
    return 0;
}

void or_pc_reset(or_pc *c)
{
    memset(c->state, 0, sizeof(c->state));
    c->history = 0x69;
    c->phase = 0;
    c->prev_y = 0.0f;
}

size_t or_pc_process(or_pc *c, const double *in, float *out, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        float x = (float)in[i];
        float d[OR_MAX_ORDER];
        float v = pc_filt(c->state, d, c->a, c->g, c->order, x, c->prev_y);
        float y = (v >= 0.0f) ? 1.0f : -1.0f;
        if (c->state_limit > 0.0f)
            for (int k = 0; k < c->order; k++) {
                if (d[k] > c->state_limit) d[k] = c->state_limit;
                else if (d[k] < -c->state_limit) d[k] = -c->state_limit;
            }
        memcpy(c->state, d, (size_t)c->order * sizeof(float));
        float yc = y - c->table[c->history][c->phase];
        float oy = (yc >= 0.0f) ? 1.0f : -1.0f;
        out[i] = oy;
        c->history = ((c->history << 1) | (oy > 0.0f ? 1u : 0u)) & 255u;
        c->phase = (c->phase + 1) & 7;
        c->prev_y = oy;
    }
    return n;
}
