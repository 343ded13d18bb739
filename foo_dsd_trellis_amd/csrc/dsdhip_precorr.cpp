// Host side of the PreCorr SDM: the prediction-table training (setup, as
// the reference does it at context init), per-channel device states, and the
// launches.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "dsdhip_internal.h"
#include "pc_types.h"

extern "C" int pc_launch(int P, int in_f32, const void *in, int64_t in_stride, int nch, int64_t n,
                         const float *table, const pc_params *prm, dsdhip_precorr_state_t *states,
                         unsigned long long *out, int64_t out_stride, hipStream_t stream);

struct dsdhip_precorr {
    int nch;
    pc_params prm;
    float table[256 * 8];
    float *d_table;
    dsdhip_precorr_state_t *d_state;
    hipStream_t stream;
};

/* the loop filter in float, operation order of src/precorr.c:20-108 */
static float pc_filter(const float *s, float *d, const float *a, const float *g, int order, float x, float y)
{
    d[0] = s[0] - g[0] * s[1] + x - y;
    float v = x + a[0] * d[0];
    int i;
    for (i = 1; i < order - 1; i++) {
        d[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        v = v + a[i] * d[i];
    }
    d[i] = s[i] + s[i - 1];
    v = v + a[i] * d[i];
    return v;
}

static void pc_initial(dsdhip_precorr_state_t *st)
{
    memset(st, 0, sizeof(*st));
    st->history = 0x69;            /* DSD silence pattern (src/precorr.c:203) */
}

extern "C" {

/* src/precorr.c:134-186: 65,536 samples of LCG noise at +-0.1 through the
 * greedy loop; table entry = mean of (y - v) per (history, phase) with at
 * least 4 hits, else 0. */
int dsdhip_precorr_train(const float *a, const float *g, int order, float *table)
{
    if (!a || !g || !table || order < 2 || order > 8) return -1;
    static float accum[256][8];
    static int hits[256][8];
    memset(accum, 0, sizeof(accum));
    memset(hits, 0, sizeof(hits));
    float state[8] = {0}, d[8];
    unsigned history = 0x69;
    int phase = 0;
    float prev_y = 0.0f;
    uint32_t rng = 0xDEADBEEFu;
    for (int i = 0; i < 65536; i++) {
        rng = rng * 1664525u + 1013904223u;
        const float x = 0.1f * ((float)(int32_t)rng / 2147483648.0f);
        const float v = pc_filter(state, d, a, g, order, x, prev_y);
        const float y = (v >= 0.0f) ? 1.0f : -1.0f;
        memcpy(state, d, sizeof(float) * (size_t)order);
        accum[history][phase] += y - v;
        hits[history][phase]++;
        history = ((history << 1) | (y > 0.0f ? 1u : 0u)) & 255u;
        phase = (phase + 1) & 7;
        prev_y = y;
    }
    for (int h = 0; h < 256; h++)
        for (int p = 0; p < 8; p++)
            table[h * 8 + p] = hits[h][p] >= 4 ? accum[h][p] / (float)hits[h][p] : 0.0f;
    return 0;
}

void dsdhip_precorr_destroy(dsdhip_precorr_t *p)
{
    if (!p) return;
    if (p->d_table) hipFree(p->d_table);
    if (p->d_state) hipFree(p->d_state);
    delete p;
}

dsdhip_precorr_t *dsdhip_precorr_create_coeffs(const double *a, const double *g, int order, int nch,
                                               float state_limit)
{
    if (!a || !g || order < 4 || order > 8 || nch < 1) return nullptr;
    if (dsdhip_ensure_device() != 0) return nullptr;
    dsdhip_precorr_t *p = new dsdhip_precorr();
    memset((void *)p, 0, sizeof(*p));
    p->nch = nch;
    p->prm.order = order;
    p->prm.state_limit = state_limit;
    for (int i = 0; i < order; i++) {
        p->prm.a[i] = (float)a[i];               /* src/precorr.c:196-199 */
        p->prm.g[i] = (float)g[i];
    }
    if (dsdhip_precorr_train(p->prm.a, p->prm.g, order, p->table) != 0 ||
        hipMalloc(&p->d_table, sizeof(p->table)) != hipSuccess ||
        hipMemcpy(p->d_table, p->table, sizeof(p->table), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&p->d_state, sizeof(dsdhip_precorr_state_t) * (size_t)nch) != hipSuccess ||
        dsdhip_precorr_reset(p) != 0) {
        dsdhip_precorr_destroy(p);
        return nullptr;
    }
    return p;
}

dsdhip_precorr_t *dsdhip_precorr_create(int ntf_index, int nch, float state_limit)
{
    const dsdhip_ntf_t *f = dsdhip_ntf_get(ntf_index);
    if (!f) return nullptr;
    return dsdhip_precorr_create_coeffs(f->a, f->g, f->order, nch, state_limit);
}

int dsdhip_precorr_reset(dsdhip_precorr_t *p)
{
    if (!p) return -1;
    std::vector<dsdhip_precorr_state_t> st((size_t)p->nch);
    for (auto &s : st) pc_initial(&s);
    return hipMemcpy(p->d_state, st.data(), sizeof(dsdhip_precorr_state_t) * st.size(), hipMemcpyHostToDevice) ==
                   hipSuccess ? 0 : -1;
}

int dsdhip_precorr_set_stream(dsdhip_precorr_t *p, void *stream)
{
    if (!p) return -1;
    p->stream = (hipStream_t)stream;
    return 0;
}

int dsdhip_precorr_table(const dsdhip_precorr_t *p, float *table)
{
    if (!p || !table) return -1;
    memcpy(table, p->table, sizeof(p->table));
    return 0;
}

int dsdhip_precorr_get_state(dsdhip_precorr_t *p, int ch, dsdhip_precorr_state_t *st)
{
    if (!p || !st || ch < 0 || ch >= p->nch) return -1;
    if (hipStreamSynchronize(p->stream) != hipSuccess) return -1;
    return hipMemcpy(st, p->d_state + ch, sizeof(*st), hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

int dsdhip_precorr_set_state(dsdhip_precorr_t *p, int ch, const dsdhip_precorr_state_t *st)
{
    if (!p || !st || ch < 0 || ch >= p->nch) return -1;
    if (hipStreamSynchronize(p->stream) != hipSuccess) return -1;
    return hipMemcpy(p->d_state + ch, st, sizeof(*st), hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

int dsdhip_precorr_process_dev(dsdhip_precorr_t *p, const double *d_in, int64_t in_stride, size_t n,
                               unsigned long long *d_bits, int64_t words_stride)
{
    if (!p || !d_in || !d_bits) return -1;
    if (n == 0) return 0;
    if (words_stride < ((int64_t)n + 63) / 64 || in_stride < (int64_t)n) return -1;
    return pc_launch(p->prm.order, 0, d_in, in_stride, p->nch, (int64_t)n, p->d_table, &p->prm, p->d_state, d_bits,
                     words_stride, p->stream);
}

int gpu_hip_precorr_process(void *ctx, const float *in, float *out, size_t count, const float *ntf_a,
                            const float *ntf_g, int order, const float pred_table[256][8],
                            const gpu_hip_precorr_state_t *init, gpu_hip_precorr_state_t *final_state)
{
    (void)ctx;
    if (!in || !out || !ntf_a || !ntf_g || !pred_table || order < 4 || order > 8) return -1;
    if (count < 8192) return -1;                  /* GPU_MIN_SAMPLES (include/gpu_compute.h:40) */
    pc_params prm;
    memset(&prm, 0, sizeof(prm));
    prm.order = order;
    for (int i = 0; i < order; i++) {
        prm.a[i] = ntf_a[i];
        prm.g[i] = ntf_g[i];
    }
    dsdhip_precorr_state_t st;
    if (init) st = *init;
    else pc_initial(&st);
    float *d_in = nullptr, *d_tab = nullptr;
    dsdhip_precorr_state_t *d_st = nullptr;
    unsigned long long *d_bits = nullptr;
    const size_t words = (count + 63) / 64;
    std::vector<unsigned long long> bits(words);
    int rc = -1;
    if (hipMalloc(&d_in, sizeof(float) * count) == hipSuccess &&
        hipMalloc(&d_tab, sizeof(float) * 2048) == hipSuccess &&
        hipMalloc(&d_st, sizeof(st)) == hipSuccess &&
        hipMalloc(&d_bits, sizeof(unsigned long long) * words) == hipSuccess &&
        hipMemcpy(d_in, in, sizeof(float) * count, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(d_tab, pred_table, sizeof(float) * 2048, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(d_st, &st, sizeof(st), hipMemcpyHostToDevice) == hipSuccess &&
        pc_launch(order, 1, d_in, (int64_t)count, 1, (int64_t)count, d_tab, &prm, d_st, d_bits, (int64_t)words,
                  0) == 0 &&
        hipMemcpy(bits.data(), d_bits, sizeof(unsigned long long) * words, hipMemcpyDeviceToHost) == hipSuccess &&
        hipMemcpy(&st, d_st, sizeof(st), hipMemcpyDeviceToHost) == hipSuccess) {
        for (size_t i = 0; i < count; i++) out[i] = ((bits[i >> 6] >> (i & 63)) & 1ull) ? 1.0f : -1.0f;
        if (final_state) *final_state = st;
        rc = 0;
    }
    if (d_in) hipFree(d_in);
    if (d_tab) hipFree(d_tab);
    if (d_st) hipFree(d_st);
    if (d_bits) hipFree(d_bits);
    return rc;
}

}  // extern "C"
