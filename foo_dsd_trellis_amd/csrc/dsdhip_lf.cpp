// Host side of the differentiable loop filter and the pre-emphasis FIR:
// argument checks, workspace carving and the backward slab schedule.  All
// buffers are device pointers owned by the caller; nothing is allocated.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "dsdhip_internal.h"

extern "C" {
int lf_launch_fwd(int P, const double *a, const double *g, double T, const double *u, int B,
                  int64_t N, int E, double *y, double *ck, hipStream_t st);
int lf_launch_recompute(int P, const double *a, const double *g, double T, const double *u,
                        int B, int64_t N, int E, const double *ck, int64_t c0, int nchunks,
                        double *S, double *YD, hipStream_t st);
int lf_launch_adj_seq(int P, const double *a, const double *g, const double *gy, int B,
                      int64_t n_slab, int64_t cnt, const double *YD, double *lam, int first,
                      int SUB, double *LAMCK, double *VB, int64_t span, hipStream_t st);
int lf_launch_adj_par(int P, const double *a, const double *g, const double *u, int B,
                      int64_t n_slab, int64_t cnt, int SUB, const double *S, const double *LAMCK,
                      const double *VB, double *gu, double *part, int nblocks, hipStream_t st);
int lf_launch_reduce(const double *part, int nparts, int w, double *out, int acc, hipStream_t st);
int pe_launch_fwd(const double *w, int T, const float *p, int B, int64_t N, double *u,
                  hipStream_t st);
int pe_launch_bwd(const double *gu, int T, const float *p, int B, int64_t N, double *part,
                  hipStream_t st);
}

namespace {

const int kParBlocks = 1024;      // blocks of the parallel adjoint pass
const int kPeTile = 64;           // samples per pre-emphasis block (PE_TN)
const int kPeMaxTaps = 128;

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Sub-chunk length of the parallel adjoint pass: gcd(E, 64) (E % 16 == 0).
int lf_sub(int E)
{
    int s = 64;
    while (E % s) s >>= 1;
    return s;
}

bool lf_args_ok(int order, int B, size_t N, int E)
{
    return order >= 3 && order <= 10 && B >= 1 && N >= 1 && E >= 16 && (E % 16) == 0;
}

struct LfWs {
    double *S, *YD, *LAMCK, *VB, *lam, *part, *gagg;
};

size_t lf_ws_layout(int B, int P, int E, int G, char *base, LfWs *w)
{
    const size_t ns = (size_t)G * E;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return (double *)p;
    };
    LfWs t;
    t.S = take(ns * P * B * 8);
    t.YD = take(ns * B * 8);
    t.LAMCK = take((ns / lf_sub(E) + 1) * P * B * 8);
    t.VB = take(ns * B * 8);
    t.lam = take((size_t)P * B * 8);
    t.part = take((size_t)kParBlocks * 2 * P * 8);
    t.gagg = take((size_t)2 * P * 8);
    if (w) *w = t;
    return off;
}

}  // namespace

extern "C" {

size_t dsdhip_lf_checkpoints(size_t N, int ck_every)
{
    if (ck_every <= 0) return 0;
    return (N + (size_t)ck_every - 1) / (size_t)ck_every;
}

size_t dsdhip_lf_workspace_bytes(int B, size_t N, int order, int ck_every, int slab_chunks)
{
    if (!lf_args_ok(order, B, N, ck_every)) return 0;
    const size_t nck = dsdhip_lf_checkpoints(N, ck_every);
    const int G = (slab_chunks <= 0 || (size_t)slab_chunks > nck) ? (int)nck : slab_chunks;
    return lf_ws_layout(B, order, ck_every, G, nullptr, nullptr);
}

int dsdhip_lf_forward(const double *d_a, const double *d_g, int order, double temperature,
                      const double *d_u, int B, size_t N, int ck_every, double *d_y,
                      double *d_ck, void *stream)
{
    if (!lf_args_ok(order, B, N, ck_every) || !d_a || !d_g || !d_u || !d_y || !d_ck) return -1;
    if (dsdhip_ensure_device()) return -1;
    return lf_launch_fwd(order, d_a, d_g, temperature, d_u, B, (int64_t)N, ck_every, d_y, d_ck,
                         (hipStream_t)stream);
}

int dsdhip_lf_backward_ex(const double *d_a, const double *d_g, int order, double temperature,
                          const double *d_u, int B, size_t N, int ck_every, const double *d_ck,
                          const double *d_gy, double *d_gu, double *d_ga, double *d_gg, void *d_ws,
                          size_t ws_bytes, int slab_chunks, int trunc_span, void *stream)
{
    if (!lf_args_ok(order, B, N, ck_every) || !d_a || !d_g || !d_u || !d_ck || !d_gy || !d_gu ||
        !d_ga || !d_gg || !d_ws)
        return -1;
    if (dsdhip_ensure_device()) return -1;
    const int P = order, E = ck_every;
    if (trunc_span < 0 || (trunc_span > 0 && (E % trunc_span != 0 || trunc_span % lf_sub(E) != 0))) return -1;
    const size_t nck = dsdhip_lf_checkpoints(N, E);
    const int G = (slab_chunks <= 0 || (size_t)slab_chunks > nck) ? (int)nck : slab_chunks;
    LfWs w;
    if (lf_ws_layout(B, P, E, G, (char *)d_ws, &w) > ws_bytes) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int64_t nslabs = ((int64_t)nck + G - 1) / G;
    for (int64_t s = nslabs - 1; s >= 0; s--) {
        const int64_t c0 = s * G;
        const int cn = (int)(((int64_t)nck - c0 < G) ? (int64_t)nck - c0 : G);
        const int64_t n_slab = c0 * E;
        const int64_t n_hi = ((c0 + cn) * E < (int64_t)N) ? (c0 + cn) * E : (int64_t)N;
        const int first = (s == nslabs - 1);
        if (lf_launch_recompute(P, d_a, d_g, temperature, d_u, B, (int64_t)N, E, d_ck, c0, cn,
                                w.S, w.YD, st))
            return -1;
        const int SUB = lf_sub(E);
        const int64_t cnt = n_hi - n_slab;
        if (lf_launch_adj_seq(P, d_a, d_g, d_gy, B, n_slab, cnt, w.YD, w.lam, first, SUB,
                              w.LAMCK, w.VB, (int64_t)trunc_span, st))
            return -1;
        const int64_t tasks = ((cnt + SUB - 1) / SUB) * ((B + 63) / 64);
        const int nblk = (int)((tasks + 3) / 4 < kParBlocks ? (tasks + 3) / 4 : kParBlocks);
        if (lf_launch_adj_par(P, d_a, d_g, d_u, B, n_slab, cnt, SUB, w.S, w.LAMCK, w.VB, d_gu,
                              w.part, nblk, st))
            return -1;
        if (lf_launch_reduce(w.part, nblk, 2 * P, w.gagg, first ? 0 : 1, st)) return -1;
    }
    if (hipMemcpyAsync(d_ga, w.gagg, (size_t)P * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return -1;
    if (hipMemcpyAsync(d_gg, w.gagg + P, (size_t)P * 8, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
        return -1;
    return 0;
}

int dsdhip_lf_backward(const double *d_a, const double *d_g, int order, double temperature,
                       const double *d_u, int B, size_t N, int ck_every, const double *d_ck,
                       const double *d_gy, double *d_gu, double *d_ga, double *d_gg, void *d_ws,
                       size_t ws_bytes, int slab_chunks, void *stream)
{
    return dsdhip_lf_backward_ex(d_a, d_g, order, temperature, d_u, B, N, ck_every, d_ck, d_gy, d_gu, d_ga,
                                 d_gg, d_ws, ws_bytes, slab_chunks, 0, stream);
}

size_t dsdhip_pe_workspace_bytes(int B, size_t N, int ntaps)
{
    if (B < 1 || N < 1 || ntaps < 1 || ntaps > kPeMaxTaps) return 0;
    const size_t blocks = ((N + kPeTile - 1) / kPeTile) * (size_t)((B + 63) / 64);
    return align256(blocks * (size_t)ntaps * 8);
}

int dsdhip_pe_forward(const double *d_w, int ntaps, const float *d_p, int B, size_t N,
                      double *d_u, void *stream)
{
    if (B < 1 || N < 1 || ntaps < 1 || ntaps > kPeMaxTaps || !d_w || !d_p || !d_u) return -1;
    if (dsdhip_ensure_device()) return -1;
    return pe_launch_fwd(d_w, ntaps, d_p, B, (int64_t)N, d_u, (hipStream_t)stream);
}

int dsdhip_pe_backward(const float *d_p, int B, size_t N, int ntaps, const double *d_gu,
                       double *d_gw, void *d_ws, size_t ws_bytes, void *stream)
{
    if (B < 1 || N < 1 || ntaps < 1 || ntaps > kPeMaxTaps || !d_p || !d_gu || !d_gw || !d_ws)
        return -1;
    if (ws_bytes < dsdhip_pe_workspace_bytes(B, N, ntaps)) return -1;
    if (dsdhip_ensure_device()) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int blocks = (int)(((N + kPeTile - 1) / kPeTile) * (size_t)((B + 63) / 64));
    if (pe_launch_bwd(d_gu, ntaps, d_p, B, (int64_t)N, (double *)d_ws, st)) return -1;
    return lf_launch_reduce((const double *)d_ws, blocks, ntaps, d_gw, 0, st);
}

}  // extern "C"
