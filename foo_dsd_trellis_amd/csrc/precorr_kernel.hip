// PreCorr SDM on the device: the greedy 1-bit quantizer through the NTF
// loop filter in single precision, followed by the prediction-table
// correction, per sample (src/precorr.c:222-268).  Each sample depends on
// the previous output, so a channel is one sequential recurrence: one wave
// per channel, the lanes fetching 64 inputs a block ahead and the wave
// advancing through them in order; the prediction table sits in LDS.
// The step is bound by its dependence chain (the previous output enters
// the first term of the loop filter's sum, which the reference adds left to
// right), about 150 ns per sample; the lanes hold the filter state so the
// independent parts take one instruction each.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pc_types.h"

typedef unsigned long long u64;

/* lane l <- lane l-1 (row_shr:1) / lane l+1 (row_shl:1) within rows of 16;
 * lanes past a row's edge read 0 */
__device__ __forceinline__ float shr1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));
}
__device__ __forceinline__ float shl1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xf, 0xf, true));
}
__device__ __forceinline__ float bcast(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

/* One wave per channel; lane k < P holds loop-filter state s[k] and its
 * coefficients, so one instruction computes every d[k].  The sum
 * v = xs + a0 d0 + a1 d1 + ... keeps the reference's left-to-right order:
 * lane k adds a_k d_k to lane k-1's partial sum, one lane step per term. */
template <int P, typename TIn>
__global__ __launch_bounds__(64) void precorr_k(const TIn *__restrict__ in, int64_t in_stride, int nch, int64_t n,
                                                const float *__restrict__ table, pc_params prm,
                                                dsdhip_precorr_state_t *__restrict__ states, u64 *__restrict__ out,
                                                int64_t out_stride)
{
    __shared__ float tab[256 * 8];
    for (int i = threadIdx.x; i < 256 * 8; i += 64) tab[i] = table[i];
    __syncthreads();
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= nch) return;
    const TIn *x = in + (int64_t)c * in_stride;
    dsdhip_precorr_state_t *st = states + c;
    const bool kin = lane < P;
    const int m0 = (lane == 0) ? -1 : 0;
    float s = kin ? st->state[lane] : 0.0f;
    const float ak = kin ? prm.a[lane] : 0.0f, gk = kin ? prm.g[lane] : 0.0f;
    float prev_y = st->prev_y;
    int history = st->history & 255, phase = st->phase & 7;
    const float lim = prm.state_limit;
    u64 *o = out + (int64_t)c * out_stride;
    /* the correction of the current step; the next step's is read for both
     * possible output bits while this step's filter runs */
    float corr = tab[history * 8 + phase];
    float xv = (lane < n) ? (float)x[lane] : 0.0f;
    for (int64_t b = 0; b < n; b += 64) {
        const int64_t nx = b + 64 + lane;
        /* next block, in flight meanwhile (converted at the block's end, so
         * nothing waits for it before then) */
        const TIn xn = x[nx < n ? nx : (n - 1)];
        const int m = (n - b < 64) ? (int)(n - b) : 64;
        int obv = 0;                                          /* lane j: output bit of sample j */
        for (int j = 0; j < m; j++) {
            const int hb = (history << 1) & 255, np = (phase + 1) & 7;
            const float c0 = tab[hb * 8 + np], c1 = tab[(hb | 1) * 8 + np];
            const float xs = bcast(xv, j);
            /* d[0] = s0 - g0 s1 + xs - prev_y; d[k] = s[k] + s[k-1] - g[k] s[k+1];
             * d[P-1] = s[P-1] + s[P-2] (src/precorr.c:230-242) */
            const float u = gk * shl1(s);
            const float t = s + shr1(s);
            const float d0 = ((s - u) + xs) - prev_y;
            /* per-lane choice by bit masks: a select here becomes a branch */
            const int dm = __float_as_int((lane == P - 1) ? t : (t - u));
            float d = __int_as_float((__float_as_int(d0) & m0) | (dm & ~m0));
            /* v = xs + a0 d0, then + a_k d_k in k order */
            const float pk = ak * d;
            float w = xs + pk;
#pragma unroll
            for (int k = 1; k < P; k++) w = shr1(w) + pk;
            const float v = bcast(w, P - 1);
            const float y = (v >= 0.0f) ? 1.0f : -1.0f;
            if (lim > 0.0f) d = (d > lim) ? lim : ((d < -lim) ? -lim : d);
            s = kin ? d : 0.0f;
            const float yc = y - corr;
            const int ob = (yc >= 0.0f) ? 1 : 0;
            obv = (lane == j) ? ob : obv;
            history = hb | ob;
            phase = np;
            corr = ob ? c1 : c0;
            prev_y = ob ? 1.0f : -1.0f;
        }
        const u64 bits = __builtin_amdgcn_ballot_w64(obv != 0);
        if (lane == 0) o[b >> 6] = bits;
        xv = (b + 64 + lane < n) ? (float)xn : 0.0f;
    }
    if (kin) st->state[lane] = s;
    if (lane == 0) {
        st->prev_y = prev_y;
        st->history = history;
        st->phase = phase;
    }
}

extern "C" int pc_launch(int P, int in_f32, const void *in, int64_t in_stride, int nch, int64_t n,
                         const float *table, const pc_params *prm, dsdhip_precorr_state_t *states, u64 *out,
                         int64_t out_stride, hipStream_t stream)
{
#define PC_CASE(PP)                                                                                          \
    if (P == PP) {                                                                                           \
        if (in_f32)                                                                                          \
            hipLaunchKernelGGL((precorr_k<PP, float>), dim3((unsigned)nch), dim3(64), 0, stream,            \
                               (const float *)in, in_stride, nch, n, table, *prm, states, out, out_stride);  \
        else                                                                                                 \
            hipLaunchKernelGGL((precorr_k<PP, double>), dim3((unsigned)nch), dim3(64), 0, stream,           \
                               (const double *)in, in_stride, nch, n, table, *prm, states, out, out_stride); \
        return hipGetLastError() == hipSuccess ? 0 : -1;                                                     \
    }
    PC_CASE(4) PC_CASE(5) PC_CASE(6) PC_CASE(7) PC_CASE(8)
#undef PC_CASE
    return -1;
}
