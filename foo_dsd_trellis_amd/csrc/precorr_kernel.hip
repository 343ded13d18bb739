// PreCorr SDM on the device: the greedy 1-bit quantizer through the NTF
// loop filter in single precision, followed by the prediction-table
// correction, per sample (src/precorr.c:222-268).  Each sample depends on
// the previous output, so a channel is one sequential recurrence: one wave
// per channel, the lanes fetching 64 inputs a block ahead and the wave
// advancing through them in order; the prediction table sits in LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pc_types.h"

typedef unsigned long long u64;

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
    float s[P], a[P], g[P];
#pragma unroll
    for (int k = 0; k < P; k++) {
        s[k] = st->state[k];
        a[k] = prm.a[k];
        g[k] = prm.g[k];
    }
    float prev_y = st->prev_y;
    int history = st->history & 255, phase = st->phase & 7;
    const float lim = prm.state_limit;
    u64 *o = out + (int64_t)c * out_stride;
    /* the correction of the current step; the next step's is read for both
     * possible output bits while this step's filter runs, so the table's
     * LDS latency stays off the sample-to-sample dependence */
    float corr = tab[history * 8 + phase];
    float xv = (lane < n) ? (float)x[lane] : 0.0f;
    for (int64_t b = 0; b < n; b += 64) {
        const int64_t nx = b + 64 + lane;
        const float xn = (nx < n) ? (float)x[nx] : 0.0f;     /* next block, in flight meanwhile */
        const int m = (n - b < 64) ? (int)(n - b) : 64;
        u64 bits = 0;
        for (int j = 0; j < m; j++) {
            const int hb = (history << 1) & 255, np = (phase + 1) & 7;
            const float c0 = tab[hb * 8 + np], c1 = tab[(hb | 1) * 8 + np];
            const float xs = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xv), j));
            float d[P];
            d[0] = s[0] - g[0] * s[1] + xs - prev_y;
            float v = xs + a[0] * d[0];
#pragma unroll
            for (int k = 1; k < P - 1; k++) {
                d[k] = s[k] + s[k - 1] - g[k] * s[k + 1];
                v = v + a[k] * d[k];
            }
            d[P - 1] = s[P - 1] + s[P - 2];
            v = v + a[P - 1] * d[P - 1];
            const float y = (v >= 0.0f) ? 1.0f : -1.0f;
            if (lim > 0.0f) {
#pragma unroll
                for (int k = 0; k < P; k++) d[k] = (d[k] > lim) ? lim : ((d[k] < -lim) ? -lim : d[k]);
            }
#pragma unroll
            for (int k = 0; k < P; k++) s[k] = d[k];
            const float yc = y - corr;
            const int ob = (yc >= 0.0f) ? 1 : 0;
            bits |= (u64)ob << j;
            history = hb | ob;
            phase = np;
            corr = ob ? c1 : c0;
            prev_y = ob ? 1.0f : -1.0f;
        }
        if (lane == 0) o[b >> 6] = bits;
        xv = xn;
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < P; k++) st->state[k] = s[k];
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
