// PreCorr SDM on the device: the greedy 1-bit quantizer through the NTF
// loop filter in single precision, followed by the prediction-table
// correction, per sample (src/precorr.c:222-268).  Each sample depends on
// the previous output, so a channel is one sequential recurrence: one wave
// per channel, advancing through 64-sample blocks whose inputs the lanes
// fetch a block ahead.
//
// The previous output enters the first term of the loop filter's sum,
// which the reference adds left to right, so the whole sum waits for it.
// Everything else of a sample does not: lane k < P holds state s[k] and
// computes d[k] (k >= 1) and its product a[k] d[k], and lane 0 the part of
// d[0] before the previous output, one instruction each for all lanes.  The
// chain d0 = A - y', v = x + a0 d0 + a1 d1 + ... + a(P-1) d(P-1), the
// quantizer and the corrected output then run in lane 0, each product
// taken from its lane by the add itself (DPP quad_perm / row_shl), so
// the chain is a run of dependent adds with nothing between them; a
// single wave issues one instruction at a time, so the rest of a sample is
// kept to a few instructions too.  The other lanes run the same
// instructions on values nobody reads (their LDS addresses stay in range).
//
// Outputs are shifted into a 32-bit register per lane (v_addc with the
// output's compare as carry): its low 8 bits are the reference's history
// byte and every 32 samples it holds 32 outputs, latest in bit 0.  The
// prediction table sits in LDS rotated so that row r holds phase
// (phase0 + r) & 7: sample j of a call (blocks are 64 samples) reads row
// (j + 1) & 7 for the next sample's two candidates, a constant offset in
// the unrolled loop, one 8-byte read a sample ahead of its use.
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
/* lane 0 (and every lane l) <- lane l + k, k = 1..7: quad_perm for k < 4,
 * row_shl:k from 4 */
template <int K>
__device__ __forceinline__ float from_lane(float v)
{
    constexpr int ctrl = (K < 4) ? (K | (K << 2) | (K << 4) | (K << 6)) : (0x100 + K);
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, true));
}
/* acc * 2 + b: the output bit enters at bit 0 (one add-with-carry, the
 * compare's lane mask as the carry) */
__device__ __forceinline__ unsigned push_bit(unsigned acc, bool b)
{
    unsigned r;
    const u64 m = __builtin_amdgcn_ballot_w64(b);
    __asm__("v_addc_co_u32_e64 %0, s[0:1], %1, %1, %2" : "=v"(r) : "v"(acc), "s"(m) : "s0", "s1");
    return r;
}

template <int P, int K>
struct pc_sum {
    /* v + a[K] d[K] + ... + a[P-1] d[P-1], in k order */
    __device__ static __forceinline__ float run(float v, float pk)
    {
        if constexpr (K >= P) {
            return v;
        } else {
            return pc_sum<P, K + 1>::run(from_lane<K>(pk) + v, pk);
        }
    }
};

/* One wave per channel; lane k < P holds loop-filter state s[k] and its
 * coefficients; lanes >= P hold zeros.  LIM: the state limit is on. */
template <int P, typename TIn, bool LIM>
__global__ __launch_bounds__(64) void precorr_k(const TIn *__restrict__ in, int64_t in_stride, int nch, int64_t n,
                                                const float *__restrict__ table, pc_params prm,
                                                dsdhip_precorr_state_t *__restrict__ states, u64 *__restrict__ out,
                                                int64_t out_stride)
{
    __shared__ float rt[8 * 256];      /* rt[r][h] = table[h][(phase0 + r) & 7] */
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= nch) return;              /* the whole workgroup (one wave) leaves together */
    dsdhip_precorr_state_t *st = states + c;
    const int phase0 = st->phase & 7;
    for (int i = lane; i < 256 * 8; i += 64) {       /* i = h * 8 + phase */
        const int h = i >> 3, r = ((i & 7) - phase0) & 7;
        rt[r * 256 + h] = table[i];
    }
    __syncthreads();
    const TIn *x = in + (int64_t)c * in_stride;
    const bool kin = lane < P;
    float s = kin ? st->state[lane] : 0.0f;
    const float ak = kin ? prm.a[lane] : 0.0f;
    /* g[P-1] is not a term of d[P-1] = s[P-1] + s[P-2]: with g 0 there and
     * s[P] held at 0, lane P-1's t - g s[P] is t - 0 = t exactly */
    const float gk = (lane < P - 1) ? prm.g[lane] : 0.0f;
    const float a0 = prm.a[0];
    const float lim = prm.state_limit;
    float prev_y = st->prev_y;
    unsigned acc = (unsigned)(st->history & 255);     /* outputs, latest in bit 0 */
    u64 *o = out + (int64_t)c * out_stride;
    float corr = rt[acc];                             /* row 0: phase0 */
    float xv = (lane < n) ? (float)x[lane] : 0.0f;

    /* one sample; R = row of the next sample's correction */
    auto step = [&](const float xs, const int R) {
        /* the next sample's correction for either output bit of this one */
        const unsigned ha = (acc << 3) & 1016u;       /* ((history << 1) & 255) * 4 */
        const float2 cc = *(const float2 *)((const char *)rt + R * 1024 + ha);
        /* d[k] = s[k] + s[k-1] - g[k] s[k+1] (k >= 1); lane 0:
         * A = s[0] - g[0] s[1] + x (src/precorr.c:20-108) */
        const float u = gk * shl1(s);
        const float t = s + shr1(s);
        const float dk = t - u;
        const float A = (s - u) + xs;
        const float pk = ak * dk;
        /* lane 0: d0 = A - y', v = x + a0 d0, then + a[k] d[k] in k order */
        const float d0 = A - prev_y;
        const float v = pc_sum<P, 1>::run(xs + a0 * d0, pk);
        const float y = (v >= 0.0f) ? 1.0f : -1.0f;
        const float yc = y - corr;
        const bool ob = yc >= 0.0f;
        /* state: d, limited as the reference limits it */
        float d = (lane == 0) ? d0 : dk;
        if (LIM) d = (d > lim) ? lim : ((d < -lim) ? -lim : d);
        s = kin ? d : 0.0f;
        acc = push_bit(acc, ob);
        corr = ob ? cc.y : cc.x;
        prev_y = ob ? 1.0f : -1.0f;
    };

    for (int64_t b = 0; b < n; b += 64) {
        const int64_t nx = b + 64 + lane;
        /* next block, in flight meanwhile */
        const TIn xn = x[nx < n ? nx : (n - 1)];
        const int m = (n - b < 64) ? (int)(n - b) : 64;
        unsigned lo = 0;
        if (m == 64) {
#pragma unroll
            for (int q = 0; q < 64; q++) {
                step(bcast(xv, q), (q + 1) & 7);
                if (q == 31) lo = acc;                /* samples 0..31, latest in bit 0 */
            }
        } else {
            for (int j = 0; j < m; j++) {
                step(bcast(xv, j), (j + 1) & 7);
                if (j == 31) lo = acc;
            }
        }
        if (lane == 0) {
            /* bit i of the word = sample i: reverse each half, keep its own samples */
            const int c0 = m < 32 ? m : 32, c1 = m - c0;
            const unsigned w0 = c0 == 32 ? lo : acc;
            u64 word = (u64)(__builtin_bitreverse32(w0) >> (32 - c0));
            if (c1 > 0) word |= (u64)(__builtin_bitreverse32(acc) >> (32 - c1)) << 32;
            o[b >> 6] = word;
        }
        xv = (b + 64 + lane < n) ? (float)xn : 0.0f;
    }
    if (kin) st->state[lane] = s;
    if (lane == 0) {
        st->prev_y = prev_y;
        st->history = (int)(acc & 255u);
        st->phase = (int)((phase0 + n) & 7);
    }
}

template <int P, typename TIn>
static int pc_go(const void *in, int64_t in_stride, int nch, int64_t n, const float *table, const pc_params *prm,
                 dsdhip_precorr_state_t *states, u64 *out, int64_t out_stride, hipStream_t stream)
{
    if (prm->state_limit > 0.0f)
        hipLaunchKernelGGL((precorr_k<P, TIn, true>), dim3((unsigned)nch), dim3(64), 0, stream, (const TIn *)in,
                           in_stride, nch, n, table, *prm, states, out, out_stride);
    else
        hipLaunchKernelGGL((precorr_k<P, TIn, false>), dim3((unsigned)nch), dim3(64), 0, stream, (const TIn *)in,
                           in_stride, nch, n, table, *prm, states, out, out_stride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int pc_launch(int P, int in_f32, const void *in, int64_t in_stride, int nch, int64_t n,
                         const float *table, const pc_params *prm, dsdhip_precorr_state_t *states, u64 *out,
                         int64_t out_stride, hipStream_t stream)
{
#define PC_CASE(PP)                                                                                         \
    if (P == PP)                                                                                            \
        return in_f32 ? pc_go<PP, float>(in, in_stride, nch, n, table, prm, states, out, out_stride, stream)  \
                      : pc_go<PP, double>(in, in_stride, nch, n, table, prm, states, out, out_stride, stream);
    PC_CASE(4) PC_CASE(5) PC_CASE(6) PC_CASE(7) PC_CASE(8)
#undef PC_CASE
    return -1;
}
