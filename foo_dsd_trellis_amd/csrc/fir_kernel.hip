/*
 * 2x polyphase half-band upsampling stage for gfx950 (FP64).
 *
 * Reference semantics: ippsFIRMR_64f with upFactor 2, phase 0, taps
 * pre-multiplied by 2 and zero-initialised delay line (src/fir.c:370-465):
 *   y[2m+p] = sum_q h2[2q+p] * x[m-q],  p in {0,1},  q ascending,
 * accumulated with fused multiply-add in ascending q (the oracle uses the
 * same order, so results are identical doubles).
 *
 * Layout: one block = 256 threads = 4 waves covers TILE = 256*R input
 * positions of one channel (blockIdx.y).  The input tile plus its Q-1
 * sample history is staged in LDS with coalesced loads; each thread keeps a
 * sliding window of R+1 inputs in registers and walks q, so every LDS value
 * is read once per q-step and reused for 2R outputs.  The 2R outputs of a
 * thread are consecutive, so they go through LDS once more and leave as
 * contiguous 16-byte lanes (a wave stores 1 KB per instruction).
 * The last stage multiplies by the engine gain (src/engine.c:872-883).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int FIR_THREADS = 256;
constexpr int FIR_R = 4;
constexpr int FIR_TILE = FIR_THREADS * FIR_R;

template <typename TIn, int T>
__global__ __launch_bounds__(FIR_THREADS) void fir_up2_kernel(
    const TIn *__restrict__ x, int64_t x_stride, const double *__restrict__ hist,
    int64_t n, const double *__restrict__ h2, double *__restrict__ y, int64_t y_stride,
    double gain, int apply_gain)
{
    constexpr int Q0 = (T + 1) / 2;   /* even-phase taps */
    constexpr int Q1 = T / 2;         /* odd-phase taps */
    constexpr int H = Q0 - 1;         /* history samples needed */
    __shared__ double tile[FIR_TILE + H];
    __shared__ double2 ost[FIR_TILE];

    const int ch = blockIdx.y;
    const TIn *xc = x + (int64_t)ch * x_stride;
    const double *hc = hist ? hist + (int64_t)ch * H : nullptr;
    double *yc = y + (int64_t)ch * y_stride;
    const int64_t m0 = (int64_t)blockIdx.x * FIR_TILE;

    for (int i = threadIdx.x; i < FIR_TILE + H; i += FIR_THREADS) {
        const int64_t g = m0 - H + i;
        double v = 0.0;
        if (g >= 0) {
            if (g < n) v = (double)xc[g];
        } else if (hc) {
            v = hc[H + g];
        }
        tile[i] = v;
    }
    __syncthreads();

    /* thread handles inputs m0 + R*t + r, r = 0..R-1 */
    const int base = FIR_R * threadIdx.x + H;   /* tile index of its first input */
    double acc0[FIR_R], acc1[FIR_R], w[FIR_R];
#pragma unroll
    for (int r = 0; r < FIR_R; r++) {
        acc0[r] = 0.0;
        acc1[r] = 0.0;
        w[r] = tile[base + r];   /* window: x[m+r-q] for q = 0 */
    }
#pragma unroll
    for (int q = 0; q < Q0; q++) {
        const double he = h2[2 * q];
        const double ho = (q < Q1) ? h2[2 * q + 1] : 0.0;
#pragma unroll
        for (int r = 0; r < FIR_R; r++) {
            acc0[r] = __builtin_fma(he, w[r], acc0[r]);
            if (q < Q1) acc1[r] = __builtin_fma(ho, w[r], acc1[r]);
        }
        if (q + 1 < Q0) {
            /* shift window down by one input: w[r] <- x[m+r-(q+1)] */
#pragma unroll
            for (int r = FIR_R - 1; r > 0; r--) w[r] = w[r - 1];
            w[0] = tile[base - (q + 1)];
        }
    }
#pragma unroll
    for (int r = 0; r < FIR_R; r++) {
        double e = acc0[r], o = acc1[r];
        if (apply_gain) {
            e = e * gain;
            o = o * gain;
        }
        double2 v;
        v.x = e;
        v.y = o;
        ost[FIR_R * threadIdx.x + r] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < FIR_R; r++) {
        const int i = r * FIR_THREADS + threadIdx.x;
        const int64_t m = m0 + i;
        if (m < n) *reinterpret_cast<double2 *>(yc + 2 * m) = ost[i];
    }
}

/* Copies the last H inputs of each channel into its history (for the next
 * streaming call); when n < H the old history is shifted. */
template <typename TIn>
__global__ void fir_hist_update(const TIn *__restrict__ x, int64_t x_stride, int64_t n,
                                double *__restrict__ hist, int H)
{
    const int ch = blockIdx.x;
    double *hc = hist + (int64_t)ch * H;
    const TIn *xc = x + (int64_t)ch * x_stride;
    __shared__ double tmp[128];
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        const int64_t g = n - H + i;   /* position relative to this call's input */
        tmp[i] = g >= 0 ? (double)xc[g] : hc[H + g];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < H; i += blockDim.x) hc[i] = tmp[i];
}

/*
 * Two consecutive 63-tap 2x stages in one pass: a block takes FU_T inputs
 * of the first stage (plus the 47 before them), computes the first stage's
 * 2*FU_T outputs and the 31 before them (the second stage's history) into
 * LDS, and writes only the second stage's 4*FU_T outputs.  Each output is
 * the same ascending-q FMA sum as fir_up2_kernel, so the result is
 * identical to running the two stages one after the other.  Block 0 takes
 * the 31 intermediate samples before the chunk from histB (the second
 * stage's delay line); the last block stores the chunk's last 31
 * intermediate samples into histB_next (copied into histB afterwards, so no
 * block reads a line another block writes).
 */
constexpr int FU_T = 512;           /* first-stage inputs per block */
constexpr int FU_H = 31;            /* history of a 63-tap stage */
constexpr int FU_XH = FU_H + 16;    /* inputs before the tile: 16 for the halo pairs, 31 for their taps */

template <typename TIn>
__global__ __launch_bounds__(FIR_THREADS) void fir_up4_kernel(
    const TIn *__restrict__ x, int64_t x_stride, const double *__restrict__ histA, int64_t n1,
    const double *__restrict__ h2A, const double *__restrict__ h2B, const double *__restrict__ histB,
    double *__restrict__ histB_next, double *__restrict__ y, int64_t y_stride, double gain, int apply_gain)
{
    constexpr int Q0 = 32, Q1 = 31;
    constexpr int NI = FU_T + 16;          /* first-stage positions computed (halo pairs first) */
    __shared__ double xin[FU_T + FU_XH];   /* positions m0 - 47 .. m0 + FU_T - 1 */
    __shared__ double inter[2 * NI];       /* intermediate positions 2 m0 - 32 .. 2 m0 + 2 FU_T - 1 */
    __shared__ double2 ost[2 * FU_T];

    const int ch = blockIdx.y;
    const TIn *xc = x + (int64_t)ch * x_stride;
    const double *hA = histA ? histA + (int64_t)ch * FU_H : nullptr;
    const double *hB = histB ? histB + (int64_t)ch * FU_H : nullptr;
    double *yc = y + (int64_t)ch * y_stride;
    const int64_t m0 = (int64_t)blockIdx.x * FU_T;
    const int64_t n2 = 2 * n1;

    for (int i = threadIdx.x; i < FU_T + FU_XH; i += FIR_THREADS) {
        const int64_t g = m0 - FU_XH + i;
        double v = 0.0;
        if (g >= 0) {
            if (g < n1) v = (double)xc[g];
        } else if (g >= -FU_H && hA) {
            v = hA[FU_H + g];
        }
        xin[i] = v;
    }
    if (m0 == 0) {
        /* intermediate positions -32 .. -1: the second stage's delay line */
        for (int k = threadIdx.x; k < 32; k += FIR_THREADS) inter[k] = (k >= 1 && hB) ? hB[k - 1] : 0.0;
    }
    __syncthreads();

    /* first stage, two positions per thread and round: position
     * j = m0 - 16 + li, taps over xin[li + 31 - q] */
    for (int pr = threadIdx.x; pr < NI / 2; pr += FIR_THREADS) {
        const int li0 = 2 * pr;
        if (m0 == 0 && li0 < 16) continue;     /* halo of block 0 comes from histB */
        double acc0[2], acc1[2], w[2];
#pragma unroll
        for (int r = 0; r < 2; r++) {
            acc0[r] = 0.0;
            acc1[r] = 0.0;
            w[r] = xin[li0 + r + FU_H];
        }
#pragma unroll
        for (int q = 0; q < Q0; q++) {
            const double he = h2A[2 * q];
            const double ho = (q < Q1) ? h2A[2 * q + 1] : 0.0;
#pragma unroll
            for (int r = 0; r < 2; r++) {
                acc0[r] = __builtin_fma(he, w[r], acc0[r]);
                if (q < Q1) acc1[r] = __builtin_fma(ho, w[r], acc1[r]);
            }
            if (q + 1 < Q0) {
                w[1] = w[0];
                w[0] = xin[li0 + FU_H - (q + 1)];
            }
        }
#pragma unroll
        for (int r = 0; r < 2; r++) {
            inter[2 * (li0 + r)] = acc0[r];
            inter[2 * (li0 + r) + 1] = acc1[r];
        }
    }
    __syncthreads();

    /* the chunk's last 31 intermediate samples for the next call */
    if (histB_next && blockIdx.x == gridDim.x - 1 && threadIdx.x < FU_H) {
        const int64_t pos = n2 - FU_H + threadIdx.x;
        histB_next[(int64_t)ch * FU_H + threadIdx.x] = inter[pos - (2 * m0 - 32)];
    }

    /* second stage: thread t takes intermediate positions 2 m0 + 4t + r */
    const int base = FIR_R * threadIdx.x + 32;
    double acc0[FIR_R], acc1[FIR_R], w[FIR_R];
#pragma unroll
    for (int r = 0; r < FIR_R; r++) {
        acc0[r] = 0.0;
        acc1[r] = 0.0;
        w[r] = inter[base + r];
    }
#pragma unroll
    for (int q = 0; q < Q0; q++) {
        const double he = h2B[2 * q];
        const double ho = (q < Q1) ? h2B[2 * q + 1] : 0.0;
#pragma unroll
        for (int r = 0; r < FIR_R; r++) {
            acc0[r] = __builtin_fma(he, w[r], acc0[r]);
            if (q < Q1) acc1[r] = __builtin_fma(ho, w[r], acc1[r]);
        }
        if (q + 1 < Q0) {
#pragma unroll
            for (int r = FIR_R - 1; r > 0; r--) w[r] = w[r - 1];
            w[0] = inter[base - (q + 1)];
        }
    }
#pragma unroll
    for (int r = 0; r < FIR_R; r++) {
        double e = acc0[r], o = acc1[r];
        if (apply_gain) {
            e = e * gain;
            o = o * gain;
        }
        double2 v;
        v.x = e;
        v.y = o;
        ost[FIR_R * threadIdx.x + r] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < FIR_R; r++) {
        const int i = r * FIR_THREADS + threadIdx.x;
        const int64_t p = 2 * m0 + i;
        if (p < n2) *reinterpret_cast<double2 *>(yc + 2 * p) = ost[i];
    }
}

__global__ void fir_copy(const double *__restrict__ src, double *__restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

} // namespace

/* Two 63-tap stages in one launch over nch channels: n1 first-stage inputs
 * -> 4 n1 outputs.  histA / histB: the two stages' delay lines (nch x 31);
 * histB is refreshed from the chunk's intermediate samples through
 * hist_tmp (nch x 31 scratch). */
extern "C" int fir_launch_pair(const void *x, int in_f32, int64_t x_stride, const double *histA, int64_t n1,
                               int nch, const double *d_h2A, const double *d_h2B, double *histB,
                               double *hist_tmp, double *y, int64_t y_stride, double gain, int apply_gain,
                               hipStream_t stream)
{
    if (n1 <= 0 || nch <= 0) return 0;
    dim3 grid((unsigned)((n1 + FU_T - 1) / FU_T), (unsigned)nch), block(FIR_THREADS);
    if (in_f32)
        hipLaunchKernelGGL((fir_up4_kernel<float>), grid, block, 0, stream, (const float *)x, x_stride, histA, n1,
                           d_h2A, d_h2B, histB, hist_tmp, y, y_stride, gain, apply_gain);
    else
        hipLaunchKernelGGL((fir_up4_kernel<double>), grid, block, 0, stream, (const double *)x, x_stride, histA,
                           n1, d_h2A, d_h2B, histB, hist_tmp, y, y_stride, gain, apply_gain);
    if (hipGetLastError() != hipSuccess) return -1;
    const int nh = nch * FU_H;
    hipLaunchKernelGGL(fir_copy, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, stream, hist_tmp, histB, nh);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* Launch one stage over nch channels.  in_f32 selects a float input (first
 * stage).  hist: nch * (Q0-1) doubles or NULL (zero history). */
extern "C" int fir_launch_stage(const void *x, int in_f32, int64_t x_stride, const double *hist,
                                int64_t n, int nch, int ntaps, const double *d_h2, double *y,
                                int64_t y_stride, double gain, int apply_gain, hipStream_t stream)
{
    if (n <= 0 || nch <= 0) return 0;
    dim3 grid((unsigned)((n + FIR_TILE - 1) / FIR_TILE), (unsigned)nch), block(FIR_THREADS);
#define FIR_GO(TIN, TT)                                                                         \
    hipLaunchKernelGGL((fir_up2_kernel<TIN, TT>), grid, block, 0, stream, (const TIN *)x,      \
                       x_stride, hist, n, d_h2, y, y_stride, gain, apply_gain)
    if (ntaps == 63) {
        if (in_f32) FIR_GO(float, 63); else FIR_GO(double, 63);
    } else if (ntaps == 127) {
        if (in_f32) FIR_GO(float, 127); else FIR_GO(double, 127);
    } else {
        return -1;
    }
#undef FIR_GO
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fir_launch_hist_update(const void *x, int in_f32, int64_t x_stride, int64_t n,
                                      int nch, double *hist, int H, hipStream_t stream)
{
    if (nch <= 0 || H <= 0 || H > 128) return H <= 0 ? 0 : -1;
    if (in_f32)
        hipLaunchKernelGGL(fir_hist_update<float>, dim3(nch), dim3(128), 0, stream,
                           (const float *)x, x_stride, n, hist, H);
    else
        hipLaunchKernelGGL(fir_hist_update<double>, dim3(nch), dim3(128), 0, stream,
                           (const double *)x, x_stride, n, hist, H);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
