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

} // namespace

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
