// Pre-SDM conditioning of the engine (src/engine.c) on the device:
//   - boxcar of a DSD stream (packed bits), times the gain ramp
//     (src/engine.c:533-546, 829-841): the window sums of +-1 samples are
//     small integers, so each output is computed independently and exactly;
//   - boxcar pre-smooth of float PCM before rate conversion
//     (src/engine.c:629-642): a running double sum whose rounding depends on
//     every earlier sample, so it advances sequentially, one wave per channel;
//   - 127-tap lowpass of a DSD stream (src/fir.c:795-873) times the ramp
//     (src/engine.c:505-525, 813-828);
//   - gain ramp (src/engine.c:646-652, 877-888) and the tanh soft clip
//     (src/engine.c:704-716).
// All arithmetic is FP64 in the reference's operation order with no
// contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;

namespace {

/* 64 consecutive samples of the virtual stream hist ++ chunk starting at
 * sample s (s >= -64; hist holds samples -64..-1 as bits 0..63, bits past
 * the chunk's end read as 0). */
__device__ __forceinline__ u64 window64(const u64 *row, u64 hist, int64_t n, int64_t s)
{
    const int64_t nw = (n + 63) >> 6;
    const int64_t w0 = (s >= 0) ? (s >> 6) : -1;
    const int sh = (int)(s - w0 * 64);
    const u64 a = (w0 < 0) ? hist : row[w0];
    const u64 b = (w0 + 1 < nw) ? row[w0 + 1] : 0ull;
    return sh ? ((a >> sh) | (b << (64 - sh))) : a;
}

/* tanh from +,-,*,/,fma,rint,ldexp only (the same sequence as the host
 * restatement): m = expm1(2|z|) by Cody-Waite reduction and a degree-13
 * Taylor series (Estrin), tanh|z| = m / (m + 2). */
__device__ double cond_tanh(double z)
{
    double x = fabs(z) * 2.0;
    if (x != x) return z;
    if (x > 40.0) x = 40.0;
    const double kd = rint(x * 0x1.71547652b82fep+0);
    double r = fma(-kd, 0x1.62e42fee00000p-1, x);
    r = fma(-kd, 0x1.a39ef35793c76p-33, r);
    const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    const double e0 = fma(0x1.5555555555555p-3, r, 0x1.0000000000000p-1);
    const double e1 = fma(0x1.1111111111111p-7, r, 0x1.5555555555555p-5);
    const double e2 = fma(0x1.a01a01a01a01ap-13, r, 0x1.6c16c16c16c17p-10);
    const double e3 = fma(0x1.71de3a556c734p-19, r, 0x1.a01a01a01a01ap-16);
    const double e4 = fma(0x1.ae64567f544e4p-26, r, 0x1.27e4fb7789f5cp-22);
    const double e5 = fma(0x1.6124613a86d09p-33, r, 0x1.1eed8eff8d898p-29);
    const double f0 = fma(e1, r2, e0), f1 = fma(e3, r2, e2), f2 = fma(e5, r2, e4);
    const double q = fma(f2, r8, fma(f1, r4, f0));
    const double p = fma(r2, q, r);
    const int k = (int)kd;
    const double m = ldexp(p, k) + (ldexp(1.0, k) - 1.0);
    const double t = m / (m + 2.0);
    return copysign(t, z);
}

}  // namespace

/* y[i] = (sum * inv_n) * (gs + gstep * i), sum = the window sum of the taps
 * samples ending at i (samples before the stream's start count 0). */
extern "C" __global__ void cond_boxcar_bits_k(const u64 *__restrict__ bits, int64_t stride, int nch, int64_t n,
                                              const u64 *__restrict__ hist, int64_t done, int taps, double inv_n,
                                              double gs, double gstep, double *__restrict__ out,
                                              int64_t out_stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nch) return;
    const u64 win = window64(bits + (int64_t)c * stride, hist[c], n, i - taps + 1);
    const u64 m = (taps == 64) ? ~0ull : ((1ull << taps) - 1ull);
    const int64_t g = done + i;                       /* global sample index */
    const int present = (g + 1 < taps) ? (int)(g + 1) : taps;
    const int ones = __popcll(win & m);
    const double sum = (double)(2 * ones - present);
    out[(int64_t)c * out_stride + i] = (sum * inv_n) * (gs + gstep * (double)i);
}

/* hist <- the last 64 samples of hist ++ chunk (after the chunk). */
extern "C" __global__ void cond_bits_hist_k(const u64 *__restrict__ bits, int64_t stride, int nch, int64_t n,
                                            u64 *__restrict__ hist)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    hist[c] = window64(bits + (int64_t)c * stride, hist[c], n, n - 64);
}

/* Running-sum boxcar over float PCM widened to double: sum -= x[i - taps];
 * sum += x[i]; y = sum * inv_n.  One wave per channel: the lanes load a
 * block of 64 inputs (and the inputs taps earlier) a block ahead, and the
 * sum advances through the block in order.  state: [c][taps + 1] = the
 * previous taps inputs (oldest first), then the sum. */
extern "C" __global__ __launch_bounds__(64) void cond_boxcar_pcm_k(const float *__restrict__ in, int64_t in_stride,
                                                                   int nch, int64_t n, int taps, double inv_n,
                                                                   double *__restrict__ state,
                                                                   double *__restrict__ out, int64_t out_stride)
{
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= nch) return;
    const float *x = in + (int64_t)c * in_stride;
    double *st = state + (int64_t)c * (taps + 1);
    double *y = out + (int64_t)c * out_stride;
    auto load = [&](int64_t i, double &xn, double &xo) {
        xn = (i < n) ? (double)x[i] : 0.0;
        const int64_t k = i - taps;
        xo = (i < n) ? ((k >= 0) ? (double)x[k] : st[k + taps]) : 0.0;
    };
    double sum = st[taps];
    double xn, xo, nn, no;
    load(lane, xn, xo);
    for (int64_t b = 0; b < n; b += 64) {
        load(b + 64 + lane, nn, no);                  /* next block, in flight meanwhile */
        double yv = 0.0;
        const int m = (n - b < 64) ? (int)(n - b) : 64;
        for (int j = 0; j < m; j++) {
            const double o = __shfl(xo, j, 64), v = __shfl(xn, j, 64);
            sum -= o;
            sum += v;
            const double r = sum * inv_n;
            yv = (lane == j) ? r : yv;
        }
        if (lane < m) y[b + lane] = yv;
        xn = nn;
        xo = no;
    }
    __syncthreads();
    /* keep the last taps inputs of hist ++ chunk */
    if (lane == 0) {
        if (n >= taps) {
            for (int j = 0; j < taps; j++) st[j] = (double)x[n - taps + j];
        } else {
            for (int j = 0; j < taps - n; j++) st[j] = st[j + n];
            for (int64_t j = 0; j < n; j++) st[taps - n + j] = (double)x[j];
        }
        st[taps] = sum;
    }
}

/* x[i] *= gs + step * i */
extern "C" __global__ void cond_ramp_k(double *__restrict__ x, int64_t stride, int nch, int64_t n, double gs,
                                       double step)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nch) return;
    double *p = x + (int64_t)c * stride + i;
    *p = *p * (gs + step * (double)i);
}

/* y[i] = (float)x[i] (round to nearest) */
extern "C" __global__ void cond_narrow_k(const double *__restrict__ x, float *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

/* soft clip above 0.95 with a tanh knee */
extern "C" __global__ void cond_clip_k(double *__restrict__ x, int64_t stride, int nch, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nch) return;
    const double th = 0.95;
    const double inv_knee = 1.0 / (1.0 - th);
    double *p = x + (int64_t)c * stride + i;
    const double v = *p;
    if (v > th)
        *p = th + (1.0 - th) * cond_tanh((v - th) * inv_knee);
    else if (v < -th)
        *p = -th - (1.0 - th) * cond_tanh((-v - th) * inv_knee);
}

/* Lowpass of a DSD stream: y[i] = (sum_k h[k] s[i-k]) * (gs + gstep * i),
 * k ascending, s = +-1 (0 before the stream's start); ntaps <= 129.  hist:
 * [c][2] = samples -128..-1 as bits (word 0 = -128..-65). */
extern "C" __global__ void cond_lowpass_bits_k(const u64 *__restrict__ bits, int64_t stride, int nch, int64_t n,
                                               const u64 *__restrict__ hist, int64_t done,
                                               const double *__restrict__ h, int ntaps, double gs, double gstep,
                                               double *__restrict__ out, int64_t out_stride)
{
    __shared__ double sh[129];
    for (int k = threadIdx.x; k < ntaps; k += blockDim.x) sh[k] = h[k];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nch) return;
    const u64 *row = bits + (int64_t)c * stride;
    const int64_t nw = (n + 63) >> 6;
    double acc = 0.0;
    for (int k = 0; k < ntaps; k++) {
        const int64_t j = i - k;
        if (done + j < 0) break;                       /* before the stream's start: s = 0 */
        u64 w;
        if (j >= 0) w = (j >> 6) < nw ? row[j >> 6] : 0ull;
        else w = hist[(int64_t)c * 2 + ((j + 128) >> 6)];
        const int b = (int)((j + 128) & 63);
        acc = acc + (((w >> b) & 1ull) ? sh[k] : -sh[k]);
    }
    out[(int64_t)c * out_stride + i] = acc * (gs + gstep * (double)i);
}

/* hist <- the last 128 samples of hist ++ chunk. */
extern "C" __global__ void cond_bits_hist2_k(const u64 *__restrict__ bits, int64_t stride, int nch, int64_t n,
                                             u64 *__restrict__ hist)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nch) return;
    const u64 *row = bits + (int64_t)c * stride;
    u64 h0 = hist[c * 2], h1 = hist[c * 2 + 1];
    /* a 192-sample virtual stream would be needed for n < 128; shift by
     * whole 64-sample steps and a remainder */
    u64 w[2];
    for (int q = 0; q < 2; q++) {
        const int64_t s = n - 128 + 64 * q;           /* first sample of new word q */
        u64 v = 0;
        for (int b = 0; b < 64; b++) {
            const int64_t j = s + b;
            int bit;
            if (j >= 0) bit = (int)((row[j >> 6] >> (j & 63)) & 1ull);
            else {
                const int64_t k = j + 128;                /* index into the old history */
                bit = (k >= 0) ? (int)((((k >> 6) ? h1 : h0) >> (k & 63)) & 1ull) : 0;
            }
            v |= (u64)bit << b;
        }
        w[q] = v;
    }
    hist[c * 2] = w[0];
    hist[c * 2 + 1] = w[1];
}
