// Uniformly partitioned overlap-save convolution (UPOLS) kernels for gfx950.
//
// Reference semantics: src/convolution.c:484-524 (process_partition) and
// :893-958 (conv_process_direct): blocks of P samples, FFT size 2P over
// [previous block | this block], a frequency-domain delay line (FDL) of the
// input spectra, Y_b = sum_p X_(b-p) . H_p, inverse FFT, output = the second
// half.  Here a whole chunk of nb blocks is processed at once: the forward
// and inverse FFTs are batched library FFTs (real-to-complex, P+1 bins), and
// the spectral multiply-accumulate over all blocks of the chunk is the
// hand-written kernel below: for every bin it is a 1-D convolution along the
// block axis, register-blocked as TB output blocks x TB partitions per pass.
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CV_TB 8

// Frames for the forward FFT: frame b of channel c = [x(b-1) | x(b)] where
// x(-1) is the previous chunk's last block (tail).  f32, [c][b][2P].
extern "C" __global__ void cv_pack(const double *__restrict__ in, int64_t in_stride,
                                   const double *__restrict__ tail, int P, int nb, int nch,
                                   float *__restrict__ frames)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // sample within the chunk
    const int c = blockIdx.y;
    const int64_t n = (int64_t)nb * P;
    if (i >= n || c >= nch) return;
    const double v = in[(int64_t)c * in_stride + i];
    const int b = (int)(i / P), j = (int)(i - (int64_t)b * P);
    float *fc = frames + (int64_t)c * nb * 2 * P;
    fc[(int64_t)b * 2 * P + P + j] = (float)v;            // this block, second half of frame b
    if (b + 1 < nb) fc[(int64_t)(b + 1) * 2 * P + j] = (float)v;   // first half of frame b+1
    if (b == 0) fc[j] = (float)tail[(int64_t)c * P + j];
}

// Tail for the next chunk: the chunk's last block.
extern "C" __global__ void cv_tail(const double *__restrict__ in, int64_t in_stride, int P, int nb,
                                   int nch, double *__restrict__ tail)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (j >= P || c >= nch) return;
    tail[(int64_t)c * P + j] = in[(int64_t)c * in_stride + (int64_t)(nb - 1) * P + j];
}

// Spectral multiply-accumulate.  S = [history (np-1 spectra) | chunk (nb
// spectra)] per channel, Y_b = sum_{p<np} S[np-1+b-p] * H_p for the nb blocks
// of the chunk.  hist: [c][np-1][K], X: [c][nb][K], H: [c][np][K], Y:
// [c][nb][K] (float2 = complex), K = P+1 bins.  One thread per (bin, tile of
// CV_TB blocks, channel); partitions in passes of CV_TB with a window of
// 2*CV_TB-1 spectra in registers.  np is a multiple of CV_TB; nb is any
// count (a partial last tile reads zeros past the chunk and stores only the
// chunk's blocks).
__device__ __forceinline__ float2 cv_S(const float2 *__restrict__ hist, const float2 *__restrict__ X,
                                       int idx, int nh, int K, int k)
{
    return idx < nh ? hist[(int64_t)idx * K + k] : X[(int64_t)(idx - nh) * K + k];
}
__device__ __forceinline__ float2 cv_Sb(const float2 *__restrict__ hist, const float2 *__restrict__ X,
                                        int idx, int nh, int nall, int K, int k)
{
    return idx < nall ? cv_S(hist, X, idx, nh, K, k) : make_float2(0.0f, 0.0f);
}

extern "C" __global__ __launch_bounds__(256) void cv_mac(const float2 *__restrict__ hist,
                                                         const float2 *__restrict__ X,
                                                         const float2 *__restrict__ H, int K, int np,
                                                         int nb, int nch, float2 *__restrict__ Y)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int b0 = blockIdx.y * CV_TB;
    const int c = blockIdx.z;
    if (k >= K || b0 >= nb || c >= nch) return;
    const int nh = np - 1;
    const float2 *hc = hist + (int64_t)c * nh * K;
    const float2 *xc = X + (int64_t)c * nb * K;
    const float2 *Hc = H + (int64_t)c * np * K;
    float ar[CV_TB], ai[CV_TB];
#pragma unroll
    for (int t = 0; t < CV_TB; t++) ar[t] = ai[t] = 0.0f;
    for (int p0 = 0; p0 < np; p0 += CV_TB) {
        // window w[u] = S[nh + b0 - p0 - (CV_TB-1) + u], u = 0 .. 2*CV_TB-2
        float2 w[2 * CV_TB - 1];
        const int base = nh + b0 - p0 - (CV_TB - 1);
#pragma unroll
        for (int u = 0; u < 2 * CV_TB - 1; u++) w[u] = cv_Sb(hc, xc, base + u, nh, nh + nb, K, k);
#pragma unroll
        for (int q = 0; q < CV_TB; q++) {
            const float2 h = Hc[(int64_t)(p0 + q) * K + k];
#pragma unroll
            for (int t = 0; t < CV_TB; t++) {
                // S[nh + b0 + t - (p0 + q)] = w[t - q + CV_TB - 1]
                const float2 s = w[t - q + CV_TB - 1];
                ar[t] = fmaf(s.x, h.x, ar[t]);
                ar[t] = fmaf(-s.y, h.y, ar[t]);
                ai[t] = fmaf(s.x, h.y, ai[t]);
                ai[t] = fmaf(s.y, h.x, ai[t]);
            }
        }
    }
    float2 *yc = Y + (int64_t)c * nb * K;
#pragma unroll
    for (int t = 0; t < CV_TB; t++)
        if (b0 + t < nb) yc[(int64_t)(b0 + t) * K + k] = make_float2(ar[t], ai[t]);
}

// FDL history for the next chunk: the last np-1 spectra of [hist | X].
extern "C" __global__ void cv_hist(const float2 *__restrict__ hist, const float2 *__restrict__ X,
                                   int K, int np, int nb, int nch, float2 *__restrict__ hist_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;                 // slot 0 .. np-2
    const int c = blockIdx.z;
    const int nh = np - 1;
    if (k >= K || j >= nh || c >= nch) return;
    const int idx = nb + j;                   // index in the concatenation (length nh + nb)
    hist_out[((int64_t)c * nh + j) * K + k] =
        cv_S(hist + (int64_t)c * nh * K, X + (int64_t)c * nb * K, idx, nh, K, k);
}

// Output: second half of every inverse-transformed frame, scaled by 1/(2P).
extern "C" __global__ void cv_unpack(const float *__restrict__ frames, int P, int nb, int nch,
                                     double *__restrict__ out, int64_t out_stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int64_t n = (int64_t)nb * P;
    if (i >= n || c >= nch) return;
    const int b = (int)(i / P), j = (int)(i - (int64_t)b * P);
    const float v = frames[((int64_t)c * nb + b) * 2 * P + P + j];
    out[(int64_t)c * out_stride + i] = (double)v * (1.0 / (2.0 * (double)P));
}

// ---- FIFO form for any chunk length (src/convolution.c:893-958) ----------
// Gather of two sources by position: out[c][i] = i < na ? a[c][i] : b[c][i - na]
// for i < n (rows of the given strides).
extern "C" __global__ void cv_cat2(const double *__restrict__ a, int64_t a_stride, int64_t na,
                                   const double *__restrict__ b, int64_t b_stride, int64_t n, int nch,
                                   double *__restrict__ out, int64_t out_stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nch) return;
    out[(int64_t)c * out_stride + i] = i < na ? a[(int64_t)c * a_stride + i] : b[(int64_t)c * b_stride + (i - na)];
}
