// Host side of the DSD wire formats: DoP detection on host frames and the
// launches of the device pack/unpack kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dsdhip_internal.h"

extern "C" {
__global__ void dop_unpack_k(const float *, int64_t, int, unsigned long long *, int64_t);
__global__ void dop_pack_k(const unsigned long long *, int64_t, int, int64_t, float *);
__global__ void dop_pack_i24_k(const unsigned long long *, int64_t, int, int64_t, int, uint8_t *);
__global__ void bytes_unpack_k(const uint8_t *, int64_t, int, int64_t, unsigned long long *, int64_t);
__global__ void bytes_pack_k(const unsigned long long *, int64_t, int, int64_t, uint8_t *, int64_t);
__global__ void bits_to_f32_k(const unsigned long long *, int64_t, int, int64_t, float *, int64_t);
__global__ void f32_to_bits_k(const float *, int64_t, int, int64_t, unsigned long long *, int64_t);
}

static int launched(void) { return hipGetLastError() == hipSuccess ? 0 : -1; }

static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" {

/* dop_detect_interleaved (src/dop.c:54-72): the first (up to) 8 frames of
 * channel 0 carry alternating markers 0x05 / 0xFA in their top byte. */
int dsdhip_dop_detect(const float *pcm24, size_t frames, int channels)
{
    if (!pcm24 || frames < 2 || channels < 1) return 0;
    const size_t check = frames < 8 ? frames : 8;
    for (size_t i = 0; i < check; i++) {
        const double d = (double)pcm24[i * (size_t)channels] * 8388608.0;
        const int32_t v = d >= 0.0 ? (int32_t)(d + 0.5) : (int32_t)(d - 0.5);
        const unsigned marker = ((uint32_t)v & 0x00FFFFFFu) >> 16;
        if (marker != ((i & 1) ? 0xFAu : 0x05u)) return 0;
    }
    return 1;
}

int dsdhip_dop_unpack_dev(const float *d_pcm24, size_t frames, int nch, unsigned long long *d_bits,
                          int64_t words_stride, void *stream)
{
    if (!d_pcm24 || !d_bits || nch < 1) return -1;
    if (frames == 0) return 0;
    const int64_t nw = ((int64_t)frames * 16 + 63) / 64;
    if (words_stride < nw) return -1;
    hipLaunchKernelGGL(dop_unpack_k, dim3(blocks_for(nw), (unsigned)nch), dim3(256), 0, (hipStream_t)stream,
                       d_pcm24, (int64_t)frames, nch, d_bits, words_stride);
    return launched();
}

int dsdhip_dop_pack_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t bit_count,
                        float *d_pcm24, void *stream)
{
    if (!d_bits || !d_pcm24 || nch < 1) return -1;
    const int64_t frames = (int64_t)(bit_count / 16);
    if (frames == 0) return 0;
    if (words_stride < ((int64_t)bit_count + 63) / 64) return -1;
    hipLaunchKernelGGL(dop_pack_k, dim3(blocks_for(frames), (unsigned)nch), dim3(256), 0, (hipStream_t)stream,
                       d_bits, words_stride, nch, frames, d_pcm24);
    return launched();
}

int dsdhip_dop_pack_i24_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t bit_count,
                            int marker_phase, unsigned char *d_out, void *stream)
{
    if (!d_bits || !d_out || nch < 1) return -1;
    const int64_t frames = (int64_t)(bit_count / 16);
    if (frames == 0) return marker_phase & 1;
    if (words_stride < ((int64_t)bit_count + 63) / 64) return -1;
    hipLaunchKernelGGL(dop_pack_i24_k, dim3(blocks_for(frames), (unsigned)nch), dim3(256), 0,
                       (hipStream_t)stream, d_bits, words_stride, nch, frames, marker_phase & 1, d_out);
    if (launched() != 0) return -1;
    return (int)((frames + (marker_phase & 1)) & 1);
}

int dsdhip_bytes_unpack_dev(const unsigned char *d_src, int64_t src_stride, int nch, size_t n_bits,
                            unsigned long long *d_bits, int64_t words_stride, void *stream)
{
    if (!d_src || !d_bits || nch < 1) return -1;
    if (n_bits == 0) return 0;
    const int64_t nw = ((int64_t)n_bits + 63) / 64;
    if (words_stride < nw || src_stride < ((int64_t)n_bits + 7) / 8) return -1;
    hipLaunchKernelGGL(bytes_unpack_k, dim3(blocks_for(nw), (unsigned)nch), dim3(256), 0, (hipStream_t)stream,
                       d_src, src_stride, nch, (int64_t)n_bits, d_bits, words_stride);
    return launched();
}

int dsdhip_bytes_pack_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t n_bits,
                          unsigned char *d_dst, int64_t dst_stride, void *stream)
{
    if (!d_bits || !d_dst || nch < 1) return -1;
    if (n_bits == 0) return 0;
    const int64_t nbytes = ((int64_t)n_bits + 7) / 8;
    if (words_stride < ((int64_t)n_bits + 63) / 64 || dst_stride < nbytes) return -1;
    hipLaunchKernelGGL(bytes_pack_k, dim3(blocks_for(nbytes), (unsigned)nch), dim3(256), 0, (hipStream_t)stream,
                       d_bits, words_stride, nch, (int64_t)n_bits, d_dst, dst_stride);
    return launched();
}

int dsdhip_bits_to_f32_dev(const unsigned long long *d_bits, int64_t words_stride, int nch, size_t n,
                           float *d_out, int64_t out_stride, void *stream)
{
    if (!d_bits || !d_out || nch < 1) return -1;
    if (n == 0) return 0;
    if (words_stride < ((int64_t)n + 63) / 64 || out_stride < (int64_t)n) return -1;
    hipLaunchKernelGGL(bits_to_f32_k, dim3(blocks_for((int64_t)n), (unsigned)nch), dim3(256), 0,
                       (hipStream_t)stream, d_bits, words_stride, nch, (int64_t)n, d_out, out_stride);
    return launched();
}

int dsdhip_f32_to_bits_dev(const float *d_in, int64_t in_stride, int nch, size_t n, unsigned long long *d_bits,
                           int64_t words_stride, void *stream)
{
    if (!d_in || !d_bits || nch < 1) return -1;
    if (n == 0) return 0;
    const int64_t nw = ((int64_t)n + 63) / 64;
    if (words_stride < nw || in_stride < (int64_t)n) return -1;
    hipLaunchKernelGGL(f32_to_bits_k, dim3(blocks_for(nw), (unsigned)nch), dim3(256), 0, (hipStream_t)stream, d_in,
                       in_stride, nch, (int64_t)n, d_bits, words_stride);
    return launched();
}

}  // extern "C"
