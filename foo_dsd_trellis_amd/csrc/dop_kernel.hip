// DSD wire formats on the device: DoP (DSD over 24-bit PCM) and raw
// MSB-first bit bytes, to and from the packed-bit streams the trellis and
// DAS produce (sample i of a channel = bit i % 64 of word i / 64, 1 = +1.0).
//
// Reference formats: src/dop.c:18-30 (float <-> int24), :74-87 (unpack:
// 16 DSD bits per frame, MSB first, in the low 16 bits of the int24),
// :89-111 (pack to float frames, marker 0x05 / 0xFA alternating from the
// first frame), :113-136 (pack to little-endian 24-bit with a marker phase
// carried between calls), :138-192 (raw bit bytes, MSB first).
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;

namespace {

__device__ __forceinline__ int32_t dop_int24(float f)
{
    const double d = (double)f * 8388608.0;
    return d >= 0.0 ? (int32_t)(d + 0.5) : (int32_t)(d - 0.5);
}

/* 16 samples starting at sample s (s a multiple of 16) as the DoP payload:
 * the first sample in bit 15 */
__device__ __forceinline__ unsigned dop_payload(const u64 *row, int64_t s)
{
    const unsigned chunk = (unsigned)(row[s >> 6] >> (s & 63)) & 0xFFFFu;
    return __builtin_bitreverse32(chunk) >> 16;
}

}  // namespace

/* DoP float frames, interleaved [frames][nch] -> packed bits [nch][stride]:
 * one thread per output word (4 frames) and channel. */
extern "C" __global__ void dop_unpack_k(const float *__restrict__ pcm, int64_t frames, int nch,
                                        u64 *__restrict__ bits, int64_t stride)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int64_t nw = (frames * 16 + 63) / 64;
    if (w >= nw || c >= nch) return;
    u64 out = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int64_t f = w * 4 + q;
        if (f >= frames) break;
        const unsigned p = (unsigned)dop_int24(pcm[f * nch + c]) & 0xFFFFu;
        const unsigned r = __builtin_bitreverse32(p) >> 16;        /* first sample to bit 0 */
        out |= (u64)r << (16 * q);
    }
    bits[(int64_t)c * stride + w] = out;
}

/* packed bits -> DoP float frames, interleaved; markers from frame 0. */
extern "C" __global__ void dop_pack_k(const u64 *__restrict__ bits, int64_t stride, int nch, int64_t frames,
                                      float *__restrict__ pcm)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (f >= frames || c >= nch) return;
    const unsigned p = dop_payload(bits + (int64_t)c * stride, f * 16);
    const unsigned marker = (f & 1) ? 0xFAu : 0x05u;
    int32_t v = (int32_t)((marker << 16) | p);
    if (v & 0x800000) v |= (int32_t)0xFF000000u;
    pcm[f * nch + c] = (float)((double)v / 8388608.0);
}

/* packed bits -> DoP little-endian 24-bit frames, interleaved
 * [frames][nch][3]; marker of frame f from (f + phase) & 1. */
extern "C" __global__ void dop_pack_i24_k(const u64 *__restrict__ bits, int64_t stride, int nch,
                                          int64_t frames, int phase, uint8_t *__restrict__ out)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (f >= frames || c >= nch) return;
    const unsigned p = dop_payload(bits + (int64_t)c * stride, f * 16);
    uint8_t *o = out + (f * nch + c) * 3;
    o[0] = (uint8_t)(p & 0xFF);
    o[1] = (uint8_t)(p >> 8);
    o[2] = ((f + phase) & 1) ? 0xFA : 0x05;
}

/* raw MSB-first bytes [nch][src_stride] -> packed bits; n_bits per channel. */
extern "C" __global__ void bytes_unpack_k(const uint8_t *__restrict__ src, int64_t src_stride, int nch,
                                          int64_t n_bits, u64 *__restrict__ bits, int64_t stride)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int64_t nw = (n_bits + 63) / 64;
    if (w >= nw || c >= nch) return;
    const uint8_t *s = src + (int64_t)c * src_stride;
    const int64_t nbytes = (n_bits + 7) / 8;
    u64 out = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int64_t b = w * 8 + q;
        if (b >= nbytes) break;
        const unsigned r = __builtin_bitreverse32((unsigned)s[b]) >> 24;
        out |= (u64)r << (8 * q);
    }
    const int64_t rem = n_bits - w * 64;
    if (rem < 64) out &= (rem > 0) ? ((1ull << rem) - 1ull) : 0ull;
    bits[(int64_t)c * stride + w] = out;
}

/* packed bits -> raw MSB-first bytes; a partial last byte keeps its bits
 * in the high positions and zeros below. */
extern "C" __global__ void bytes_pack_k(const u64 *__restrict__ bits, int64_t stride, int nch, int64_t n_bits,
                                        uint8_t *__restrict__ dst, int64_t dst_stride)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int64_t nbytes = (n_bits + 7) / 8;
    if (b >= nbytes || c >= nch) return;
    const u64 w = bits[(int64_t)c * stride + (b >> 3)];
    unsigned v = (unsigned)(w >> ((b & 7) * 8)) & 0xFFu;
    const int64_t rem = n_bits - b * 8;
    if (rem < 8) v &= (1u << rem) - 1u;
    dst[(int64_t)c * dst_stride + b] = (uint8_t)(__builtin_bitreverse32(v) >> 24);
}

/* packed bits -> +-1.0f samples [nch][out_stride] (the FIR chain's input). */
extern "C" __global__ void bits_to_f32_k(const u64 *__restrict__ bits, int64_t stride, int nch, int64_t n,
                                         float *__restrict__ out, int64_t out_stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= nch) return;
    const u64 w = bits[(int64_t)c * stride + (i >> 6)];
    out[(int64_t)c * out_stride + i] = ((w >> (i & 63)) & 1ull) ? 1.0f : -1.0f;
}

/* +-1.0f samples [nch][in_stride] -> packed bits (x >= 0 -> 1, as the
 * reference's DSD paths read their float input, src/engine.c:515). */
extern "C" __global__ void f32_to_bits_k(const float *__restrict__ in, int64_t in_stride, int nch, int64_t n,
                                         u64 *__restrict__ bits, int64_t stride)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int64_t nw = (n + 63) / 64;
    if (w >= nw || c >= nch) return;
    const float *x = in + (int64_t)c * in_stride;
    u64 out = 0;
    for (int b = 0; b < 64; b++) {
        const int64_t i = w * 64 + b;
        if (i >= n) break;
        if (x[i] >= 0.0f) out |= 1ull << b;
    }
    bits[(int64_t)c * stride + w] = out;
}
