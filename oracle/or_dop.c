/*
 * Oracle restatement of the DSD wire formats.  TEST INFRASTRUCTURE ONLY.
 * Follows src/dop.c: float <-> int24 :18-30, dop_detect(_interleaved)
 * :31-72, dop_unpack :74-87, dop_pack :89-111, dop_pack_i24 :113-136,
 * bits_unpack / bits_pack :138-192.  Pinned against the reference's own
 * dop.c (compiled into oracle/_ref) and the known answers of
 * test/test_dop.c.
 */
#include "dsd_oracle.h"

static int32_t to_i24(float f)
{
    double d = (double)f * 8388608.0;
    return d >= 0.0 ? (int32_t)(d + 0.5) : (int32_t)(d - 0.5);
}

int or_dop_detect(const float *pcm, size_t frames, int channels)
{
    if (!pcm || frames < 2 || channels < 1) return 0;
    size_t check = frames < 8 ? frames : 8;
    for (size_t i = 0; i < check; i++) {
        uint32_t u = (uint32_t)to_i24(pcm[i * (size_t)channels]) & 0x00FFFFFFu;
        if ((u >> 16) != ((i & 1) ? 0xFAu : 0x05u)) return 0;
    }
    return 1;
}

void or_dop_unpack(const float *pcm, float *bits, size_t frames)
{
    for (size_t i = 0; i < frames; i++) {
        unsigned p = (unsigned)to_i24(pcm[i]) & 0xFFFFu;
        for (int b = 15; b >= 0; b--) *bits++ = ((p >> b) & 1u) ? 1.0f : -1.0f;
    }
}

void or_dop_pack(const float *bits, float *pcm, size_t bit_count)
{
    for (size_t i = 0; i < bit_count / 16; i++) {
        unsigned p = 0;
        for (int b = 15; b >= 0; b--)
            if (*bits++ >= 0.0f) p |= 1u << b;
        int32_t v = (int32_t)(((i & 1) ? 0xFAu : 0x05u) << 16 | p);
        if (v & 0x800000) v |= (int32_t)0xFF000000u;
        pcm[i] = (float)((double)v / 8388608.0);
    }
}

int or_dop_pack_i24(const float *bits, uint8_t *out, size_t bit_count, int phase)
{
    size_t frames = bit_count / 16;
    for (size_t i = 0; i < frames; i++) {
        unsigned p = 0;
        for (int b = 15; b >= 0; b--)
            if (*bits++ >= 0.0f) p |= 1u << b;
        out[i * 3 + 0] = (uint8_t)(p & 0xFF);
        out[i * 3 + 1] = (uint8_t)(p >> 8);
        out[i * 3 + 2] = ((i + (size_t)phase) & 1) ? 0xFA : 0x05;
    }
    return (int)((frames + (size_t)phase) & 1);
}

void or_bits_unpack(const uint8_t *src, float *dst, size_t n_bits)
{
    for (size_t i = 0; i < n_bits; i++) dst[i] = ((src[i / 8] >> (7 - i % 8)) & 1) ? 1.0f : -1.0f;
}

void or_bits_pack(const float *src, uint8_t *dst, size_t n_bits)
{
    for (size_t i = 0; i < (n_bits + 7) / 8; i++) dst[i] = 0;
    for (size_t i = 0; i < n_bits; i++)
        if (src[i] >= 0.0f) dst[i / 8] |= (uint8_t)(1u << (7 - i % 8));
}
