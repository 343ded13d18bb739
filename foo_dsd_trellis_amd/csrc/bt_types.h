/* Device-resident descriptors of the batched DAS pipeline, shared by the
 * kernels (batch_kernel.hip) and the host library (dsdhip_batch.cpp). */
#ifndef BT_TYPES_H
#define BT_TYPES_H

#include <stdint.h>

#define BT_MAXQ 32        /* stitch candidates per boundary (das_stitch) */

struct bt_track {
    const float *pcm;       /* [nch][pcm_stride] frames */
    int64_t pcm_stride;
    int64_t frames;
    unsigned long long *out;               /* [nch][out_stride] packed output */
    int64_t out_stride;
    int64_t out_words;
    int64_t out_pos;        /* bits written so far (every channel) */
};

struct bt_cfg {
    int ntracks, nch, smax, L, P, reposition;
    int64_t CF;             /* PCM frames per chunk */
    int64_t R;              /* DSD samples per PCM frame */
    int64_t seg_len, O;
    int64_t fstride;        /* doubles per stream in the FIR output buffer */
    int64_t seg_words;      /* words per (stream, segment) output region */
    int64_t asm_words;      /* output words one chunk can touch per stream */
    int nstreams;
};

struct bt_state {
    int64_t chunk;
    int32_t queue;
    int32_t err;            /* bit 0: plan, 1: trellis job, 2: stitch precondition, 3: output full */
};

#endif
