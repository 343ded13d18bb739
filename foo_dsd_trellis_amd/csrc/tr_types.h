/*
 * Device-resident trellis state and job descriptors shared by the HIP
 * kernels and the host library.
 *
 * tr_ctx is the GPU form of the reference's sdm_context_t
 * (include/trellis.h:31-59): the active survivor list of the current bank
 * (state, cost, path, traceback bit, history), the per-slot path-hash links
 * of both banks (which persist between steps in the reference and can be
 * followed again, src/trellis.c:599-619), and the per-slot path/cost of both
 * banks (read when another context that was copied from this one follows a
 * link into it, src/trellis.c:1064-1092).
 */
#ifndef TR_TYPES_H
#define TR_TYPES_H

#include <stdint.h>

#define TR_MAXK 32
#define TR_MAXP 8
#define TR_MAXL 2048
#define TR_HW_MAX (TR_MAXL / 64)
#define TR_SLOTS 64

/* path-hash link encoding (int8): -1 = none, 0..63 = slot of the same bank
 * of this context, 64..127 = slot (v - 64) of the same bank of the context
 * this one was copied from (its "snapshot"). */
#define TR_LINK_NONE (-1)
#define TR_LINK_FOREIGN 64

typedef struct tr_ctx {
    int32_t num_cands, pos, pending, idx;
    int32_t draining, has_foreign, err, pad;
    double st[TR_MAXK][TR_MAXP];
    double cost[TR_MAXK];
    uint32_t path[TR_MAXK];
    uint8_t next[TR_MAXK];
    uint8_t slot[TR_MAXK];      /* slot of each survivor in the active bank: the
                                   child index 2*parent_rank + branch of the
                                   step that made it (the node act[] points at) */
    int8_t link[2][TR_SLOTS];
    uint32_t node_path[2][TR_SLOTS];
    double node_cost[2][TR_SLOTS];
    uint64_t hist[TR_MAXK][TR_HW_MAX];
} tr_ctx;

typedef struct tr_params {
    double a[TR_MAXP];
    double g[TR_MAXP];
    double state_limit;
    int32_t P, K, L, B;
    uint32_t mask;
    int32_t force_slow;     /* validation: decide every step with the literal prune */
    int32_t fma_o8;         /* order 8: the 4-candidate FMA filter of the AVX2 build */
    int32_t wave_only;      /* K <= 2 too on the one-wave-per-job kernel (no lane-per-job form) */
} tr_params;

/* One unit of trellis work: run `n_in` inputs through a context.
 * Outputs produced at input index >= keep_from are stored packed (bit = 1
 * for +1.0) starting at bit 0 of `out`. */
typedef struct tr_job {
    const double *in;
    uint64_t *out;
    int64_t n_in;
    int64_t keep_from;
    int32_t ctx_in;        /* context to start from */
    int32_t ctx_out;       /* context to store the final state into, -1: none */
    int32_t snap;          /* context followed by foreign links, -1: none */
    int32_t copy;          /* start from a copy of ctx_in (links become foreign) */
    int32_t seed;          /* overwrite every survivor state with seed_st, cost 0 */
    int32_t skip;          /* unused slot of a fixed job table: not run */
    double seed_st[TR_MAXP];
    int64_t kept;          /* out: outputs stored */
    int64_t slow_steps;    /* out: steps decided by the literal prune */
    int32_t err;           /* out: 1 non-finite metric, 2 unsupported link chain */
    int32_t pad2;
} tr_job;

/* Host-fed launch (tr_run_feed): the host fills job descriptors in pinned,
 * coherent host memory while the launch runs and publishes how many are
 * complete; the wave of job j waits until the host has published it.  Host
 * side (coherent pinned memory, written by the host only): */
typedef struct tr_feed_host {
    int32_t ready;         /* descriptors 0..ready-1 are complete */
    int32_t abort;         /* nonzero: leave without taking further jobs */
} tr_feed_host;

/* Device side (device memory, reset by tr_feed_init before each launch): */
typedef struct tr_feed_dev {
    int32_t done;          /* jobs run to the end */
    int32_t ready;         /* last published count seen by any wave */
    int32_t abort;         /* the host's abort flag, seen by a wave */
    int32_t err;           /* 1: a wave gave up waiting (time limit) */
    uint64_t last_poll;    /* wall clock of the last read of the host word */
    uint64_t t0;           /* wall clock when the first wave started */
} tr_feed_dev;

#endif
