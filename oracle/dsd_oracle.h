/*
 * CPU oracle for the PCM->DSD hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's algorithm, used as the
 * checker for the HIP implementation.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links or calls it.
 *
 * Pinning status per function (how each is checked, see tests/):
 *   or_ntf_*            table vs the compiled reference ntf.c (oracle/_ref)
 *                       and committed fixtures.
 *   or_sdm_*            bit-exact vs the compiled reference trellis.c on
 *                       seeded inputs (fixtures + live comparison), plus the
 *                       reference's own invariants (test/test_trellis.c).
 *   or_estimate_state   bit-exact vs reference sdm_estimate_state.
 *   or_fir_*            closed form (impulse response == taps, DC gain,
 *                       passband/stopband), long-double truth; the exact IPP
 *                       sample values are parity unpinned (IPP is absent).
 *   or_das_*            invariants (S=1 / Par2 seg0-first == sequential,
 *                       output count, stitch brute force on tiny cases);
 *                       the reference DAS driver (src/dsp_plugin.c) does not
 *                       build here, so multi-segment stitching beyond those
 *                       invariants is parity unpinned.
 */
#ifndef DSD_ORACLE_H
#define DSD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_ORDER 10
#define OR_MAXK 32
#define OR_MAXL 2048
#define OR_NODES 128           /* child slots per bank (4 levels x 32) */
#define OR_HASH 128
#define OR_HWORDS (OR_MAXL / 64)

typedef struct {
    double a[OR_MAX_ORDER];
    double g[OR_MAX_ORDER];
    int order;
    unsigned freq;
    const char *name;
} or_ntf_t;

int or_ntf_count(void);
const or_ntf_t *or_ntf_get(int index);
/* Index of the filter with reference id (NTF_CLANS_4=0 .. NTF_SDM_8=9) for a
 * DSD rate, -1 if none (src/ntf.c:814-830). */
int or_ntf_lookup(int id, unsigned rate);
/* Auto-selected filter index for a rate (src/ntf.c:839-855). */
int or_ntf_auto(unsigned rate);

/* One trellis node (src/trellis.h:18-28 semantics; history is carried by
 * value per survivor, see or_ctx.hist). */
typedef struct or_node {
    double st[OR_MAX_ORDER];
    double cost;
    uint32_t path;
    uint8_t next;
    uint8_t parent;            /* position of the parent in the previous act list */
    struct or_node *plist;     /* hash chain link, persists like the reference's */
} or_node;

typedef struct {
    or_node node[2][OR_NODES];
    or_node *act[2][OR_MAXK + 1];
    or_node *head[OR_HASH];
    uint64_t hist[2][OR_MAXK][OR_HWORDS];   /* history of act[bank][k] */
    uint8_t next_new[OR_MAXK];
    const or_ntf_t *f;
    unsigned K, L;
    uint32_t mask;
    unsigned num_cands, pos, pending, draining, idx;
    double state_limit;
    int fma_o8;                /* 4-wide FMA filter of the shipped AVX2 build */
    uint64_t conv_fail, cands_collapse, next_filter_drops, total_children;
} or_ctx;

/* One observed prune step: children (cost before normalisation), the
 * survivors and the hash registrations made by the sequential prune. */
typedef struct {
    int m, K, bank, n;
    double cost[64];
    uint32_t path[64];
    uint8_t next[64];
    int act[OR_MAXK];
    uint64_t reg;
    int plist[64];     /* registered: same-bank lane, -1 NULL, -3 elsewhere */
} or_prune_obs;
typedef void (*or_prune_observer_fn)(const or_prune_obs *);
extern or_prune_observer_fn or_prune_observer;

/* Contexts initialised while this is set use the order-8 4-wide FMA filter
 * of the reference's AVX2 build (src/trellis.c:348-533, 904-928) for the
 * first 4*floor(n/4) candidates of every step; parity pinned against the
 * reference's own trellis.c compiled with that path selected. */
extern int or_fma_o8_default;
int or_sdm_init(or_ctx *c, const or_ntf_t *f, int depth, int cands, int lat);
void or_sdm_reset(or_ctx *c);
/* One trellis step; returns the decided bit (1 => +1.0). */
int or_sdm_sample(or_ctx *c, double x);
size_t or_sdm_process(or_ctx *c, const double *in, float *out, size_t count);
size_t or_sdm_drain(or_ctx *c, float *out, size_t max_out);
void or_sdm_copy_state(or_ctx *dst, const or_ctx *src);
void or_sdm_best_state(const or_ctx *c, double *out);
size_t or_sdm_segment(or_ctx *c, const double *in, float *out, size_t count,
                      size_t discard);

void or_estimate_state(const or_ntf_t *f, const double *init, const double *in,
                       size_t count, double state_limit, int levels,
                       double *out);

/* ---- FIR chain ---- */
#define OR_FIR_MAX_STAGES 9
double or_bessel_i0(double x);
void or_design_halfband(double *h, int ntaps, double beta);
/* Stage plan for a 2^n upsampling chain: returns number of stages or -1. */
int or_fir_plan(unsigned fs_in, unsigned fs_out, int *ntaps);
/* One 2x polyphase upsampling stage with zero history:
 * y[2m+p] = sum_q fma-ascending h2[2q+p] * x[m-q], h2 = 2*h. */
void or_fir_stage(const double *h2, int ntaps, const double *hist, int nhist,
                  const double *in, size_t n, double *out);
/* Whole chain from float PCM, output multiplied by gain (skipped if 1.0). */
size_t or_fir_chain(unsigned fs_in, unsigned fs_out, const float *in,
                    size_t n, double gain, double *out);
/* Same chain in long double with the exact pairwise-free sum (truth). */
size_t or_fir_chain_d(unsigned fs_in, unsigned fs_out, const double *in,
                      size_t n, double gain, double *out);
size_t or_fir_chain_ld(unsigned fs_in, unsigned fs_out, const float *in,
                       size_t n, double gain, long double *out);

/* ---- DAS (src/dsp_plugin.c:1314-2761, generalised to any segment count) -- */
typedef struct {
    int nseg;
    size_t overlap;
    int seg0_first;           /* 1: seg0 -> seg1 direct -> estimate; 0: all parallel */
    int reposition;
    size_t *nom_start;        /* [nseg] per plan (shared by channels) */
    size_t *nom_size;         /* [nseg] */
} or_das_plan_t;

/* Nominal plan: equal split of F into segments of `seg_len` (last takes the
 * rest) or into `nseg` parts when seg_len == 0. */
int or_das_plan_init(or_das_plan_t *p, size_t F, size_t seg_len, int nseg,
                     size_t overlap, int seg0_first, int reposition);
void or_das_plan_free(or_das_plan_t *p);
/* Boundary reposition on channel-0 data (src/dsp_plugin.c:1963-2034). */
void or_das_reposition(or_das_plan_t *p, const double *ch0, size_t F);

/* Runs the DAS pipeline over one chunk of `nch` channels of trellis input
 * (each F samples, channel-major).  `persist` holds one context per channel
 * and is updated like the reference (copy-back of the last segment).
 * out receives per channel the assembled +-1 stream; returns its length. */
/* host threads of or_das_run's phases (default 1; results do not change) */
void or_das_set_threads(int n);
size_t or_das_run(const or_das_plan_t *p, or_ctx *persist, int nch,
                  const double *in, size_t F, float *out, int *stitch_pos);

/* Every intermediate of one stitch search: half window, density map (dm
 * when not NULL, ovl_len entries), peak and its first position, the
 * candidates found by the scan (scanned) and in total (ncand, the peak-run
 * candidate last), their position, run, density, largest decoded
 * difference and quality, and the chosen index. */
typedef struct {
    int half_w, peak, peak_pos, scanned, ncand, best;
    int pos[33], run[33], density[33];
    double quality[33], max_decoded[33];
    int *dm;
} or_stitch_detail_t;
int or_das_stitch_ex(const float *prev, const float *cur, int ovl_len, int lat, or_stitch_detail_t *det);
/* Candidate choice of src/dsp_plugin.c:2641-2652. */
int or_das_pick(int n, const double *quality, const int *density);
/* Stitch search for one boundary (src/dsp_plugin.c:2452-2640):
 * returns the chosen position in [0, ovl_len). */
int or_das_stitch(const float *prev_ovl, const float *this_ovl, int ovl_len,
                  int lat, double *quality_out, int *ncand_out);
const double *or_das_qfir(void);

/* Differentiable loop filter and pre-emphasis (or_lf.c); arrays [B][N]. */
double or_lf_tanh(double z);
void or_lf_forward(const double *a, const double *g, int P, double T, const double *u, int B,
                   size_t N, double *y);
void or_lf_backward(const double *a, const double *g, int P, double T, const double *u, int B,
                    size_t N, const double *gy, double *gu, double *ga, double *gg);
void or_lf_backward_tr(const double *a, const double *g, int P, double T, const double *u, int B,
                       size_t N, const double *gy, double *gu, double *ga, double *gg, size_t trunc_every);
void or_pe_forward(const double *w, int T, const float *p, int B, size_t N, double *u);
void or_pe_backward(int T, const float *p, int B, size_t N, const double *gu, double *gw);

#ifdef __cplusplus
}
#endif
/* DSD wire formats (or_dop.c): src/dop.c restated; float +-1.0 streams
 * as the reference's API uses them.  Pinned against the compiled dop.c
 * and test/test_dop.c's known answers. */
int or_dop_detect(const float *pcm, size_t frames, int channels);
void or_dop_unpack(const float *pcm, float *bits, size_t frames);
void or_dop_pack(const float *bits, float *pcm, size_t bit_count);
int or_dop_pack_i24(const float *bits, uint8_t *out, size_t bit_count, int phase);
void or_bits_unpack(const uint8_t *src, float *dst, size_t n_bits);
void or_bits_pack(const float *src, uint8_t *dst, size_t n_bits);

/* Engine pre-SDM conditioning (or_cond.c), one channel. */
typedef struct {
    int taps;                 /* boxcar taps */
    double ring[64];
    int pos;
    double sum;
    int lp_n;                 /* lowpass taps (0: none) */
    double lp[127];
    double dly[127];          /* lowpass delay line, dly[0] newest */
} or_cond;
int or_lowpass_taps(unsigned rate, double *h);
void or_cond_init(or_cond *c, int taps, unsigned lowpass_rate);
void or_cond_dsd_boxcar(or_cond *c, const float *in, size_t n, double gs, double ge, double *out);
void or_cond_pcm_boxcar(or_cond *c, const float *in, size_t n, double *out);
void or_cond_dsd_lowpass(or_cond *c, const float *in, size_t n, double gs, double ge, double *out);
void or_gain_ramp(double *x, size_t n, double gs, double ge);
void or_soft_clip(double *x, size_t n);

/* PreCorr SDM (or_precorr.c), one channel. */
typedef struct {
    float a[8], g[8];
    int order;
    float state_limit;
    float table[256][8];
    float state[OR_MAX_ORDER];
    unsigned history;
    int phase;
    float prev_y;
} or_pc;
int or_pc_init(or_pc *c, int filter_index, float state_limit);
void or_pc_reset(or_pc *c);
size_t or_pc_process(or_pc *c, const double *in, float *out, size_t n);

#endif
