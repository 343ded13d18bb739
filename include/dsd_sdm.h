/*
 * dsd_sdm.h — the reference's trellis-SDM and NTF C surface
 * (include/trellis.h:61-118, include/ntf.h:23-36 of foo_dsd_trellis),
 * exported by libdsdhip.so with the same names, signatures and context
 * layout, so a caller built against those headers links against this
 * library instead of src/trellis.c + src/ntf.c.
 *
 * The context stays caller-owned memory in the reference's layout, and
 * callers may keep reading and writing its fields directly (the DAS driver
 * reads trellis[idx & 1].act[0]->state and seeds every act[c]->state and
 * cost, src/dsp_plugin.c:2251-2253, 2274-2281, 2329-2338).
 * sdm_process_block and sdm_drain run the trellis on the GPU: the context is
 * converted to the device form (survivors, both banks' per-node path,
 * metric and hash-chain link, history bits), the samples run through one
 * trellis job, and the result is written back into the same struct.
 * Hash-chain links that lead into another context (left by
 * sdm_context_copy_state, which copies the source's chain pointers as they
 * are) are followed as the reference follows them: the nodes they reach
 * are read from that memory when the call starts.
 *
 * Layout: every type below has the reference's field order, types and
 * sizes (sdm_context_t is 46,752 bytes on LP64).
 *
 * Not reproduced: the four diagnostic counters (conv_fail, cands_collapse,
 * next_filter_drops, total_children) are not advanced by the GPU calls, and
 * nodes that are not survivors keep only what the trellis reads from them
 * again (path, metric, chain link).  mb_levels must be 2 (the reference's
 * default, the only branching the plugin uses); loop-filter orders 4..8.
 * A call that cannot run (no GPU, unsupported parameters) prints the reason
 * on stderr, leaves the context unchanged and returns 0.
 *
 * The _fast variants (the reference's build without /fp:precise, whose
 * bits depend on MSVC's optimiser) run the precise arithmetic.
 */
#ifndef DSD_SDM_H
#define DSD_SDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAX_NTF_ORDER 10
#define SDM_TRELLIS_MAX_ORDER 32
#define SDM_TRELLIS_MAX_NUM 32
#define SDM_TRELLIS_MAX_LAT 2048
#define PATH_HASH_SIZE 128
#define SDM_MB_LEVELS 4

/* include/dsd_types.h:56-69 */
typedef enum {
    NTF_AUTO = -1,
    NTF_CLANS_4 = 0,
    NTF_SDM_4,
    NTF_CLANS_5,
    NTF_SDM_5,
    NTF_CLANS_6,
    NTF_SDM_6,
    NTF_CLANS_7,
    NTF_SDM_7,
    NTF_CLANS_8,
    NTF_SDM_8,
    NTF_COUNT
} ntf_filter_id_t;

/* include/ntf.h:14-20: CRFB loop-filter coefficients */
typedef struct {
    double a[MAX_NTF_ORDER];
    double g[MAX_NTF_ORDER];
    int order;
    unsigned freq;
    const char *name;
} ntf_filter_t;

/* include/trellis.h:18-28: one trellis node */
typedef struct sdm_state {
    double state[MAX_NTF_ORDER];   /* loop-filter integrators */
    double cost;                   /* accumulated squared error */
    uint32_t path;                 /* last bits, trellis_mask wide */
    uint8_t next;                  /* bit at the traceback position */
    uint8_t hist;                  /* index of its history buffer */
    uint8_t hist_used;
    struct sdm_state *parent;
    struct sdm_state *path_list;   /* next node of its path-hash chain */
} sdm_state_t;

/* include/trellis.h:30-35: one bank of nodes and its ordered survivor list */
typedef struct {
    sdm_state_t sdm[SDM_MB_LEVELS * SDM_TRELLIS_MAX_NUM];
    sdm_state_t *act[SDM_TRELLIS_MAX_NUM];
} sdm_trellis_t;

/* include/trellis.h:37-59: per-channel context */
typedef struct {
    sdm_trellis_t trellis[2];
    sdm_state_t *path_hash[PATH_HASH_SIZE];
    uint8_t hist_free[2 * SDM_TRELLIS_MAX_NUM];
    unsigned hist_fnum;
    uint32_t trellis_mask;
    uint32_t trellis_num;
    uint32_t trellis_lat;
    unsigned num_cands;
    unsigned mb_levels;
    unsigned pos;
    unsigned pending;
    unsigned draining;
    unsigned idx;
    const ntf_filter_t *filter;
    double prev_y;
    uint64_t conv_fail;
    uint64_t cands_collapse;
    uint64_t next_filter_drops;
    uint64_t total_children;
    double state_limit;
    uint8_t hist[2 * SDM_TRELLIS_MAX_NUM][SDM_TRELLIS_MAX_LAT / 8];
} sdm_context_t;

/* ---- include/trellis.h:61-118 ---- */
/* 0 on success, -1 on bad arguments (src/trellis.c:977-1011) */
int sdm_context_init(sdm_context_t *ctx, const ntf_filter_t *filter, int trellis_depth, int trellis_cands,
                     int trellis_lat);
/* outputs produced (+-1.0f); the first trellis_lat - pending inputs give none
 * (src/trellis.c:1013-1040).  GPU. */
size_t sdm_process_block(sdm_context_t *ctx, const double *in, float *out, size_t count);
/* up to max_out outputs from zero input, flushing an unfilled latency first
 * (src/trellis.c:1042-1062).  GPU. */
size_t sdm_drain(sdm_context_t *ctx, float *out, size_t max_out);
/* struct copy with the survivor lists re-pointed into dst; chain links keep
 * pointing into src; hash heads and counters cleared (src/trellis.c:1064-1092) */
void sdm_context_copy_state(sdm_context_t *dst, const sdm_context_t *src);
/* squared distance of the best survivors' states, 1e30 if either has none
 * (src/trellis.c:1094-1109) */
double sdm_state_distance(const sdm_context_t *a, const sdm_context_t *b);
/* back to one survivor at the origin, parameters and mb_levels kept
 * (src/trellis.c:1189-1224) */
void sdm_context_reset(sdm_context_t *ctx);
void sdm_context_free(sdm_context_t *ctx);

int sdm_context_init_fast(sdm_context_t *ctx, const ntf_filter_t *filter, int trellis_depth, int num_cands,
                          int trellis_lat);
size_t sdm_process_block_fast(sdm_context_t *ctx, const double *in, float *out, size_t count);
size_t sdm_drain_fast(sdm_context_t *ctx, float *out, size_t max_out);
void sdm_context_copy_state_fast(sdm_context_t *dst, const sdm_context_t *src);
double sdm_state_distance_fast(const sdm_context_t *a, const sdm_context_t *b);
void sdm_context_reset_fast(sdm_context_t *ctx);
void sdm_context_free_fast(sdm_context_t *ctx);

/* greedy 16-level estimate of the loop-filter state after `count` inputs
 * (src/trellis.c:1118-1123); host threads, bit-exact */
void sdm_estimate_state(const ntf_filter_t *filter, const double *init_state, const double *in, size_t count,
                        double state_limit, double *out_state);
/* the same with num_levels quantizer levels (src/trellis.c:1125-1187) */
void sdm_estimate_state_multibit(const ntf_filter_t *filter, const double *init_state, const double *in,
                                 size_t count, double state_limit, int num_levels, double *out_state);

/* ---- include/ntf.h:23-36 (this library's coefficient table) ---- */
const ntf_filter_t *ntf_get_filter(ntf_filter_id_t id, unsigned dsd_rate);
const ntf_filter_t *ntf_auto_select(unsigned dsd_rate);
const ntf_filter_t *ntf_auto_select_precorr(unsigned dsd_rate);
int ntf_get_filter_count(void);
const ntf_filter_t *ntf_get_by_index(int index);

/* Round trip of ctx through the device form, on the host only: what
 * sdm_process_block does around its GPU job, with no job.  ctx must be
 * unchanged in every field the trellis reads again.  0 on success. */
int dsdhip_sdm_roundtrip(sdm_context_t *ctx);

#ifdef __cplusplus
}
#endif

#endif /* DSD_SDM_H */
