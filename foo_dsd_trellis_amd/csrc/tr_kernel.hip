/*
 * Trellis sigma-delta modulator for gfx950: one wave64 per trellis context.
 *
 * Lane j holds child j of the current step (parent = survivor j/2, branch
 * bit j&1), so a full K=32 trellis fills the wave.  Per input sample:
 *   1. expand   : each lane runs its parent's CRFB loop filter in the
 *                 reference's scalar operation order (FP64, no contraction,
 *                 src/trellis.c:37-191) and forms the child metric;
 *   2. prune    : a wave-parallel formulation of the reference's sequential
 *                 candidate sort (src/trellis.c:621-731): order masks by
 *                 (metric, index) from an LDS broadcast sweep, majority/
 *                 traceback filter, then the early skip and the
 *                 duplicate-path rules in closed form on lane masks where
 *                 every same-path group holds at most two candidates (a
 *                 fixed-point resolution otherwise), and a final rank;
 *                 steps that fall outside what the parallel form
 *                 decides (a bucket-0 lookup after the full-list write past
 *                 act[31] at K=32, paths wider than 8 bits) step the
 *                 sequential sort literally, child by child, with the list
 *                 spread over the lanes;
 *   3. commit   : survivors write state/metric/path/history for the next
 *                 step into the other LDS bank; the traceback bit of the
 *                 best survivor is the output (src/trellis.c:938-972).
 * History is carried by value per survivor (HW 64-bit words of L bits).
 */
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "tr_types.h"

typedef unsigned long long u64;
#define GPTR(T) T __attribute__((address_space(1))) *

namespace {

__device__ __forceinline__ double rl_d(double v, int l)
{
    u64 b = (u64)__double_as_longlong(v);
    unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((u64)hi << 32) | lo));
}

__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ u64 ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int ff1(u64 m) { return (int)__builtin_ctzll(m); }
/* this lane's bit of a wave-uniform mask, used as the lane predicate itself */
__device__ __forceinline__ bool lane_in(u64 m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
__device__ __forceinline__ long long bits(double d) { return __double_as_longlong(d); }

/* count of set bits of `m` below this lane */
__device__ __forceinline__ int mbcnt(u64 m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

/* Exclusive prefix OR over the lanes of a 64-bit value whose set bits are
 * distinct across lanes (inclusive scan by DPP row shifts and row
 * broadcasts, then the lane's own bits removed). */
__device__ __forceinline__ unsigned scan_or32(unsigned x)
{
    x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);   /* row_shr:1 */
    x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);   /* row_shr:2 */
    x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);   /* row_shr:4 */
    x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);   /* row_shr:8 */
    x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  /* row_bcast:15 */
    x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  /* row_bcast:31 */
    return x;
}
__device__ __forceinline__ u64 wave_excl_or(u64 v)
{
    const u64 inc = ((u64)scan_or32((unsigned)(v >> 32)) << 32) | scan_or32((unsigned)v);
    return inc & ~v;
}

/* The workgroup is one wave: its LDS instructions execute in issue order,
 * so cross-lane exchange through LDS only needs the compiler not to move
 * memory operations across this point. */
__device__ __forceinline__ void wsync() { __asm__ volatile("" ::: "memory"); }

/* acc * 2 + (k < c): the compare result feeds the add-with-carry directly
 * (two VALU instructions per element). */
__device__ __forceinline__ unsigned shl_add_lt(unsigned acc, double k, double c)
{
    unsigned r;
    __asm__("v_cmp_lt_f64_e32 vcc, %1, %2\n\t"
            "v_addc_co_u32_e32 %0, vcc, %3, %3, vcc"
            : "=v"(r)
            : "v"(k), "v"(c), "v"(acc)
            : "vcc");
    return r;
}

/* four steps of shl_add_ltu (keys k.w, k.z, k.y, k.x in that order) in one
 * asm statement: the compiler pads every asm statement with a wait state */
__device__ __forceinline__ unsigned shl_add_ltu4(unsigned acc, uint4 k, unsigned c)
{
    __asm__("v_cmp_lt_u32_e32 vcc, %1, %5\n\t"
            "v_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
            "v_cmp_lt_u32_e32 vcc, %2, %5\n\t"
            "v_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
            "v_cmp_lt_u32_e32 vcc, %3, %5\n\t"
            "v_addc_co_u32_e32 %0, vcc, %0, %0, vcc\n\t"
            "v_cmp_lt_u32_e32 vcc, %4, %5\n\t"
            "v_addc_co_u32_e32 %0, vcc, %0, %0, vcc"
            : "+v"(acc)
            : "v"(k.w), "v"(k.z), "v"(k.y), "v"(k.x), "v"(c)
            : "vcc");
    return acc;
}

/* acc * 2 + (k <= c) */
__device__ __forceinline__ unsigned shl_add_le(unsigned acc, double k, double c)
{
    unsigned r;
    __asm__("v_cmp_le_f64_e32 vcc, %1, %2\n\t"
            "v_addc_co_u32_e32 %0, vcc, %3, %3, vcc"
            : "=v"(r)
            : "v"(k), "v"(c), "v"(acc)
            : "vcc");
    return r;
}

/* acc * 2 + (k < c) for unsigned 32-bit keys */
__device__ __forceinline__ unsigned shl_add_ltu(unsigned acc, unsigned k, unsigned c)
{
    unsigned r;
    __asm__("v_cmp_lt_u32_e32 vcc, %1, %2\n\t"
            "v_addc_co_u32_e32 %0, vcc, %3, %3, vcc"
            : "=v"(r)
            : "v"(k), "v"(c), "v"(acc)
            : "vcc");
    return r;
}

/* Strict-order mask on the metrics' high words: for non-negative doubles
 * the order of the bit patterns is the order of the values, so where all
 * high words differ this is the exact strict order.  16 keys per batch
 * (four 128-bit loads).  With ALL_LOADS every batch is loaded before the
 * first compare (one LDS round trip, 4 registers a key: for a wave that
 * has its SIMD to itself); otherwise batch by batch (registers for 16
 * keys; the other waves of the SIMD cover the round trips). */
template <int NJ, bool ALL_LOADS = false>
__device__ __forceinline__ u64 lt_sweep_hi(const uint32_t *key, unsigned c)
{
    const uint4 *k4 = reinterpret_cast<const uint4 *>(key);
    unsigned w[2] = {0u, 0u};
    if constexpr (ALL_LOADS) {
        uint4 kv[NJ / 4];
#pragma unroll
        for (int i = NJ / 4 - 1; i >= 0; i--) kv[i] = k4[i];
#pragma unroll
        for (int b = NJ / 16 - 1; b >= 0; b--) {
            unsigned acc = w[b >> 1];
#pragma unroll
            for (int i = 3; i >= 0; i--) acc = shl_add_ltu4(acc, kv[b * 4 + i], c);
            w[b >> 1] = acc;
        }
    } else {
#pragma unroll
        for (int b = NJ / 16 - 1; b >= 0; b--) {
            uint4 kv[4];
#pragma unroll
            for (int i = 3; i >= 0; i--) kv[i] = k4[b * 4 + i];
            unsigned acc = w[b >> 1];
#pragma unroll
            for (int i = 3; i >= 0; i--) acc = shl_add_ltu4(acc, kv[i], c);
            w[b >> 1] = acc;
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    return ((u64)w[1] << 32) | w[0];
}

/* Strict-order mask against the first NJ metrics of the LDS key array
 * (broadcast reads).  Keys are read 16 at a time with 128-bit loads and
 * consumed in load order, so one LDS round trip is exposed per 16 keys. */
template <int NJ>
__device__ __forceinline__ u64 lt_sweep(const double *key, double c)
{
    const double2 *k2 = reinterpret_cast<const double2 *>(key);
    unsigned w[2] = {0u, 0u};
#pragma unroll
    for (int b = NJ / 8 - 1; b >= 0; b--) {
        double2 kv[4];
#pragma unroll
        for (int i = 3; i >= 0; i--) kv[i] = k2[b * 4 + i];
        unsigned acc = w[b >> 2];
#pragma unroll
        for (int i = 3; i >= 0; i--) {
            acc = shl_add_lt(acc, kv[i].y, c);
            acc = shl_add_lt(acc, kv[i].x, c);
        }
        w[b >> 2] = acc;
        __builtin_amdgcn_sched_barrier(0);
    }
    return ((u64)w[1] << 32) | w[0];
}

/* Order mask "at most c" against the first NJ metrics, read like lt_sweep. */
template <int NJ>
__device__ __forceinline__ u64 le_sweep(const double *key, double c)
{
    const double2 *k2 = reinterpret_cast<const double2 *>(key);
    unsigned w[2] = {0u, 0u};
#pragma unroll
    for (int b = NJ / 8 - 1; b >= 0; b--) {
        double2 kv[4];
#pragma unroll
        for (int i = 3; i >= 0; i--) kv[i] = k2[b * 4 + i];
        unsigned acc = w[b >> 2];
#pragma unroll
        for (int i = 3; i >= 0; i--) {
            acc = shl_add_le(acc, kv[i].y, c);
            acc = shl_add_le(acc, kv[i].x, c);
        }
        w[b >> 2] = acc;
        __builtin_amdgcn_sched_barrier(0);
    }
    return ((u64)w[1] << 32) | w[0];
}

/* LDS working set of one context: the per-slot records of both banks
 * (followed by the path-hash chains), the prune's tables, and the lane
 * holding each survivor.  The survivors' own records (loop-filter state,
 * metric, path, traceback bit, history) stay in the registers of the lane
 * that produced them and are fetched by the next step's children with
 * ds_bpermute, which keeps LDS per wave small (more waves per CU). */
template <int P, int HW>
struct TrLds {
    double ncost[2][64];
    uint32_t npath[2][64];
    uint32_t khi[64];              /* high words of this step's child metrics */
    u64 gtab[256 + 64];            /* + one scratch entry per lane */
    int8_t link[2][64];
    uint8_t rank[64];
    int8_t head[128];
    uint8_t src[TR_MAXK];          /* lane holding the survivor of each rank, x 4
                                      (the bpermute byte address) */
};

/* value of another lane's register (all lanes active) */
/* from: byte address of the source lane (lane * 4) */
__device__ __forceinline__ int bp_i(int v, int from) { return __builtin_amdgcn_ds_bpermute(from, v); }
__device__ __forceinline__ double bp_d(double v, int from)
{
    const u64 b = (u64)__double_as_longlong(v);
    const unsigned lo = (unsigned)bp_i((int)(unsigned)b, from);
    const unsigned hi = (unsigned)bp_i((int)(unsigned)(b >> 32), from);
    return __longlong_as_double((long long)(((u64)hi << 32) | lo));
}
__device__ __forceinline__ u64 bp_u64(u64 v, int from)
{
    const unsigned lo = (unsigned)bp_i((int)(unsigned)v, from);
    const unsigned hi = (unsigned)bp_i((int)(unsigned)(v >> 32), from);
    return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uni_ll(long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u64)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((u64)v >> 32));
    return (long long)(((u64)hi << 32) | lo);
}
__device__ __forceinline__ long long rl_ll(long long v, int l)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u64)v, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((u64)v >> 32), l);
    return (long long)(((u64)hi << 32) | lo);
}

/* The reference's sequential candidate sort (src/trellis.c:582-731) stepped
 * literally, one child after the other in index order, with every lane
 * taking part: lane k holds entry k of the active list (child and metric
 * bits), so the insertion point is a ballot and a shift is one lane
 * exchange.  The path-hash chains are walked as the reference walks them,
 * including links into the context this one was copied from and the write
 * of the 33rd list entry onto bucket 0's head in bank 1 (act[] has 32
 * entries and the hash heads of bank 1 follow it).  Used for the steps the
 * parallel prune does not decide.  Writes rank[] (255 = pruned) and returns
 * the number of survivors. */
template <int P, int HW>
__device__ int wide_prune(TrLds<P, HW> &s, int nb, int m, int K, const tr_ctx *snap, int *err,
                          double cost, uint32_t path, int cnext, u64 LT, u64 N1, int lane, bool live)
{
    const u64 ALL = (m == 64) ? ~0ull : ((1ull << m) - 1ull);
    const long long mycb = bits(cost);
    if (live) s.head[path & 127u] = (int8_t)-1;
    wsync();
    /* first strict minimum; majority of the inherited bits (tie -> 0) and
     * the cheapest child of the majority */
    const int min0 = ff1(ballot(live && (LT & ALL) == 0));
    const int v1 = __popcll(N1 & ALL);
    const int maj = (v1 > m - v1) ? 1 : 0;
    int mn = min0;
    if ((int)((N1 >> min0) & 1) != maj) {
        const u64 MJ = maj ? (N1 & ALL) : (ALL & ~N1);
        const u64 bmm = ballot(live && ((MJ >> lane) & 1) && (LT & MJ) == 0);
        if (bmm) {
            const int bm = ff1(bmm);
            if (rl_d(cost, bm) < rl_d(cost, min0) * 1.1) mn = bm;
        }
    }
    const u64 CAND = ((N1 >> mn) & 1) ? (N1 & ALL) : (ALL & ~N1);

    int n = 0;                  /* list length (uniform) */
    int aidx = 0;               /* lane k < n: child listed at k */
    long long acb = 0;          /* and its metric bits */
    u64 rest = CAND;
    while (rest) {
        const int i = ff1(rest);
        rest &= rest - 1;
        const long long ci = rl_ll(mycb, i);
        if (n == K && rl_ll(acb, n - 1) <= ci) continue;
        const uint32_t pi = (uint32_t)__builtin_amdgcn_readlane((int)path, i);
        const int b = (int)(pi & 127u);
        /* lookup along the hash chain (src/trellis.c:599-619) */
        int t = -1;
        long long tcb = 0;
        int v = uni(s.head[b]);
        int hops = 0;
        while (v != -1) {
            if (++hops > 2 * TR_SLOTS + 2) { *err = 2; break; }   /* links only point downwards */
            uint32_t vp;
            long long vc;
            int vl;
            if (v < TR_LINK_FOREIGN) {
                vp = (uint32_t)uni((int)s.npath[nb][v]);
                vc = rl_ll(bits(s.ncost[nb][v]), 0);
                vl = uni(s.link[nb][v]);
            } else {
                if (!snap) { *err = 2; break; }
                const int sv = v - TR_LINK_FOREIGN;
                vp = (uint32_t)uni((int)snap->node_path[nb][sv]);
                vc = rl_ll(bits(snap->node_cost[nb][sv]), 0);
                vl = uni(snap->link[nb][sv]);
                if (vl >= TR_LINK_FOREIGN) { *err = 2; break; }
                if (vl >= 0) vl += TR_LINK_FOREIGN;
            }
            if (vp == pi) { t = v; tcb = vc; break; }
            v = vl;
        }
        int tp = n;             /* list position of the registrant (n: not listed) */
        bool guarded = true;
        if (t < 0) {
            /* register at the head of the bucket */
            const int h = uni(s.head[b]);
            if (lane == 0) {
                s.link[nb][i] = (int8_t)h;
                s.head[b] = (int8_t)i;
            }
            wsync();
            guarded = false;
        } else {
            if (tcb <= ci) continue;
            if (t < TR_LINK_FOREIGN) {
                const u64 hit = ballot(lane < n && aidx == t);
                if (hit) tp = ff1(hit);
            }
            if (tp < n) {
                /* remove the registrant's entry */
                const int ai = __shfl_down(aidx, 1, 64);
                const long long ab = (long long)(((u64)(unsigned)__shfl_down((int)((u64)acb >> 32), 1, 64) << 32) |
                                                 (unsigned)__shfl_down((int)(unsigned)(u64)acb, 1, 64));
                if (lane >= tp) {
                    aidx = ai;
                    acb = ab;
                }
                n--;
            }
        }
        /* stable insertion behind equal metrics */
        const int j = __popcll(ballot(lane < n && acb <= ci));
        if (!guarded && n == TR_MAXK && nb == 1) {
            /* the shift writes act[32], which is bank 1's bucket-0 head */
            const int r = __builtin_amdgcn_readlane(aidx, TR_MAXK - 1);
            if (lane == 0) s.head[0] = (int8_t)r;
            wsync();
        }
        {
            const int ai = __shfl_up(aidx, 1, 64);
            const long long ab = (long long)(((u64)(unsigned)__shfl_up((int)((u64)acb >> 32), 1, 64) << 32) |
                                             (unsigned)__shfl_up((int)(unsigned)(u64)acb, 1, 64));
            if (lane > j) {
                aidx = ai;
                acb = ab;
            } else if (lane == j) {
                aidx = i;
                acb = ci;
            }
        }
        if (n < K) n++;
    }
    s.rank[lane] = 255;
    wsync();
    if (lane < n) s.rank[aidx] = (uint8_t)lane;
    wsync();
    return n;
}

} // namespace

/* One job on one wave; s is the wave's own LDS region.  Nothing here
 * synchronises beyond the wave, so several waves of a workgroup may run
 * different jobs (and different numbers of them). */
template <int P, int HW, bool F8 = false, bool UJ = false, int LIM = 2, bool LONE = false>
__device__ __forceinline__ void tr_job_body(const tr_params &prm, tr_job *jb, tr_ctx *ctxs,
                                            TrLds<P, HW> &s, const int lane)
{
    /* UJ: the descriptor sits in LDS, whose loads the compiler does not
     * treat as wave-uniform; its fields are then made uniform explicitly */
    auto fi = [](int v) { return UJ ? uni(v) : v; };
    auto fl = [](long long v) { return UJ ? uni_ll(v) : v; };
    const int ctx_in = fi(jb->ctx_in), snap_i = fi(jb->snap);
    const tr_ctx *c0 = &ctxs[ctx_in];
    const tr_ctx *snap = snap_i >= 0 ? &ctxs[snap_i] : nullptr;
    const int K = prm.K, L = prm.L;
    const uint32_t mask = prm.mask;
    const bool groups_ok = (prm.B <= 8) && !prm.force_slow;
    int err = 0;

    /* ---- load the context ---- */
    int num_cands = c0->num_cands, pos = c0->pos, pending = c0->pending, idx = c0->idx;
    const int draining = c0->draining;
    /* survivor record held by this lane, and its rank (255: none) */
    double rst[P], rcost = 0.0;
    uint32_t rpath = 0;
    int rnext = 0, myr = 255;
    u64 rhist[HW];
#pragma unroll
    for (int q = 0; q < P; q++) rst[q] = 0.0;
#pragma unroll
    for (int w = 0; w < HW; w++) rhist[w] = 0;
    {
        if (lane < num_cands) {
            const bool sd = fi(jb->seed) != 0;
#pragma unroll
            for (int q = 0; q < P; q++)
                rst[q] = sd ? __longlong_as_double(fl(bits(jb->seed_st[q]))) : c0->st[lane][q];
#pragma unroll
            for (int w = 0; w < HW; w++) rhist[w] = c0->hist[lane][w];
            rcost = sd ? 0.0 : c0->cost[lane];
            rpath = c0->path[lane];
            rnext = c0->next[lane];
            myr = lane;
        }
        if (lane < TR_MAXK) s.src[lane] = (uint8_t)(lane << 2);
        for (int b = 0; b < 2; b++) {
            int8_t l = c0->link[b][lane];
            if (fi(jb->copy)) {
                if (l >= TR_LINK_FOREIGN) err = 2;
                else if (l >= 0) l = (int8_t)(l + TR_LINK_FOREIGN);
            }
            s.link[b][lane] = l;
            s.npath[b][lane] = c0->node_path[b][lane];
            s.ncost[b][lane] = c0->node_cost[b][lane];
        }
        for (int i = lane; i < 256 + 64; i += 64) s.gtab[i] = 0;
    }
    wsync();

    const double a0 = prm.a[0];
    const u64 lowerm = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    /* the history bit a child writes is its path's low bit, the branch bit;
     * kept opaque so the update below stays one bit-field insert per word */
    u64 brm = ((mask & 1u) && (lane & 1)) ? ~0ull : 0ull;
    __asm__ volatile("" : "+v"(brm));
    /* global address space: these accesses count against vmcnt only, so the
     * prefetched block stays in flight across the step's LDS waits */
    const GPTR(double) in = (const GPTR(double))uni_ll((long long)jb->in);
    GPTR(u64) out = (GPTR(u64))uni_ll((long long)jb->out);
    /* wave-uniform: the step loop's tests stay scalar */
    const int64_t n_in = uni_ll(jb->n_in), keep_from = uni_ll(jb->keep_from);
    /* src/trellis.c:1013-1040: the first L - pending steps only fill the
     * traceback (pending counts up to L); after that the output of step t
     * is kept from keep_from on.  So step t keeps its output exactly when
     * t >= t_keep, and pending ends at min(L, pending + n_in). */
    const int64_t t_fill = pending < L ? (int64_t)(L - pending) : 0;
    const int64_t t_keep = t_fill > keep_from ? t_fill : keep_from;
    /* steps left before the first kept one (counted down: a 64-bit zero
     * test is scalar, an ordered 64-bit compare is not) */
    int64_t kc = t_keep > 0 ? t_keep : 0;
    u64 acc = 0;
    int nacc = 0;
    int64_t kept = 0, slow_steps = 0;
    /* input in blocks of 64 samples, one lane each; the next block is
     * loaded a block ahead */
    double xv = 0.0, xn = 0.0;
    if (n_in > 0) xv = in[lane < n_in ? lane : n_in - 1];

    for (int64_t t = 0; t < n_in; t++) {
        const int tl = (int)(t & 63);
        if (tl == 0) {
            if (t > 0) xv = xn;
            const int64_t ti = t + 64 + lane;
            xn = in[ti < n_in ? ti : n_in - 1];    /* lanes past the end are never read */
        }
        const double x = rl_d(xv, tl);

        const int nb = idx ^ 1;
        const int n = num_cands, m = 2 * n;
        const bool live = lane < m;
        const int k = lane >> 1, br = lane & 1;
        int next_pos = pos + 1;
        if (next_pos == L) next_pos = 0;

        /* ---- 1. expand ---- */
        /* byte address (lane * 4) of the lane holding parent k; entries
         * k >= n are stale lane addresses, read and unused */
        const int from = s.src[k];
        double sp[P], d[P];
#pragma unroll
        for (int q = 0; q < P; q++) sp[q] = bp_d(rst[q], from);
        const double pcost = bp_d(rcost, from);
        const uint32_t ppath = (uint32_t)bp_i((int)rpath, from);
        const int pnext = bp_i(rnext, from);
        u64 phist[HW];
#pragma unroll
        for (int w = 0; w < HW; w++) phist[w] = bp_u64(rhist[w], from);
        u64 hnext = phist[0];
#pragma unroll
        for (int w = 1; w < HW; w++)
            if (w == (next_pos >> 6)) hnext = phist[w];
        const int nnew = (int)((hnext >> (next_pos & 63)) & 1ull);

        d[0] = sp[0] - prm.g[0] * sp[1] + x - 0.0;
        double v = x + prm.a[0] * d[0];
#pragma unroll
        for (int q = 1; q < P - 1; q++) {
            d[q] = sp[q] + sp[q - 1] - prm.g[q] * sp[q + 1];
            v = v + prm.a[q] * d[q];
        }
        d[P - 1] = sp[P - 1] + sp[P - 2];
        v = v + prm.a[P - 1] * d[P - 1];
        if constexpr (F8) {
            /* the reference's AVX2 build runs candidates four at a time while
             * four remain, with a fused order-8 filter (src/trellis.c:395-421,
             * 904-927): every d[i] fuses its g term, v accumulates by FMA.
             * Computed beside the scalar form and selected per lane. */
            const bool f = k < (n & ~3);
            double df[P];
            df[0] = __fma_rn(-prm.g[0], sp[1], sp[0] + x);
#pragma unroll
            for (int q = 1; q < P - 1; q++) df[q] = __fma_rn(-prm.g[q], sp[q + 1], sp[q] + sp[q - 1]);
            df[P - 1] = d[P - 1];
            double vf = __fma_rn(prm.a[0], df[0], x);
#pragma unroll
            for (int q = 1; q < P; q++) vf = __fma_rn(prm.a[q], df[q], vf);
#pragma unroll
            for (int q = 0; q < P - 1; q++) d[q] = f ? df[q] : d[q];
            v = f ? vf : v;
        }
        d[0] = d[0] + (br ? -1.0 : 1.0);
        /* LIM: 0 no state limit, 1 a state limit, 2 read from prm (the
         * kernels that dispatch on it get a step loop without the test) */
        if (LIM == 1 || (LIM == 2 && prm.state_limit > 0.0)) {
            /* src/trellis.c:179-187, as selects (no per-lane branches) */
            const double lim = prm.state_limit, nlim = -prm.state_limit;
#pragma unroll
            for (int q = 0; q < P; q++) {
                const double v0 = d[q];
                d[q] = (v0 > lim) ? lim : ((v0 < nlim) ? nlim : v0);
            }
        }
        const double e = v + (br ? -a0 : a0);
        const double cost = pcost + e * e;
        const uint32_t path = ((ppath << 1) | (uint32_t)br) & mask;
        const int cnext = pnext;
        /* masks of live lanes are formed as ballot(test) & ALL: a ballot of a
         * plain compare is the compare's own lane mask */
        const u64 ALL = (m == 64) ? ~0ull : ((1ull << m) - 1ull);
        if (ballot(!__builtin_isfinite(cost)) & ALL) err = 1;

        const unsigned chi = (unsigned)((u64)bits(cost) >> 32);
        if (live) {
            s.npath[nb][lane] = path;
            s.ncost[nb][lane] = cost;      /* also the sweep's keys: entries >= m are never counted */
            s.khi[lane] = chi;
        }
        wsync();

        /* ---- 2. prune ---- */
        /* LT bit j: child j has a strictly smaller metric (bits >= m are
         * don't-care; every use masks them off).  First on the high words
         * (metrics are >= 0; a set sign bit takes the full compare); equal
         * high words give equal strict ranks, and only then the full 64-bit
         * sweep runs. */
        u64 LT;
        bool tie = false;
        const u64 N1 = ballot(cnext != 0) & ALL;
        /* equal keys: two children with one strict rank */
        auto has_tie = [&](u64 lt) {
            const int r = __popcll(lt & ALL) & 63;   /* a live lane's is <= m - 1 */
            if (live) s.rank[r] = (uint8_t)lane;
            wsync();
            const bool t = (ballot(s.rank[r] != lane) & ALL) != 0;
            wsync();
            return t;
        };
        /* candidates: first strict minimum, majority of next with the best
         * majority member's 1.1 override, then the next filter */
        auto pick_cand = [&]() -> u64 {
            const int min0 = ff1(ballot((LT & ALL) == 0) & ALL);
            const int v1 = __popcll(N1);
            const int maj = (v1 > m - v1) ? 1 : 0;
            int mn = min0;
            if (((N1 >> min0) ^ (u64)maj) & 1) {
                const u64 MJ = maj ? N1 : (ALL & ~N1);
                const u64 bmm = ballot((LT & MJ) == 0) & MJ;
                if (bmm) {
                    const int bm = ff1(bmm);
                    if (rl_d(cost, bm) < rl_d(cost, min0) * 1.1) mn = bm;
                }
            }
            return ((N1 >> mn) & 1) ? N1 : (ALL & ~N1);
        };
        /* same-path groups among candidates (paths are < 256 here) in one
         * LDS exchange, which also carries the tie test when rank_too: G,
         * the lane's group, G2 the group of the other path of its hash
         * bucket (path ^ 128) */
        u64 CAND = 0, G = 0, G2 = 0;
        auto group_pass = [&](bool rank_too) -> bool {
            const bool cand = lane_in(CAND);
            const int gi = cand ? (int)path : 256 + lane;
            const int r = __popcll(LT & ALL) & 63;
            if (rank_too && live) s.rank[r] = (uint8_t)lane;
            atomicOr(&s.gtab[gi], 1ull << lane);
            wsync();
            const bool t = rank_too && (ballot(s.rank[r] != lane) & ALL) != 0;
            G = s.gtab[gi] & (cand ? ~0ull : 0ull);
            G2 = cand ? s.gtab[gi ^ 128] : 0ull;
            wsync();
            /* every member takes its own bit out again: the entry is 0 */
            atomicXor(&s.gtab[gi], 1ull << lane);
            return t;
        };
        bool grouped = false;
        {
            const bool neg = (ballot((int)chi < 0) & ALL) != 0;
            if (!neg) {
                LT = (m <= 16) ? lt_sweep_hi<16, LONE>(s.khi, chi)
                   : (m <= 32) ? lt_sweep_hi<32, LONE>(s.khi, chi) : lt_sweep_hi<64, LONE>(s.khi, chi);
                if (groups_ok) {
                    /* candidates and groups from these masks, the tie test in
                     * the same exchange; with a tie all of it is redone below */
                    CAND = pick_cand();
                    tie = group_pass(true);
                    grouped = !tie;
                } else {
                    tie = has_tie(LT);
                }
            }
            if (neg || tie) {
                const double *keys = s.ncost[nb];
                LT = (m <= 16) ? lt_sweep<16>(keys, cost)
                   : (m <= 32) ? lt_sweep<32>(keys, cost) : lt_sweep<64>(keys, cost);
                tie = has_tie(LT);
                if (tie) {
                    /* Every rule of the sequential sort compares a child with
                     * earlier ones as "cost <= mine" and with later ones as
                     * "cost < mine" (first strict minimum, skip when the
                     * full list's last is <= mine, drop a duplicate whose
                     * registrant is <= mine, insertion behind equals), so it
                     * follows the total order (metric, index): equal metrics
                     * rank by index. */
                    const u64 LE = (m <= 16) ? le_sweep<16>(keys, cost)
                                 : (m <= 32) ? le_sweep<32>(keys, cost) : le_sweep<64>(keys, cost);
                    LT = (LE & lowerm) | (LT & ~lowerm);
                    tie = false;
                }
            }
        }
        bool slow = !groups_ok || tie;

        int nsurv = 0, myrank = 255;
        if (!slow) {
            if (!grouped) {
                CAND = pick_cand();
                group_pass(false);
            }
            const bool cand = lane_in(CAND);

            /* Closed form where every same-path group holds at most two
             * candidates (two children share a path only when their
             * parents' paths differ in the oldest bit alone).  The
             * survivors are then the first K, in (metric, index) order, of
             * the candidates whose partner does not precede them in that
             * order.  The early skip only decides which candidates
             * register in the hash (the links of the new entries): a
             * candidate passes when fewer than K earlier candidates still
             * listed when it is reached precede it; a pair's earlier member
             * leaves the list at its better partner, a later member whose
             * partner is better never enters it.  At K = 32 in bank 1 the
             * write past act[31] can precede a bucket-0 lookup; steps where
             * that is possible (a passing bucket-0 candidate with more than
             * K passing ones before it), and groups of three or more, go to
             * the fixed point below. */
            bool done = false;
            if ((ballot(__popcll(G) > 2) & CAND) == 0) {
                /* pm: the partner's lane bit (a group holds at most two), 0
                 * without one; every per-lane test below is a plain compare
                 * whose ballot is its lane mask, and the rest is scalar */
                const u64 pm = G & ~(1ull << lane);
                const u64 DROP = ballot((LT & pm) != 0);         /* partner precedes: dropped */
                const u64 EARLY = ballot(pm != 0) & ballot(pm <= lowerm);   /* partner has the lower index */
                const u64 E = DROP & EARLY;
                /* this lane takes its earlier, worse partner off the list */
                const u64 LATE = wave_excl_or(lane_in(EARLY & ~DROP) ? pm : 0ull);
                const u64 X = CAND & ~E;
                const u64 PS = ballot(__popcll(LT & lowerm & ~LATE & X) < K) & CAND;
                bool decline = false;
                if (K == TR_MAXK && nb == 1)
                    decline = (PS & ballot((path & 127u) == 0) & ballot(mbcnt(PS) > K)) != 0;
                if (!decline) {
                    const u64 REG = PS & ~(EARLY & ballot((PS & pm) != 0));
                    const u64 tw = REG & G2;
                    if (lane_in(REG)) s.link[nb][lane] = (tw & lowerm) ? (int8_t)ff1(tw) : (int8_t)-1;
                    const u64 F = CAND & ~DROP;
                    const int rk = __popcll(LT & F);
                    myrank = lane_in(F & ballot(rk < K)) ? rk : 255;
                    const int nf = __popcll(F);
                    nsurv = nf < K ? nf : K;
                    done = true;
                }
            }
            if (!done) {
                /* fixed point of early skip + duplicate-path resolution */
                const u64 LTc = LT & CAND & lowerm;
                u64 DS = 0, KM = 0;
                /* ktgt: in the first killer of a registrant, that registrant */
                int ktgt = 64, R = 64;
                bool pass = false, kil = false;
                u64 KILLERS = 0;
                int it;
                for (it = 0; it < 64; it++) {
                    /* registrants already removed when this lane is reached:
                     * those whose first killer comes before it */
                    u64 kb = 0;
                    if (KM) kb = wave_excl_or(ktgt < 64 ? (1ull << ktgt) : 0ull);
                    const int cnt = __popcll(LTc & ~DS & ~kb);
                    pass = cand && cnt < K;
                    const u64 nPASS = ballot(pass);
                    const u64 gp = G & nPASS;
                    R = gp ? ff1(gp) : 64;
                    const bool dup = pass && R != lane;
                    const bool rlt = dup && ((LT >> (R & 63)) & 1);
                    kil = dup && !rlt;
                    const u64 nDS = ballot(rlt);
                    KILLERS = ballot(kil);
                    /* every killer of a registrant is in the registrant's path
                     * group (R is the group's first passing lane): the group's
                     * first killer removes it */
                    const u64 gk = G & KILLERS;
                    const u64 nK = ballot(pass && R == lane && gk != 0);
                    const int nkg = (kil && lane == ff1(gk)) ? R : 64;
                    /* the iteration map depends on (DS, KM, ktgt) only.  Removals
                     * only lower the counts, so every pass set contains the first
                     * one (computed with no removals): when that first set is all
                     * of CAND it never changes and this state is the fixed point */
                    const bool same = (it == 0 && nPASS == CAND) ||
                                      (nDS == DS && nK == KM && ballot(nkg != ktgt) == 0);
                    DS = nDS;
                    KM = nK;
                    ktgt = nkg;
                    if (same) break;
                }
                if (it >= 64) slow = true;

                const bool reg = pass && R == lane;
                const u64 REG = ballot(reg);
                if (!slow && K == TR_MAXK && nb == 1) {
                    const u64 gk = G & KILLERS;
                    const u64 FIRSTK = ballot(kil && gk && lane == ff1(gk));
                    const u64 INC = REG | (KILLERS & ~FIRSTK);
                    const u64 CL = ballot(reg && mbcnt(INC) >= K);
                    if (CL) {
                        const int fc = ff1(CL);
                        if (ballot(pass && (path & 127u) == 0 && lane > fc)) slow = true;
                    }
                }
                if (!slow) {
                    /* hash links of the new registrations: the previous
                     * registrant of the same bucket in index order.  With paths
                     * below 256 a bucket holds at most two registrants, one for
                     * each of the paths p and p ^ 128 */
                    const u64 tw = REG & G2;
                    if (reg) s.link[nb][lane] = (tw & lowerm) ? (int8_t)ff1(tw) : (int8_t)-1;
                    const u64 F = CAND & ~DS & ~KM;
                    const int rk = __popcll(LT & F);
                    myrank = lane_in(F & ballot(rk < K)) ? rk : 255;
                    const int nf = __popcll(F);
                    nsurv = nf < K ? nf : K;
                }
            }
        }
        if (slow) {
            slow_steps++;
            int e2 = 0;
            nsurv = wide_prune<P, HW>(s, nb, m, K, snap, &e2, cost, path, cnext, LT, N1, lane, live);
            if (e2) err = e2;
            myrank = live ? s.rank[lane] : 255;
        }

        /* ---- 3. commit ---- */
        const bool surv = live && myrank < nsurv;
        const u64 B0 = ballot(myrank == 0) & (nsurv > 0 ? ALL : 0ull);
        const int r0 = ff1(B0);
        const int outbit = (int)((N1 >> r0) & 1);
        const double minc = rl_d(cost, r0);
        /* every lane takes its child's record; only survivors' records are
         * referenced afterwards (through src[] and myr), so the previous
         * record is dead here and shares registers with the new one */
#pragma unroll
        for (int q = 0; q < P; q++) rst[q] = d[q];
        const double nc = cost - minc;
        rcost = nc;
        rpath = path;
        rnext = nnew;
        {
            const int pw = pos >> 6;
            const u64 pbit = 1ull << (pos & 63);
#pragma unroll
            for (int w = 0; w < HW; w++) {
                u64 hv = phist[w];
                /* one word: pos < L <= 64 is always in it */
                if (HW == 1 || w == pw) hv = (hv & ~pbit) | (brm & pbit);
                rhist[w] = hv;
            }
        }
        if (surv) {
            s.ncost[nb][lane] = nc;
            s.src[myrank] = (uint8_t)(lane << 2);
        }
        myr = surv ? myrank : 255;
        wsync();
        num_cands = nsurv;
        pos = next_pos;
        idx = nb;

        /* output bookkeeping (src/trellis.c:1013-1040) */
        if (kc != 0) {
            kc--;
        } else {
            acc |= (u64)outbit << nacc;
            if (++nacc == 64) {
                if (lane == 0) out[kept >> 6] = acc;
                kept += 64;
                acc = 0;
                nacc = 0;
            }
        }
    }
    if (pending < L) pending = (int)((int64_t)pending + n_in < (int64_t)L ? (int64_t)pending + n_in : (int64_t)L);
    if (nacc) {
        if (lane == 0) out[kept >> 6] = acc;
        kept += nacc;
    }

    /* ---- store the final context ---- */
    const int ctx_out = fi(jb->ctx_out);
    if (ctx_out >= 0) {
        tr_ctx *co = &ctxs[ctx_out];
        if (myr < num_cands) {
#pragma unroll
            for (int q = 0; q < P; q++) co->st[myr][q] = rst[q];
#pragma unroll
            for (int w = 0; w < HW; w++) co->hist[myr][w] = rhist[w];
            co->cost[myr] = rcost;
            co->path[myr] = rpath;
            co->next[myr] = (uint8_t)rnext;
            /* after a step the survivor sits in its child's lane, 2*parent + branch */
            co->slot[myr] = n_in > 0 ? (uint8_t)lane : c0->slot[lane];
        }
        bool fr = false;
        for (int b = 0; b < 2; b++) {
            co->link[b][lane] = s.link[b][lane];
            co->node_path[b][lane] = s.npath[b][lane];
            co->node_cost[b][lane] = s.ncost[b][lane];
            fr |= s.link[b][lane] >= TR_LINK_FOREIGN;
        }
        const bool anyf = ballot(fr) != 0;
        if (lane == 0) {
            co->num_cands = num_cands;
            co->pos = pos;
            co->pending = pending;
            co->idx = idx;
            co->draining = draining;
            co->has_foreign = anyf ? 1 : 0;
            co->err = err;
        }
    }
    if (lane == 0) {
        jb->kept = kept;
        jb->slow_steps = slow_steps;
        jb->err = err;
    }
}

/* Occupancy target: with the survivor records in registers a wave needs
 * 6 KB of LDS, so registers set the waves per SIMD; up to 128-sample
 * histories the kernel fits 80 VGPRs (6 waves per SIMD, 24 per CU) with a
 * few spills outside the step loop.  Wider histories keep their registers. */
template <int HW>
struct TrWaves {
    static constexpr int value = HW <= 2 ? 6 : 1;
};

/* One workgroup (one wave) per job. */
template <int P, int HW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TrWaves<HW>::value, 8))) void tr_run_jobs(tr_params prm, tr_job *jobs, tr_ctx *ctxs)
{
    __shared__ TrLds<P, HW> s;
    tr_job *jb = &jobs[blockIdx.x];
    if (jb->skip) return;
    if (prm.state_limit > 0.0) tr_job_body<P, HW, false, false, 1>(prm, jb, ctxs, s, threadIdx.x);
    else tr_job_body<P, HW, false, false, 0>(prm, jb, ctxs, s, threadIdx.x);
}

/* The same for launches of a few jobs (segment 0 of a seg0-first step, a
 * plain stream), each of which has a SIMD to itself: registers are not
 * rationed for occupancy, and the sweep loads all its keys at once. */
template <int P, int HW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 8))) void tr_run_jobs_lone(tr_params prm, tr_job *jobs, tr_ctx *ctxs)
{
    __shared__ TrLds<P, HW> s;
    tr_job *jb = &jobs[blockIdx.x];
    if (jb->skip) return;
    if (prm.state_limit > 0.0) tr_job_body<P, HW, false, false, 1, true>(prm, jb, ctxs, s, threadIdx.x);
    else tr_job_body<P, HW, false, false, 0, true>(prm, jb, ctxs, s, threadIdx.x);
}

/* The same with the order-8 FMA filter of the reference's AVX2 build. */
template <int HW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TrWaves<HW>::value, 8))) void tr_run_jobs_f8(tr_params prm, tr_job *jobs, tr_ctx *ctxs)
{
    __shared__ TrLds<8, HW> s;
    tr_job *jb = &jobs[blockIdx.x];
    if (jb->skip) return;
    tr_job_body<8, HW, true>(prm, jb, ctxs, s, threadIdx.x);
}

/* Persistent form: WPB waves per workgroup, each taking the next job slot
 * from a device counter until `njobs` slots are handed out; slots marked
 * `skip` are passed over.  Every wave leaves once the counter is past the
 * end, so the grid drains.  The counter is reset by the caller. */
template <int P, int HW, int WPB>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(TrWaves<HW>::value, 8))) void tr_run_queue(tr_params prm, tr_job *jobs, int njobs,
                                                         int *counter, tr_ctx *ctxs)
{
    __shared__ TrLds<P, HW> s[WPB];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (;;) {
        int j = 0;
        if (lane == 0) j = atomicAdd(counter, 1);
        j = __builtin_amdgcn_readfirstlane(j);
        if (j >= njobs) break;
        tr_job *jb = &jobs[j];
        if (jb->skip) continue;
        tr_job_body<P, HW>(prm, jb, ctxs, s[w], lane);
    }
}

/* Host-fed queue: the wait of one wave (lane 0) for job j.  Returns 0 once
 * the host has published j, 1 on the host's abort flag or when the wait
 * exceeds `limit` wall-clock ticks or `max_iter` rounds (dc->err is then
 * set); both bounds are checked on every round.  Waves read the device copy
 * of the published count; the host word is read by one wave at a time, at
 * most once per `poll` ticks (a compare-and-swap on the time of the last
 * read), which raises the device copy for every waiting wave. */
__device__ int feed_wait(const tr_feed_host *hc, tr_feed_dev *dc, int j, uint64_t poll, uint64_t limit,
                         uint64_t max_iter)
{
    const uint64_t t_start = (uint64_t)wall_clock64();
    for (uint64_t iter = 0;; iter++) {
        const int r = __hip_atomic_load(&dc->ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (r > j &&
            __hip_atomic_load(&hc->ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) > j)
            return 0;
        if (__hip_atomic_load(&dc->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return 1;
        const uint64_t now = (uint64_t)wall_clock64();
        if ((now > t_start && now - t_start > limit) || iter >= max_iter) {
            __hip_atomic_fetch_or(&dc->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return 1;
        }
        uint64_t last = __hip_atomic_load(&dc->last_poll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (now > last && now - last >= poll &&
            __hip_atomic_compare_exchange_strong(&dc->last_poll, &last, now, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            const int h = __hip_atomic_load(&hc->ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            const int a = __hip_atomic_load(&hc->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_fetch_max(&dc->ready, h, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (a) {
                __hip_atomic_store(&dc->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return 1;
            }
            if (h > j) return 0;
        }
        __builtin_amdgcn_s_sleep(16);
    }
}

static_assert(sizeof(tr_job) % 8 == 0, "tr_job is copied as 64-bit words");
static_assert(offsetof(tr_job, slow_steps) == offsetof(tr_job, kept) + 8 &&
                  offsetof(tr_job, err) == offsetof(tr_job, kept) + 16 && sizeof(tr_job) == offsetof(tr_job, kept) + 24,
              "the results are the descriptor's last three words");

/* Host-fed launch, one wave per workgroup and one job per workgroup:
 * workgroup j waits until the host has published descriptor j (in coherent
 * pinned host memory), copies it into LDS, runs it, and stores the
 * descriptor with its results into the device table `jobs` (read back by
 * the host after the launch).  The host publishes in index order and
 * workgroups are dispatched in index order, so the waiting workgroups are
 * the next ones to run; a wave leaves early on the host's abort flag or
 * after a wait longer than `limit`, so the grid always drains. */
/* The host-fed launch holds at most a few waves per CU (its LDS request is
 * padded to 6 per CU by default): registers for two waves per SIMD, which
 * the sweep spends on loading all its keys at once. */
template <int HW>
struct TrFeedWaves {
    static constexpr int value = HW <= 2 ? 2 : 1;
};

template <int P, int HW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(TrFeedWaves<HW>::value, 8))) void tr_run_feed(
    tr_params prm, const tr_job *hjobs, tr_job *jobs, const tr_feed_host *hc, tr_feed_dev *dc, tr_ctx *ctxs,
    uint64_t poll, uint64_t limit, uint64_t max_iter)
{
    __shared__ TrLds<P, HW> s;
    __shared__ tr_job jd;
    constexpr int NW = (int)(sizeof(tr_job) / 8);
    const int lane = threadIdx.x;
    const int j = blockIdx.x;
    int st = 0;
    if (lane == 0) st = feed_wait(hc, dc, j, poll, limit, max_iter);
    if (__builtin_amdgcn_readfirstlane(st)) return;
    /* the jobs published last set the launch's end: they take issue
     * priority over the earlier ones that share their SIMD */
    const int nj = (int)gridDim.x;
    if (j >= nj - nj / 32) __builtin_amdgcn_s_setprio(3);
    else if (j >= nj - nj / 8) __builtin_amdgcn_s_setprio(2);
    else if (j >= nj - nj / 4) __builtin_amdgcn_s_setprio(1);
    /* lane w copies word w of the descriptor into LDS */
    if (lane < NW)
        reinterpret_cast<uint64_t *>(&jd)[lane] = __hip_atomic_load(
            reinterpret_cast<const uint64_t *>(&hjobs[j]) + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    wsync();
    if (!uni(jd.skip)) {
        if (prm.state_limit > 0.0) tr_job_body<P, HW, false, true, 1, true>(prm, &jd, ctxs, s, lane);
        else tr_job_body<P, HW, false, true, 0, true>(prm, &jd, ctxs, s, lane);
    }
    wsync();
    if (lane < NW) reinterpret_cast<uint64_t *>(&jobs[j])[lane] = reinterpret_cast<const uint64_t *>(&jd)[lane];
    if (lane == 0) __hip_atomic_fetch_add(&dc->done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

/* resets the device side of a host-fed queue (a kernel, on the launch's stream) */
__global__ void tr_feed_init(tr_feed_dev *dc)
{
    if (threadIdx.x == 0) {
        dc->done = 0;
        dc->ready = 0;
        dc->abort = 0;
        dc->err = 0;
        dc->last_poll = 0;
        dc->t0 = 0;
    }
}

#define TF_INST(P) \
    template __global__ void tr_run_feed<P, 1>(tr_params, const tr_job *, tr_job *, const tr_feed_host *, \
                                              tr_feed_dev *, tr_ctx *, uint64_t, uint64_t, uint64_t); \
    template __global__ void tr_run_feed<P, 2>(tr_params, const tr_job *, tr_job *, const tr_feed_host *, \
                                              tr_feed_dev *, tr_ctx *, uint64_t, uint64_t, uint64_t);
TF_INST(4)
TF_INST(5)
TF_INST(6)
TF_INST(7)
TF_INST(8)

/* the few-job form where the occupancy target applies (TrWaves above 1) */
#define TL_INST(P) \
    template __global__ void tr_run_jobs_lone<P, 1>(tr_params, tr_job *, tr_ctx *); \
    template __global__ void tr_run_jobs_lone<P, 2>(tr_params, tr_job *, tr_ctx *);
TL_INST(4)
TL_INST(5)
TL_INST(6)
TL_INST(7)
TL_INST(8)

#define TR_INST(P, HW) \
    template __global__ void tr_run_jobs<P, HW>(tr_params, tr_job *, tr_ctx *); \
    template __global__ void tr_run_queue<P, HW, 1>(tr_params, tr_job *, int, int *, tr_ctx *); \
    template __global__ void tr_run_queue<P, HW, 2>(tr_params, tr_job *, int, int *, tr_ctx *); \
    template __global__ void tr_run_queue<P, HW, 4>(tr_params, tr_job *, int, int *, tr_ctx *);
#define TR_INST_P(P) TR_INST(P, 1) TR_INST(P, 2) TR_INST(P, 4) TR_INST(P, 8) TR_INST(P, 16) TR_INST(P, 32)
TR_INST_P(4)
TR_INST_P(5)
TR_INST_P(6)
TR_INST_P(7)
TR_INST_P(8)
template __global__ void tr_run_jobs_f8<1>(tr_params, tr_job *, tr_ctx *);
template __global__ void tr_run_jobs_f8<2>(tr_params, tr_job *, tr_ctx *);
template __global__ void tr_run_jobs_f8<4>(tr_params, tr_job *, tr_ctx *);
template __global__ void tr_run_jobs_f8<8>(tr_params, tr_job *, tr_ctx *);
template __global__ void tr_run_jobs_f8<16>(tr_params, tr_job *, tr_ctx *);
template __global__ void tr_run_jobs_f8<32>(tr_params, tr_job *, tr_ctx *);

extern "C" int tr_lane_eligible(const tr_params *prm);
extern "C" int tr_lane_launch(const tr_params *prm, tr_job *d_jobs, int njobs, tr_ctx *d_ctxs,
                              hipStream_t stream);

/* Host-side launcher table (P in 4..8, HW in {1,2,4,8,16,32}).  At K <= 2
 * the jobs run one per lane (tr_lane_kernel.hip) unless prm asks for the
 * one-wave-per-job kernel. */
/* compute units of the current device */
static int tr_cu_count()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        return 0;
    return n;
}

extern "C" int tr_launch_jobs(const tr_params *prm, tr_job *d_jobs, int njobs,
                              tr_ctx *d_ctxs, hipStream_t stream)
{
    if (njobs <= 0) return 0;
    if (tr_lane_eligible(prm)) return tr_lane_launch(prm, d_jobs, njobs, d_ctxs, stream);
    const int P = prm->P;
    const int hw_need = (prm->L + 63) / 64;
    int HW = 1;
    while (HW < hw_need) HW *= 2;
    if (P < 4 || P > 8 || HW > 32) return -1;
    dim3 grid((unsigned)njobs), block(64);
    if (prm->fma_o8 && P == 8) {
#define TF_CASE(HH) \
        if (HW == HH) { \
            hipLaunchKernelGGL((tr_run_jobs_f8<HH>), grid, block, 0, stream, *prm, d_jobs, d_ctxs); \
            return hipGetLastError() == hipSuccess ? 0 : -1; }
        TF_CASE(1) TF_CASE(2) TF_CASE(4) TF_CASE(8) TF_CASE(16) TF_CASE(32)
        return -1;
    }
    /* a launch of at most one job per CU gives each job a SIMD of its own */
    if (HW <= 2 && njobs <= tr_cu_count()) {
#define TL_CASE(PP, HH) \
        if (P == PP && HW == HH) { \
            hipLaunchKernelGGL((tr_run_jobs_lone<PP, HH>), grid, block, 0, stream, *prm, d_jobs, d_ctxs); \
            return hipGetLastError() == hipSuccess ? 0 : -1; }
        TL_CASE(4, 1) TL_CASE(4, 2) TL_CASE(5, 1) TL_CASE(5, 2) TL_CASE(6, 1) TL_CASE(6, 2)
        TL_CASE(7, 1) TL_CASE(7, 2) TL_CASE(8, 1) TL_CASE(8, 2)
        return -1;
    }
#define TR_CASE(PP, HH) \
    if (P == PP && HW == HH) { \
        hipLaunchKernelGGL((tr_run_jobs<PP, HH>), grid, block, 0, stream, *prm, d_jobs, d_ctxs); \
        return hipGetLastError() == hipSuccess ? 0 : -1; }
#define TR_CASE_P(PP) TR_CASE(PP, 1) TR_CASE(PP, 2) TR_CASE(PP, 4) TR_CASE(PP, 8) TR_CASE(PP, 16) TR_CASE(PP, 32)
    TR_CASE_P(4) TR_CASE_P(5) TR_CASE_P(6) TR_CASE_P(7) TR_CASE_P(8)
    return -1;
}

/* Persistent launch: nblocks workgroups of 64*wpb threads over njobs slots.
 * At K <= 2 the slots run one per lane instead (the counter is not used). */
extern "C" int tr_launch_queue(const tr_params *prm, tr_job *d_jobs, int njobs, int *d_counter,
                               tr_ctx *d_ctxs, int nblocks, int wpb, hipStream_t stream)
{
    if (njobs <= 0) return 0;
    if (tr_lane_eligible(prm)) return tr_lane_launch(prm, d_jobs, njobs, d_ctxs, stream);
    const int P = prm->P;
    const int hw_need = (prm->L + 63) / 64;
    int HW = 1;
    while (HW < hw_need) HW *= 2;
    if (P < 4 || P > 8 || HW > 32 || nblocks < 1) return -1;
    if (wpb != 1 && wpb != 2 && wpb != 4) return -1;
    if (prm->fma_o8 && P == 8) return -1;       /* the FMA filter runs in tr_run_jobs only */
    dim3 grid((unsigned)nblocks), block(64 * wpb);
#define TQ_CASE(PP, HH, WW) \
    if (P == PP && HW == HH && wpb == WW) { \
        hipLaunchKernelGGL((tr_run_queue<PP, HH, WW>), grid, block, 0, stream, *prm, d_jobs, njobs, \
                           d_counter, d_ctxs); \
        return hipGetLastError() == hipSuccess ? 0 : -1; }
#define TQ_CASE_W(PP, HH) TQ_CASE(PP, HH, 1) TQ_CASE(PP, HH, 2) TQ_CASE(PP, HH, 4)
#define TQ_CASE_P(PP) TQ_CASE_W(PP, 1) TQ_CASE_W(PP, 2) TQ_CASE_W(PP, 4) TQ_CASE_W(PP, 8) \
    TQ_CASE_W(PP, 16) TQ_CASE_W(PP, 32)
    TQ_CASE_P(4) TQ_CASE_P(5) TQ_CASE_P(6) TQ_CASE_P(7) TQ_CASE_P(8)
    return -1;
}

/* Whether the host-fed launch covers prm: the one-wave kernel without the
 * FMA filter, histories up to 128 samples. */
extern "C" int tr_feed_eligible(const tr_params *prm)
{
    return !tr_lane_eligible(prm) && !(prm->fma_o8 && prm->P == 8) && prm->P >= 4 && prm->P <= 8 &&
           prm->L >= 1 && prm->L <= 128;
}

/* Host-fed launch: resets dc on `stream`, then one one-wave workgroup per
 * entry of h_jobs (coherent pinned host memory, visible to the device at
 * that address) runs it once the host has published it through hc.
 * poll / limit in wall-clock ticks, max_iter in wait rounds; lds_pad bytes
 * of dynamic LDS per workgroup (unused) cap the resident workgroups per CU.
 * The launch waits for the host: nothing the host needs to publish may be
 * queued behind it (streams can share a hardware queue, so queue such work
 * before the launch, on the same stream). */
extern "C" int tr_launch_feed(const tr_params *prm, const tr_job *h_jobs, tr_job *d_jobs, int njobs,
                              const tr_feed_host *hc, tr_feed_dev *dc, tr_ctx *d_ctxs, uint64_t poll,
                              uint64_t limit, uint64_t max_iter, size_t lds_pad, hipStream_t stream)
{
    if (njobs <= 0) return 0;
    if (!tr_feed_eligible(prm)) return -1;
    const int P = prm->P, HW = prm->L > 64 ? 2 : 1;
    hipLaunchKernelGGL(tr_feed_init, dim3(1), dim3(64), 0, stream, dc);
    if (hipGetLastError() != hipSuccess) return -1;
    dim3 grid((unsigned)njobs), block(64);
#define TFL_CASE(PP, HH) \
    if (P == PP && HW == HH) { \
        hipLaunchKernelGGL((tr_run_feed<PP, HH>), grid, block, lds_pad, stream, *prm, h_jobs, d_jobs, hc, dc, \
                           d_ctxs, poll, limit, max_iter); \
        return hipGetLastError() == hipSuccess ? 0 : -1; }
    TFL_CASE(4, 1) TFL_CASE(5, 1) TFL_CASE(6, 1) TFL_CASE(7, 1) TFL_CASE(8, 1)
    TFL_CASE(4, 2) TFL_CASE(5, 2) TFL_CASE(6, 2) TFL_CASE(7, 2) TFL_CASE(8, 2)
    return -1;
}

/* Static LDS bytes of the host-fed kernel for prm (-1 if not covered). */
extern "C" int tr_feed_static_lds(const tr_params *prm)
{
    const int P = prm->P, HW = prm->L > 64 ? 2 : 1;
    hipFuncAttributes fa;
#define TFS_CASE(PP, HH) \
    if (P == PP && HW == HH) { \
        if (hipFuncGetAttributes(&fa, (const void *)tr_run_feed<PP, HH>) != hipSuccess) return -1; \
        return (int)fa.sharedSizeBytes; }
    TFS_CASE(4, 1) TFS_CASE(5, 1) TFS_CASE(6, 1) TFS_CASE(7, 1) TFS_CASE(8, 1)
    TFS_CASE(4, 2) TFS_CASE(5, 2) TFS_CASE(6, 2) TFS_CASE(7, 2) TFS_CASE(8, 2)
    return -1;
}

/* Resident workgroups (one wave each) per CU of tr_run_jobs for prm. */
extern "C" int tr_occupancy(const tr_params *prm)
{
    const int P = prm->P;
    const int hw_need = (prm->L + 63) / 64;
    int HW = 1;
    while (HW < hw_need) HW *= 2;
    int n = 0;
#define TO_CASE(PP, HH) \
    if (P == PP && HW == HH) { \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)tr_run_jobs<PP, HH>, 64, 0) != hipSuccess) return -1; \
        return n; }
#define TO_CASE_P(PP) TO_CASE(PP, 1) TO_CASE(PP, 2) TO_CASE(PP, 4) TO_CASE(PP, 8) TO_CASE(PP, 16) TO_CASE(PP, 32)
    TO_CASE_P(4) TO_CASE_P(5) TO_CASE_P(6) TO_CASE_P(7) TO_CASE_P(8)
    return -1;
}
