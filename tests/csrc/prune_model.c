/*
 * CPU model of the HIP trellis kernel's wave-parallel prune ("bulk path"),
 * written with 64-bit lane masks exactly as the kernel computes them, and a
 * driver that runs it beside the literal sequential prune of the oracle on
 * every step of a trellis run.  Used by tests/test_prune_model.py to show
 * that the bulk formulation reproduces the reference's sequential
 * sdm_sort_cands (src/trellis.c:621-731) whenever it does not request the
 * literal fallback.  TEST INFRASTRUCTURE ONLY.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include "../../oracle/dsd_oracle.h"

typedef uint64_t u64;

static inline u64 lower_mask(int i) { return i >= 64 ? ~0ull : ((1ull << i) - 1); }
static inline int ff1(u64 m) { return m ? __builtin_ctzll(m) : -1; }
static inline int pc(u64 m) { return __builtin_popcountll(m); }
static inline int64_t b64(double d) { int64_t i; memcpy(&i, &d, 8); return i; }

/* Result of the bulk prune. */
typedef struct {
    int slow;            /* 1: the literal path must decide this step */
    int n;               /* survivors */
    int act[OR_MAXK];    /* child lane of each survivor, ascending cost */
    u64 reg;             /* registered children (new hash entries) */
    int plist[64];       /* for registered children: previous same-bucket registrant or -1 */
    int iters;
} bulk_res;

/*
 * m children with cost/path/next; K survivors max; bank of the sort.
 * Mirrors the kernel: strict less-than masks, candidate filter, fixed-point
 * resolution of early skips and duplicate paths, final rank.
 */
int bulk_prune(int m, const double *cost, const uint32_t *path, const uint8_t *next,
               int K, int bank, bulk_res *r)
{
    u64 ALL = (m == 64) ? ~0ull : ((1ull << m) - 1);
    u64 LT[64];
    memset(r, 0, sizeof(*r));
    for (int i = 0; i < m; i++) {
        u64 mk = 0;
        for (int j = 0; j < m; j++)
            if (b64(cost[j]) < b64(cost[i])) mk |= 1ull << j;
        LT[i] = mk;
    }
    /* ties: strict ranks must be a permutation */
    u64 seen = 0;
    for (int i = 0; i < m; i++) seen |= 1ull << pc(LT[i] & ALL);
    if (pc(seen) != m) { r->slow = 1; return 0; }

    int min0 = -1;
    for (int i = 0; i < m; i++) if ((LT[i] & ALL) == 0) min0 = i;
    u64 N1 = 0;
    for (int i = 0; i < m; i++) if (next[i] & 1) N1 |= 1ull << i;
    int v1 = pc(N1);
    int maj = v1 > m - v1;
    int mn = min0;
    if ((int)next[min0] != maj) {
        u64 MJ = maj ? N1 : (ALL & ~N1);
        int bm = -1;
        for (int i = 0; i < m; i++) if (((MJ >> i) & 1) && (LT[i] & MJ) == 0) bm = i;
        if (bm >= 0 && cost[bm] < cost[min0] * 1.1) mn = bm;
    }
    int M = next[mn];
    u64 CAND = 0;
    for (int i = 0; i < m; i++) if (next[i] == M) CAND |= 1ull << i;

    /* same-path groups among candidates (table of masks indexed by path) */
    u64 G[64];
    for (int i = 0; i < m; i++) {
        G[i] = 0;
        if (!((CAND >> i) & 1)) continue;
        for (int j = 0; j < m; j++)
            if (((CAND >> j) & 1) && path[j] == path[i]) G[i] |= 1ull << j;
    }

    u64 PASS = 0, DS = 0, KILLED = 0;
    int killt[64];           /* time (killer lane) of each killed registrant */
    for (int i = 0; i < 64; i++) killt[i] = 64;
    int it;
    for (it = 0; it < 66; it++) {
        u64 npass = 0;
        for (int c = 0; c < m; c++) {
            if (!((CAND >> c) & 1)) continue;
            u64 kb = 0;
            for (int d = 0; d < m; d++) if (killt[d] < c) kb |= 1ull << d;
            int cnt = pc(LT[c] & CAND & lower_mask(c) & ~DS & ~kb);
            if (cnt < K) npass |= 1ull << c;
        }
        u64 nds = 0, killers = 0;
        int nkt[64];
        for (int i = 0; i < 64; i++) nkt[i] = 64;
        for (int s = 0; s < m; s++) {
            if (!((npass >> s) & 1)) continue;
            int R = ff1(G[s] & npass);
            if (R == s) continue;
            if ((LT[s] >> R) & 1) nds |= 1ull << s;
            else {
                killers |= 1ull << s;
                if (s < nkt[R]) nkt[R] = s;
            }
        }
        int same = (npass == PASS) && (nds == DS);
        for (int i = 0; i < 64 && same; i++) if (nkt[i] != killt[i]) same = 0;
        PASS = npass;
        DS = nds;
        memcpy(killt, nkt, sizeof(killt));
        KILLED = 0;
        for (int i = 0; i < 64; i++) if (killt[i] < 64) KILLED |= 1ull << i;
        (void)killers;
        if (same) break;
    }
    r->iters = it;
    if (it >= 66) { r->slow = 1; return 0; }

    /* registrants: passing candidates that are the first passing member */
    u64 REG = 0, INC = 0;
    for (int s = 0; s < m; s++) {
        if (!((PASS >> s) & 1)) continue;
        int R = ff1(G[s] & PASS);
        if (R == s) { REG |= 1ull << s; INC |= 1ull << s; }
        else if (!((DS >> s) & 1) && killt[R] != s) INC |= 1ull << s;   /* later killer: NF */
    }
    if (K == OR_MAXK && bank == 1) {
        int fc = -1;
        for (int c = 0; c < m; c++)
            if (((REG >> c) & 1) && pc(INC & lower_mask(c)) >= K) { fc = c; break; }
        if (fc >= 0) {
            for (int c = fc + 1; c < m; c++)
                if (((PASS >> c) & 1) && (path[c] & 127) == 0) { r->slow = 1; return 0; }
        }
    }
    r->reg = REG;
    for (int s = 0; s < m; s++) {
        r->plist[s] = -1;
        if (!((REG >> s) & 1)) continue;
        for (int j = s - 1; j >= 0; j--)
            if (((REG >> j) & 1) && (path[j] & 127) == (path[s] & 127)) { r->plist[s] = j; break; }
    }
    u64 F = CAND & ~DS & ~KILLED;
    int nF = pc(F);
    r->n = nF < K ? nF : K;
    for (int j = 0; j < m; j++) {
        if (!((F >> j) & 1)) continue;
        int rk = pc(LT[j] & F);
        if (rk < K) r->act[rk] = j;
    }
    return 0;
}

/*
 * Closed form of the same prune when every same-path group among the
 * candidates has at most two members (two children share a path only when
 * their parents' paths differ in the oldest bit alone, and parents' paths
 * are distinct unless an earlier step kept two entries of one path):
 *   - a member whose partner precedes it in (metric, index) order is
 *     dropped; the survivors are the first K of the rest in that order;
 *   - the early skip decides only which members register in the hash (and
 *     so the links of the new entries): a candidate passes when fewer than
 *     K earlier candidates that are still listed when it is reached precede
 *     it in (metric, index) order.  A pair's earlier member leaves the list
 *     when its better partner is reached; a pair's later member never
 *     enters it when its partner is better;
 *   - at K = 32 in bank 1 a lookup in bucket 0 after an insertion into the
 *     full list follows the entry written past act[31]: declined (the
 *     fixed point decides) whenever a passing candidate with a bucket-0
 *     path has more than K passing candidates before it.
 * Returns 1 when the closed form declines (groups of three or more, the
 * bucket-0 condition), else fills r like bulk_prune.
 */
int pair_prune(int m, const double *cost, const uint32_t *path, const uint8_t *next, int K, int bank,
               bulk_res *r)
{
    u64 ALL = (m == 64) ? ~0ull : ((1ull << m) - 1);
    u64 LT[64];
    memset(r, 0, sizeof(*r));
    for (int i = 0; i < m; i++) {
        u64 mk = 0;
        for (int j = 0; j < m; j++)
            if (b64(cost[j]) < b64(cost[i]) || (b64(cost[j]) == b64(cost[i]) && j < i)) mk |= 1ull << j;
        LT[i] = mk;
    }
    int min0 = -1;
    for (int i = 0; i < m; i++) if ((LT[i] & ALL) == 0) min0 = i;
    u64 N1 = 0;
    for (int i = 0; i < m; i++) if (next[i] & 1) N1 |= 1ull << i;
    int v1 = pc(N1);
    int maj = v1 > m - v1;
    int mn = min0;
    if ((int)next[min0] != maj) {
        u64 MJ = maj ? N1 : (ALL & ~N1);
        int bm = -1;
        for (int i = 0; i < m; i++) if (((MJ >> i) & 1) && (LT[i] & MJ) == 0) bm = i;
        if (bm >= 0 && cost[bm] < cost[min0] * 1.1) mn = bm;
    }
    int M = next[mn];
    u64 CAND = 0;
    for (int i = 0; i < m; i++) if (next[i] == M) CAND |= 1ull << i;
    u64 G[64], G2[64];
    for (int i = 0; i < m; i++) {
        G[i] = G2[i] = 0;
        if (!((CAND >> i) & 1)) continue;
        for (int j = 0; j < m; j++) {
            if (((CAND >> j) & 1) && path[j] == path[i]) G[i] |= 1ull << j;
            if (((CAND >> j) & 1) && path[j] == (path[i] ^ 128u)) G2[i] |= 1ull << j;
        }
        if (pc(G[i]) > 2) return 1;
    }
    /* partners, drops */
    int pl[64];
    u64 DROP = 0, E = 0, LATE[64];
    for (int i = 0; i < m; i++) {
        pl[i] = -1;
        LATE[i] = 0;
        if (!((CAND >> i) & 1)) continue;
        u64 pm = G[i] & ~(1ull << i);
        if (pm) pl[i] = ff1(pm);
    }
    for (int i = 0; i < m; i++) {
        if (pl[i] < 0) continue;
        if ((LT[i] >> pl[i]) & 1) {
            DROP |= 1ull << i;
            if (pl[i] < i) E |= 1ull << i;
        }
    }
    /* exclusive prefix over lanes p of the earlier, worse partner's bit */
    u64 acc = 0;
    for (int i = 0; i < m; i++) {
        LATE[i] = acc;
        if (pl[i] >= 0 && pl[i] < i && !((DROP >> i) & 1)) acc |= 1ull << pl[i];
    }
    u64 PASS = 0;
    for (int i = 0; i < m; i++) {
        if (!((CAND >> i) & 1)) continue;
        int cnt = pc(LT[i] & CAND & lower_mask(i) & ~(E | LATE[i]));
        if (cnt < K) PASS |= 1ull << i;
    }
    u64 REG = 0;
    for (int i = 0; i < m; i++) {
        if (!((PASS >> i) & 1)) continue;
        if (pl[i] >= 0 && pl[i] < i && ((PASS >> pl[i]) & 1)) continue;
        REG |= 1ull << i;
    }
    if (K == OR_MAXK && bank == 1) {
        for (int z = 0; z < m; z++)
            if (((PASS >> z) & 1) && (path[z] & 127) == 0 && pc(PASS & lower_mask(z)) > K) return 1;
    }
    r->reg = REG;
    for (int s = 0; s < m; s++) {
        r->plist[s] = -1;
        if (!((REG >> s) & 1)) continue;
        u64 tw = REG & G2[s] & lower_mask(s);
        r->plist[s] = tw ? ff1(tw) : -1;
    }
    u64 F = CAND & ~DROP;
    int nF = pc(F);
    r->n = nF < K ? nF : K;
    for (int j = 0; j < m; j++) {
        if (!((F >> j) & 1)) continue;
        int rk = pc(LT[j] & F);
        if (rk < K) r->act[rk] = j;
    }
    return 0;
}

/* ---- driver: compare the bulk model with the sequential prune ---- */
static unsigned long long g_cmp[8];   /* steps, slow, mismatches, iters_sum, max_iters */
static void observe(const or_prune_obs *o)
{
    bulk_res r;
    bulk_prune(o->m, o->cost, o->path, o->next, o->K, o->bank, &r);
    g_cmp[0]++;
    if (r.slow) { g_cmp[1]++; return; }
    g_cmp[3] += (unsigned long long)r.iters;
    if ((unsigned long long)r.iters > g_cmp[4]) g_cmp[4] = (unsigned long long)r.iters;
    int bad = (r.n != o->n) || (r.reg != o->reg);
    for (int k = 0; k < o->n && !bad; k++) if (r.act[k] != o->act[k]) bad = 1;
    for (int i = 0; i < o->m && !bad; i++)
        if (((o->reg >> i) & 1) && r.plist[i] != o->plist[i]) bad = 1;
    if (bad) g_cmp[2]++;
}

/* the closed form against the sequential prune: steps, declined, mismatches */
static unsigned long long g_pair[4];
static void observe_pair(const or_prune_obs *o)
{
    bulk_res r;
    g_pair[0]++;
    if (pair_prune(o->m, o->cost, o->path, o->next, o->K, o->bank, &r)) { g_pair[1]++; return; }
    int bad = (r.n != o->n) || (r.reg != o->reg);
    for (int k = 0; k < o->n && !bad; k++) if (r.act[k] != o->act[k]) bad = 1;
    for (int i = 0; i < o->m && !bad; i++)
        if (((o->reg >> i) & 1) && r.plist[i] != o->plist[i]) bad = 1;
    if (bad) g_pair[2]++;
}

void pair_model_run(void *ctx, const double *x, size_t n, unsigned long long *stats)
{
    memset(g_pair, 0, sizeof(g_pair));
    or_prune_observer = observe_pair;
    float *out = (float *)malloc(n * sizeof(float));
    or_sdm_process((or_ctx *)ctx, x, out, n);
    free(out);
    or_prune_observer = NULL;
    memcpy(stats, g_pair, 3 * sizeof(unsigned long long));
}

/* Runs the oracle over x with the observer attached; fills stats[5]. */
void prune_model_run(void *ctx, const double *x, size_t n, unsigned long long *stats)
{
    memset(g_cmp, 0, sizeof(g_cmp));
    or_prune_observer = observe;
    float *out = (float *)malloc(n * sizeof(float));
    or_sdm_process((or_ctx *)ctx, x, out, n);
    free(out);
    or_prune_observer = NULL;
    memcpy(stats, g_cmp, 5 * sizeof(unsigned long long));
}
