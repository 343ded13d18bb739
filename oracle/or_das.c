/*
 * Oracle restatement of Density-Aligned Stitching (DAS).  TEST
 * INFRASTRUCTURE ONLY (see dsd_oracle.h).  Compile with -ffp-contract=off.
 *
 * Follows src/dsp_plugin.c of the reference, generalised from its per-chunk
 * limit of 8 segments to any count:
 *   nominal split                   :1941-1957
 *   boundary reposition (ch0)       :1963-2034
 *   seg0-first: seg0 on persistent, seg1 = exact continuation, 16-level
 *   estimate for seg>=2             :2185-2266 (+ src/threadpool.c:345-366)
 *   full-parallel seeding           :2268-2283
 *   segment dispatch                :2288-2363 (+ src/engine.c:915-938)
 *   stitch search on ch0            :2411-2681
 *   apply to other channels         :2683-2713
 *   copy-back of the last segment   :2715-2724
 * One difference in execution order: in full-parallel mode the reference
 * runs seg0 on the persistent context concurrently with the segments that
 * were copied from it; here the copied segments run first, which is the
 * same as them reading the persistent context as it was when copied.
 */
#include "dsd_oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

int or_das_plan_init(or_das_plan_t *p, size_t F, size_t seg_len, int nseg,
                     size_t overlap, int seg0_first, int reposition)
{
    memset(p, 0, sizeof(*p));
    int S;
    if (seg_len > 0) {
        S = (int)(F / seg_len);
        if (S < 1) S = 1;
    } else {
        S = nseg < 1 ? 1 : nseg;
        size_t min_seg = overlap * 4;
        if (min_seg < 256) min_seg = 256;
        while (S > 1 && F / (size_t)S < min_seg) S--;
    }
    p->nseg = S;
    p->overlap = overlap;
    p->seg0_first = seg0_first;
    p->reposition = reposition;
    p->nom_start = (size_t *)calloc((size_t)S, sizeof(size_t));
    p->nom_size = (size_t *)calloc((size_t)S, sizeof(size_t));
    /* sizes by the reference's rule for S segments (src/dsp_plugin.c:1944-
     * 1953); a segment length only sets the count */
    size_t pos = 0;
    size_t base = F / (size_t)S, rem = F % (size_t)S;
    for (int s = 0; s < S; s++) {
        p->nom_start[s] = pos;
        p->nom_size[s] = (s == S - 1) ? F - pos : base + ((size_t)s < rem ? 1 : 0);
        pos += p->nom_size[s];
    }
    return S;
}

void or_das_plan_free(or_das_plan_t *p)
{
    free(p->nom_start);
    free(p->nom_size);
    memset(p, 0, sizeof(*p));
}

static double rms_at(const double *x, size_t c, size_t w)
{
    double sum_sq = 0.0;
    for (size_t i = c - w / 2; i < c + w / 2; i++) sum_sq += x[i] * x[i];
    return sqrt(sum_sq / (double)w);
}

void or_das_reposition(or_das_plan_t *p, const double *ch0, size_t F)
{
    const int S = p->nseg;
    if (S <= 2) return;
    const size_t base = F / (size_t)S;
    size_t max_shift = base * 2 / 5;
    if (max_shift > 200000) max_shift = 200000;
    const size_t w = 2048;
    const double loud = 0.005;
    const size_t O = p->overlap;
    for (int seg = 2; seg < S; seg++) {
        const size_t b = p->nom_start[seg];
        if (b < w || b + w >= F) continue;
        const double cur = rms_at(ch0, b, w);
        if (cur > loud) continue;
        size_t best = b;
        double best_rms = cur;
        const size_t step = w / 2;
        for (size_t off = step; off <= max_shift; off += step) {
            for (int dir = -1; dir <= 1; dir += 2) {
                size_t cand = (dir < 0) ? (b > off ? b - off : 0) : b + off;
                if (cand < w / 2 || cand + w / 2 >= F) continue;
                if (cand <= p->nom_start[seg - 1] + O * 2) continue;
                if (seg < S - 1 && cand >= p->nom_start[seg + 1] - O * 2) continue;
                const double r = rms_at(ch0, cand, w);
                if (r > best_rms) {
                    best_rms = r;
                    best = cand;
                }
            }
            if (best_rms > loud) break;
        }
        if (best != b) {
            p->nom_start[seg] = best;
            p->nom_size[seg - 1] = best - p->nom_start[seg - 1];
            p->nom_size[seg] = (seg < S - 1) ? p->nom_start[seg + 1] - best : F - best;
        }
    }
}

/* ---- stitch search ---- */
static double g_qfir[256];
static int g_qfir_n = 0;

const double *or_das_qfir(void)
{
    if (g_qfir_n == 0) {
        const int N = 255;
        const double fc = 20000.0 / (5644800.0 / 2.0);
        double sum = 0;
        for (int i = 0; i < N; i++) {
            double n = i - 127.0;
            if (n > -0.001 && n < 0.001)
                g_qfir[i] = 2.0 * fc;
            else
                g_qfir[i] = sin(2.0 * 3.14159265358979 * fc * n) / (3.14159265358979 * n);
            double x = 2.0 * i / (N - 1) - 1.0;
            double kb = 8.0;
            double arg = kb * sqrt(1.0 - x * x > 0 ? 1.0 - x * x : 0);
            double i0a = 1.0, term = 1.0;
            for (int m = 1; m <= 20; m++) {
                term *= (arg / 2.0 / m) * (arg / 2.0 / m);
                i0a += term;
            }
            double i0k = 1.0;
            term = 1.0;
            for (int m = 1; m <= 20; m++) {
                term *= (kb / 2.0 / m) * (kb / 2.0 / m);
                i0k += term;
            }
            g_qfir[i] *= i0a / i0k;
            sum += g_qfir[i];
        }
        for (int i = 0; i < N; i++) g_qfir[i] /= sum;
        g_qfir_n = N;
    }
    return g_qfir;
}

typedef struct { int pos, run, density; double quality; } qcand_t;

/* Candidate choice (src/dsp_plugin.c:2641-2652): keep the first; a later
 * candidate replaces the kept one when its quality is below 0.8 of the
 * kept one's, or below 1.2 of it with a strictly higher density. */
int or_das_pick(int n, const double *quality, const int *density)
{
    int best = 0;
    for (int ci = 1; ci < n; ci++) {
        double rq = quality[ci] / (quality[best] + 1e-15);
        if (rq < 0.8) best = ci;
        else if (rq < 1.2 && density[ci] > density[best]) best = ci;
    }
    return best;
}

int or_das_stitch_ex(const float *prev, const float *cur, int ovl_len, int lat, or_stitch_detail_t *det)
{
    int half_w = lat;
    if (half_w > ovl_len / 2) half_w = ovl_len / 2;
    if (half_w < 4) half_w = 4;

    int *dm = (int *)calloc((size_t)(ovl_len > 0 ? ovl_len : 1), sizeof(int));
    int peak = 0;
    {
        int c = 0;
        int ie = half_w < ovl_len ? half_w : ovl_len;
        for (int w = 0; w < ie; w++)
            if (prev[w] == cur[w]) c++;
        if (ovl_len > 0) dm[0] = c;
        if (c > peak) peak = c;
        for (int p = 1; p < ovl_len; p++) {
            int add = p + half_w - 1;
            if (add >= 0 && add < ovl_len && prev[add] == cur[add]) c++;
            int rem = p - half_w - 1;
            if (rem >= 0 && rem < ovl_len && prev[rem] == cur[rem]) c--;
            dm[p] = c;
            if (c > peak) peak = c;
        }
    }
    const int floor_d = peak * 9 / 10;
    const int min_sp = half_w;
    qcand_t q[32];
    int nq = 0;
    for (int p = 0; p < ovl_len - 3 && nq < 32;) {
        if (prev[p] != cur[p] || dm[p] < floor_d) { p++; continue; }
        int rs = p, rl = 0;
        while (p < ovl_len && prev[p] == cur[p]) { rl++; p++; }
        if (rl < 4) continue;
        if (nq > 0 && rs - q[nq - 1].pos < min_sp) continue;
        q[nq].pos = rs;
        q[nq].run = rl;
        q[nq].density = dm[rs];
        q[nq].quality = 1e30;
        nq++;
    }
    int scanned = nq, peak_pos = 0;
    {
        int pp = 0;
        for (int p = 1; p < ovl_len; p++)
            if (dm[p] > dm[pp]) pp = p;
        peak_pos = pp;
        int prp = pp, prl = 1;
        for (int mr = 8; mr >= 1 && prl < 4; mr >>= 1) {
            for (int r = 0; r <= half_w && prl < 4; r++) {
                int tryp[2] = {pp - r, pp + r};
                for (int t = 0; t < 2 && prl < 4; t++) {
                    int tp = tryp[t];
                    if (tp < 0 || tp + mr > ovl_len) continue;
                    int rl = 0;
                    for (int k = 0; k < mr; k++) {
                        if (prev[tp + k] == cur[tp + k]) rl++;
                        else break;
                    }
                    if (rl >= mr) { prp = tp; prl = mr; }
                }
            }
        }
        int dup = 0;
        for (int ci = 0; ci < nq; ci++)
            if (abs(q[ci].pos - prp) < min_sp) { dup = 1; break; }
        if (!dup && nq < 32) {
            q[nq].pos = prp;
            q[nq].run = prl;
            q[nq].density = ovl_len > 0 ? dm[prp] : 0;
            q[nq].quality = 1e30;
            nq++;
        }
    }
    if (nq == 0) {
        q[0].pos = 0; q[0].run = 1; q[0].density = 0; q[0].quality = 0;
        nq = 1;
    }
    double maxd[32] = {0};
    if (nq == 1) q[0].quality = 0;
    if (nq > 1) {
        const double *qf = or_das_qfir();
        const int fh = 127;
        for (int ci = 0; ci < nq; ci++) {
            int sp = q[ci].pos;
            double max_d = 0;
            for (int j = -512; j < 512; j++) {
                double val = 0;
                for (int k = 0; k < 255; k++) {
                    int idx = sp + j - fh + k;
                    if (idx >= sp && idx >= 0 && idx < ovl_len) {
                        double diff = (double)cur[idx] - (double)prev[idx];
                        val += qf[k] * diff;
                    }
                }
                double ad = fabs(val);
                if (ad > max_d) max_d = ad;
            }
            maxd[ci] = max_d;
            double dr = (double)q[ci].density / (double)(peak > 0 ? peak : 1);
            if (dr < 0.5) dr = 0.5;
            q[ci].quality = max_d / dr;
        }
    }
    double qual[32];
    int den[32];
    for (int ci = 0; ci < nq; ci++) {
        qual[ci] = q[ci].quality;
        den[ci] = q[ci].density;
    }
    const int best = or_das_pick(nq, qual, den);
    if (det) {
        det->half_w = half_w;
        det->peak = peak;
        det->peak_pos = peak_pos;
        det->scanned = scanned;
        det->ncand = nq;
        det->best = best;
        for (int ci = 0; ci < nq; ci++) {
            det->pos[ci] = q[ci].pos;
            det->run[ci] = q[ci].run;
            det->density[ci] = q[ci].density;
            det->quality[ci] = q[ci].quality;
            det->max_decoded[ci] = maxd[ci];
        }
        if (det->dm)
            for (int p = 0; p < ovl_len; p++) det->dm[p] = dm[p];
    }
    free(dm);
    return q[best].pos;
}

int or_das_stitch(const float *prev, const float *cur, int ovl_len, int lat,
                  double *quality_out, int *ncand_out)
{
    or_stitch_detail_t det;
    memset(&det, 0, sizeof(det));
    const int pos = or_das_stitch_ex(prev, cur, ovl_len, lat, &det);
    if (quality_out) *quality_out = det.quality[det.best];
    if (ncand_out) *ncand_out = det.ncand;
    return pos;
}

/* ---- the pipeline ---- */

/* Host threads for or_das_run's phases (the reference's threadpool runs
 * the blocks of one phase concurrently and waits between phases,
 * src/dsp_plugin.c:2185-2369).  Results do not depend on the count. */
static int g_das_threads = 1;

void or_das_set_threads(int n) { g_das_threads = n < 1 ? 1 : n; }

typedef struct {
    void (*fn)(void *, int);
    void *arg;
    int n;
    int next;
    pthread_mutex_t mu;
} or_par_t;

static void *par_worker(void *v)
{
    or_par_t *p = (or_par_t *)v;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        const int i = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (i >= p->n) break;
        p->fn(p->arg, i);
    }
    return NULL;
}

static void par_for(int n, void (*fn)(void *, int), void *arg)
{
    int nt = g_das_threads < n ? g_das_threads : n;
    if (nt <= 1) {
        for (int i = 0; i < n; i++) fn(arg, i);
        return;
    }
    or_par_t p = {fn, arg, n, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t th[256];
    if (nt > 256) nt = 256;
    for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, par_worker, &p);
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
}

typedef struct {
    const or_das_plan_t *p;
    or_ctx *persist, *tmp;
    int nch, S, temps;
    const double *in;
    size_t F;
    float **sb;
    size_t *so;
    double *est;
} das_job_t;

static void task_seg0(void *v, int ch)
{
    das_job_t *j = (das_job_t *)v;
    const or_das_plan_t *p = j->p;
    size_t cnt = p->nom_size[0];
    if (!p->seg0_first && j->S > 1 && p->nom_size[0] + p->overlap <= j->F) cnt += p->overlap;
    j->so[ch * j->S] = or_sdm_segment(&j->persist[ch], j->in + (size_t)ch * j->F, j->sb[ch * j->S], cnt, 0);
}

/* seg0-first: seg1 on a copy of the persistent context (after seg0), or
 * the estimate for seg >= 2 (task index nch + ch) */
static void task_seg1_or_estimate(void *v, int i)
{
    das_job_t *j = (das_job_t *)v;
    const or_das_plan_t *p = j->p;
    const int S = j->S;
    const size_t O = p->overlap, F = j->F;
    if (i < j->nch) {
        const int ch = i;
        or_ctx *t = &j->tmp[ch * j->temps];
        or_sdm_copy_state(t, &j->persist[ch]);
        size_t ns = p->nom_start[1], nz = p->nom_size[1], cnt = nz;
        if (S > 2 && ns + nz + O <= F) cnt += O;
        j->so[ch * S + 1] = or_sdm_segment(t, j->in + (size_t)ch * F + ns, j->sb[ch * S + 1], cnt, 0);
        return;
    }
    const int ch = i - j->nch;
    double st[OR_MAX_ORDER];
    or_sdm_best_state(&j->persist[ch], st);
    const size_t es = p->nom_start[1];
    const double *data = j->in + (size_t)ch * F + es;
    const size_t cnt = F - es;
    size_t pos = 0;
    for (int seg = 1; seg < S; seg++) {
        size_t bnd = seg >= 2 ? p->nom_start[seg] - es : 0;
        if (bnd > pos && bnd <= cnt) {
            or_estimate_state(j->persist[ch].f, st, data + pos, bnd - pos, j->persist[ch].state_limit, 16, st);
            pos = bnd;
        }
        memcpy(&j->est[(ch * S + seg) * OR_MAX_ORDER], st, sizeof(st));
    }
}

/* segment first + k (k over channels x segments) on a copy of the
 * persistent context as it is now, seeded with its estimate */
static void task_segment(void *v, int k)
{
    das_job_t *j = (das_job_t *)v;
    const or_das_plan_t *p = j->p;
    const int first = p->seg0_first ? 2 : 1;
    const int per = j->S - first;
    const int ch = k / per, seg = first + k % per;
    const size_t O = p->overlap, F = j->F;
    size_t ns = p->nom_start[seg], nz = p->nom_size[seg];
    int ext = seg < j->S - 1;
    size_t is = ns >= O ? ns - O : 0;
    size_t cnt = ns + nz - is;
    if (ext && ns + nz + O <= F) cnt += O;
    size_t disc = ns >= O ? O - j->persist[ch].L : 0;
    or_ctx *t = &j->tmp[ch * j->temps + seg - 1];
    or_sdm_copy_state(t, &j->persist[ch]);
    const double *e = &j->est[(ch * j->S + seg) * OR_MAX_ORDER];
    int bank = t->idx & 1;
    for (unsigned c = 0; c < t->num_cands; c++) {
        or_node *nd = t->act[bank][c];
        if (nd) {
            for (int q = 0; q < t->f->order; q++) nd->st[q] = e[q];
            nd->cost = 0.0;
        }
    }
    j->so[ch * j->S + seg] = or_sdm_segment(t, j->in + (size_t)ch * F + is, j->sb[ch * j->S + seg], cnt, disc);
}

size_t or_das_run(const or_das_plan_t *p, or_ctx *persist, int nch,
                  const double *in, size_t F, float *out, int *stitch_pos)
{
    const int S = p->nseg;
    const size_t O = p->overlap;
    const int seg0_first = p->seg0_first;
    const int temps = S - 1;
    or_ctx *tmp = temps > 0 ? (or_ctx *)calloc((size_t)(nch * temps), sizeof(or_ctx)) : NULL;
    float **sb = (float **)calloc((size_t)(nch * S), sizeof(float *));
    size_t *so = (size_t *)calloc((size_t)(nch * S), sizeof(size_t));
    double *est = (double *)calloc((size_t)(nch * S * OR_MAX_ORDER), sizeof(double));
    for (int i = 0; i < nch * S; i++) {
        int seg = i % S;
        sb[i] = (float *)malloc((p->nom_size[seg] + 2 * O + 4096) * sizeof(float));
    }
    das_job_t job = {p, persist, tmp, nch, S, temps, in, F, sb, so, est};

    if (seg0_first) {
        par_for(nch, task_seg0, &job);
        /* seg1 (if any) and the estimate for seg >= 2 (if any), concurrently */
        if (S > 2) par_for(2 * nch, task_seg1_or_estimate, &job);
        else if (S > 1) par_for(nch, task_seg1_or_estimate, &job);
    } else {
        for (int ch = 0; ch < nch; ch++) {
            double st[OR_MAX_ORDER];
            or_sdm_best_state(&persist[ch], st);
            for (int seg = 1; seg < S; seg++)
                memcpy(&est[(ch * S + seg) * OR_MAX_ORDER], st, sizeof(st));
        }
    }

    /* parallel segments (copies of persistent as it is now) */
    const int first = seg0_first ? 2 : 1;
    if (S > first) par_for(nch * (S - first), task_segment, &job);
    if (!seg0_first) par_for(nch, task_seg0, &job);

    /* stitch positions from channel 0 */
    int *sp = (int *)calloc((size_t)S, sizeof(int));
    size_t total = 0;
    for (int ch = 0; ch < nch; ch++) {
        float *o = out + (size_t)ch * F;
        size_t s0 = so[ch * S];
        if (!seg0_first && s0 > p->nom_size[0]) s0 = p->nom_size[0];
        memcpy(o, sb[ch * S], s0 * sizeof(float));
        size_t wp = s0;
        for (int seg = 1; seg < S; seg++) {
            size_t sz = so[ch * S + seg];
            if (sz == 0) { if (ch == 0) sp[seg] = 0; continue; }
            if (seg == 1 && seg0_first) {
                memcpy(o + wp, sb[ch * S + seg], sz * sizeof(float));
                wp += sz;
                continue;
            }
            size_t pst = wp >= O ? wp - O : 0;
            if (ch == 0) {
                size_t ovl = wp - pst;
                if (ovl > sz) ovl = sz;
                if (ovl > O) ovl = O;
                sp[seg] = or_das_stitch(o + pst, sb[ch * S + seg], (int)ovl,
                                        (int)persist[0].L, NULL, NULL);
            }
            size_t at = pst + (size_t)sp[seg];
            size_t cc = sz - (size_t)sp[seg];
            memcpy(o + at, sb[ch * S + seg] + sp[seg], cc * sizeof(float));
            wp = at + cc;
        }
        if (ch == 0) total = wp;
    }
    if (stitch_pos) memcpy(stitch_pos, sp, (size_t)S * sizeof(int));

    if (S > 1)
        for (int ch = 0; ch < nch; ch++)
            or_sdm_copy_state(&persist[ch], &tmp[ch * temps + S - 2]);

    /* The persistent contexts now link into the segment contexts; detach
     * those links before the segment contexts are released (they are only
     * followed by a later DAS call on the same contexts). */
    for (int ch = 0; ch < nch && S > 1; ch++)
        for (int b = 0; b < 2; b++)
            for (int i = 0; i < OR_NODES; i++) persist[ch].node[b][i].plist = NULL;

    for (int i = 0; i < nch * S; i++) free(sb[i]);
    free(sb);
    free(so);
    free(est);
    free(sp);
    free(tmp);
    return total;
}
