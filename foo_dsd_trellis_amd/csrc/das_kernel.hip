/*
 * DAS support kernels for gfx950 (src/dsp_plugin.c:1963-2713 semantics).
 *
 *  das_sumsq        sum of squares of 2048-sample windows (boundary
 *                   reposition, :1973-2019); sequential ascending sum per
 *                   window so the doubles match a scalar loop exactly;
 *                   sqrt and the comparisons are finished on the host.
 *  das_estimate     16-level greedy loop-filter estimate between segment
 *                   boundaries (src/trellis.c:1128-1187, src/threadpool.c:
 *                   345-366): inherently sequential, one lane per channel.
 *  das_stitch       per boundary: match density map, candidate runs, the
 *                   peak-density run, and for each candidate the largest
 *                   |255-tap decode of (this - prev)| over +-512 positions
 *                   (:2452-2640).  The final ratio/selection is done on the
 *                   host with the same doubles.
 *  das_assemble     gather of the kept segment spans into the output bit
 *                   stream (:2411-2713), disjoint spans, one word per lane.
 *  bits_to_f32      packed bits -> +-1.0f (the reference's output format).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;

#define DAS_MAXQ 32
#define DAS_EVAL 512
#define DAS_QTAPS 255

extern "C" __global__ void das_sumsq(const double *__restrict__ x, const int64_t *__restrict__ centers,
                                     int n, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = centers[i];
    double s = 0.0;
    for (int64_t k = c - 1024; k < c + 1024; k++) {
        const double v = x[k];
        s = s + v * v;
    }
    out[i] = s;
}

/* C (int) conversion as executed by x86-64 cvttsd2si: out-of-range and NaN
 * give INT_MIN. */
__device__ __forceinline__ int x86_int(double q)
{
    if (!(q < 2147483648.0) || !(q > -2147483649.0)) return (int)0x80000000u;
    return (int)q;
}

/* One lane per channel.  x: channel ch starts at x + ch*stride + start;
 * bnd[s] (s = 0..nb-1) are boundary offsets relative to start, ascending;
 * est receives nb states of 10 doubles per channel. */
extern "C" __global__ void das_estimate(const double *__restrict__ x, int64_t stride, int64_t start,
                                        int64_t count, const int64_t *__restrict__ bnd, int nb,
                                        const double *__restrict__ init, const double *__restrict__ a,
                                        const double *__restrict__ g, int P, double state_limit,
                                        int levels, int nch, double *__restrict__ est)
{
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch >= nch) return;
    double s[10], ns[10];
    for (int i = 0; i < 10; i++) s[i] = (i < P) ? init[ch * 10 + i] : 0.0;
    const double *xc = x + (int64_t)ch * stride + start;
    const double step = 2.0 / (double)(levels - 1);
    int64_t pos = 0;
    for (int b = 0; b < nb; b++) {
        const int64_t e = bnd[b];
        if (e > pos && e <= count) {
            for (int64_t n = pos; n < e; n++) {
                const double xv = xc[n];
                ns[0] = s[0] - g[0] * s[1] + xv;
                double v = xv + a[0] * ns[0];
                for (int i = 1; i < P - 1; i++) {
                    ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
                    v = v + a[i] * ns[i];
                }
                ns[P - 1] = s[P - 1] + s[P - 2];
                v = v + a[P - 1] * ns[P - 1];
                double y;
                if (levels <= 2) {
                    y = (v * a[0] > 0.0) ? 1.0 : -1.0;
                } else {
                    const double yi = v / a[0];
                    int idx = x86_int((yi + 1.0) / step + 0.5);
                    if (idx < 0) idx = 0;
                    if (idx > levels - 1) idx = levels - 1;
                    y = -1.0 + (double)idx * step;
                }
                ns[0] = ns[0] - y;
                if (state_limit > 0.0)
                    for (int i = 0; i < P; i++) {
                        if (ns[i] > state_limit) ns[i] = state_limit;
                        else if (ns[i] < -state_limit) ns[i] = -state_limit;
                    }
                for (int i = 0; i < P; i++) s[i] = ns[i];
            }
            pos = e;
        }
        for (int i = 0; i < 10; i++) est[((int64_t)ch * nb + b) * 10 + i] = (i < P) ? s[i] : 0.0;
    }
}

/* bit i of a packed stream */
__device__ __forceinline__ int bit_at(const u64 *w, int64_t i) { return (int)((w[i >> 6] >> (i & 63)) & 1ull); }

/*
 * One block (256 threads) per boundary.  For boundary b: prev stream bits
 * start at prev_w[b] (a bit offset prev_off[b] into the packed buffer),
 * this stream at cur_w[b]; ovl[b] positions, half window hw.
 * Outputs per boundary: ncand, peak, and per candidate pos / density /
 * max |decode| (the host applies the density ratio and selection).
 */
/* first set bit of `w` at or after position p (limit if none before it) */
__device__ __forceinline__ int next_set(const u64 *w, int p, int limit)
{
    if (p >= limit) return limit;
    int i = p >> 6;
    u64 x = w[i] & (~0ull << (p & 63));
    while (!x) {
        if (++i * 64 >= limit) return limit;
        x = w[i];
    }
    const int r = i * 64 + __builtin_ctzll(x);
    return r < limit ? r : limit;
}

extern "C" __global__ __launch_bounds__(256) void das_stitch(
    const u64 *__restrict__ bits, const int64_t *__restrict__ prev_bit,
    const int64_t *__restrict__ cur_bit, const int *__restrict__ ovl_len, int hw_lat,
    const double *__restrict__ qfir, int *__restrict__ o_ncand, int *__restrict__ o_peak,
    int *__restrict__ o_pos, int *__restrict__ o_dens, double *__restrict__ o_maxd)
{
    constexpr int MAXO = 8192;   /* overlap up to 64*128 */
    constexpr int NW = MAXO / 64;
    constexpr int NT = 256;
    __shared__ uint8_t match[MAXO];
    __shared__ uint16_t psum[MAXO + 1];    /* matches before position i */
    __shared__ uint16_t mlist[MAXO];       /* mismatch positions, ascending */
    __shared__ u64 mword[NW], qword[NW], cword[NW];
    __shared__ int part[NT];
    __shared__ int cpos[DAS_MAXQ + 1], cden[DAS_MAXQ + 1];
    __shared__ int ncand, peak, ppos;
    __shared__ int rmax[NT / 64], rarg[NT / 64];
    __shared__ double qf[DAS_QTAPS];
    __shared__ double redmax[DAS_MAXQ + 1][4];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ovl = ovl_len[b];
    if (ovl <= 0 || ovl > MAXO) {   /* unused boundary slot of a fixed table (overlaps are <= 64*128) */
        if (tid == 0) {
            o_ncand[b] = 0;
            o_peak[b] = 0;
        }
        return;
    }
    const int64_t pb = prev_bit[b], cbit = cur_bit[b];
    int hw = hw_lat;
    if (hw > ovl / 2) hw = ovl / 2;
    if (hw < 4) hw = 4;
    const int nw = (ovl + 63) >> 6;

    for (int i = tid; i < DAS_QTAPS; i += NT) qf[i] = qfir[i];
    for (int i = tid; i < ovl; i += NT) match[i] = (uint8_t)(bit_at(bits, pb + i) == bit_at(bits, cbit + i));
    /* this stream's overlap bits, and the match bits, as words */
    for (int wi = wv; wi < nw; wi += NT / 64) {
        const int p = wi * 64 + lane;
        const bool in = p < ovl;
        const int pc = in ? p : 0;
        const int cur = bit_at(bits, cbit + pc);
        const u64 cb = __builtin_amdgcn_ballot_w64(in && cur);
        const u64 mb = __builtin_amdgcn_ballot_w64(in && bit_at(bits, pb + pc) == cur);
        if (lane == 0) {
            cword[wi] = cb;
            mword[wi] = mb;
        }
    }
    __syncthreads();
    /* prefix counts of matches: per-thread chunks, then a scan of the 256
     * chunk sums in wave 0 */
    const int C = (ovl + NT - 1) / NT;
    const int c0 = tid * C, c1 = (c0 + C < ovl) ? c0 + C : ovl;
    {
        int cnt = 0;
        for (int i = c0; i < c1; i++) cnt += match[i];
        part[tid] = cnt;
    }
    __syncthreads();
    if (wv == 0) {
        const int a0 = part[4 * lane], a1 = part[4 * lane + 1], a2 = part[4 * lane + 2], a3 = part[4 * lane + 3];
        const int tot = a0 + a1 + a2 + a3;
        int inc = tot;
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(inc, off, 64);
            if (lane >= off) inc += o;
        }
        const int ex = inc - tot;
        part[4 * lane] = ex;
        part[4 * lane + 1] = ex + a0;
        part[4 * lane + 2] = ex + a0 + a1;
        part[4 * lane + 3] = ex + a0 + a1 + a2;
    }
    __syncthreads();
    {
        int c = part[tid];
        for (int i = c0; i < c1; i++) {
            psum[i] = (uint16_t)c;
            if (!match[i]) mlist[i - c] = (uint16_t)i;
            c += match[i];
        }
        if (c1 == ovl && c0 < c1) psum[ovl] = (uint16_t)c;
    }
    __syncthreads();
    /* density at p: matches in [max(0,p-hw), min(ovl, p+hw)) -- the
     * reference's sliding window (:2465-2482) */
    auto dm = [&](int p) {
        int lo = p - hw; if (lo < 0) lo = 0;
        int hi = p + hw; if (hi > ovl) hi = ovl;
        return (int)psum[hi] - (int)psum[lo];
    };
    /* peak density and its first position */
    {
        int best = -1, arg = 0;
        for (int p = tid; p < ovl; p += NT) {
            const int d = dm(p);
            if (d > best) { best = d; arg = p; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const int ob = __shfl_down(best, off, 64), oa = __shfl_down(arg, off, 64);
            if (ob > best || (ob == best && oa < arg)) { best = ob; arg = oa; }
        }
        if (lane == 0) { rmax[wv] = best; rarg[wv] = arg; }
        __syncthreads();
        if (tid == 0) {
            int bb = rmax[0], ba = rarg[0];
            for (int w2 = 1; w2 < NT / 64; w2++)
                if (rmax[w2] > bb || (rmax[w2] == bb && rarg[w2] < ba)) { bb = rmax[w2]; ba = rarg[w2]; }
            peak = bb < 0 ? 0 : bb;
            ppos = ba;
        }
        __syncthreads();
    }
    const int pk = peak;
    const int fl = pk * 9 / 10;
    /* run starts that qualify: matching, density >= 90 % of the peak */
    for (int wi = wv; wi < nw; wi += NT / 64) {
        const int p = wi * 64 + lane;
        const u64 qb = __builtin_amdgcn_ballot_w64(p < ovl - 3 && match[p] && dm(p) >= fl);
        if (lane == 0) qword[wi] = qb;
    }
    __syncthreads();
    if (tid == 0) {
        /* candidate runs in scan order (:2484-2530): from each qualifying
         * position the run takes every following match; runs shorter than 4
         * or closer than hw to the last candidate are passed over */
        int nq = 0;
        int p = 0;
        while (nq < DAS_MAXQ) {
            p = next_set(qword, p, ovl - 3);
            if (p >= ovl - 3) break;
            const int rs = p;
            /* first non-match at or after p */
            int e = p;
            {
                int i = e >> 6;
                u64 x = ~mword[i] & (~0ull << (e & 63));
                while (!x && ++i < nw) x = ~mword[i];
                e = x ? i * 64 + __builtin_ctzll(x) : ovl;
                if (e > ovl) e = ovl;
            }
            const int rl = e - rs;
            p = e;
            if (rl < 4) continue;
            if (nq > 0 && rs - cpos[nq - 1] < hw) continue;
            cpos[nq] = rs;
            cden[nq] = dm(rs);
            nq++;
        }
        const int pp = ppos;
        int prp = pp, prl = 1;
        for (int mr = 8; mr >= 1 && prl < 4; mr >>= 1)
            for (int r = 0; r <= hw && prl < 4; r++)
                for (int t = 0; t < 2 && prl < 4; t++) {
                    const int tp = t ? pp + r : pp - r;
                    if (tp < 0 || tp + mr > ovl) continue;
                    int rl = 0;
                    for (int k = 0; k < mr; k++) { if (match[tp + k]) rl++; else break; }
                    if (rl >= mr) { prp = tp; prl = mr; }
                }
        int dup = 0;
        for (int ci = 0; ci < nq; ci++) {
            int dd = cpos[ci] - prp;
            if (dd < 0) dd = -dd;
            if (dd < hw) { dup = 1; break; }
        }
        if (!dup && nq < DAS_MAXQ) {
            cpos[nq] = prp;
            cden[nq] = ovl > 0 ? dm(prp) : 0;
            nq++;
        }
        if (nq == 0) { cpos[0] = 0; cden[0] = 0; nq = 1; }
        ncand = nq;
    }
    __syncthreads();
    const int nq = ncand;
    /* decode of (this - prev) from each candidate on: lanes over the 1024
     * shifts; a shift sums q[k] * (+-2) over the mismatching positions of
     * its 255-tap window from the candidate on, in ascending k (the
     * matching ones contribute zero terms) */
    for (int ci = 0; ci < nq; ci++) {
        double mx = 0.0;
        if (nq > 1) {
            const int sp = cpos[ci];
            for (int jj = tid; jj < 2 * DAS_EVAL; jj += NT) {
                const int j = jj - DAS_EVAL;
                const int base = sp + j - 127;             /* position of tap 0 */
                int lo = base > sp ? base : sp;
                int hi = base + DAS_QTAPS;
                if (hi > ovl) hi = ovl;
                double val = 0.0;
                if (lo < hi) {
                    const int m1 = hi - (int)psum[hi];
                    for (int mi = lo - (int)psum[lo]; mi < m1; mi++) {
                        const int idx = mlist[mi];
                        const double diff = ((cword[idx >> 6] >> (idx & 63)) & 1ull) ? 2.0 : -2.0;
                        val = val + qf[idx - base] * diff;
                    }
                }
                const double ad = fabs(val);
                if (ad > mx) mx = ad;
            }
        }
        /* block max */
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_down(mx, off, 64);
            if (o > mx) mx = o;
        }
        if (lane == 0) redmax[ci][wv] = mx;
    }
    __syncthreads();
    if (tid < nq) {
        const int ci = tid;
        double mx = 0.0;
        for (int w = 0; w < 4; w++) if (redmax[ci][w] > mx) mx = redmax[ci][w];
        o_pos[b * (DAS_MAXQ + 1) + ci] = cpos[ci];
        o_dens[b * (DAS_MAXQ + 1) + ci] = cden[ci];
        o_maxd[b * (DAS_MAXQ + 1) + ci] = mx;
    }
    if (tid == 0) {
        o_ncand[b] = nq;
        o_peak[b] = peak;
    }
}

/* Span table: span s copies src bits [src_bit[s], +len[s]) to output bits
 * [dst_bit[s], +len[s]) of channel ch_of[s]; spans of one channel are
 * disjoint and ascending.  One lane per output word. */
extern "C" __global__ void das_assemble(const u64 *__restrict__ src, const int64_t *__restrict__ src_bit,
                                        const int64_t *__restrict__ dst_bit,
                                        const int64_t *__restrict__ len, int nspan_per_ch,
                                        u64 *__restrict__ out, int64_t out_words_per_ch,
                                        int64_t total_bits)
{
    const int ch = blockIdx.y;
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= out_words_per_ch) return;
    const int64_t b0 = w * 64;
    const int base = ch * nspan_per_ch;
    /* last span with dst_bit <= b0 (binary search) */
    int lo = 0, hi = nspan_per_ch - 1, s = 0;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (dst_bit[base + mid] <= b0) { s = mid; lo = mid + 1; } else hi = mid - 1;
    }
    /* the word's bits come in runs, one per span piece: each run is one
     * funnel shift of (at most) two source words; bits no span covers are 0 */
    u64 word = 0;
    const int64_t end = (b0 + 64 < total_bits) ? b0 + 64 : total_bits;
    int64_t ob = b0;
    while (ob < end) {
        while (s + 1 < nspan_per_ch && dst_bit[base + s + 1] <= ob) s++;
        const int64_t d0 = dst_bit[base + s], l = len[base + s];
        if (ob < d0 || ob >= d0 + l) {      /* gap: up to the next span start */
            int64_t nxt = (ob < d0) ? d0 : ((s + 1 < nspan_per_ch) ? dst_bit[base + s + 1] : end);
            ob = (nxt > ob && nxt < end) ? nxt : end;
            continue;
        }
        const int64_t run_end = (d0 + l < end) ? d0 + l : end;
        const int n = (int)(run_end - ob);  /* 1..64 */
        const int64_t sb = src_bit[base + s] + (ob - d0);
        const int64_t sw = sb >> 6;
        const int off = (int)(sb & 63);
        u64 v = src[sw] >> off;
        if (off != 0 && n > 64 - off) v |= src[sw + 1] << (64 - off);
        if (n < 64) v &= (1ull << n) - 1ull;
        word |= v << (ob - b0);
        ob = run_end;
    }
    out[(int64_t)ch * out_words_per_ch + w] = word;
}

extern "C" __global__ void bits_to_f32(const u64 *__restrict__ bits, int64_t words_per_ch,
                                       int64_t n, float *__restrict__ out, int64_t out_stride)
{
    const int ch = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 w = bits[(int64_t)ch * words_per_ch + (i >> 6)];
    out[(int64_t)ch * out_stride + i] = ((w >> (i & 63)) & 1ull) ? 1.0f : -1.0f;
}

/* Appends n packed bits of each channel (src row ch, bit 0 on) at bit
 * position pos of the channel's destination row; bits of dst below pos keep
 * their value, bits from pos + n to the end of the last written word become
 * 0.  One lane per destination word touched. */
extern "C" __global__ void bits_append(const u64 *__restrict__ src, int64_t src_words,
                                       u64 *__restrict__ dst, int64_t dst_stride,
                                       int64_t dst_words, int64_t pos, int64_t n)
{
    const int ch = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = pos / 64 + i;
    if (n <= 0 || w * 64 >= pos + n || w >= dst_words) return;
    const u64 *s = src + (int64_t)ch * src_words;
    u64 *d = dst + (int64_t)ch * dst_stride;
    u64 word = (w * 64 < pos) ? d[w] : 0ull;
    for (int b = 0; b < 64; b++) {
        const int64_t g = w * 64 + b;            /* destination bit */
        if (g < pos) continue;
        const u64 mask = 1ull << b;
        if (g >= pos + n) {
            word &= ~mask;
            continue;
        }
        const int64_t sb = g - pos;              /* source bit, < n <= 64 * src_words */
        const u64 bit = (s[sb >> 6] >> (sb & 63)) & 1ull;
        word = (word & ~mask) | (bit << b);
    }
    d[w] = word;
}
