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
extern "C" __global__ __launch_bounds__(256) void das_stitch(
    const u64 *__restrict__ bits, const int64_t *__restrict__ prev_bit,
    const int64_t *__restrict__ cur_bit, const int *__restrict__ ovl_len, int hw_lat,
    const double *__restrict__ qfir, int *__restrict__ o_ncand, int *__restrict__ o_peak,
    int *__restrict__ o_pos, int *__restrict__ o_dens, double *__restrict__ o_maxd)
{
    constexpr int MAXO = 8192;   /* overlap up to 64*128 */
    __shared__ uint8_t match[MAXO];
    __shared__ int psum[MAXO + 1];
    __shared__ int cpos[DAS_MAXQ + 1], cden[DAS_MAXQ + 1];
    __shared__ int ncand, peak;
    __shared__ double qf[DAS_QTAPS];
    __shared__ double redmax[DAS_MAXQ + 1][4];
    const int b = blockIdx.x;
    const int ovl = ovl_len[b];
    if (ovl <= 0) {            /* unused boundary slot of a fixed table */
        if (threadIdx.x == 0) {
            o_ncand[b] = 0;
            o_peak[b] = 0;
        }
        return;
    }
    const int64_t pb = prev_bit[b], cbit = cur_bit[b];
    int hw = hw_lat;
    if (hw > ovl / 2) hw = ovl / 2;
    if (hw < 4) hw = 4;

    for (int i = threadIdx.x; i < DAS_QTAPS; i += blockDim.x) qf[i] = qfir[i];
    for (int i = threadIdx.x; i < ovl; i += blockDim.x)
        match[i] = (uint8_t)(bit_at(bits, pb + i) == bit_at(bits, cbit + i));
    __syncthreads();
    /* prefix sums of matches (serial in one lane; ovl <= 8192) */
    if (threadIdx.x == 0) {
        int c = 0;
        psum[0] = 0;
        for (int i = 0; i < ovl; i++) { c += match[i]; psum[i + 1] = c; }
    }
    __syncthreads();
    /* density at p: matches in [max(0,p-hw), min(ovl, p+hw)) — the
     * reference's sliding window (:2465-2482) */
    auto dm = [&](int p) {
        int lo = p - hw; if (lo < 0) lo = 0;
        int hi = p + hw; if (hi > ovl) hi = ovl;
        return psum[hi] - psum[lo];
    };
    if (threadIdx.x == 0) {
        int pk = 0;
        for (int p = 0; p < ovl; p++) if (dm(p) > pk) pk = dm(p);
        const int fl = pk * 9 / 10;
        int nq = 0;
        for (int p = 0; p < ovl - 3 && nq < DAS_MAXQ;) {
            if (!match[p] || dm(p) < fl) { p++; continue; }
            int rs = p, rl = 0;
            while (p < ovl && match[p]) { rl++; p++; }
            if (rl < 4) continue;
            if (nq > 0 && rs - cpos[nq - 1] < hw) continue;
            cpos[nq] = rs;
            cden[nq] = dm(rs);
            nq++;
        }
        int pp = 0;
        for (int p = 1; p < ovl; p++) if (dm(p) > dm(pp)) pp = p;
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
        peak = pk;
    }
    __syncthreads();
    const int nq = ncand;
    /* decode of (this - prev) from each candidate on: lanes over (cand, j) */
    for (int ci = 0; ci < nq; ci++) {
        double mx = 0.0;
        if (nq > 1) {
            const int sp = cpos[ci];
            for (int jj = threadIdx.x; jj < 2 * DAS_EVAL; jj += blockDim.x) {
                const int j = jj - DAS_EVAL;
                double val = 0.0;
                int k0 = sp - (sp + j - 127);   /* first k with idx >= sp */
                if (k0 < 0) k0 = 0;
                for (int k = k0; k < DAS_QTAPS; k++) {
                    const int idx = sp + j - 127 + k;
                    if (idx >= ovl) break;
                    if (idx < 0 || match[idx]) continue;
                    const double diff = bit_at(bits, cbit + idx) ? 2.0 : -2.0;
                    val = val + qf[k] * diff;
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
        if ((threadIdx.x & 63) == 0) redmax[ci][threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x < nq) {
        const int ci = threadIdx.x;
        double mx = 0.0;
        for (int w = 0; w < 4; w++) if (redmax[ci][w] > mx) mx = redmax[ci][w];
        o_pos[b * (DAS_MAXQ + 1) + ci] = cpos[ci];
        o_dens[b * (DAS_MAXQ + 1) + ci] = cden[ci];
        o_maxd[b * (DAS_MAXQ + 1) + ci] = mx;
    }
    if (threadIdx.x == 0) {
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
    u64 word = 0;
    for (int i = 0; i < 64; i++) {
        const int64_t ob = b0 + i;
        if (ob >= total_bits) break;
        while (s + 1 < nspan_per_ch && dst_bit[base + s + 1] <= ob) s++;
        const int64_t d0 = dst_bit[base + s];
        if (ob >= d0 && ob < d0 + len[base + s]) {
            const int64_t sb = src_bit[base + s] + (ob - d0);
            word |= (u64)((src[sb >> 6] >> (sb & 63)) & 1ull) << i;
        }
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
