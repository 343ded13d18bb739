// Differentiable CRFB loop filter (tanh-relaxed quantizer) and causal
// pre-emphasis FIR: forward, checkpointed recompute and adjoint kernels.
//
// Loop filter step per row, state s[P] (structure of src/trellis.c:37-151,
// relaxed quantizer as in training_pipeline/diff_sdm.py:61-100):
//   ns0 = s0 - g0*s1 + x;  v = x + a0*ns0
//   nsi = si + s(i-1) - gi*s(i+1), v = v + ai*nsi      (i = 1..P-2)
//   ns(P-1) = s(P-1) + s(P-2),      v = v + a(P-1)*ns(P-1)
//   y = tanh(v*T);  s = ns;  s0 -= y
//
// Device layouts are time-major so a wavefront's 64 lanes (64 rows) touch
// 512 contiguous bytes per sample: u, y, gy, gu are [N][B]; checkpoints
// [nck][P][B]; recompute workspace S [n][P][B], YD [n][B]; adjoint
// workspace LAM [n][P][B], VB [n][B].
//
// The sequential forward and the sequential adjoint sweep (the lambda
// recurrence) carry only what the recurrence needs; everything that is
// parallel over time (pre-emphasis, state recompute from checkpoints, the
// coefficient-gradient reductions, dL/du) runs in separate wide kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LF_U 16            // samples per prefetch block in the sequential kernels

// Keeps a uniform value in a VGPR: FP64 constants used from SGPRs are
// rematerialised per use (two s_mov each) and spill to VGPR lanes once the
// unrolled loops hold many of them.
__device__ __forceinline__ double vreg(double x)
{
    __asm__ volatile("" : "+v"(x));
    return x;
}

// Constants of the tanh below, held in VGPRs for the whole kernel.
struct LfTanhK {
    double c2, c3, c4, c5, c6, c7, c8, c9, c10, c11, c12, c13, invln2, ln2hi, ln2lo, cap, one;
    __device__ void init()
    {
        c2 = vreg(0x1.0000000000000p-1);  c3 = vreg(0x1.5555555555555p-3);
        c4 = vreg(0x1.5555555555555p-5);  c5 = vreg(0x1.1111111111111p-7);
        c6 = vreg(0x1.6c16c16c16c17p-10); c7 = vreg(0x1.a01a01a01a01ap-13);
        c8 = vreg(0x1.a01a01a01a01ap-16); c9 = vreg(0x1.71de3a556c734p-19);
        c10 = vreg(0x1.27e4fb7789f5cp-22); c11 = vreg(0x1.ae64567f544e4p-26);
        c12 = vreg(0x1.1eed8eff8d898p-29); c13 = vreg(0x1.6124613a86d09p-33);
        invln2 = vreg(0x1.71547652b82fep+0); ln2hi = vreg(0x1.62e42fee00000p-1);
        ln2lo = vreg(0x1.a39ef35793c76p-33); cap = vreg(40.0); one = vreg(1.0);
    }
};

// tanh(v*T) from correctly rounded operations only (+,-,*,/,fma,rint,
// ldexp), so that host and device produce identical bits:
// tanh|z| = m/(m+2), m = expm1(2|z|) by Cody-Waite reduction
// (2|z| = k ln2 + r) and a degree-13 Taylor series of expm1(r) evaluated
// with Estrin's scheme; 2|z| is capped at 40 (tanh = 1 in double there).
// T2 = 2T > 0, so 2|v*T| = |v|*T2 exactly and the sign of z is that of v.
__device__ __forceinline__ double lf_tanh_vt(double v, double T2, const LfTanhK &K)
{
    double x = fabs(v) * T2;
    x = (x > K.cap) ? K.cap : x;                 // NaN stays NaN
    const double kd = rint(x * K.invln2);
    double r = fma(-kd, K.ln2hi, x);
    r = fma(-kd, K.ln2lo, r);
    const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    const double e0 = fma(K.c3, r, K.c2), e1 = fma(K.c5, r, K.c4), e2 = fma(K.c7, r, K.c6);
    const double e3 = fma(K.c9, r, K.c8), e4 = fma(K.c11, r, K.c10), e5 = fma(K.c13, r, K.c12);
    const double f0 = fma(e1, r2, e0), f1 = fma(e3, r2, e2), f2 = fma(e5, r2, e4);
    const double q = fma(f2, r8, fma(f1, r4, f0));
    const double p = fma(r2, q, r);
    const int k = (kd == kd) ? (int)kd : 0;
    const double m = ldexp(p, k) + (ldexp(K.one, k) - K.one);
    const double t = m / (m + 2.0);
    return copysign(t, v);
}

template <int P>
struct LfCoef {
    double a[P], g[P];
};

template <int P>
__device__ __forceinline__ void lf_load_coef(LfCoef<P> &c, const double *__restrict__ a,
                                             const double *__restrict__ g)
{
#pragma unroll
    for (int i = 0; i < P; i++) {
        c.a[i] = vreg(a[i]);
        c.g[i] = vreg(g[i]);
    }
}

// One step; s is updated in place, returns y.
template <int P>
__device__ __forceinline__ double lf_step(double (&s)[P], const LfCoef<P> &c, double T2,
                                          const LfTanhK &K, double x)
{
    double ns[P];
    ns[0] = s[0] - c.g[0] * s[1] + x;
    double v = x + c.a[0] * ns[0];
#pragma unroll
    for (int i = 1; i < P - 1; i++) {
        ns[i] = s[i] + s[i - 1] - c.g[i] * s[i + 1];
        v = v + c.a[i] * ns[i];
    }
    ns[P - 1] = s[P - 1] + s[P - 2];
    v = v + c.a[P - 1] * ns[P - 1];
    const double y = lf_tanh_vt(v, T2, K);
    ns[0] = ns[0] - y;
#pragma unroll
    for (int i = 0; i < P; i++) s[i] = ns[i];
    return y;
}

// ---------------------------------------------------------------------------
// Forward: one lane per row, sequential over time.  Full blocks of LF_U
// samples run without per-sample checks; the inputs of the next block are
// loaded while the current one computes.  The state before sample c*E is
// written to ck[c] (E is a multiple of LF_U).
template <int P>
__global__ __launch_bounds__(64) void lf_fwd(const double *__restrict__ a,
                                             const double *__restrict__ g, double T,
                                             const double *__restrict__ u, int B, int64_t N,
                                             int E, double *__restrict__ y,
                                             double *__restrict__ ck)
{
    const int b = blockIdx.x * 64 + threadIdx.x;
    if (b >= B) return;
    LfCoef<P> c;
    lf_load_coef<P>(c, a, g);
    LfTanhK K;
    K.init();
    const double T2 = vreg(2.0 * T);
    double s[P];
#pragma unroll
    for (int i = 0; i < P; i++) s[i] = 0.0;
    const int64_t nfull = N / LF_U;
    const int blocks_per_ck = E / LF_U;
    const double *up = u + b;
    double *yp = y + b;
    double *ckp = ck + b;
    double cur[LF_U], nxt[LF_U];
    if (nfull > 0) {
#pragma unroll
        for (int j = 0; j < LF_U; j++) cur[j] = up[(int64_t)j * B];
    }
    int to_ck = 0;
    for (int64_t blk = 0; blk < nfull; blk++) {
        if (blk + 1 < nfull) {
            const double *q = up + (blk + 1) * LF_U * (int64_t)B;
#pragma unroll
            for (int j = 0; j < LF_U; j++) nxt[j] = q[(int64_t)j * B];
        }
        if (to_ck == 0) {
#pragma unroll
            for (int i = 0; i < P; i++) ckp[(int64_t)i * B] = s[i];
            ckp += (int64_t)P * B;
            to_ck = blocks_per_ck;
        }
        to_ck--;
#pragma unroll
        for (int j = 0; j < LF_U; j++) {
            yp[0] = lf_step<P>(s, c, T2, K, cur[j]);
            yp += B;
        }
#pragma unroll
        for (int j = 0; j < LF_U; j++) cur[j] = nxt[j];
    }
    for (int64_t n = nfull * LF_U; n < N; n++) {
        if (n % E == 0) {
#pragma unroll
            for (int i = 0; i < P; i++) ckp[(int64_t)i * B] = s[i];
            ckp += (int64_t)P * B;
        }
        yp[0] = lf_step<P>(s, c, T2, K, up[n * B]);
        yp += B;
    }
}

// ---------------------------------------------------------------------------
// Recompute of chunks [c0, c0 + gridDim.x) from their checkpoints: stores
// the state before each sample and T*(1 - y^2).  Rows x chunks in parallel.
template <int P>
__global__ __launch_bounds__(64) void lf_recompute(const double *__restrict__ a,
                                                   const double *__restrict__ g, double T,
                                                   const double *__restrict__ u, int B,
                                                   int64_t N, int E,
                                                   const double *__restrict__ ck, int64_t c0,
                                                   double *__restrict__ S,
                                                   double *__restrict__ YD)
{
    const int b = blockIdx.y * 64 + threadIdx.x;
    if (b >= B) return;
    const int64_t cidx = c0 + blockIdx.x;
    const int64_t nb = cidx * E;
    if (nb >= N) return;
    const int64_t cnt = (nb + E < N) ? E : N - nb;
    LfCoef<P> c;
    lf_load_coef<P>(c, a, g);
    LfTanhK K;
    K.init();
    const double T2 = vreg(2.0 * T);
    double s[P];
    const double *src = ck + cidx * P * (int64_t)B + b;
#pragma unroll
    for (int i = 0; i < P; i++) s[i] = src[(int64_t)i * B];
    const double *up = u + nb * B + b;
    double *Sp = S + (int64_t)blockIdx.x * E * P * B + b;
    double *Yp = YD + (int64_t)blockIdx.x * E * B + b;
    const int64_t nfull = cnt / LF_U;
    double cur[LF_U], nxt[LF_U];
    if (nfull > 0) {
#pragma unroll
        for (int j = 0; j < LF_U; j++) cur[j] = up[(int64_t)j * B];
    }
    for (int64_t blk = 0; blk < nfull; blk++) {
        if (blk + 1 < nfull) {
            const double *q = up + (blk + 1) * LF_U * (int64_t)B;
#pragma unroll
            for (int j = 0; j < LF_U; j++) nxt[j] = q[(int64_t)j * B];
        }
#pragma unroll
        for (int j = 0; j < LF_U; j++) {
#pragma unroll
            for (int i = 0; i < P; i++) Sp[(int64_t)i * B] = s[i];
            Sp += (int64_t)P * B;
            const double yv = lf_step<P>(s, c, T2, K, cur[j]);
            Yp[0] = (1.0 - yv * yv) * T;
            Yp += B;
        }
#pragma unroll
        for (int j = 0; j < LF_U; j++) cur[j] = nxt[j];
    }
    for (int64_t n = nfull * LF_U; n < cnt; n++) {
#pragma unroll
        for (int i = 0; i < P; i++) Sp[(int64_t)i * B] = s[i];
        Sp += (int64_t)P * B;
        const double yv = lf_step<P>(s, c, T2, K, up[n * B]);
        Yp[0] = (1.0 - yv * yv) * T;
        Yp += B;
    }
}

// Adjoint step: lambda (dL/ds after the sample) -> lambda before it, given
// vbar.  nb = lambda + a*vbar; lambda_j = nb_j + nb_(j+1) - g_(j-1) nb_(j-1).
template <int P>
__device__ __forceinline__ void lf_lambda_step(double (&l)[P], double (&nb)[P], const LfCoef<P> &c,
                                               double vbar)
{
#pragma unroll
    for (int i = 0; i < P; i++) nb[i] = fma(c.a[i], vbar, l[i]);
    l[0] = nb[0] + nb[1];
#pragma unroll
    for (int i = 1; i < P - 1; i++) l[i] = fma(-c.g[i - 1], nb[i - 1], nb[i] + nb[i + 1]);
    l[P - 1] = fma(-c.g[P - 2], nb[P - 2], nb[P - 1]);
}

// ---------------------------------------------------------------------------
// Sequential adjoint sweep over the cnt samples of a slab starting at
// n_slab, newest first.  Carries only the lambda recurrence: stores vbar per
// sample and lambda at the top of every SUB-sample sub-chunk (LAMCK[q] =
// lambda after sample min(cnt, (q+1)SUB) - 1).  lam carries lambda between
// slab launches ([P][B]).  SUB is a multiple of LF_U.
// span > 0: truncated adjoint, lambda restarts from zero at the top of
// every span-sample interval (span a multiple of SUB), so blockIdx.y runs
// the intervals of the slab independently and lam is not used.
template <int P>
__global__ __launch_bounds__(64) void lf_adj_seq(const double *__restrict__ a,
                                                 const double *__restrict__ g,
                                                 const double *__restrict__ gy, int B,
                                                 int64_t n_slab, int64_t cnt,
                                                 const double *__restrict__ YD,
                                                 double *__restrict__ lam, int first, int SUB,
                                                 double *__restrict__ LAMCK,
                                                 double *__restrict__ VB, int64_t span)
{
    const int b = blockIdx.x * 64 + threadIdx.x;
    if (b >= B) return;
    const int64_t lo = span > 0 ? (int64_t)blockIdx.y * span : 0;
    if (lo >= cnt) return;
    const int64_t cn = span > 0 ? (cnt - lo < span ? cnt - lo : span) : cnt;
    LfCoef<P> c;
    lf_load_coef<P>(c, a, g);
    double l[P], nb[P];
#pragma unroll
    for (int i = 0; i < P; i++) l[i] = (first || span > 0) ? 0.0 : lam[(int64_t)i * B + b];
    const double *gyp = gy + (n_slab + lo) * B + b;
    const double *ydp = YD + lo * B + b;
    double *vbl = VB + lo * B + b;
    auto store_ck = [&](int64_t m) {
        double *d = LAMCK + ((lo + m) / SUB) * P * (int64_t)B + b;
#pragma unroll
        for (int i = 0; i < P; i++) d[(int64_t)i * B] = l[i];
    };
    store_ck(cn - 1);
    const int64_t nfull = cn / LF_U;
    for (int64_t m = cn - 1; m >= nfull * LF_U; m--) {       // partial top block
        const double vbar = (gyp[m * B] - l[0]) * ydp[m * B];
        vbl[m * B] = vbar;
        lf_lambda_step<P>(l, nb, c, vbar);
    }
    double cg[LF_U], cy[LF_U], ng[LF_U], ny[LF_U];
    if (nfull > 0) {
        const int64_t m0 = (nfull - 1) * LF_U;
#pragma unroll
        for (int j = 0; j < LF_U; j++) {
            cg[j] = gyp[(m0 + j) * B];
            cy[j] = ydp[(m0 + j) * B];
        }
    }
    for (int64_t k = nfull - 1; k >= 0; k--) {
        if (k > 0) {
            const int64_t m0 = (k - 1) * LF_U;
#pragma unroll
            for (int j = 0; j < LF_U; j++) {
                ng[j] = gyp[(m0 + j) * B];
                ny[j] = ydp[(m0 + j) * B];
            }
        }
        const int64_t mtop = k * LF_U + LF_U - 1;
        if ((lo + mtop + 1) % SUB == 0) store_ck(mtop);
        double *vbp = vbl + mtop * B;
#pragma unroll
        for (int j = LF_U - 1; j >= 0; j--) {
            const double vbar = (cg[j] - l[0]) * cy[j];
            vbp[0] = vbar;
            vbp -= B;
            lf_lambda_step<P>(l, nb, c, vbar);
        }
#pragma unroll
        for (int j = 0; j < LF_U; j++) {
            cg[j] = ng[j];
            cy[j] = ny[j];
        }
    }
    if (span == 0) {
#pragma unroll
        for (int i = 0; i < P; i++) lam[(int64_t)i * B + b] = l[i];
    }
}

// ---------------------------------------------------------------------------
// Parallel pass over a slab: one wave per (sub-chunk, 64 rows) task reruns
// the lambda recurrence down its SUB samples from LAMCK and the stored vbar,
// writes dL/du, and accumulates dL/da, dL/dg; per-block partial sums (2P
// values, fixed-order reduction afterwards).
template <int P>
__global__ __launch_bounds__(256) void lf_adj_par(const double *__restrict__ a,
                                                  const double *__restrict__ g,
                                                  const double *__restrict__ u, int B,
                                                  int64_t n_slab, int64_t cnt, int SUB,
                                                  const double *__restrict__ S,
                                                  const double *__restrict__ LAMCK,
                                                  const double *__restrict__ VB,
                                                  double *__restrict__ gu,
                                                  double *__restrict__ part)
{
    LfCoef<P> c;
    lf_load_coef<P>(c, a, g);
    double ga[P], gg[P];
#pragma unroll
    for (int i = 0; i < P; i++) ga[i] = gg[i] = 0.0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bgroups = (B + 63) / 64;
    const int64_t nsub = (cnt + SUB - 1) / SUB;
    const int64_t total = nsub * bgroups;
    for (int64_t t = (int64_t)blockIdx.x * 4 + wv; t < total; t += (int64_t)gridDim.x * 4) {
        const int64_t q = t / bgroups;
        const int b = (int)(t - q * bgroups) * 64 + lane;
        if (b >= B) continue;
        double l[P], nb[P];
        const double *lc = LAMCK + q * P * (int64_t)B + b;
#pragma unroll
        for (int i = 0; i < P; i++) l[i] = lc[(int64_t)i * B];
        const int64_t mlo = q * SUB;
        const int64_t mhi = (mlo + SUB < cnt) ? mlo + SUB : cnt;
        for (int64_t m = mhi - 1; m >= mlo; m--) {
            const double *Sd = S + m * P * (int64_t)B + b;
            double s[P];
#pragma unroll
            for (int i = 0; i < P; i++) s[i] = Sd[(int64_t)i * B];
            const double x = u[(n_slab + m) * B + b];
            const double vbar = VB[m * B + b];
            double ns[P];
            ns[0] = s[0] - c.g[0] * s[1] + x;
#pragma unroll
            for (int i = 1; i < P - 1; i++) ns[i] = s[i] + s[i - 1] - c.g[i] * s[i + 1];
            ns[P - 1] = s[P - 1] + s[P - 2];
#pragma unroll
            for (int i = 0; i < P; i++) ga[i] = fma(vbar, ns[i], ga[i]);
            lf_lambda_step<P>(l, nb, c, vbar);
#pragma unroll
            for (int i = 0; i < P - 1; i++) gg[i] = fma(-nb[i], s[i + 1], gg[i]);
            gu[(n_slab + m) * B + b] = vbar + nb[0];
        }
    }
    __shared__ double red[2 * P][256];
#pragma unroll
    for (int i = 0; i < P; i++) {
        red[i][threadIdx.x] = ga[i];
        red[P + i][threadIdx.x] = gg[i];
    }
    __syncthreads();
    if (threadIdx.x < 2 * P) {
        double acc = 0.0;
        for (int k = 0; k < 256; k++) acc += red[threadIdx.x][k];
        part[(int64_t)blockIdx.x * 2 * P + threadIdx.x] = acc;
    }
}

// Sums nparts partial vectors of width w in fixed order; out = (acc ? out : 0) + sum.
__global__ void lf_reduce_parts(const double *__restrict__ part, int nparts, int w,
                                double *__restrict__ out, int acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w) return;
    double s = 0.0;
    for (int k = 0; k < nparts; k++) s += part[(int64_t)k * w + i];
    out[i] = acc ? out[i] + s : s;
}

// ---------------------------------------------------------------------------
// Pre-emphasis forward: u[n][b] = sum_k w[k] * p[b][n-k] (ascending k, fma).
// Tile: 64 rows x PE_TN samples staged in LDS (odd pitch: conflict-free
// column reads); lane = row, wave = sample phase; output rows coalesced.
#define PE_TN 64
#define PE_MAXT 128
__global__ __launch_bounds__(256) void pe_fwd(const double *__restrict__ w, int T,
                                              const float *__restrict__ p, int B, int64_t N,
                                              double *__restrict__ u)
{
    constexpr int PITCH = PE_TN + PE_MAXT - 1 | 1;
    __shared__ float tile[64 * PITCH];
    __shared__ double wl[PE_MAXT];
    const int64_t n0 = (int64_t)blockIdx.x * PE_TN;
    const int b0 = blockIdx.y * 64;
    const int W = PE_TN + T - 1;
    for (int i = threadIdx.x; i < T; i += 256) wl[i] = w[i];
    for (int i = threadIdx.x; i < 64 * W; i += 256) {
        const int r = i / W, col = i - r * W;
        const int64_t n = n0 - (T - 1) + col;
        const int b = b0 + r;
        tile[r * PITCH + col] = (b < B && n >= 0 && n < N) ? p[(int64_t)b * N + n] : 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int b = b0 + lane;
    if (b >= B) return;
    const float *row = tile + lane * PITCH;
    for (int j = wv; j < PE_TN; j += 4) {
        const int64_t n = n0 + j;
        if (n >= N) break;
        double acc = 0.0;
        for (int k = 0; k < T; k++) acc = fma(wl[k], (double)row[j + T - 1 - k], acc);
        u[n * B + b] = acc;
    }
}

// Pre-emphasis tap gradient: per block partial sums over a tile of samples
// and 64 rows, gw_part[blk][k] = sum gu[n][b] * p[b][n-k].
__global__ __launch_bounds__(256) void pe_bwd(const double *__restrict__ gu, int T,
                                              const float *__restrict__ p, int B, int64_t N,
                                              double *__restrict__ part)
{
    constexpr int PITCH = PE_TN + PE_MAXT - 1 | 1;
    __shared__ float tile[64 * PITCH];
    __shared__ double red[4][PE_MAXT];
    const int64_t n0 = (int64_t)blockIdx.x * PE_TN;
    const int b0 = blockIdx.y * 64;
    const int W = PE_TN + T - 1;
    for (int i = threadIdx.x; i < 64 * W; i += 256) {
        const int r = i / W, col = i - r * W;
        const int64_t n = n0 - (T - 1) + col;
        const int b = b0 + r;
        tile[r * PITCH + col] = (b < B && n >= 0 && n < N) ? p[(int64_t)b * N + n] : 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int b = b0 + lane;
    const float *row = tile + lane * PITCH;
    // wave wv handles taps k = wv, wv+4, ...; lanes = rows, summed below
    for (int k = wv; k < T; k += 4) {
        double acc = 0.0;
        if (b < B) {
            for (int j = 0; j < PE_TN; j++) {
                const int64_t n = n0 + j;
                if (n >= N) break;
                acc = fma(gu[n * B + b], (double)row[j + T - 1 - k], acc);
            }
        }
        // wave reduction over the 64 rows, fixed order via LDS
        __shared__ double wred[4][64];
        // LDS operations of one wavefront complete in issue order; the asm
        // statements keep the compiler from reordering across them.
        wred[wv][lane] = acc;
        __asm__ volatile("" ::: "memory");
        if (lane == 0) {
            double s = 0.0;
            for (int r = 0; r < 64; r++) s += wred[wv][r];
            red[wv][k >> 2] = s;
        }
        __asm__ volatile("" ::: "memory");
    }
    __syncthreads();
    const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    for (int k = threadIdx.x; k < T; k += 256) part[blk * T + k] = red[k & 3][k >> 2];
}

// ---------------------------------------------------------------------------
extern "C" {

int lf_launch_fwd(int P, const double *a, const double *g, double T, const double *u, int B,
                  int64_t N, int E, double *y, double *ck, hipStream_t st)
{
    dim3 grid((B + 63) / 64), blk(64);
    switch (P) {
#define LF_CASE(PP) case PP: hipLaunchKernelGGL(lf_fwd<PP>, grid, blk, 0, st, a, g, T, u, B, N, E, y, ck); break;
        LF_CASE(3) LF_CASE(4) LF_CASE(5) LF_CASE(6) LF_CASE(7) LF_CASE(8) LF_CASE(9) LF_CASE(10)
#undef LF_CASE
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lf_launch_recompute(int P, const double *a, const double *g, double T, const double *u,
                        int B, int64_t N, int E, const double *ck, int64_t c0, int nchunks,
                        double *S, double *YD, hipStream_t st)
{
    dim3 grid(nchunks, (B + 63) / 64), blk(64);
    switch (P) {
#define LF_CASE(PP) case PP: hipLaunchKernelGGL(lf_recompute<PP>, grid, blk, 0, st, a, g, T, u, B, N, E, ck, c0, S, YD); break;
        LF_CASE(3) LF_CASE(4) LF_CASE(5) LF_CASE(6) LF_CASE(7) LF_CASE(8) LF_CASE(9) LF_CASE(10)
#undef LF_CASE
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lf_launch_adj_seq(int P, const double *a, const double *g, const double *gy, int B,
                      int64_t n_slab, int64_t cnt, const double *YD, double *lam, int first,
                      int SUB, double *LAMCK, double *VB, int64_t span, hipStream_t st)
{
    if (span > 0 && (span % SUB) != 0) return -1;
    const int64_t nint = span > 0 ? (cnt + span - 1) / span : 1;
    if (nint > 65535) return -1;
    dim3 grid((B + 63) / 64, (unsigned)nint), blk(64);
    switch (P) {
#define LF_CASE(PP) case PP: hipLaunchKernelGGL(lf_adj_seq<PP>, grid, blk, 0, st, a, g, gy, B, n_slab, cnt, YD, lam, first, SUB, LAMCK, VB, span); break;
        LF_CASE(3) LF_CASE(4) LF_CASE(5) LF_CASE(6) LF_CASE(7) LF_CASE(8) LF_CASE(9) LF_CASE(10)
#undef LF_CASE
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lf_launch_adj_par(int P, const double *a, const double *g, const double *u, int B,
                      int64_t n_slab, int64_t cnt, int SUB, const double *S, const double *LAMCK,
                      const double *VB, double *gu, double *part, int nblocks, hipStream_t st)
{
    dim3 grid(nblocks), blk(256);
    switch (P) {
#define LF_CASE(PP) case PP: hipLaunchKernelGGL(lf_adj_par<PP>, grid, blk, 0, st, a, g, u, B, n_slab, cnt, SUB, S, LAMCK, VB, gu, part); break;
        LF_CASE(3) LF_CASE(4) LF_CASE(5) LF_CASE(6) LF_CASE(7) LF_CASE(8) LF_CASE(9) LF_CASE(10)
#undef LF_CASE
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lf_launch_reduce(const double *part, int nparts, int w, double *out, int acc, hipStream_t st)
{
    hipLaunchKernelGGL(lf_reduce_parts, dim3((w + 63) / 64), dim3(64), 0, st, part, nparts, w, out, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pe_launch_fwd(const double *w, int T, const float *p, int B, int64_t N, double *u,
                  hipStream_t st)
{
    if (T < 1 || T > PE_MAXT) return -1;
    dim3 grid((unsigned)((N + PE_TN - 1) / PE_TN), (B + 63) / 64), blk(256);
    hipLaunchKernelGGL(pe_fwd, grid, blk, 0, st, w, T, p, B, N, u);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int pe_launch_bwd(const double *gu, int T, const float *p, int B, int64_t N, double *part,
                  hipStream_t st)
{
    if (T < 1 || T > PE_MAXT) return -1;
    dim3 grid((unsigned)((N + PE_TN - 1) / PE_TN), (B + 63) / 64), blk(256);
    hipLaunchKernelGGL(pe_bwd, grid, blk, 0, st, gu, T, p, B, N, part);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
