/*
 * Differentiable CRFB loop filter with a tanh-relaxed quantizer, and the
 * causal pre-emphasis FIR in front of it: forward and adjoint in plain FP64.
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP kernels).
 *
 * Loop filter step (structure of src/trellis.c:37-151, relaxed quantizer of
 * training_pipeline/diff_sdm.py:61-100), per row, state s[P] starting at 0:
 *   ns0 = s0 - g0*s1 + x
 *   v   = x + a0*ns0
 *   nsi = si + s(i-1) - gi*s(i+1)      i = 1..P-2,   v = v + ai*nsi
 *   ns(P-1) = s(P-1) + s(P-2),                         v = v + a(P-1)*ns(P-1)
 *   y   = lf_tanh(v*T)       (T = 1/tau, tau = 0.1 -> T = 10)
 *   s   = ns, s0 -= y
 * Pre-emphasis: u[n] = sum_{k<Tn} w[k] * p[n-k]  (zero history), summed in
 * ascending k with fused multiply-add.
 *
 * Adjoint of the loop filter, lambda = dL/ds after the step, gy = dL/dy:
 *   ybar = gy - lambda0,  vbar = ybar * T * (1 - y*y)
 *   nsbar_i = lambda_i + a_i*vbar
 *   dL/da_i += vbar*ns_i,  dL/dg_i += -nsbar_i*s(i+1)  (i <= P-2)
 *   dL/dx = vbar + nsbar_0
 *   lambda_j <- nsbar_j + nsbar_(j+1) [j<=P-2] - g(j-1)*nsbar_(j-1) [j>=1]
 *
 * Pinning: the forward against a torch float64 statement of diff_sdm.py's
 * recurrence; the adjoint against torch float64 autograd of that statement
 * and against central finite differences (tests/test_lf_oracle.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "dsd_oracle.h"

/* tanh from +,-,*,/,fma, rint and ldexp only, so that the same sequence of
 * correctly rounded operations gives the same bits on the host and on the
 * GPU: tanh|z| = m/(m+2) with m = expm1(2|z|); expm1 by Cody-Waite range
 * reduction x = k ln2 + r, |r| <= ln2/2, Taylor series of expm1(r) to r^13
 * evaluated by Estrin's scheme, m = 2^k expm1(r) + (2^k - 1).  Max error
 * against libm tanh is a few ulp (tests/test_lf_oracle.py). */
double or_lf_tanh(double z)
{
    static const double c[14] = {0, 0, 0x1.0000000000000p-1, 0x1.5555555555555p-3,
        0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10,
        0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19,
        0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
        0x1.6124613a86d09p-33};
    double x = fabs(z) * 2.0;
    if (x != x) return z;
    if (x > 40.0) x = 40.0;
    double kd = rint(x * 0x1.71547652b82fep+0);
    double r = fma(-kd, 0x1.62e42fee00000p-1, x);
    r = fma(-kd, 0x1.a39ef35793c76p-33, r);
    double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    double e0 = fma(c[3], r, c[2]), e1 = fma(c[5], r, c[4]), e2 = fma(c[7], r, c[6]);
    double e3 = fma(c[9], r, c[8]), e4 = fma(c[11], r, c[10]), e5 = fma(c[13], r, c[12]);
    double f0 = fma(e1, r2, e0), f1 = fma(e3, r2, e2), f2 = fma(e5, r2, e4);
    double q = fma(f2, r8, fma(f1, r4, f0));
    double p = fma(r2, q, r);
    int k = (int)kd;
    double m = ldexp(p, k) + (ldexp(1.0, k) - 1.0);
    double t = m / (m + 2.0);
    return copysign(t, z);
}

static double lf_step(double *s, const double *a, const double *g, int P, double T, double x,
                      double *ns_out)
{
    double ns[OR_MAX_ORDER];
    int i;
    ns[0] = s[0] - g[0] * s[1] + x;
    double v = x + a[0] * ns[0];
    for (i = 1; i < P - 1; i++) {
        ns[i] = s[i] + s[i - 1] - g[i] * s[i + 1];
        v = v + a[i] * ns[i];
    }
    ns[P - 1] = s[P - 1] + s[P - 2];
    v = v + a[P - 1] * ns[P - 1];
    double y = or_lf_tanh(v * T);
    if (ns_out) memcpy(ns_out, ns, sizeof(double) * (size_t)P);
    ns[0] = ns[0] - y;
    memcpy(s, ns, sizeof(double) * (size_t)P);
    return y;
}

void or_lf_forward(const double *a, const double *g, int P, double T, const double *u, int B,
                   size_t N, double *y)
{
    for (int b = 0; b < B; b++) {
        double s[OR_MAX_ORDER] = {0};
        for (size_t n = 0; n < N; n++)
            y[(size_t)b * N + n] = lf_step(s, a, g, P, T, u[(size_t)b * N + n], NULL);
    }
}

void or_lf_backward(const double *a, const double *g, int P, double T, const double *u, int B,
                    size_t N, const double *gy, double *gu, double *ga, double *gg)
{
    or_lf_backward_tr(a, g, P, T, u, B, N, gy, gu, ga, gg, 0);
}

/* trunc_every > 0: the adjoint restarts from zero at the top of every
 * trunc_every-sample interval (the state at each interval start held
 * constant, truncated backpropagation through time) */
void or_lf_backward_tr(const double *a, const double *g, int P, double T, const double *u, int B,
                       size_t N, const double *gy, double *gu, double *ga, double *gg, size_t trunc_every)
{
    double *S = (double *)malloc(sizeof(double) * (size_t)P * (N ? N : 1));
    double *NS = (double *)malloc(sizeof(double) * (size_t)P * (N ? N : 1));
    double *Y = (double *)malloc(sizeof(double) * (N ? N : 1));
    for (int i = 0; i < P; i++) ga[i] = gg[i] = 0.0;
    for (int b = 0; b < B; b++) {
        const double *ub = u + (size_t)b * N;
        double s[OR_MAX_ORDER] = {0};
        for (size_t n = 0; n < N; n++) {
            memcpy(S + n * P, s, sizeof(double) * (size_t)P);
            Y[n] = lf_step(s, a, g, P, T, ub[n], NS + n * P);
        }
        double lam[OR_MAX_ORDER] = {0}, nb[OR_MAX_ORDER];
        for (size_t m = N; m-- > 0;) {
            if (trunc_every && (m + 1) % trunc_every == 0)
                for (int i = 0; i < P; i++) lam[i] = 0.0;
            const double *sn = S + m * P, *nsn = NS + m * P;
            double ybar = gy[(size_t)b * N + m] - lam[0];
            double vbar = ybar * ((1.0 - Y[m] * Y[m]) * T);
            for (int i = 0; i < P; i++) {
                nb[i] = lam[i] + a[i] * vbar;
                ga[i] += vbar * nsn[i];
            }
            for (int i = 0; i < P - 1; i++) gg[i] += -nb[i] * sn[i + 1];
            gu[(size_t)b * N + m] = vbar + nb[0];
            for (int j = 0; j < P; j++) {
                double t = nb[j];
                if (j + 1 < P) t = t + nb[j + 1];
                if (j >= 1) t = t - g[j - 1] * nb[j - 1];
                lam[j] = t;
            }
        }
    }
    free(S);
    free(NS);
    free(Y);
}

void or_pe_forward(const double *w, int T, const float *p, int B, size_t N, double *u)
{
    for (int b = 0; b < B; b++)
        for (size_t n = 0; n < N; n++) {
            double acc = 0.0;
            for (int k = 0; k < T; k++) {
                double x = (n >= (size_t)k) ? (double)p[(size_t)b * N + n - k] : 0.0;
                acc = fma(w[k], x, acc);
            }
            u[(size_t)b * N + n] = acc;
        }
}

void or_pe_backward(int T, const float *p, int B, size_t N, const double *gu, double *gw)
{
    for (int k = 0; k < T; k++) {
        double acc = 0.0;
        for (int b = 0; b < B; b++)
            for (size_t n = (size_t)k; n < N; n++)
                acc += gu[(size_t)b * N + n] * (double)p[(size_t)b * N + n - k];
        gw[k] = acc;
    }
}
