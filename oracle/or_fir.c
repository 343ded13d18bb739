/*
 * Oracle restatement of the half-band 2x upsampling chain.  TEST
 * INFRASTRUCTURE ONLY (see dsd_oracle.h).  Compile with -ffp-contract=off.
 *
 * Follows the reference:
 *   bessel_I0 / design_halfband_kaiser   src/fir.c:65-98 (beta 12, normalised)
 *   get_stage_ntaps                      src/fir.c:131-150 (127 taps for the
 *                                        first stage of a chain ending at
 *                                        DSD512 or DSD512/48)
 *   taps x factor for upsampling         src/fir.c:386-390
 *   zero delay lines                     src/fir.c:419-426
 *   ippsFIRMR_64f up by 2, phase 0       src/fir.c:455-465
 *   chain + gain                         src/fir.c:678-725, src/engine.c:872-883
 * The per-output accumulation order is fixed here as "fused multiply-add,
 * ascending polyphase tap index" (IPP's own summation order is not
 * published, so its exact sample values are parity unpinned); the HIP
 * kernels use the same order, so both produce identical doubles.
 */
#include "dsd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

double or_bessel_i0(double x)
{
    double sum = 1.0, term = 1.0;
    double q = (x / 2.0) * (x / 2.0);
    for (int k = 1; k <= 25; k++) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-20 * sum) break;
    }
    return sum;
}

void or_design_halfband(double *h, int ntaps, double beta)
{
    int c = (ntaps - 1) / 2;
    double i0b = or_bessel_i0(beta), sum = 0.0;
    for (int n = 0; n < ntaps; n++) {
        double x = (double)(n - c) / 2.0;
        double sv = (fabs(x) < 1e-15) ? 1.0 : sin(M_PI * x) / (M_PI * x);
        double t = (double)(n - c) / c;
        double arg = 1.0 - t * t;
        double w = or_bessel_i0(beta * sqrt(arg > 0.0 ? arg : 0.0)) / i0b;
        h[n] = sv * w;
        sum += h[n];
    }
    for (int n = 0; n < ntaps; n++) h[n] /= sum;
}

int or_fir_plan(unsigned fs_in, unsigned fs_out, int *ntaps)
{
    if (fs_in == 0 || fs_out <= fs_in || fs_out % fs_in) return -1;
    unsigned r = fs_out / fs_in;
    if (r & (r - 1)) return -1;
    int st = 0;
    while (r > 1) { st++; r >>= 1; }
    if (st > OR_FIR_MAX_STAGES) return -1;
    if (fs_out / fs_in == 8 && (fs_out == 22579200u || fs_out == 24576000u))
        return -1;   /* 4x+2x special case: DSD64 input only, not a PCM path */
    for (int i = 0; i < st; i++)
        ntaps[i] = (st >= 2 && i == 0 && (fs_out == 22579200u || fs_out == 24576000u)) ? 127 : 63;
    return st;
}

void or_fir_stage(const double *h2, int ntaps, const double *hist, int nhist,
                  const double *in, size_t n, double *out)
{
    int Q0 = (ntaps + 1) / 2;   /* taps of phase 0 (even k) */
    int Q1 = ntaps / 2;         /* taps of phase 1 (odd k) */
    for (size_t m = 0; m < n; m++) {
        for (int p = 0; p < 2; p++) {
            int Q = p ? Q1 : Q0;
            double acc = 0.0;
            for (int q = 0; q < Q; q++) {
                long idx = (long)m - q;
                double xv;
                if (idx >= 0) xv = in[idx];
                else if (-idx <= nhist) xv = hist[nhist + idx];
                else xv = 0.0;
                acc = fma(h2[2 * q + p], xv, acc);
            }
            out[2 * m + p] = acc;
        }
    }
}

size_t or_fir_chain(unsigned fs_in, unsigned fs_out, const float *in,
                    size_t n, double gain, double *out)
{
    int ntaps[OR_FIR_MAX_STAGES];
    int st = or_fir_plan(fs_in, fs_out, ntaps);
    if (st < 0) return 0;
    double *a = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) a[i] = (double)in[i];
    size_t cnt = n;
    for (int s = 0; s < st; s++) {
        double h[255], h2[255];
        or_design_halfband(h, ntaps[s], 12.0);
        for (int k = 0; k < ntaps[s]; k++) h2[k] = h[k] * 2.0;
        double *b = (s == st - 1) ? out : (double *)malloc(2 * cnt * sizeof(double));
        or_fir_stage(h2, ntaps[s], NULL, 0, a, cnt, b);
        free(a);
        a = b;
        cnt *= 2;
    }
    if (gain != 1.0)
        for (size_t i = 0; i < cnt; i++) out[i] *= gain;
    return cnt;
}

/* The chain with double input (fir_chain_process_d of a boxcar-smoothed
 * signal, src/engine.c:620-645). */
size_t or_fir_chain_d(unsigned fs_in, unsigned fs_out, const double *in,
                      size_t n, double gain, double *out)
{
    int ntaps[OR_FIR_MAX_STAGES];
    int st = or_fir_plan(fs_in, fs_out, ntaps);
    if (st < 0) return 0;
    double *a = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) a[i] = in[i];
    size_t cnt = n;
    for (int s = 0; s < st; s++) {
        double h[255], h2[255];
        or_design_halfband(h, ntaps[s], 12.0);
        for (int k = 0; k < ntaps[s]; k++) h2[k] = h[k] * 2.0;
        double *b = (s == st - 1) ? out : (double *)malloc(2 * cnt * sizeof(double));
        or_fir_stage(h2, ntaps[s], NULL, 0, a, cnt, b);
        free(a);
        a = b;
        cnt *= 2;
    }
    if (gain != 1.0)
        for (size_t i = 0; i < cnt; i++) out[i] *= gain;
    return cnt;
}

size_t or_fir_chain_ld(unsigned fs_in, unsigned fs_out, const float *in,
                       size_t n, double gain, long double *out)
{
    int ntaps[OR_FIR_MAX_STAGES];
    int st = or_fir_plan(fs_in, fs_out, ntaps);
    if (st < 0) return 0;
    long double *a = (long double *)malloc(n * sizeof(long double));
    for (size_t i = 0; i < n; i++) a[i] = (long double)in[i];
    size_t cnt = n;
    for (int s = 0; s < st; s++) {
        double h[255];
        or_design_halfband(h, ntaps[s], 12.0);
        long double *b = (s == st - 1) ? out : (long double *)malloc(2 * cnt * sizeof(long double));
        for (size_t m = 0; m < cnt; m++) {
            for (int p = 0; p < 2; p++) {
                long double acc = 0.0L;
                for (int k = p; k < ntaps[s]; k += 2) {
                    long idx = (long)m - (k - p) / 2;
                    if (idx >= 0) acc += (long double)h[k] * 2.0L * a[idx];
                }
                b[2 * m + p] = acc;
            }
        }
        free(a);
        a = b;
        cnt *= 2;
    }
    if (gain != 1.0)
        for (size_t i = 0; i < cnt; i++) out[i] *= (long double)gain;
    return cnt;
}
