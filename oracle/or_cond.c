/*
 * Oracle restatement of the engine's pre-SDM conditioning.  TEST
 * INFRASTRUCTURE ONLY.  Follows src/engine.c: channel init of the boxcar
 * taps :131-147, same-rate boxcar :829-841 and lowpass :813-828 branches of
 * engine_process_fir_gain, the PCM boxcar pre-smooth :629-642, the gain
 * ramp :646-652 / :877-888, the soft clip :704-716; the lowpass design and
 * filter of src/fir.c:795-873 (IPP FIRSR restated as a direct-form sum in
 * ascending tap order).  Pinning: exact integer window sums for the DSD
 * boxcar, window semantics of the PCM boxcar against exact sums, an
 * independent numpy Kaiser design and long-double convolution for the
 * lowpass, closed-form properties of the clip; the IPP summation order of
 * the lowpass and the rounding sequence of the running PCM sum are what
 * the reference's code states and are not checked against a build of it
 * (engine.c needs Win32 and IPP).
 */
#include <math.h>
#include <string.h>

#include "dsd_oracle.h"

static double or_i0(double x)
{
    double sum = 1.0, term = 1.0, q = (x / 2.0) * (x / 2.0);
    for (int k = 1; k <= 25; k++) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-20 * sum) break;
    }
    return sum;
}

int or_lowpass_taps(unsigned rate, double *h)
{
    const int ntaps = 127, M = ntaps - 1;
    double fc = 50000.0 / ((double)rate / 2.0);
    if (fc > 0.5) fc = 0.5;
    double i0b = or_i0(10.0);
    for (int n = 0; n <= M; n++) {
        double x = (double)n - (double)M / 2.0;
        double sinc = (fabs(x) < 1e-10) ? 2.0 * fc : sin(2.0 * M_PI * fc * x) / (M_PI * x);
        double arg = 1.0 - (2.0 * n / M - 1.0) * (2.0 * n / M - 1.0);
        h[n] = sinc * (or_i0(10.0 * sqrt(fabs(arg))) / i0b);
    }
    double sum = 0.0;
    for (int n = 0; n < ntaps; n++) sum += h[n];
    if (sum > 0.0)
        for (int n = 0; n < ntaps; n++) h[n] /= sum;
    return ntaps;
}

void or_cond_init(or_cond *c, int taps, unsigned lowpass_rate)
{
    memset(c, 0, sizeof(*c));
    c->taps = taps;
    if (lowpass_rate) c->lp_n = or_lowpass_taps(lowpass_rate, c->lp);
}

void or_cond_dsd_boxcar(or_cond *c, const float *in, size_t n, double gs, double ge, double *out)
{
    double inv_n = 1.0 / (double)c->taps;
    double gstep = (n > 1) ? (ge - gs) / (double)(n - 1) : 0.0;
    for (size_t i = 0; i < n; i++) {
        double s = in[i] >= 0.0f ? 1.0 : -1.0;
        c->sum -= c->ring[c->pos];
        c->ring[c->pos] = s;
        c->sum += s;
        c->pos = (c->pos + 1) % c->taps;
        out[i] = c->sum * inv_n * (gs + gstep * (double)i);
    }
}

void or_cond_pcm_boxcar(or_cond *c, const float *in, size_t n, double *out)
{
    double inv_n = 1.0 / (double)c->taps;
    for (size_t i = 0; i < n; i++) {
        double s = (double)in[i];
        c->sum -= c->ring[c->pos];
        c->ring[c->pos] = s;
        c->sum += s;
        c->pos = (c->pos + 1) % c->taps;
        out[i] = c->sum * inv_n;
    }
}

void or_cond_dsd_lowpass(or_cond *c, const float *in, size_t n, double gs, double ge, double *out)
{
    double gstep = (n > 1) ? (ge - gs) / (double)(n - 1) : 0.0;
    int T = c->lp_n;
    for (size_t i = 0; i < n; i++) {
        /* delay line: dly[0] = newest */
        memmove(c->dly + 1, c->dly, sizeof(double) * (size_t)(T - 1));
        c->dly[0] = in[i] >= 0.0f ? 1.0 : -1.0;
        double acc = 0.0;
        for (int k = 0; k < T; k++) acc = acc + c->lp[k] * c->dly[k];
        out[i] = acc * (gs + gstep * (double)i);
    }
}

void or_gain_ramp(double *x, size_t n, double gs, double ge)
{
    double step = (n > 1) ? (ge - gs) / (double)(n - 1) : 0.0;
    for (size_t i = 0; i < n; i++) x[i] *= gs + step * (double)i;
}

void or_soft_clip(double *x, size_t n)
{
    const double th = 0.95;
    const double inv_knee = 1.0 / (1.0 - th);
    for (size_t i = 0; i < n; i++) {
        double v = x[i];
        if (v > th)
            x[i] = th + (1.0 - th) * or_lf_tanh((v - th) * inv_knee);
        else if (v < -th)
            x[i] = -th - (1.0 - th) * or_lf_tanh((-v - th) * inv_knee);
    }
}
