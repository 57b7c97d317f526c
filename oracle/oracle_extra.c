/*
 * oracle_extra.c -- test entry points of the CPU oracle (TEST INFRASTRUCTURE).
 * Exposes the deterministic binary32 functions of ofmath.h so tests can sweep
 * them against binary64 libm and mpmath.
 */
#include <stddef.h>
#include "ofmath.h"

enum { OF_LOG = 1, OF_SIN, OF_COS, OF_ATAN2, OF_ACOS, OF_POW, OF_EXP };

void orc_math(int fn, const float *a, const float *b, float *out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        switch (fn) {
            case OF_LOG: out[i] = of_logf(a[i]); break;
            case OF_SIN: out[i] = of_sinf(a[i]); break;
            case OF_COS: out[i] = of_cosf(a[i]); break;
            case OF_ATAN2: out[i] = of_atan2f(a[i], b[i]); break;
            case OF_ACOS: out[i] = of_acosf(a[i]); break;
            case OF_POW: out[i] = of_powf_pos(a[i], b[i]); break;
            case OF_EXP: out[i] = of_expf(a[i]); break;
            default: out[i] = 0.0f;
        }
    }
}

/* Pulse-pair moments of R gates over M pulses, binary64 throughout (the
 * estimator's definition: lag-0 powers, lag-1 autocorrelation of H, H-V
 * cross-correlation; see include/rsd.h rsd_moments).  iq[k][4 g + c] with
 * c = (Ih, Qh, Iv, Qv); out[4 g + (Z, V, ZDR, rho)]. */
void orc_moments(const double *iq, int M, int R, double lambda, double prt, double gain, double range_start,
                 double range_delta, double *out) {
    for (int g = 0; g < R; g++) {
        double ph = 0.0, pv = 0.0, r1r = 0.0, r1i = 0.0, cr = 0.0, ci = 0.0;
        for (int k = 0; k < M; k++) {
            const double *q = iq + (size_t)k * 4 * R + 4 * g;
            double hr = q[0] * gain, hi = q[1] * gain, vr = q[2] * gain, vi = q[3] * gain;
            ph += hr * hr + hi * hi;
            pv += vr * vr + vi * vi;
            cr += hr * vr + hi * vi;
            ci += hi * vr - hr * vi;
            if (k + 1 < M) {
                const double *n = iq + (size_t)(k + 1) * 4 * R + 4 * g;
                double nr = n[0] * gain, ni = n[1] * gain;
                r1r += hr * nr + hi * ni;
                r1i += hr * ni - hi * nr;
            }
        }
        ph /= M;
        pv /= M;
        cr /= M;
        ci /= M;
        double rg = range_start + g * range_delta;
        out[4 * g + 0] = 10.0 * log10(ph) + 20.0 * log10(fabs(rg));
        out[4 * g + 1] = -lambda / (4.0 * M_PI * prt) * atan2(r1i, r1r);
        out[4 * g + 2] = 10.0 * log10(ph / pv);
        out[4 * g + 3] = sqrt(cr * cr + ci * ci) / sqrt(ph * pv);
    }
}
