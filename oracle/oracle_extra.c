/*
 * oracle_extra.c -- test entry points of the CPU oracle (TEST INFRASTRUCTURE).
 * Exposes the deterministic binary32 functions of ofmath.h so tests can sweep
 * them against binary64 libm and mpmath.
 */
#include <stddef.h>
#include "ofmath.h"

enum { OF_LOG1P = 0, OF_LOG, OF_SIN, OF_COS, OF_ATAN2, OF_ACOS };

void orc_math(int fn, const float *a, const float *b, float *out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        switch (fn) {
            case OF_LOG1P: out[i] = of_log1pf(a[i]); break;
            case OF_LOG: out[i] = of_logf(a[i]); break;
            case OF_SIN: out[i] = of_sinf(a[i]); break;
            case OF_COS: out[i] = of_cosf(a[i]); break;
            case OF_ATAN2: out[i] = of_atan2f(a[i], b[i]); break;
            case OF_ACOS: out[i] = of_acosf(a[i]); break;
            default: out[i] = 0.0f;
        }
    }
}
