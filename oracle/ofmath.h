/*
 * ofmath.h -- deterministic binary32 elementary functions for the CPU oracle.
 *
 * TEST INFRASTRUCTURE: this file belongs to the oracle.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may build or call it.
 *
 * The reference evaluates atan2/acos/sincos/cos/log with the OpenCL device's
 * built-ins (rs.cl:218-219, 275-276, 292-305, 326-344, 848-857, 894, 1066), whose
 * accuracy is implementation-defined.  The oracle fixes one documented
 * algorithm per function so that the GPU kernels, which implement the same
 * operation sequence, reproduce positions, respawn decisions and gate indices
 * bit for bit.  (The stretched wind index's log1p, rs.cl:151, goes through the
 * index tables of rs_oracle.c instead.)
 *
 *   log         : fdlibm-style reduction u = m 2^k, m in [sqrt(1/2), sqrt(2)),
 *                 s = f/(2+f), log(1+f) = f - (hfsq - s (hfsq + R(s^2))) with R the
 *                 atanh series 2/3 z + 2/5 z^2 + 2/7 z^3 + 2/9 z^4 (truncation error
 *                 below 3e-9 relative), ln2 split in a 16-bit head and a tail.
 *   sincos      : Cody-Waite reduction by pi/2 in three parts (each part the binary32
 *                 rounding of what is left of pi/2), Taylor polynomials of sin to r^9
 *                 and cos to r^10 on |r| <= pi/4.
 *   atan2       : ratio a = min/max in [0, 1]; a > tan(pi/8) is mapped by
 *                 (a - 1)/(a + 1) + pi/4; atan(t) = t + t z P(z), z = t^2, with P a
 *                 degree-4 weighted least-squares fit on Chebyshev nodes of
 *                 [0, tan(pi/8)] (max relative error 7.3e-10 before rounding).
 *   acos        : atan2(sqrt((1 - x)(1 + x)), x).
 *
 * Every product-sum is written either as an explicit fmaf() or as separate
 * binary32 operations; the file is compiled with -ffp-contract=off.
 */
#ifndef OFMATH_H
#define OFMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#define OF_PIO2_1   0x1.921fb6p+0f
#define OF_PIO2_2  -0x1.777a5cp-25f
#define OF_PIO2_3  -0x1.ee59dap-50f
#define OF_2_OVER_PI 0x1.45f306p-1f
#define OF_PI_HI    0x1.921fb6p+1f
#define OF_PI_LO   -0x1.777a5cp-24f
#define OF_PIO2_HI  0x1.921fb6p+0f
#define OF_PIO2_LO -0x1.777a5cp-25f
#define OF_PIO4_HI  0x1.921fb6p-1f
#define OF_PIO4_LO -0x1.777a5cp-26f
#define OF_TANPI8   0x1.a8279ap-2f
#define OF_LN2_HI   0x1.62e400p-1f
#define OF_LN2_LO   0x1.7f7d1cp-20f
#define OF_SQRT2    0x1.6a09e6p+0f

static inline float of_bits2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t of_f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* log(1 + f) for f in [sqrt(1/2) - 1, sqrt(2) - 1], plus k ln2 and a tail term. */
static inline float of_log_kernel(float f, int k, float tail) {
    float s = f / (2.0f + f);
    float z = s * s;
    float R = z * fmaf(z, fmaf(z, fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f), 0x1.99999ap-2f), 0x1.555556p-1f);
    float hfsq = 0.5f * (f * f);
    float kf = (float)k;
    float lo = fmaf(kf, OF_LN2_LO, tail);
    return kf * OF_LN2_HI - ((hfsq - (s * (hfsq + R) + lo)) - f);
}

/* Split a positive normal u into m 2^k with m in [sqrt(1/2), sqrt(2)). */
static inline float of_split(float u, int *k) {
    uint32_t iu = of_f2bits(u);
    int e = (int)((iu >> 23) & 0xffu) - 127;
    float m = of_bits2f((iu & 0x007fffffu) | 0x3f800000u);
    if (m > OF_SQRT2) { m = 0.5f * m; e += 1; }
    *k = e;
    return m;
}

static inline float of_logf(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -INFINITY : NAN;
    if (isinf(x)) return x;
    int k;
    float scale = 0.0f;
    if (x < 0x1p-126f) { x = x * 0x1p+24f; scale = 24.0f; }
    float m = of_split(x, &k);
    float r = of_log_kernel(m - 1.0f, k, 0.0f);
    return (scale != 0.0f) ? r - scale * 0x1.62e430p-1f : r;
}

/* sin and cos of x; exact reduction for |x| below about 2^20. */
static inline void of_sincosf(float x, float *s_out, float *c_out) {
    float q = rintf(x * OF_2_OVER_PI);
    float r = fmaf(-q, OF_PIO2_1, x);
    r = fmaf(-q, OF_PIO2_2, r);
    r = fmaf(-q, OF_PIO2_3, r);
    float z = r * r;
    float ps = fmaf(z, fmaf(z, fmaf(z, 0x1.71de3ap-19f, -0x1.a01a02p-13f), 0x1.111112p-7f), -0x1.555556p-3f);
    float s = fmaf(r * z, ps, r);
    float pc = fmaf(z, fmaf(z, fmaf(z, fmaf(z, -0x1.27e4fcp-22f, 0x1.a01a02p-16f), -0x1.6c16c2p-10f), 0x1.555556p-5f), -0.5f);
    float c = fmaf(z, pc, 1.0f);
    int n = ((int)q) & 3;
    switch (n) {
        case 0: *s_out = s; *c_out = c; break;
        case 1: *s_out = c; *c_out = -s; break;
        case 2: *s_out = -s; *c_out = -c; break;
        default: *s_out = -c; *c_out = s; break;
    }
}

static inline float of_sinf(float x) { float s, c; of_sincosf(x, &s, &c); return s; }
static inline float of_cosf(float x) { float s, c; of_sincosf(x, &s, &c); return c; }

/* atan(a) for a in [0, 1]. */
static inline float of_atan01(float a) {
    float base_hi = 0.0f, base_lo = 0.0f, t = a;
    if (a > OF_TANPI8) {
        t = (a - 1.0f) / (a + 1.0f);
        base_hi = OF_PIO4_HI;
        base_lo = OF_PIO4_LO;
    }
    float z = t * t;
    float p = fmaf(z, fmaf(z, fmaf(z, fmaf(z, -0x1.f31592p-5f, 0x1.b22d32p-4f), -0x1.23ba22p-3f), 0x1.9991e6p-3f), -0x1.55554ap-2f);
    float at = fmaf(t * z, p, t);
    return base_hi + (at + base_lo);
}

static inline float of_atan2f(float y, float x) {
    if (isnan(x) || isnan(y)) return x + y;
    float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) {
        r = signbit(x) ? OF_PI_HI : 0.0f;
        return copysignf(r, y);
    }
    if (isinf(ax) || isinf(ay)) {
        if (isinf(ax) && isinf(ay)) r = signbit(x) ? 3.0f * OF_PIO4_HI : OF_PIO4_HI;
        else if (isinf(ay)) r = OF_PIO2_HI;
        else r = signbit(x) ? OF_PI_HI : 0.0f;
        return copysignf(r, y);
    }
    int swap = ay > ax;
    float a = swap ? ax / ay : ay / ax;
    r = of_atan01(a);
    if (swap) r = OF_PIO2_HI - (r - OF_PIO2_LO);
    if (signbit(x)) r = OF_PI_HI - (r - OF_PI_LO);
    return copysignf(r, y);
}

static inline float of_acosf(float x) {
    float s = sqrtf((1.0f - x) * (1.0f + x));
    return of_atan2f(s, x);
}

/* acos of a cosine by construction (unit vector or quaternion component, a
 * coordinate over the norm bounding it): arguments rounded beyond [-1, 1] are
 * taken as +-1 (the reference's acos would give NaN there); NaN stays NaN. */
static inline float of_acos_unit(float x) { return of_acosf(x > 1.0f ? 1.0f : (x < -1.0f ? -1.0f : x)); }

/* OpenCL-style helpers with a fixed evaluation order. */
static inline float of_clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float of_mixf(float x, float y, float a) { return x + (y - x) * a; }
static inline float of_fractf(float x, float *ip) {
    float fl = floorf(x);
    *ip = fl;
    return fminf(x - fl, 0x1.fffffep-1f);
}
static inline float of_signf(float x) {
    if (isnan(x)) return 0.0f;
    if (x > 0.0f) return 1.0f;
    if (x < 0.0f) return -1.0f;
    return x;
}
static inline float of_len2(float x, float y) { return sqrtf(x * x + y * y); }
static inline float of_len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
static inline float of_len4(float x, float y, float z, float w) { return sqrtf(((x * x + y * y) + z * z) + w * w); }
static inline float of_radiansf(float d) { return d * 0x1.1df46ap-6f; }

/* exp(x): Cody-Waite reduction x = k ln2 + r (ln2 in a 16-bit head and a tail),
 * |r| <= ln2 / 2, Taylor polynomial of e^r to r^8, scaled by 2^k. */
static inline float of_expf(float x) {
    const float k = rintf(x * 0x1.715476p+0f);
    float r = fmaf(-k, OF_LN2_HI, x);
    r = fmaf(-k, OF_LN2_LO, r);
    float p = 1.0f / 40320.0f;
    p = fmaf(p, r, 1.0f / 5040.0f);
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}

/* pow(a, y) for a > 0 as exp(y log a). */
static inline float of_powf_pos(float a, float y) { return of_expf(y * of_logf(a)); }

#endif
