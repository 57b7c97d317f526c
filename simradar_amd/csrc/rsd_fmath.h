// rsd_fmath.h -- deterministic binary32 elementary functions for the gfx950 kernels.
//
// The position update of every scatterer feeds the next pulse's respawn test,
// so the advection arithmetic must not depend on which math library evaluates
// it.  These functions use only IEEE-754 binary32 add/mul/fma, the correctly
// rounded division and square root (HIP's defaults for '/' and sqrtf), floor and
// rint, in a fixed order (the translation unit is compiled with
// -ffp-contract=off; every fused multiply-add is an explicit __builtin_fmaf):
//
//   log         : u = m 2^k with m in [sqrt(1/2), sqrt(2)); s = f / (2 + f);
//                 log(1 + f) = f - (hfsq - s (hfsq + R)), R = z (2/3 + z (2/5 +
//                 z (2/7 + z 2/9))), z = s^2; ln2 split into a 16-bit head and a
//                 tail.
//   sincos      : three-part Cody-Waite reduction by pi/2, Taylor polynomials of
//                 sin (to r^9) and cos (to r^10) on |r| <= pi/4.
//   atan2       : a = min/max, a > tan(pi/8) mapped by (a - 1)/(a + 1) + pi/4,
//                 atan(t) = t + t z P(z), P a degree-4 least-squares fit on
//                 [0, tan(pi/8)] (relative error 7.3e-10 before rounding).
//   acos        : atan2(sqrt((1 - x)(1 + x)), x).
//
// Constants are the binary32 roundings of the exact values (pi/2 split, ln 2
// split, 1/n!, 2/(2n+1)) or of the fitted coefficients, written in hex.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RF_PIO2_1   0x1.921fb6p+0f
#define RF_PIO2_2  -0x1.777a5cp-25f
#define RF_PIO2_3  -0x1.ee59dap-50f
#define RF_2_OVER_PI 0x1.45f306p-1f
#define RF_PI_HI    0x1.921fb6p+1f
#define RF_PI_LO   -0x1.777a5cp-24f
#define RF_PIO2_HI  0x1.921fb6p+0f
#define RF_PIO2_LO -0x1.777a5cp-25f
#define RF_PIO4_HI  0x1.921fb6p-1f
#define RF_PIO4_LO -0x1.777a5cp-26f
#define RF_TANPI8   0x1.a8279ap-2f
#define RF_LN2_HI   0x1.62e400p-1f
#define RF_LN2_LO   0x1.7f7d1cp-20f
#define RF_SQRT2    0x1.6a09e6p+0f
#define RF_2PI      0x1.921fb6p+2f

__device__ __forceinline__ float rf_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ float rf_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float rf_div(float a, float b) { return a / b; }

__device__ __forceinline__ float rf_log_kernel(float f, int k, float tail) {
    float s = rf_div(f, 2.0f + f);
    float z = s * s;
    float R = z * rf_fma(z, rf_fma(z, rf_fma(z, 0x1.c71c72p-3f, 0x1.24924ap-2f), 0x1.99999ap-2f), 0x1.555556p-1f);
    float hfsq = 0.5f * (f * f);
    float kf = (float)k;
    float lo = rf_fma(kf, RF_LN2_LO, tail);
    return kf * RF_LN2_HI - ((hfsq - (s * (hfsq + R) + lo)) - f);
}

__device__ __forceinline__ float rf_split(float u, int *k) {
    uint32_t iu = __float_as_uint(u);
    int e = (int)((iu >> 23) & 0xffu) - 127;
    float m = __uint_as_float((iu & 0x007fffffu) | 0x3f800000u);
    if (m > RF_SQRT2) { m = 0.5f * m; e += 1; }
    *k = e;
    return m;
}

// log(x) for positive normal x (the respawn draws are at least 2^-31).
__device__ __forceinline__ float rf_logf(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_inff() : __builtin_nanf("");
    if (__builtin_isinf(x)) return x;
    int k;
    float scale = 0.0f;
    if (x < 0x1p-126f) { x = x * 0x1p+24f; scale = 24.0f; }
    float m = rf_split(x, &k);
    float r = rf_log_kernel(m - 1.0f, k, 0.0f);
    return (scale != 0.0f) ? r - scale * 0x1.62e430p-1f : r;
}

// Faster forms for quantities that only enter the pulse sums (checked against
// the oracle within the IQ tolerance, not bit for bit): hardware reciprocal
// (about 1 ulp) and acos from asin(x) = x + x^3 P(x^2) on |x| <= 1/2 (P a
// degree-4 least-squares fit, relative error 6.4e-9) with the half-angle
// reduction acos(c) = 2 asin(sqrt((1 - c) / 2)) for |c| > 1/2.
__device__ __forceinline__ float rf_rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ float rf_sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

// atan2 with the ratio taken through the hardware reciprocal.
__device__ __forceinline__ float rf_atan2_fast(float y, float x) {
    float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    float mx = __builtin_fmaxf(ax, ay), mn = __builtin_fminf(ax, ay);
    float a = mx > 0.0f ? mn * rf_rcp_fast(mx) : 0.0f;
    float t = a, base_hi = 0.0f, base_lo = 0.0f;
    if (a > RF_TANPI8) {
        t = (a - 1.0f) * rf_rcp_fast(a + 1.0f);
        base_hi = RF_PIO4_HI;
        base_lo = RF_PIO4_LO;
    }
    float z = t * t;
    float p = rf_fma(z, rf_fma(z, rf_fma(z, rf_fma(z, -0x1.f31592p-5f, 0x1.b22d32p-4f), -0x1.23ba22p-3f), 0x1.9991e6p-3f), -0x1.55554ap-2f);
    float r = base_hi + (rf_fma(t * z, p, t) + base_lo);
    if (ay > ax) r = RF_PIO2_HI - (r - RF_PIO2_LO);
    if (__builtin_signbit(x)) r = RF_PI_HI - (r - RF_PI_LO);
    return __builtin_copysignf(r, y);
}

__device__ __forceinline__ float rf_asin_half(float x) {
    float z = x * x;
    float p = rf_fma(z, rf_fma(z, rf_fma(z, rf_fma(z, 0x1.567920p-5f, 0x1.8fd26ep-6f), 0x1.741946p-5f), 0x1.33063ap-4f), 0x1.5555bap-3f);
    return rf_fma(x * z, p, x);
}

__device__ __forceinline__ float rf_acos_fast(float c) {
    c = __builtin_fminf(__builtin_fmaxf(c, -1.0f), 1.0f);
    float a = __builtin_fabsf(c);
    if (a <= 0.5f) return RF_PIO2_HI - (rf_asin_half(c) - RF_PIO2_LO);
    float s = 2.0f * rf_asin_half(rf_sqrt(0.5f * (1.0f - a)));
    return c > 0.0f ? s : RF_PI_HI - (s - RF_PI_LO);
}

__device__ __forceinline__ void rf_sincos_poly(float r, int n, float *s_out, float *c_out) {
    float z = r * r;
    float ps = rf_fma(z, rf_fma(z, rf_fma(z, 0x1.71de3ap-19f, -0x1.a01a02p-13f), 0x1.111112p-7f), -0x1.555556p-3f);
    float s = rf_fma(r * z, ps, r);
    float pc = rf_fma(z, rf_fma(z, rf_fma(z, rf_fma(z, -0x1.27e4fcp-22f, 0x1.a01a02p-16f), -0x1.6c16c2p-10f), 0x1.555556p-5f), -0.5f);
    float c = rf_fma(z, pc, 1.0f);
    // quadrant n mod 4: (s, c), (c, -s), (-s, -c), (-c, s) by selects and sign
    // flips (a divergent switch would run every case under masks)
    const uint32_t q = (uint32_t)n;
    const float a = (q & 1u) ? c : s, b = (q & 1u) ? s : c;
    *s_out = __uint_as_float(__float_as_uint(a) ^ ((q & 2u) << 30));
    *c_out = __uint_as_float(__float_as_uint(b) ^ (((q + 1u) & 2u) << 30));
}

__device__ __forceinline__ void rf_sincosf(float x, float *s_out, float *c_out) {
    float q = __builtin_rintf(x * RF_2_OVER_PI);
    float r = rf_fma(-q, RF_PIO2_1, x);
    r = rf_fma(-q, RF_PIO2_2, r);
    r = rf_fma(-q, RF_PIO2_3, r);
    rf_sincos_poly(r, (int)q, s_out, c_out);
}

// sin(2 pi t), cos(2 pi t) for a phase t given in turns: the quarter-turn
// reduction t - n/4 is exact, so only the final scaling by 2 pi rounds.
__device__ __forceinline__ void rf_sincos_turns(float t, float *s_out, float *c_out) {
    float n = __builtin_rintf(4.0f * t);
    float f = rf_fma(n, -0.25f, t);
    rf_sincos_poly(f * RF_2PI, (int)n, s_out, c_out);
}

__device__ __forceinline__ float rf_cosf(float x) { float s, c; rf_sincosf(x, &s, &c); return c; }

__device__ __forceinline__ float rf_atan01(float a) {
    float base_hi = 0.0f, base_lo = 0.0f, t = a;
    if (a > RF_TANPI8) {
        t = rf_div(a - 1.0f, a + 1.0f);
        base_hi = RF_PIO4_HI;
        base_lo = RF_PIO4_LO;
    }
    float z = t * t;
    float p = rf_fma(z, rf_fma(z, rf_fma(z, rf_fma(z, -0x1.f31592p-5f, 0x1.b22d32p-4f), -0x1.23ba22p-3f), 0x1.9991e6p-3f), -0x1.55554ap-2f);
    float at = rf_fma(t * z, p, t);
    return base_hi + (at + base_lo);
}

__device__ __forceinline__ float rf_atan2f(float y, float x) {
    if (__builtin_isnan(x) || __builtin_isnan(y)) return x + y;
    float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) {
        r = __builtin_signbit(x) ? RF_PI_HI : 0.0f;
        return __builtin_copysignf(r, y);
    }
    if (__builtin_isinf(ax) || __builtin_isinf(ay)) {
        if (__builtin_isinf(ax) && __builtin_isinf(ay)) r = __builtin_signbit(x) ? 3.0f * RF_PIO4_HI : RF_PIO4_HI;
        else if (__builtin_isinf(ay)) r = RF_PIO2_HI;
        else r = __builtin_signbit(x) ? RF_PI_HI : 0.0f;
        return __builtin_copysignf(r, y);
    }
    bool swap = ay > ax;
    float a = swap ? rf_div(ax, ay) : rf_div(ay, ax);
    r = rf_atan01(a);
    if (swap) r = RF_PIO2_HI - (r - RF_PIO2_LO);
    if (__builtin_signbit(x)) r = RF_PI_HI - (r - RF_PI_LO);
    return __builtin_copysignf(r, y);
}

__device__ __forceinline__ float rf_acosf(float x) {
    float s = rf_sqrt((1.0f - x) * (1.0f + x));
    return rf_atan2f(s, x);
}

// acos of a quantity that is a cosine by construction (a component of a unit
// vector or unit quaternion, a coordinate over the norm that bounds it):
// rounding can carry it a few ulp beyond [-1, 1], where acos is NaN (and a NaN
// orientation or velocity never leaves the state); such arguments are taken as
// +-1, a NaN stays NaN.
__device__ __forceinline__ float rf_acos_unit(float x) {
    return rf_acosf(x > 1.0f ? 1.0f : (x < -1.0f ? -1.0f : x));
}

__device__ __forceinline__ float rf_clamp(float x, float lo, float hi) { return __builtin_fminf(__builtin_fmaxf(x, lo), hi); }
__device__ __forceinline__ float rf_mix(float x, float y, float a) { return x + (y - x) * a; }
__device__ __forceinline__ float rf_fract(float x, float *ip) {
    float fl = __builtin_floorf(x);
    *ip = fl;
    return __builtin_fminf(x - fl, 0x1.fffffep-1f);
}
__device__ __forceinline__ float rf_sign(float x) {
    if (__builtin_isnan(x)) return 0.0f;
    if (x > 0.0f) return 1.0f;
    if (x < 0.0f) return -1.0f;
    return x;
}
__device__ __forceinline__ float rf_len2(float x, float y) { return rf_sqrt(x * x + y * y); }
__device__ __forceinline__ float rf_len3(float x, float y, float z) { return rf_sqrt((x * x + y * y) + z * z); }
__device__ __forceinline__ float rf_len4(float x, float y, float z, float w) { return rf_sqrt(((x * x + y * y) + z * z) + w * w); }
__device__ __forceinline__ float rf_radians(float d) { return d * 0x1.1df46ap-6f; }

// exp(x): Cody-Waite reduction x = k ln2 + r (ln2 in a 16-bit head and a tail,
// so k * head is exact for |k| < 256), |r| <= ln2 / 2, Taylor polynomial of e^r
// to r^8 (truncation below 3e-10 relative), scaled by 2^k.
__device__ __forceinline__ float rf_expf(float x) {
    const float k = __builtin_rintf(x * 0x1.715476p+0f);
    float r = __builtin_fmaf(-k, 0x1.62e400p-1f, x);
    r = __builtin_fmaf(-k, 0x1.7f7d1cp-20f, r);
    float p = 1.0f / 40320.0f;
    p = __builtin_fmaf(p, r, 1.0f / 5040.0f);
    p = __builtin_fmaf(p, r, 1.0f / 720.0f);
    p = __builtin_fmaf(p, r, 1.0f / 120.0f);
    p = __builtin_fmaf(p, r, 1.0f / 24.0f);
    p = __builtin_fmaf(p, r, 1.0f / 6.0f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_ldexpf(p, (int)k);
}

// pow(a, y) for a > 0 as exp(y log a); relative error about |y log a| 2^-24.
__device__ __forceinline__ float rf_powf_pos(float a, float y) { return rf_expf(y * rf_logf(a)); }
