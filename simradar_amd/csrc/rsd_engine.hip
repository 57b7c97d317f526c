// rsd_engine.hip -- gfx950 scatterer engine behind include/rsd.h.
//
// Data layout in HBM (one engine = one GPU):
//   pos, vel, ori, tum, rcs, aux : float4[N]   (structure of float4 streams)
//   rnd                          : uint4[N]     (four 31-bit LCG lanes)
//   type ranges                  : [0, n_rain) rain, then each debris type
//   wind                         : two float4[nz][ny][nx] LES slots (u, v, w, t)
//   adm cd/cm, rcs re/im         : float4[ny][nx] per table
//   small tables                 : drop RCS (96 float4), angular / range weights,
//                                  binary32 gate ranges
//   partials                     : double[4 R][rows] per-block gate sums, one
//                                  column per (component, gate)
//
// Rain keeps its radar cross-section out of memory: it is a function of the
// drop radius and the position's elevation (rs.cl:268-280) and is recomputed in
// registers every pulse.  The one case the reference keeps a value that the
// position no longer determines is a respawn (rs.cl:685-692 leaves rcs as it
// was); the engine then writes that value to the rcs stream and marks the
// scatterer by the sign bit of pos.w (the radius, always >= 0).  The flag is
// cleared by the next advance that does not respawn, and downloads return
// |pos.w| and the reconstructed rcs, so the streams a caller sees are the
// reference's.
//
// One pulse = one fused kernel (advance of the previous PRT, if pending, then
// the signal of every scatterer folded into its gates; table-driven debris types
// in leading blocks of the rain launch, surrogate-driven types in a kernel of
// their own on a second stream) plus a finalize kernel that sums the per-block
// gate partials in binary64.
//
// Spatial order: every sort_interval pulses the rain streams (pos, vel, rcs,
// rnd) are reordered by range (shells of 1/8 gate) and, inside a shell, by
// a Morton index of 16^3 cells of the domain box, with id[j] = the reference
// index of stored scatterer j.  Each scatterer keeps its own state and random
// stream, so the physics is unchanged; the scatterers a wave visits then fall
// into the same few range gates (summed in registers, see WaveWin) and
// neighbouring lanes gather neighbouring wind texels.  Downloads scatter through
// id and uploads first restore the reference order, so callers never see it.
// aux (read only by downloads) stays in the reference order.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hipcub/hipcub.hpp>

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rsd.h"
#include "rsd_fmath.h"
#include "rsd_mlp.h"
#include "rsd_mlp_host.h"

// Sample tiles of 16 the debris kernel's network forward carries at once (1, 2
// or 4): fewer tiles hold fewer activation registers and read each weight tile
// from LDS more often.
#ifndef RSD_MLP_TILES
#define RSD_MLP_TILES 4
#endif
#ifndef RSD_MLP_XR
#define RSD_MLP_XR 4                // network outputs exchanged per sample per round (12 or 4)
#endif
#define RSD_MLP_SCRW (RSD_MLP_XR > 8 ? RSD_MLP_XR : 8)   // scratch floats per sample (inputs: 8)
#ifndef RSD_DEBRIS_SURR_WAVES
#define RSD_DEBRIS_SURR_WAVES 3     // waves per SIMD the surrogate debris kernel is compiled for
#endif

#define RSD_MIN_HEIGHT 2.0f
#define RSD_PI_F 0x1.921fb6p+1f
#define RSD_PI_2_F 0x1.921fb6p+0f
#define RSD_SQRT1_2_F 0x1.6a09e6p-1f

// State stream access; NT: non-temporal (the lines are not kept in the caches,
// which then hold the tables for the scatterers still to come).
typedef float rsd_v4 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ float4 ld4(const float4 *p) {
    if (!NT) return *p;
    const rsd_v4 v = __builtin_nontemporal_load((const rsd_v4 *)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
template <bool NT>
__device__ __forceinline__ void st4(float4 *p, float4 x) {
    if (!NT) {
        *p = x;
        return;
    }
    const rsd_v4 v = {x.x, x.y, x.z, x.w};
    __builtin_nontemporal_store(v, (rsd_v4 *)p);
}

struct F16 { float s[16]; };
struct F4 { float s[4]; };
// A 3-D table held padded: (nx + 3) x (ny + 3) x (nz + 3) texels, the table at
// offset (1, 1, 1) with its edge texels repeated once before and twice after
// along every axis (RSD_PAD_*), so that the clamped texel pair of the sampler
// is always k, k + 1 of the padded volume.  px, pxy: padded row and plane
// lengths; fx, fy, fz: nx, ny, nz as floats.  d2 != NULL: a second volume of the
// same shape (the next LES frame) and the weight of the linear time
// interpolation between the two samples.
struct Tex3 {
    const float4 *d;
    int nx, ny, nz;
    uint32_t px, pxy;
    float fx, fy, fz;
    const float4 *d2;
    float alpha;
};
#define RSD_PAD_LO 1
#define RSD_PAD_HI 2
struct Tex2 { const float4 *d; int nx, ny; };

// ---------------------------------------------------------------------------
// Table sampling: OpenCL linear filter, unnormalised coordinates, clamp to edge
// (rs.cl:8).  Texel pair floor(u - 0.5), +1, clamped; weight (u - 0.5) - floor;
// lerp(p, q, a) = fma(a, q, fma(-a, p, p)); x first, then y, then z.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tex_axis(float u, int n, int &i0, int &i1, float &a) {
    float uu = u - 0.5f;
    float fl = __builtin_floorf(uu);
    a = uu - fl;
    float fc = __builtin_fminf(__builtin_fmaxf(fl, -1.0f), (float)n);
    int k = (int)fc;
    i0 = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
    i1 = k + 1 < 0 ? 0 : (k + 1 > n - 1 ? n - 1 : k + 1);
}

__device__ __forceinline__ float lerpf_(float p, float q, float t) { return rf_fma(t, q, rf_fma(-t, p, p)); }
__device__ __forceinline__ float4 lerp4(float4 p, float4 q, float t) {
    return make_float4(lerpf_(p.x, q.x, t), lerpf_(p.y, q.y, t), lerpf_(p.z, q.z, t), lerpf_(p.w, q.w, t));
}

// Trilinear sample in two halves: gather3d issues the eight texel reads and
// lerp3d combines them, so that independent work can run while the reads are
// in flight.  Texel order t[4 z + 2 y + x].
struct Gather3 { float4 t[8]; float a, b, c; };

// tex_axis on a padded axis: the padded index of the lower texel (the clamped
// floor(u - 0.5) + RSD_PAD_LO) and the weight; the pair is (k, k + 1).
__device__ __forceinline__ uint32_t tex_axis_padded(float u, float nf, float &a) {
    const float uu = u - 0.5f;
    const float fl = __builtin_floorf(uu);
    a = uu - fl;
    return (uint32_t)((int)__builtin_fminf(__builtin_fmaxf(fl, -1.0f), nf) + RSD_PAD_LO);
}

__device__ __forceinline__ void gather3d(const Tex3 &t, float u, float v, float w, Gather3 &g) {
    const uint32_t i = tex_axis_padded(u, t.fx, g.a);
    const uint32_t j = tex_axis_padded(v, t.fy, g.b);
    const uint32_t k = tex_axis_padded(w, t.fz, g.c);
    const float4 *b0 = t.d + (k * t.pxy + j * t.px + i);
    const float4 *b1 = b0 + t.px;
    const float4 *b2 = b0 + t.pxy;
    const float4 *b3 = b2 + t.px;
    g.t[0] = b0[0]; g.t[1] = b0[1];
    g.t[2] = b1[0]; g.t[3] = b1[1];
    g.t[4] = b2[0]; g.t[5] = b2[1];
    g.t[6] = b3[0]; g.t[7] = b3[1];
}

// Trilinear texels of two volumes stored as texel pairs (pair[2 t] = volume 1,
// pair[2 t + 1] = volume 2, the padded layout of t): the eight corners of both
// volumes in eight 32-byte reads, each pair within one cache line, so the
// second volume's texels come with the first's.
__device__ __forceinline__ void gather3d_pair(const Tex3 &t, const float4 *pair, float u, float v, float w, Gather3 &g,
                                              Gather3 &g2) {
    const uint32_t i = tex_axis_padded(u, t.fx, g.a);
    const uint32_t j = tex_axis_padded(v, t.fy, g.b);
    const uint32_t k = tex_axis_padded(w, t.fz, g.c);
    g2.a = g.a;
    g2.b = g.b;
    g2.c = g.c;
    const float4 *b0 = pair + 2u * (k * t.pxy + j * t.px + i);
    const float4 *b1 = b0 + 2u * t.px;
    const float4 *b2 = b0 + 2u * t.pxy;
    const float4 *b3 = b2 + 2u * t.px;
    g.t[0] = b0[0]; g2.t[0] = b0[1]; g.t[1] = b0[2]; g2.t[1] = b0[3];
    g.t[2] = b1[0]; g2.t[2] = b1[1]; g.t[3] = b1[2]; g2.t[3] = b1[3];
    g.t[4] = b2[0]; g2.t[4] = b2[1]; g.t[5] = b2[2]; g2.t[5] = b2[3];
    g.t[6] = b3[0]; g2.t[6] = b3[1]; g.t[7] = b3[2]; g2.t[7] = b3[3];
}

__device__ __forceinline__ float4 lerp3d(const Gather3 &g) {
    float4 x00 = lerp4(g.t[0], g.t[1], g.a);
    float4 x10 = lerp4(g.t[2], g.t[3], g.a);
    float4 x01 = lerp4(g.t[4], g.t[5], g.a);
    float4 x11 = lerp4(g.t[6], g.t[7], g.a);
    return lerp4(lerp4(x00, x10, g.b), lerp4(x01, x11, g.b), g.c);
}

__device__ __forceinline__ float4 sample3d(const Tex3 &t, float u, float v, float w) {
    Gather3 g;
    gather3d(t, u, v, w, g);
    const float4 a = lerp3d(g);
    if (!t.d2) return a;
    Tex3 t2 = t;
    t2.d = t.d2;
    gather3d(t2, u, v, w, g);
    return lerp4(a, lerp3d(g), t.alpha);
}

__device__ __forceinline__ float4 sample2d(const Tex2 &t, float u, float v) {
    int i0, i1, j0, j1;
    float a, b;
    tex_axis(u, t.nx, i0, i1, a);
    tex_axis(v, t.ny, j0, j1, b);
    const size_t nx = (size_t)t.nx;
    const float4 *D = t.d;
    float4 y0 = lerp4(D[j0 * nx + i0], D[j0 * nx + i1], a);
    float4 y1 = lerp4(D[j1 * nx + i0], D[j1 * nx + i1], a);
    return lerp4(y0, y1, b);
}

// Stretched-grid index tables.  The index of a stretched LES axis is
// c = s8 + copysign(s0 log1p(s4 |r|), r) (rs.cl:130-151), r the position
// relative to the box centre in x and y and to z = 0 in z.  P(x) = s0 log1p(s4 x)
// is evaluated from RSD_ITAB_N cubic pieces per axis over x in [0, X], X the
// box's extent from the origin on that axis (with a margin): piece j covers x =
// (j + f) / ih, f in [0, 1], and holds the binary32 monomial coefficients
// (a0, a1, a2, a3) of the cubic interpolating binary64 P at the four Chebyshev
// nodes of f (rsd_wind_index_table builds them on the host).  Evaluation:
// t = |r| ih, j = min(floor(t), N - 1), f = t - j, P = a0 + f (a1 + f (a2 + f a3))
// by fma, c = copysign(P, r) + s8.  Rows: x at 0, y at N, z at 2 N.
#define RSD_ITAB_N 256
#define RSD_ITAB_ROWS (3 * RSD_ITAB_N)

struct WIdx {
    const float4 *t;        // RSD_ITAB_ROWS rows in device memory
    float ih[3];            // pieces per metre
    float jm[3];            // last piece index (N - 1)
};

__device__ __forceinline__ float itab_eval(float r, const float4 *T, float ih, float jm, float base) {
    const float t = __builtin_fabsf(r) * ih;
    const float jf = __builtin_fminf(__builtin_floorf(t), jm);
    const float f = t - jf;
    const float4 a = T[(int)jf];
    const float p = rf_fma(rf_fma(rf_fma(a.w, f, a.z), f, a.y), f, a.x);
    return __builtin_copysignf(p, r) + base;
}

__device__ __forceinline__ void itab_stage(float4 *dst, const float4 *src) {
    for (int j = threadIdx.x; j < RSD_ITAB_ROWS; j += blockDim.x) dst[j] = src[j];
}

// rs.cl:123-176: LES table coordinates of a position.  Stretched tables (slot
// 7 = 7) go through the index tables T (LDS or global) with the scales of w;
// STRETCHED: the launch has checked that the table is stretched, so only that
// branch is compiled, and (cx, cy) is the domain centre sd[8..9] + 0.5
// sd[12..13] rounded as below.
template <bool STRETCHED = false>
__device__ __forceinline__ float4 wind_index(float4 pos, const F16 &d, const F16 &sd, const float4 *T, const WIdx &w,
                                             float cx = 0.0f, float cy = 0.0f) {
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t spacing = __float_as_uint(d.s[7]);
    if (STRETCHED || spacing == 7u) {
        if (!STRETCHED) {
            cx = sd.s[8] + 0.5f * sd.s[12];
            cy = sd.s[9] + 0.5f * sd.s[13];
        }
        c.x = itab_eval(pos.x - cx, T, w.ih[0], w.jm[0], d.s[8]);
        c.y = itab_eval(pos.y - cy, T + RSD_ITAB_N, w.ih[1], w.jm[1], d.s[9]);
        c.z = itab_eval(pos.z - 0.0f, T + 2 * RSD_ITAB_N, w.ih[2], w.jm[2], d.s[10]);
    } else if (spacing == 0u) {
        c.x = rf_fma(pos.x, d.s[0], d.s[4]);
        c.y = rf_fma(pos.y, d.s[1], d.s[5]);
        c.z = rf_fma(pos.z, d.s[2], d.s[6]);
    }
    return c;
}

// rs.cl:35-44
__device__ __forceinline__ float4 rand4(uint4 &s) {
    const float n = 1.0f / 2147483647.0f;
    s.x = (s.x * 16807u) & 0x7FFFFFFFu;
    s.y = (s.y * 16807u) & 0x7FFFFFFFu;
    s.z = (s.z * 16807u) & 0x7FFFFFFFu;
    s.w = (s.w * 16807u) & 0x7FFFFFFFu;
    return make_float4((float)s.x * n, (float)s.y * n, (float)s.z * n, (float)s.w * n);
}

__device__ __forceinline__ bool is_outside(float4 p, const F16 &sd) {
    float ex = sd.s[8] + sd.s[12], ey = sd.s[9] + sd.s[13], ez = sd.s[10] + sd.s[14];
    return (p.x <= sd.s[8]) | (p.y <= sd.s[9]) | (p.z <= sd.s[10]) | (p.x >= ex) | (p.y >= ey) | (p.z >= ez);
}

// rs.cl:268-280 with cos^2(atan2(z, rho)) = rho^2 / (rho^2 + z^2): the table
// entry (alpha_x, alpha_z) of a drop radius (truncating index), then the
// elevation mix of the position.
__device__ __forceinline__ uint32_t drop_index(float radius, const F4 &td) {
    return (uint32_t)rf_clamp(rf_fma(radius, td.s[0], td.s[1]), 0.0f, td.s[2]);
}

__device__ __forceinline__ float4 drop_mix(float4 p, float4 xz) {
    float rho2 = p.x * p.x + p.y * p.y;
    float cb2 = rho2 * rf_rcp_fast(rho2 + p.z * p.z);
    return make_float4(xz.x, xz.y, xz.x + (xz.z - xz.x) * cb2, xz.y + (xz.w - xz.y) * cb2);
}

__device__ __forceinline__ float4 ellipsoid_rcs(float4 p, const float4 *table, const F4 &td) {
    return drop_mix(p, table[drop_index(p.w, td)]);
}

// rs.cl:54-85
__device__ __forceinline__ float4 qmul(float4 l, float4 r) {
    return make_float4(((l.w * r.x - l.z * r.y) + l.y * r.z) + l.x * r.w,
                       ((l.w * r.y + l.z * r.x) + l.y * r.w) - l.x * r.z,
                       ((l.w * r.z + l.z * r.w) - l.y * r.x) + l.x * r.y,
                       ((l.w * r.w - l.z * r.z) - l.y * r.y) - l.x * r.x);
}
__device__ __forceinline__ float4 qconj(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }
__device__ __forceinline__ float4 qrot(float4 v, float4 q) { return qmul(qmul(q, v), qconj(q)); }
__device__ __forceinline__ float4 norm4(float4 q) {
    float n = rf_len4(q.x, q.y, q.z, q.w);
    return make_float4(rf_div(q.x, n), rf_div(q.y, n), rf_div(q.z, n), rf_div(q.w, n));
}

// rs.cl:282-359
__device__ __forceinline__ float4 debris_rcs(float4 pos, float4 ori, const Tex2 &rr, const Tex2 &ri, const F16 &d, uint32_t concept) {
    float el = rf_atan2f(pos.z, rf_len2(pos.x, pos.y));
    float az = rf_atan2f(pos.x, pos.y);
    float se, ce, sa, ca;
    rf_sincosf(0.5f * (el + RSD_PI_2_F), &se, &ce);
    rf_sincosf(0.5f * az, &sa, &ca);
    float4 oc = make_float4((-se * ca + ce * sa) * RSD_SQRT1_2_F, (ce * ca + se * sa) * RSD_SQRT1_2_F,
                            (se * ca + ce * sa) * RSD_SQRT1_2_F, (ce * ca + (-se) * sa) * RSD_SQRT1_2_F);
    float4 R = qmul(ori, oc);
    float4 q = make_float4(R.x, R.z, -R.y, R.w);
    float alpha, beta, gamma;
    float beta_arg = ((q.w * q.w + q.z * q.z) - q.y * q.y) - q.x * q.x;
    if (beta_arg > 0.999847f || beta_arg < -0.999847f) {
        alpha = 0.0f;
        beta = 0.0f;
        gamma = (rf_sign(q.z) * rf_acos_unit(q.w)) * 2.0f;
    } else {
        alpha = rf_atan2f(q.y * q.z - q.w * q.x, q.x * q.z + q.w * q.y);
        beta = rf_acosf(((q.w * q.w + q.z * q.z) - q.y * q.y) - q.x * q.x);
        gamma = rf_atan2f(q.y * q.z + q.w * q.x, q.w * q.y - q.x * q.z);
    }
    float u = rf_fma(alpha, d.s[0], d.s[4]);
    float v = rf_fma(beta, d.s[1], d.s[5]);
    float4 re = sample2d(rr, u, v);
    float4 im = sample2d(ri, u, v);
    float sg, cg;
    rf_sincosf(gamma, &sg, &cg);
    float4 ss;
    ss.x = cg * (cg * re.x - re.z * sg) - sg * (cg * re.z - re.y * sg);
    ss.y = cg * (cg * im.x - im.z * sg) - sg * (cg * im.z - im.y * sg);
    ss.z = cg * (cg * re.y + re.z * sg) + sg * (cg * re.z + re.x * sg);
    ss.w = cg * (cg * im.y + im.z * sg) + sg * (cg * im.z + im.x * sg);
    if (concept & 1u) {
        float hv_r = cg * (cg * re.z + re.x * sg) - sg * (cg * re.y + re.z * sg);
        float hv_i = cg * (cg * im.z + im.x * sg) - sg * (cg * im.y + im.z * sg);
        float vh_r = cg * (cg * re.z - re.y * sg) + sg * (cg * re.x - re.z * sg);
        float vh_i = cg * (cg * im.z - im.y * sg) + sg * (cg * im.x - im.z * sg);
        ss.x += vh_r;
        ss.y += vh_i;
        ss.z += hv_r;
        ss.w += hv_i;
    }
    return ss;
}

// Radar-relative orientation R = ori (x) o_conj(el, az) of rs.cl:287-315, with
// the sign fixed so that R.w >= 0 (R and -R are the same rotation).  The
// half-angle cosines and sines come from the position's direction without
// atan2 and sincos: with rho = |(x, y)|, r = |p| and theta = el + pi / 2,
// (cos, sin)(theta / 2) is along (r - z, rho) = along (rho, r + z), and with
// phi = az = atan2(x, y), (cos, sin)(phi / 2) is along (rho + y, x) = along
// (|x|, sign(x) (rho - y)); of each pair the form without cancellation is
// taken.  The surrogate's inputs only (the table path keeps rs.cl's angles).
__device__ __forceinline__ void half_angle(float u, float v, float &c, float &s) {
    const float n = rf_len2(u, v);
    c = rf_div(u, n);
    s = rf_div(v, n);
}

__device__ __forceinline__ float4 radar_rel_quat(float4 pos, float4 ori) {
    const float rho = rf_len2(pos.x, pos.y), r = rf_len2(rho, pos.z);
    float se, ce, sa, ca;
    if (r == 0.0f) {            // at the radar: el = az = 0
        ce = se = RSD_SQRT1_2_F;
    } else if (pos.z <= 0.0f) {
        half_angle(r - pos.z, rho, ce, se);
    } else {
        half_angle(rho, r + pos.z, ce, se);
    }
    if (rho == 0.0f) {          // on the vertical: az = atan2(x, +-0) = +-0
        ca = 1.0f;
        sa = __builtin_copysignf(0.0f, pos.x);
    } else if (pos.y >= 0.0f) {
        half_angle(rho + pos.y, pos.x, ca, sa);
    } else {
        half_angle(__builtin_fabsf(pos.x), __builtin_copysignf(rho - pos.y, pos.x), ca, sa);
    }
    float4 oc = make_float4((-se * ca + ce * sa) * RSD_SQRT1_2_F, (ce * ca + se * sa) * RSD_SQRT1_2_F,
                            (se * ca + ce * sa) * RSD_SQRT1_2_F, (ce * ca + (-se) * sa) * RSD_SQRT1_2_F);
    float4 R = qmul(ori, oc);
    if (R.w < 0.0f) R = make_float4(-R.x, -R.y, -R.z, -R.w);
    return R;
}

// The quantities the surrogate network learns, from the tables (rs.cl:207-243,
// 316-347): cd, cm at the body-frame flow direction u, and the gamma-rotated
// amplitudes hh, vv, hv (re, im) at the radar-relative orientation R.
__device__ __forceinline__ void table_targets(float4 R, float4 uh, const Tex2 &cdt, const Tex2 &cmt, const F16 &ad,
                                              const Tex2 &rr, const Tex2 &ri, const F16 &d, float t[12]) {
    float beta = rf_acos_unit(uh.x);
    float alpha = rf_atan2f(uh.z, uh.y);
    if (alpha < 0.0f) {
        alpha = RSD_PI_F + alpha;
        beta = -beta;
    }
    float4 cd = sample2d(cdt, rf_fma(beta, ad.s[0], ad.s[4]), rf_fma(alpha, ad.s[1], ad.s[5]));
    float4 cm = sample2d(cmt, rf_fma(beta, ad.s[0], ad.s[4]), rf_fma(alpha, ad.s[1], ad.s[5]));
    float4 q = make_float4(R.x, R.z, -R.y, R.w);
    float a, b, g;
    float beta_arg = ((q.w * q.w + q.z * q.z) - q.y * q.y) - q.x * q.x;
    if (beta_arg > 0.999847f || beta_arg < -0.999847f) {
        a = 0.0f;
        b = 0.0f;
        g = (rf_sign(q.z) * rf_acos_unit(q.w)) * 2.0f;
    } else {
        a = rf_atan2f(q.y * q.z - q.w * q.x, q.x * q.z + q.w * q.y);
        b = rf_acosf(beta_arg);
        g = rf_atan2f(q.y * q.z + q.w * q.x, q.w * q.y - q.x * q.z);
    }
    float4 re = sample2d(rr, rf_fma(a, d.s[0], d.s[4]), rf_fma(b, d.s[1], d.s[5]));
    float4 im = sample2d(ri, rf_fma(a, d.s[0], d.s[4]), rf_fma(b, d.s[1], d.s[5]));
    float sg, cg;
    rf_sincosf(g, &sg, &cg);
    t[0] = cd.x; t[1] = cd.y; t[2] = cd.z;
    t[3] = cm.x; t[4] = cm.y; t[5] = cm.z;
    t[6] = cg * (cg * re.x - re.z * sg) - sg * (cg * re.z - re.y * sg);
    t[7] = cg * (cg * im.x - im.z * sg) - sg * (cg * im.z - im.y * sg);
    t[8] = cg * (cg * re.y + re.z * sg) + sg * (cg * re.z + re.x * sg);
    t[9] = cg * (cg * im.y + im.z * sg) + sg * (cg * im.z + im.x * sg);
    t[10] = cg * (cg * re.z + re.x * sg) - sg * (cg * re.y + re.z * sg);
    t[11] = cg * (cg * im.z + im.x * sg) - sg * (cg * im.y + im.z * sg);
}

// Debris RCS from the network's amplitudes (rs.cl:340-356 after the lookups;
// the cross-polar term vh equals hv for the symmetric scattering matrix).
__device__ __forceinline__ float4 rcs_from_amps(const float *t, uint32_t concept) {
    float4 ss = make_float4(t[6], t[7], t[8], t[9]);
    if (concept & 1u) {
        ss.x += t[10];
        ss.y += t[11];
        ss.z += t[10];
        ss.w += t[11];
    }
    return ss;
}

// ---------------------------------------------------------------------------
// Pulse formation arguments shared by all scatterer kernels.
// ---------------------------------------------------------------------------
struct PulseK {
    F16 sd;                 // sim_desc of the pulse (beam in slots 0-2)
    double two_over_lambda; // turns per metre of two-way path
    const float *angw; F4 angw_desc;
    const float *rangew; F4 rangew_desc;
    int rangew_n;           // entries of the range weight table
    const float *gate_r;    // binary32 gate ranges (start + repeated adds)
    int R;
    float gate_start, gate_inv_delta;
    float sup_lo, sup_hi;   // range-offset interval outside which the weight is 0
    int dense;              // 1: every gate may be weighted
    int tri;                // 1: the range weight table is the triangle (0, 1, 0)
    float tri_s0, tri_s1;   // its desc slots 0 and 1 - 1 (see tri_weight)
    int fold;               // RSD_FOLD_* strategy of the gate sums
    uint32_t phase_mode;
    double *partials;       // [4 R][rows]: column j = component j / R, gate j % R
    int row0;               // first partial row of this launch
    int rows;               // rows of the partial buffer (column length)
};

__device__ __forceinline__ float table_lerp(float x, const float *W, const F4 &d) {
    // rs.cl:1067-1093 / 1192-1206: clamp, fract, mix of two table entries
    float x0 = rf_clamp(rf_fma(x, d.s[0], d.s[1]), 0.0f, d.s[2]);
    float x1 = rf_clamp(rf_fma(x, d.s[0], d.s[1] + 1.0f), 0.0f, d.s[2]);
    float ip;
    float dec = rf_fract(x0, &ip);
    return rf_mix(W[(uint32_t)ip], W[(uint32_t)__builtin_floorf(x1)], dec);
}

// Angle between the beam and a position: acos(b . p / |p|) of rs.cl:1066 in
// the well-conditioned form atan2(|b x p|, b . p) (near the axis the cosine
// form turns one ulp into 3.5e-4 rad).
__device__ __forceinline__ float beam_angle(float4 p, const F16 &sd) {
    float cx = sd.s[1] * p.z - sd.s[2] * p.y;
    float cy = sd.s[2] * p.x - sd.s[0] * p.z;
    float cz = sd.s[0] * p.y - sd.s[1] * p.x;
    return rf_atan2_fast(rf_sqrt_fast((cx * cx + cy * cy) + cz * cz), (sd.s[0] * p.x + sd.s[1] * p.y) + sd.s[2] * p.z);
}

// Signal x angular weight of one scatterer (rs.cl:1050-1110, 1175-1180).
// With HILO the position is the pair hi + lo (binary32 each) and the range that
// enters the phase is formed in binary64 from it.  F64: the launch's phase mode
// is RSD_PHASE_F64.
template <bool HILO = false, bool F64 = false>
__device__ __forceinline__ bool scat_signal(float4 p, float4 rcs, const PulseK &k, const float *angw, float4 &sa,
                                            float &r, float4 lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f)) {
    double r64 = 0.0;
    if (HILO) {
        const double x = (double)p.x + (double)lo.x, y = (double)p.y + (double)lo.y, z = (double)p.z + (double)lo.z;
        r64 = sqrt((x * x + y * y) + z * z);
        r = (float)r64;
    } else {
        r = rf_len3(p.x, p.y, p.z);
    }
    const float inv_r = rf_rcp_fast(r);
    float wa = table_lerp(beam_angle(p, k.sd), angw, k.angw_desc);
    if (wa == 0.0f) return false;
    float atten = inv_r * inv_r;
    float ss, cc;
    if (HILO) {
        double turns = r64 * k.two_over_lambda;
        float t = (float)(turns - floor(turns));
        rf_sincos_turns(t, &ss, &cc);
    } else if (F64 || k.phase_mode == RSD_PHASE_F64) {
        double turns = (double)r * k.two_over_lambda;
        float t = (float)(turns - floor(turns));
        rf_sincos_turns(t, &ss, &cc);
    } else {
        rf_sincosf(r * k.sd.s[4], &ss, &cc);
    }
    sa.x = ((rcs.x * cc - rcs.y * (-ss)) * atten) * wa;
    sa.y = ((rcs.y * cc + rcs.x * (-ss)) * atten) * wa;
    sa.z = ((rcs.z * cc - rcs.w * (-ss)) * atten) * wa;
    sa.w = ((rcs.w * cc + rcs.z * (-ss)) * atten) * wa;
    return true;
}

// Wave-wide binary64 sum through DPP row shifts and row broadcasts (GFX9
// cross-lane operations): after row_shr 1/2/4/8 lane 15 of each row holds the
// row sum, row_bcast 15 / 31 fold rows 0+1 and 2+3 into lane 63.
template <int CTRL, int ROW_MASK, bool BOUND>
__device__ __forceinline__ double dpp64(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, BOUND);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, BOUND);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum64(double v) {
    v += dpp64<0x111, 0xf, true>(v);
    v += dpp64<0x112, 0xf, true>(v);
    v += dpp64<0x114, 0xf, true>(v);
    v += dpp64<0x118, 0xf, true>(v);
    v += dpp64<0x142, 0xa, false>(v);
    v += dpp64<0x143, 0xc, false>(v);
    int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Wave-wide binary32 sum, same DPP pattern.
template <int CTRL, int ROW_MASK, bool BOUND>
__device__ __forceinline__ float dpp32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, BOUND));
}

__device__ __forceinline__ float wave_sum32(float v) {
    v += dpp32<0x111, 0xf, true>(v);
    v += dpp32<0x112, 0xf, true>(v);
    v += dpp32<0x114, 0xf, true>(v);
    v += dpp32<0x118, 0xf, true>(v);
    v += dpp32<0x142, 0xa, false>(v);
    v += dpp32<0x143, 0xc, false>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ int wave_min_i(int v) {
    for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
    for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m));
    return v;
}

#define RSD_WAVE_GATES 16

// Fold one scatterer per lane into the block's binary64 gate sums
// (rs.cl:1182-1217), gate-sparse: only gates whose range offset lies inside the
// range weight's support.  Called by every lane of the wave (on = has a
// contribution).  RSD_FOLD_LANE: every lane adds its own terms to LDS.  The wave
// strategies apply when the wave's gates span at most RSD_WAVE_GATES (spatially
// sorted scatterers): each gate's four components are summed across the wave
// (binary64 or binary32 partial sums of 64 terms) and one lane adds them.
__device__ __forceinline__ void fold_gates(bool on, float r, float4 sa, const PulseK &k, double *acc) {
    int klo = 0x3fffffff, khi = -1;
    if (on) {
        float klo_f, khi_f;
        if (k.dense) {
            klo_f = 0.0f;
            khi_f = (float)(k.R - 1);
        } else {
            klo_f = __builtin_fmaxf(__builtin_ceilf((r - k.sup_hi - k.gate_start) * k.gate_inv_delta), 0.0f);
            khi_f = __builtin_fminf(__builtin_floorf((r - k.sup_lo - k.gate_start) * k.gate_inv_delta), (float)(k.R - 1));
        }
        if (klo_f <= khi_f) {
            klo = (int)klo_f;
            khi = (int)khi_f;
        }
    }
    const int R = k.R;
    if (k.fold != RSD_FOLD_LANE) {
        const int kmin = wave_min_i(klo), kmax = wave_max_i(khi);
        if (kmax < kmin) return;
        if (kmax - kmin < RSD_WAVE_GATES) {
            const bool lead = (threadIdx.x & 63) == 0;
            if (k.fold == RSD_FOLD_WAVE32) {
                for (int g = kmin; g <= kmax; g++) {
                    float wr = 0.0f;
                    if (g >= klo && g <= khi) wr = table_lerp(r - k.gate_r[g], k.rangew, k.rangew_desc);
                    float s0 = wave_sum32(wr * sa.x), s1 = wave_sum32(wr * sa.y);
                    float s2 = wave_sum32(wr * sa.z), s3 = wave_sum32(wr * sa.w);
                    if (lead) {
                        atomicAdd(&acc[g], (double)s0);
                        atomicAdd(&acc[R + g], (double)s1);
                        atomicAdd(&acc[2 * R + g], (double)s2);
                        atomicAdd(&acc[3 * R + g], (double)s3);
                    }
                }
                return;
            }
            for (int g = kmin; g <= kmax; g++) {
                float wr = 0.0f;
                if (g >= klo && g <= khi) wr = table_lerp(r - k.gate_r[g], k.rangew, k.rangew_desc);
                double s0 = wave_sum64(wr == 0.0f ? 0.0 : (double)(wr * sa.x));
                double s1 = wave_sum64(wr == 0.0f ? 0.0 : (double)(wr * sa.y));
                double s2 = wave_sum64(wr == 0.0f ? 0.0 : (double)(wr * sa.z));
                double s3 = wave_sum64(wr == 0.0f ? 0.0 : (double)(wr * sa.w));
                if (lead) {
                    atomicAdd(&acc[g], s0);
                    atomicAdd(&acc[R + g], s1);
                    atomicAdd(&acc[2 * R + g], s2);
                    atomicAdd(&acc[3 * R + g], s3);
                }
            }
            return;
        }
    }
    for (int g = klo; g <= khi; g++) {
        float wr = table_lerp(r - k.gate_r[g], k.rangew, k.rangew_desc);
        if (wr == 0.0f) continue;
        atomicAdd(&acc[g], (double)(wr * sa.x));
        atomicAdd(&acc[R + g], (double)(wr * sa.y));
        atomicAdd(&acc[2 * R + g], (double)(wr * sa.z));
        atomicAdd(&acc[3 * R + g], (double)(wr * sa.w));
    }
}

#define RSD_LDS_DROP_MAX 256      // drop RCS entries staged in LDS by rain pulse launches
#define RSD_LDS_ANGW_MAX 1024     // angular weight entries staged in LDS
#define RSD_WIN_EMPTY (-(1 << 24))
// Gates of the wave-shared window of range-sorted rain: a scatterer's support
// under the triangle weight spans two gates (three within rounding of a gate
// range) and a launch's range-ordered scatterers move on through the gates, so
// eight gates keep nearly every lane in the window (four send too many lanes
// to the per-lane atomics and run slower, five waves per SIMD included).
#ifndef RSD_WWIN
#define RSD_WWIN 8
#endif

// Gates whose range offset lies inside the weight's support (klo > khi: none).
__device__ __forceinline__ void gate_support(float r, const PulseK &k, int &klo, int &khi) {
    float klo_f, khi_f;
    if (k.dense) {
        klo_f = 0.0f;
        khi_f = (float)(k.R - 1);
    } else {
        klo_f = __builtin_fmaxf(__builtin_ceilf((r - k.sup_hi - k.gate_start) * k.gate_inv_delta), 0.0f);
        khi_f = __builtin_fminf(__builtin_floorf((r - k.sup_lo - k.gate_start) * k.gate_inv_delta), (float)(k.R - 1));
    }
    if (klo_f <= khi_f) {
        klo = (int)klo_f;
        khi = (int)khi_f;
    } else {
        klo = 0x3fffffff;
        khi = -1;
    }
}

// Range weight of offset x from the table in LDS, held as pairs (W[i],
// W[min(i + 1, n - 1)] - W[i]): the same clamp / fract / mix as table_lerp.
// The upper entry is W[j], j = floor(clamp(fma(x, s0, s1 + 1))), as there.
// j is i + 1 (i at the clamped top, or with dec = 0, where the difference does
// not matter) except where x0 rounds to just below a node m and x1 rounds up
// to m + 1: the table read then mixes W[m - 1] with W[m + 1], and so does this
// (the rare lane reads W[j]).  Only used where the table ends in zeros, so
// offsets outside the support weigh exactly 0.
__device__ __forceinline__ float lds_weight(float x, const float2 *W, const F4 &d) {
    const float x0 = rf_clamp(rf_fma(x, d.s[0], d.s[1]), 0.0f, d.s[2]);
    const float x1 = rf_clamp(rf_fma(x, d.s[0], d.s[1] + 1.0f), 0.0f, d.s[2]);
    const float ip = __builtin_floorf(x0);
    const float dec = __builtin_fminf(x0 - ip, 0x1.fffffep-1f);
    const int i = (int)ip, j = (int)__builtin_floorf(x1);
    const float2 p = W[i];
    float dw = p.y;
    if (j > i + 1) dw = W[j].x - p.x;
    return p.x + dw * dec;
}

// The triangle table (0, 1, 0) with max index 2 (RS_set_range_weight_to_triangle,
// rs.c:2620-2623) evaluated in closed form: x0 = fma(x, s0, s1) in [0, 2] gives
// 1 - |x0 - 1| (the table's mix up to rounding), 0 outside.
__device__ __forceinline__ float tri_weight(float x, const PulseK &k) {
    return __builtin_fmaxf(0.0f, 1.0f - __builtin_fabsf(rf_fma(x, k.tri_s0, k.tri_s1)));
}

// LDS layout of the window folds: gate sums acc[4 R] (binary64), then the range
// weight table as the pairs lds_weight reads, then the R gate ranges.
__device__ __forceinline__ float2 *win_weights(double *acc, const PulseK &k) { return (float2 *)(acc + 4 * k.R); }
__device__ __forceinline__ float *win_gates(double *acc, const PulseK &k) {
    return (float *)(win_weights(acc, k) + k.rangew_n);
}

__device__ __forceinline__ void stage_window(double *acc, const PulseK &k) {
    float2 *wl = win_weights(acc, k);
    float *gr = win_gates(acc, k);
    const int n = k.rangew_n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const float w0 = k.rangew[j], w1 = k.rangew[min(j + 1, n - 1)];
        wl[j] = make_float2(w0, w1 - w0);
    }
    for (int j = threadIdx.x; j < k.R; j += blockDim.x) gr[j] = k.gate_r[j];
    __syncthreads();
}

// Fold of scatterers in no spatial order (debris): each lane adds its terms for
// the gates of its support straight to the block's binary64 gate sums in LDS
// (the range weight from the staged table pairs).
__device__ __forceinline__ void lane_fold(bool on, float r, float4 sa, const PulseK &k, double *acc, const float2 *wl,
                                          const float *grl) {
    int klo = 0x3fffffff, khi = -1;
    if (on) gate_support(r, k, klo, khi);
    const int R = k.R;
    for (int g = klo; g <= khi; g++) {
        const float wr = lds_weight(r - grl[g], wl, k.rangew_desc);
        if (wr != 0.0f) {
            atomicAdd(&acc[g], (double)(wr * sa.x));
            atomicAdd(&acc[R + g], (double)(wr * sa.y));
            atomicAdd(&acc[2 * R + g], (double)(wr * sa.z));
            atomicAdd(&acc[3 * R + g], (double)(wr * sa.w));
        }
    }
}

// ---------------------------------------------------------------------------
// Wave-shared window (rain streams in range order): the 64 lanes of a wave
// share one window of RSD_WWIN consecutive gates [kb, kb + RSD_WWIN), kb
// wave-uniform.  Each lane keeps binary32 partial sums of its scatterers' terms
// for the 32 (gate, component) slots; the wave reduces them in one transposed
// pass when the window moves or the launch ends.  Scatterers whose gates fall
// outside the window (respawned ones, until the next reorder) add their terms
// to LDS directly.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_perm(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Transposed sum over the wave of 32 values per lane: afterwards lanes 2j and
// 2j + 1 hold the sum over all 64 lanes of a[j].  Each step pairs every lane
// with one partner, keeps half of its values and adds the partner's copy of
// that half: l ^ 32 (permlane32 swap), l ^ 16 (permlane16 swap), l ^ 15 (DPP
// row_mirror), l ^ 7 (row_half_mirror), l ^ 2 and l ^ 1 (quad_perm).
// 70 instructions for the 32 sums instead of 32 separate wave reductions.
__device__ __forceinline__ float wave_transpose_sum(const float (&a)[32]) {
    float b[16], c[8], d[4], e[2];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(a[j]), __float_as_uint(a[j + 16]), false, false);
        b[j] = __uint_as_float(s[0]) + __uint_as_float(s[1]);      // value j + 16 bit5(l)
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(b[j]), __float_as_uint(b[j + 8]), false, false);
        c[j] = __uint_as_float(s[0]) + __uint_as_float(s[1]);      // value j + 8 bit4 + 16 bit5
    }
    const unsigned lane = __lane_id();
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float t0 = c[j] + dpp_perm<0x140>(c[j]);
        const float t1 = c[j + 4] + dpp_perm<0x140>(c[j + 4]);
        d[j] = (lane & 8u) ? t1 : t0;                               // value j + 4 bit3 + ...
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const float t0 = d[j] + dpp_perm<0x141>(d[j]);
        const float t1 = d[j + 2] + dpp_perm<0x141>(d[j + 2]);
        e[j] = (lane & 4u) ? t1 : t0;                               // value j + 2 bit2 + ...
    }
    const float t0 = e[0] + dpp_perm<0x4e>(e[0]);
    const float t1 = e[1] + dpp_perm<0x4e>(e[1]);
    const float f = (lane & 2u) ? t1 : t0;                          // value l >> 1
    return f + dpp_perm<0xb1>(f);
}

struct WaveWin {
    float a[4 * RSD_WWIN];  // slot 4 s + c: gate kb + s, component c
    int kb;                 // wave-uniform first gate; RSD_WIN_EMPTY: no window
};

// Bitwise OR over the wave (row_shr 1/2/4/8 within each row of 16 lanes, then
// row_bcast 15 / 31 into lane 63), returned wave-uniform.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_or(uint32_t v) {
    return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
    v = dpp_or<0x111, 0xf>(v);
    v = dpp_or<0x112, 0xf>(v);
    v = dpp_or<0x114, 0xf>(v);
    v = dpp_or<0x118, 0xf>(v);
    v = dpp_or<0x142, 0xa>(v);
    v = dpp_or<0x143, 0xc>(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ void wwin_clear(WaveWin &w) {
#pragma unroll
    for (int j = 0; j < 4 * RSD_WWIN; j++) w.a[j] = 0.0f;
}

// Terms of one scatterer straight into the LDS gate sums.
__device__ __forceinline__ void lane_fold(float r, float4 sa, int klo, int khi, const PulseK &k, double *acc,
                                          const float2 *wl, const float *grl) {
    const int R = k.R;
    for (int g = klo; g <= khi; g++) {
        const float wr = lds_weight(r - grl[g], wl, k.rangew_desc);
        if (wr == 0.0f) continue;
        atomicAdd(&acc[g], (double)(wr * sa.x));
        atomicAdd(&acc[R + g], (double)(wr * sa.y));
        atomicAdd(&acc[2 * R + g], (double)(wr * sa.z));
        atomicAdd(&acc[3 * R + g], (double)(wr * sa.w));
    }
}

__device__ __forceinline__ void wwin_flush(const WaveWin &w, const PulseK &k, double *acc) {
    if (w.kb == RSD_WIN_EMPTY) return;
    float a32[32];      // the transposed sum takes 32 values: the window's, then zeros
#pragma unroll
    for (int j = 0; j < 32; j++) a32[j] = j < 4 * RSD_WWIN ? w.a[j] : 0.0f;
    const float v = wave_transpose_sum(a32);
    const unsigned lane = __lane_id();
    const int g = w.kb + (int)(lane >> 3), c = (int)((lane >> 1) & 3u);
    if (!(lane & 1u) && g >= 0 && g < k.R && v != 0.0f) atomicAdd(&acc[c * k.R + g], (double)v);
}

// One scatterer per lane (on: it has a contribution).  The window is placed
// once per launch, at the first step with contributions, one gate below the
// median of the first, middle and last contributing lanes' first gates (a lane
// far from the others, e.g. a respawned scatterer, cannot misplace it); a
// launch's range-ordered scatterers then stay inside it, and those that do not
// fold directly.  The slots a lane touches need no mask: outside a scatterer's
// support the table weight is exactly 0 (the window is only used for tables
// that end in zeros), and lanes outside the window carry a zero signal; so the
// wave visits the union of its window lanes' slots (one OR over the wave of
// per-lane slot bits) and every lane adds its term there.  A visited slot lies
// inside some lane's [klo, khi], hence inside [0, R).
__device__ __forceinline__ int median3(int a, int b, int c) { return max(min(a, b), min(max(a, b), c)); }

template <bool TRI>
__device__ __forceinline__ void wwin_fold(WaveWin &w, bool on, float r, float4 sa, const PulseK &k, double *acc,
                                          const float2 *wl, const float *grl) {
    int klo = 0x3fffffff, khi = -1;
    if (on) gate_support(r, k, klo, khi);
    on = klo <= khi;
    const unsigned long long on_m = __ballot(on);
    if (on_m == 0) return;
    if (w.kb == RSD_WIN_EMPTY) {
        const int l0 = __builtin_ctzll(on_m), l2 = 63 - __builtin_clzll(on_m);
        const int l1 = ((on_m >> 32) & 1ull) ? 32 : l0;
        w.kb = median3(__builtin_amdgcn_readlane(klo, l0), __builtin_amdgcn_readlane(klo, l1),
                       __builtin_amdgcn_readlane(klo, l2)) - 1;
    }
    const bool fit = on && klo >= w.kb && khi < w.kb + RSD_WWIN;
    if (on && !fit) lane_fold(r, sa, klo, khi, k, acc, wl, grl);
    if (!fit) sa = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t used = wave_or(fit ? ((2u << (khi - klo)) - 1u) << (klo - w.kb) : 0u);
#pragma unroll
    for (int s = 0; s < RSD_WWIN; s++) {
        if (!(used & (1u << s))) continue;
        const int g = w.kb + s;
        const float x = r - grl[g];
        const float wr = TRI ? tri_weight(x, k) : lds_weight(x, wl, k.rangew_desc);
        w.a[4 * s + 0] = __builtin_fmaf(wr, sa.x, w.a[4 * s + 0]);
        w.a[4 * s + 1] = __builtin_fmaf(wr, sa.y, w.a[4 * s + 1]);
        w.a[4 * s + 2] = __builtin_fmaf(wr, sa.z, w.a[4 * s + 2]);
        w.a[4 * s + 3] = __builtin_fmaf(wr, sa.w, w.a[4 * s + 3]);
    }
}

// Blocks are dispatched round-robin over the 8 XCDs, so L = (b mod 8) (G / 8) +
// b / 8 numbers the blocks of each XCD consecutively.
__device__ __forceinline__ unsigned xcd_order(unsigned b, unsigned G) {
    return (G & 7u) == 0u ? (b & 7u) * (G >> 3) + (b >> 3) : b;
}

// Contiguous range of scatterers of block b: chunk L owns [L chunk, (L + 1)
// chunk).  XCD-contiguous (L = xcd_order): each XCD takes one contiguous part
// of the streams and keeps its L2 on one region of the tables (debris, batch
// items).  Interleaved (L = b): consecutive chunks go to different XCDs, so
// every XCD gets a share of every part of the range-sorted rain streams and
// the XCDs finish together (the cost per scatterer varies along the streams).
// Returns xcd_order(b), the block's partial-sum row: the rows of one XCD stay
// adjacent either way.
__device__ __forceinline__ unsigned block_range_at(size_t begin, size_t end, unsigned b, unsigned G, size_t &lo,
                                                   size_t &hi, bool interleave = false) {
    const unsigned row = xcd_order(b, G);
    const unsigned L = interleave ? b : row;
    const size_t n = end - begin;
    size_t chunk = (n + G - 1) / G;
    chunk = (chunk + 63) / 64 * 64;     // whole waves; every block but the last equal
    lo = begin + (size_t)L * chunk;
    hi = lo + chunk;
    if (lo > end) lo = end;
    if (hi > end) hi = end;
    return row;
}

__device__ __forceinline__ void block_range(size_t begin, size_t end, size_t &lo, size_t &hi) {
    (void)block_range_at(begin, end, blockIdx.x, gridDim.x, lo, hi);
}

__device__ __forceinline__ void acc_clear(double *acc, int n) {
    for (int j = threadIdx.x; j < n; j += blockDim.x) acc[j] = 0.0;
    __syncthreads();
}

// Row `row` of the transposed partial buffer: one value per column.  The blocks
// of one XCD own consecutive rows, so its L2 merges them into whole lines.
__device__ __forceinline__ void acc_flush_row(const double *acc, const PulseK &k, int row) {
    __syncthreads();
    for (int j = threadIdx.x; j < 4 * k.R; j += blockDim.x) k.partials[(size_t)j * k.rows + row] = acc[j];
}

// ---------------------------------------------------------------------------
// Rain (type 0): el_atts (rs.cl:653-739) or bg_atts (rs.cl:533-580), fused with
// the pulse.  Template flags select advance and/or pulse.
// ---------------------------------------------------------------------------
// A pulse's partial-sum columns (rows binary64 values each, one per block of
// its launch) and the output slot(s) of their sums.
struct FinK {
    const double *partials;
    double *out, *out2;
    int rows, n4r;
};

struct RainK {
    float4 *pos, *vel, *rcs;
    float4 *plo;            // low parts of the positions (hi/lo mode), else NULL
    uint4 *rnd;
    unsigned long long n;
    Tex3 wind;
    F16 les;
    F16 sd;                 // sim_desc of the advance
    const float4 *drop;
    F4 drop_desc;
    uint32_t dragged;
    WIdx widx;              // stretched-grid index tables of the wind slot
    const float4 *wpair;    // time-interpolated wind: both volumes as texel pairs (specialised loop)
    PulseK pk;
    // Pulse launches stage the drop RCS table (drop_n entries) and the angular
    // weight table (angw_n entries) in LDS at byte offset tab_off; tab_lds = 0:
    // both are read from global memory.  In LDS their reads do not wait behind
    // the wind texel reads in flight (global reads complete in issue order).
    int tab_lds, drop_n, angw_n;
    unsigned tab_off;
    // Host-rounded constants of the specialised launch (k_rain FAST): the far
    // corner sd[8..10] + sd[12..14] of the box and the domain centre sd[8..9] +
    // 0.5 sd[12..13] of the stretched wind index.
    float hx, hy, hz, cx, cy;
    // The previous pulse's finalize, run by the last fin_blocks blocks of this
    // launch (after the rain blocks); fin_blocks = 0: none.
    FinK fin;
    unsigned fin_blocks;
};

// Byte offset of the staged rain tables in the dynamic LDS of a pulse launch:
// after the 4 R binary64 gate sums and, with the window fold, the weight pairs
// and gate ranges; 16-byte aligned.
__host__ __device__ __forceinline__ size_t rain_tab_offset(int R, int win, int rangew_n) {
    size_t o = (size_t)4 * R * sizeof(double);
    if (win) o += (size_t)rangew_n * sizeof(float2) + (size_t)R * sizeof(float);
    return (o + 15) & ~(size_t)15;
}

__device__ __forceinline__ void stage_rain_tables(double *acc, const RainK &a) {
    if (a.tab_lds) {
        float4 *dt = (float4 *)((char *)acc + a.tab_off);
        float *at = (float *)(dt + a.drop_n);
        for (int j = threadIdx.x; j < a.drop_n; j += blockDim.x) dt[j] = a.drop[j];
        for (int j = threadIdx.x; j < a.angw_n; j += blockDim.x) at[j] = a.pk.angw[j];
    }
    __syncthreads();
}

// (hi, lo) += v dt carried as a pair of binary32 values: the product's rounding
// error from fma (TwoProduct), the sum's from TwoSum, then renormalised so that
// |lo| <= ulp(hi) / 2.
__device__ __forceinline__ void hilo_step(float &hi, float &lo, float v, float dt) {
    const float d = v * dt;
    const float de = __builtin_fmaf(v, dt, -d);
    const float s = hi + d;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (d - bp);
    const float l = lo + (err + de);
    hi = s + l;
    lo = l - (hi - s);
}

// One rain step in three parts (el_atts, rs.cl:653-739, when dragged; bg_atts,
// rs.cl:533-580, otherwise): rain_move updates the position and respawns a
// scatterer that left the box (returns true; vel = 0 then and the step ends);
// rain_gather reads the wind texels at the new position; rain_drag turns them
// into the new velocity.  HILO: positions are hi + lo pairs (lo in its own
// stream); the physics reads hi.  FAST: the launch is dragged (concept D), the
// box's far corner comes rounded from the host.
template <bool HILO = false, bool FAST = false>
__device__ __forceinline__ bool rain_move(const RainK &a, size_t i, float4 &pos, float4 &vel, float4 &lo) {
    const F16 &sd = a.sd;
    const float dt = sd.s[11];
    const bool dragged = FAST || a.dragged;
    if (dragged) {
        if (HILO) {
            hilo_step(pos.x, lo.x, vel.x, dt);
            hilo_step(pos.y, lo.y, vel.y, dt);
            hilo_step(pos.z, lo.z, vel.z, dt);
        } else {
            pos.x = rf_fma(vel.x, dt, pos.x);
            pos.y = rf_fma(vel.y, dt, pos.y);
            pos.z = rf_fma(vel.z, dt, pos.z);
        }
        pos.w = rf_fma(vel.w, 0.0f, pos.w);
    } else {
        if (HILO) {
            hilo_step(pos.x, lo.x, vel.x, dt);
            hilo_step(pos.y, lo.y, vel.y, dt);
            hilo_step(pos.z, lo.z, vel.z, dt);
        } else {
            pos.x = pos.x + vel.x * dt;
            pos.y = pos.y + vel.y * dt;
            pos.z = pos.z + vel.z * dt;
        }
        pos.w = pos.w + vel.w * 0.0f;
    }
    const bool out = FAST ? (pos.x <= sd.s[8]) | (pos.y <= sd.s[9]) | (pos.z <= sd.s[10]) | (pos.x >= a.hx) |
                                (pos.y >= a.hy) | (pos.z >= a.hz)
                          : is_outside(pos, sd);
    if (out) {
        uint4 seed = a.rnd[i];
        float4 r = rand4(seed);
        a.rnd[i] = seed;
        if (dragged) {
            pos.x = rf_fma(r.x, sd.s[12], sd.s[8]);
            pos.y = rf_fma(r.y, sd.s[13], sd.s[9]);
            pos.z = rf_fma(r.z, sd.s[14], sd.s[10]);
        } else {
            pos.x = r.x * sd.s[12] + sd.s[8];
            pos.y = r.y * sd.s[13] + sd.s[9];
            pos.z = r.z * sd.s[14] + sd.s[10];
        }
        vel = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return true;
    }
    return false;
}

template <bool TI = false, bool FAST = false>
__device__ __forceinline__ void rain_gather(const RainK &a, float4 pos, const float4 *itab, Gather3 &g, Gather3 *g2 = nullptr) {
    float4 c = wind_index<FAST>(pos, a.les, a.sd, itab, a.widx, a.cx, a.cy);
    if (TI && FAST) {
        gather3d_pair(a.wind, a.wpair, c.x, c.y, c.z, g, *g2);
        return;
    }
    gather3d(a.wind, c.x, c.y, c.z, g);
    if (TI) {
        Tex3 t2 = a.wind;
        t2.d = a.wind.d2;
        gather3d(t2, c.x, c.y, c.z, *g2);
    }
}

template <bool FAST = false>
__device__ __forceinline__ void rain_drag(const RainK &a, float4 pos, float4 &vel, float4 bg) {
    if (!FAST && !a.dragged) {
        vel = bg;
        return;
    }
    const float dt = a.sd.s[11];
    const uint32_t concept = __float_as_uint(a.sd.s[5]);
    const float aom = rf_div(0.003006012f, pos.w);
    float4 dv = make_float4(bg.x - vel.x, bg.y - vel.y, bg.z - vel.z, bg.w - vel.w);
    float dva = rf_len3(dv.x, dv.y, dv.z);
    if (dva > 1.0e-3f) {
        float re = (6.7308e4f * (2.0f * pos.w)) * dva;
        float cd = (rf_div(24.0f, re) + rf_div(6.0f, 1.0f + rf_sqrt(re))) + 0.4f;
        float k = (((0.5f * 1.225f) * cd) * aom) * dva;
        float4 du = make_float4(k * dv.x + 0.0f, k * dv.y + 0.0f, k * dv.z + (-9.8f), k * dv.w + 0.0f);
        vel.x = vel.x + du.x * dt;
        vel.y = vel.y + du.y * dt;
        vel.z = vel.z + du.z * dt;
        vel.w = vel.w + du.w * 0.0f;
        // |v| > max(1, 3 |v_bg|) (rs.cl:716) on squared norms
        float v2 = ((vel.x * vel.x + vel.y * vel.y) + vel.z * vel.z) + vel.w * vel.w;
        float b2 = ((bg.x * bg.x + bg.y * bg.y) + bg.z * bg.z) + bg.w * bg.w;
        if ((concept & 8u) && v2 > __builtin_fmaxf(1.0f, 9.0f * b2)) {
            vel.x = bg.x + 0.0f * dt;
            vel.y = bg.y + 0.0f * dt;
            vel.z = bg.z + (-9.8f) * dt;
            vel.w = bg.w + 0.0f * 0.0f;
        }
    } else {
        vel.x = vel.x + 0.0f * dt;
        vel.y = vel.y + 0.0f * dt;
        vel.z = vel.z + (-9.8f) * dt;
        vel.w = vel.w + 0.0f * 0.0f;
    }
}

// Scatterers [lo, hi) of one rain population: advance and/or pulse terms into
// the block's LDS gate sums acc (cleared by the caller).  The hi/lo variant
// (96 B of state per scatterer, meant for populations far beyond the 256 MB
// Infinity Cache) streams its state non-temporally.  FAST: the common launch
// (dragged, stretched wind index with non-negative growth, drop and angular
// tables in LDS, binary64 phase) compiled without the branches for the others,
// which keeps the loop's uniform values in scalar registers.
template <bool ADV, bool PULSE, bool HILO, bool WIN, bool TRI, bool TI, bool FAST = false>
__device__ __forceinline__ void rain_range_t(const RainK &a, size_t lo, size_t hi, double *acc, const float4 *itab) {
    const float2 *wl = win_weights(acc, a.pk);
    const float *grl = win_gates(acc, a.pk);
    const bool tl = PULSE && (FAST || a.tab_lds);
    const float4 *dropt = tl ? (const float4 *)((const char *)acc + a.tab_off) : a.drop;
    const float *angt = tl ? (const float *)(dropt + a.drop_n) : a.pk.angw;
    WaveWin win;
    win.kb = RSD_WIN_EMPTY;
    wwin_clear(win);
    // pos / vel of the lane's next scatterer are loaded one step ahead
    const float4 idle = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
    float4 pn = idle, vn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (lo + threadIdx.x < hi) {
        pn = ld4<HILO>(&a.pos[lo + threadIdx.x]);
        if (ADV) vn = ld4<HILO>(&a.vel[lo + threadIdx.x]);
    }
    for (size_t base = lo; base < hi; base += blockDim.x) {
        const size_t i = base + threadIdx.x;
        const bool live = i < hi;
        float4 p = pn;
        const float4 v_in = vn;
        if (i + blockDim.x < hi) {
            pn = ld4<HILO>(&a.pos[i + blockDim.x]);
            if (ADV) vn = ld4<HILO>(&a.vel[i + blockDim.x]);
        }
        bool stale = __builtin_signbit(p.w);
        p.w = __builtin_fabsf(p.w);
        float4 rcs;
        bool have_rcs = false;
        float4 plo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (HILO && live) plo = ld4<HILO>(&a.plo[i]);
        // Advance: the new position first; the wind texels of that position
        // are read next and consumed only after the pulse terms (which need
        // the position, not the velocity), so the reads overlap that work.
        float4 v = v_in;
        bool moved = false;     // live, advanced and not respawned
        Gather3 g, g2;
        if (ADV && live) {
            const float4 p_in = p;
            bool respawned = rain_move<HILO, FAST>(a, i, p, v, plo);
            if (HILO) st4<HILO>(&a.plo[i], plo);
            if (respawned) {
                if (!stale) {
                    rcs = ellipsoid_rcs(p_in, dropt, a.drop_desc);
                    a.rcs[i] = rcs;
                    stale = true;
                    have_rcs = true;
                }
            } else {
                stale = false;
                moved = true;
                rain_gather<TI, FAST>(a, p, itab, g, &g2);
            }
            st4<HILO>(&a.pos[i], make_float4(p.x, p.y, p.z, stale ? -p.w : p.w));
        }
        if (PULSE) {
            float4 sa = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float r = 0.0f;
            bool on = false;
            if (live) {
                if (!have_rcs) rcs = stale ? a.rcs[i] : ellipsoid_rcs(p, dropt, a.drop_desc);
                on = scat_signal<HILO, FAST>(p, rcs, a.pk, angt, sa, r, plo);
            }
            if (WIN) wwin_fold<TRI>(win, on, r, sa, a.pk, acc, wl, grl);
            else fold_gates(on, r, sa, a.pk, acc);
        }
        if (ADV && live) {
            if (moved) rain_drag<FAST>(a, p, v, TI ? lerp4(lerp3d(g), lerp3d(g2), a.wind.alpha) : lerp3d(g));
            st4<HILO>(&a.vel[i], v);
        }
    }
    if (WIN) wwin_flush(win, a.pk, acc);
}

// Position and velocity streams through buffer descriptors (gfx9 raw buffers:
// 32-bit byte offsets from a scalar base, range-checked against the stream's
// size; word 3 selects 32-bit data).  A lane past the end passes
// RSD_BUF_OOB: its load returns zeros and its store is dropped, so these
// accesses need no branch around them.
typedef unsigned int rsd_u4 __attribute__((ext_vector_type(4)));
#define RSD_BUF_OOB 0x80000000u
#define RSD_BUF_MAX_BYTES 0x7ffffff0ull    // streams the descriptors cover (with RSD_BUF_OOB beyond)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t stream_rsrc(const float4 *p, unsigned long long n) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, 0, (int)(n * sizeof(float4)), 0x00020000);
}

__device__ __forceinline__ float4 bld4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const rsd_u4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}

__device__ __forceinline__ void bst4(__amdgpu_buffer_rsrc_t r, uint32_t off, float4 x) {
    const rsd_u4 v = {__float_as_uint(x.x), __float_as_uint(x.y), __float_as_uint(x.z), __float_as_uint(x.w)};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, 0);
}

// The specialised fused advance + pulse over binary32 positions: the same
// steps and results as rain_range_t<true, true, false, WIN, TRI, TI, true>,
// with the next scatterer's pos / vel reads, the wind texel reads and the pos /
// vel writes issued by every lane (a dead lane's go out of range, a respawned
// lane's texels are read and not used), so no branch surrounds them and each
// wait on a read waits for that read alone, not for the writes after it.
// Stream indices fit 32 bits (the launch checks RSD_BUF_MAX_BYTES).
template <bool WIN, bool TRI, bool TI>
__device__ __forceinline__ void rain_range_buf(const RainK &a, size_t lo, size_t hi, double *acc, const float4 *itab) {
    const float2 *wl = win_weights(acc, a.pk);
    const float *grl = win_gates(acc, a.pk);
    const float4 *dropt = (const float4 *)((const char *)acc + a.tab_off);
    const float *angt = (const float *)(dropt + a.drop_n);
    const __amdgpu_buffer_rsrc_t rp = stream_rsrc(a.pos, a.n), rv = stream_rsrc(a.vel, a.n);
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    WaveWin win;
    win.kb = RSD_WIN_EMPTY;
    wwin_clear(win);
    const uint32_t lo32 = (uint32_t)lo, hi32 = (uint32_t)hi, bd = blockDim.x;
    const uint32_t i0 = lo32 + threadIdx.x;
    float4 pn = bld4(rp, i0 < hi32 ? i0 * 16u : RSD_BUF_OOB);
    float4 vn = bld4(rv, i0 < hi32 ? i0 * 16u : RSD_BUF_OOB);
    for (uint32_t base = lo32; base < hi32; base += bd) {
        const uint32_t i = base + threadIdx.x;
        const bool live = i < hi32;
        const uint32_t off = live ? i * 16u : RSD_BUF_OOB;
        float4 p = pn, v = vn;
        const uint32_t nx = i + bd;
        const uint32_t noff = nx < hi32 ? nx * 16u : RSD_BUF_OOB;
        pn = bld4(rp, noff);
        vn = bld4(rv, noff);
        bool stale = __builtin_signbit(p.w);
        p.w = __builtin_fabsf(p.w);
        float4 rcs = zero, plo = zero;
        bool have_rcs = false, moved = false;
        if (live) {
            const float4 p_in = p;
            if (rain_move<false, true>(a, (size_t)i, p, v, plo)) {
                // respawned: the pulse takes the rcs of the position it left
                // (stored now, or kept from a respawn the step before)
                if (!stale) {
                    rcs = ellipsoid_rcs(p_in, dropt, a.drop_desc);
                    a.rcs[i] = rcs;
                    stale = true;
                } else {
                    rcs = a.rcs[i];
                }
                have_rcs = true;
            } else {
                stale = false;
                moved = true;
            }
        }
        Gather3 g, g2;
        rain_gather<TI, true>(a, p, itab, g, &g2);
        bst4(rp, off, make_float4(p.x, p.y, p.z, stale ? -p.w : p.w));
        float4 sa = zero;
        float r = 0.0f;
        bool on = false;
        if (live) {
            if (!have_rcs) rcs = ellipsoid_rcs(p, dropt, a.drop_desc);     // moved: not stale
            on = scat_signal<false, true>(p, rcs, a.pk, angt, sa, r);
        }
        if (WIN) wwin_fold<TRI>(win, on, r, sa, a.pk, acc, wl, grl);
        else fold_gates(on, r, sa, a.pk, acc);
        if (moved) rain_drag<true>(a, p, v, TI ? lerp4(lerp3d(g), lerp3d(g2), a.wind.alpha) : lerp3d(g));
        bst4(rv, off, v);
    }
    if (WIN) wwin_flush(win, a.pk, acc);
}

// The fold strategy is a template flag so that the loop carrying the window
// registers has no second fold path (which makes the compiler copy the 32
// window registers at every step).
template <bool ADV, bool PULSE, bool HILO, bool TI = false, bool FAST = false>
__device__ __forceinline__ void rain_range(const RainK &a, size_t lo, size_t hi, double *acc, const float4 *itab,
                                           bool use_win) {
    if constexpr (FAST && ADV && PULSE && !HILO) {
        if (use_win && a.pk.tri) rain_range_buf<true, true, TI>(a, lo, hi, acc, itab);
        else if (use_win) rain_range_buf<true, false, TI>(a, lo, hi, acc, itab);
        else rain_range_buf<false, false, TI>(a, lo, hi, acc, itab);
    } else {
        if (use_win && a.pk.tri) rain_range_t<ADV, PULSE, HILO, true, true, TI, FAST>(a, lo, hi, acc, itab);
        else if (use_win) rain_range_t<ADV, PULSE, HILO, true, false, TI, FAST>(a, lo, hi, acc, itab);
        else rain_range_t<ADV, PULSE, HILO, false, false, TI, FAST>(a, lo, hi, acc, itab);
    }
}

// Many independent rain domains in one persistent launch.  items[t] = (domain,
// chunk): chunk c of domain d covers scatterers [c chunk, (c + 1) chunk) and owns
// partial row c of that domain.  Block b takes items b, b + G, b + 2G, ...
// The descriptors are read through a restrict pointer, so the compiler proves
// that no store of the loop clobbers them and reads them with scalar loads, as
// it reads kernel arguments (without it each field is a vector load and a
// readfirstlane, and the batch ran at 0.84x the per-domain launches).  Keeping
// each domain on one XCD (its items walked by that XCD's blocks only) measured
// 1% slower than this plain round-robin, and equal shares of one flat range
// over all domains slower again.  FAST: every domain is a launch the
// specialised rain loop covers (rain_fast).
template <bool ADV, bool PULSE, bool FAST = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_rain_batch(
    const RainK *__restrict__ doms, const int2 *__restrict__ items, int n_items, int chunk) {
    extern __shared__ double acc[];
    const unsigned G = gridDim.x, b = blockIdx.x;
    for (int t = (int)b; t < n_items; t += (int)G) {
        const int2 it = items[t];
        const RainK &a = doms[it.x];
        if (PULSE) acc_clear(acc, 4 * a.pk.R);
        const bool use_win = PULSE && a.pk.fold == RSD_FOLD_WINDOW;
        if (PULSE) stage_rain_tables(acc, a);
        if (use_win) stage_window(acc, a.pk);
        const size_t lo = (size_t)it.y * (size_t)chunk;
        const size_t hi = lo + (size_t)chunk < a.n ? lo + (size_t)chunk : a.n;    // a.n = 0: the domain sits this round out
        rain_range<ADV, PULSE, false, false, FAST>(a, lo, hi > lo ? hi : lo, acc, a.widx.t, use_win);
        if (PULSE) acc_flush_row(acc, a.pk, it.y);
        __syncthreads();
    }
}

// Finalize of a batch: block j belongs to the domain d with first[d] <= j <
// first[d + 1] and sums that domain's partial rows for output j - first[d].
// Binary64 sum of column j (rows values) by one 256-thread block, in a fixed
// order: per thread r = t, t + 256, ... (loads issued 8 at a time), then a DPP
// wave sum, then the four wave sums in order.
__device__ __forceinline__ double column_sum(const double *col, int rows) {
    __shared__ double red[4];
    double s = 0.0;
    int r = threadIdx.x;
    for (; r + 7 * 256 < rows; r += 8 * 256) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = col[r + u * 256];
#pragma unroll
        for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; r < rows; r += 256) s += col[r];
    s = wave_sum64(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

// Column j of a pulse's partial sums into its output slot(s): (Ih, Qh, Iv, Qv)
// of gate j % R, component j / R.
__device__ __forceinline__ void finalize_column(const FinK &a, int j) {
    if (!a.out) return;     // a batched domain launched on its own this round
    const int R = a.n4r / 4;
    const double s = column_sum(a.partials + (size_t)j * a.rows, a.rows);
    if (threadIdx.x == 0) {
        a.out[4 * (j % R) + j / R] = s;
        if (a.out2) a.out2[4 * (j % R) + j / R] = s;
    }
}

__global__ __launch_bounds__(256) void k_finalize_batch(const FinK *__restrict__ f, const int *__restrict__ first, int nd) {
    int d = 0;
    while (d + 1 < nd && first[d + 1] <= (int)blockIdx.x) d++;
    finalize_column(f[d], (int)blockIdx.x - first[d]);
}

// ---------------------------------------------------------------------------
// Debris (types >= 1): db_atts (rs.cl:744-907; respawns placed uniformly in
// the box, or through the flux field of the step's wind slot under concept C)
// and db_rcs (rs.cl:909-919), fused with the pulse.
// ---------------------------------------------------------------------------
struct DebrisK {
    float4 *pos, *ori, *vel, *tum, *rcs;
    uint4 *rnd;
    unsigned long long begin, end;
    Tex3 wind;
    F16 les;
    F16 sd;
    Tex2 adm_cd, adm_cm;
    F16 adm_desc;
    Tex2 rcs_re, rcs_im;    // table of the advance
    F16 rcs_desc;
    Tex2 prcs_re, prcs_im;  // table of the pulse's refresh
    F16 prcs_desc;
    const float *dff;       // inverse CDF of the debris flux field, or null
    F16 dff_desc;
    uint32_t refresh;
    uint32_t pulse_concept;
    uint32_t reuse_rcs;     // advance and refresh use the same table and concept
    WIdx widx;              // stretched-grid index tables of the wind slot
    MlpDev net;             // surrogate network of this type (k_debris<.., true>)
    PulseK pk;
};

// Respawn position under the debris flux field (concept C, rs.cl:795-836).
// r.x goes through the inverse CDF of the cell flux (dd s15 = entries - 1) to
// a fractional cell index c; the row is floor(c / row length) (dd s13), the
// column the rest of c, and r.y spreads the body over the row.  The cell is
// then placed on the stretched grid (dd s14 = 1: size a (1 - q^|i|) / (1 - q)
// from the centre, dd s0-1 = a / (1 - q), s4-5 = q, s8-9 = last cell index),
// stretched over the box (dd s14 = 2: dd s0-1 = 1 / cells) or on the uniform
// grid (dd s0-1 = spacing); the height is MIN_HEIGHT above 0.
__device__ __forceinline__ void dff_place(const float *icdf, const F16 &dd, const F16 &sd, float4 r, float4 &pos) {
    const float sf = dd.s[15];
    float i0, i1;
    const float f0 = rf_clamp(rf_fma(r.x, sf, 0.0f), 0.0f, sf);
    const float f1 = rf_clamp(rf_fma(r.x, sf, 1.0f), 0.0f, sf);
    const float dec = rf_fract(f0, &i0);
    rf_fract(f1, &i1);
    const float c = rf_mix(icdf[(uint32_t)i0], icdf[(uint32_t)i1], dec);
    float y = __builtin_floorf(rf_div(c, dd.s[13]));
    float x = rf_fma(y, -dd.s[13], c);
    y = y + r.y;
    if (dd.s[14] == 1.0f) {
        const float rx = x - rf_fma(0.5f, dd.s[8], 0.5f);
        const float ry = y - rf_fma(0.5f, dd.s[9], 0.5f);
        x = rf_fma(rf_powf_pos(dd.s[4], __builtin_fabsf(rx)), -dd.s[0], dd.s[0]);
        y = rf_fma(rf_powf_pos(dd.s[5], __builtin_fabsf(ry)), -dd.s[1], dd.s[1]);
        x = __builtin_copysignf(x, rx) + rf_fma(0.5f, sd.s[12], sd.s[8]);
        y = __builtin_copysignf(y, ry) + rf_fma(0.5f, sd.s[13], sd.s[9]);
    } else if (dd.s[14] == 2.0f) {
        x = rf_fma(x, sd.s[12] * dd.s[0], sd.s[8]);
        y = rf_fma(y, sd.s[13] * dd.s[1], sd.s[9]);
    } else {
        x = rf_fma(x, dd.s[0], sd.s[8]);
        y = rf_fma(y, dd.s[1], sd.s[9]);
    }
    pos.x = x;
    pos.y = y;
    pos.z = RSD_MIN_HEIGHT;
}

// Position / orientation update and respawn of db_atts (rs.cl:775-873);
// returns true on respawn (state reset, rcs 0).
__device__ __forceinline__ bool debris_move(const DebrisK &a, size_t i, float4 &pos, float4 &ori, float4 &vel, float4 &tum,
                                            float4 &rcs) {
    const F16 &sd = a.sd;
    const float dt = sd.s[11];
    ori = norm4(qmul(ori, tum));
    pos.x = pos.x + vel.x * dt;
    pos.y = pos.y + vel.y * dt;
    pos.z = pos.z + vel.z * dt;
    pos.w = pos.w + vel.w * 0.0f;
    if (is_outside(pos, sd)) {
        uint4 seed = a.rnd[i];
        float4 r = rand4(seed);
        if ((__float_as_uint(sd.s[5]) & 128u) && a.dff) {
            dff_place(a.dff, a.dff_desc, sd, r, pos);
        } else {
            pos.x = rf_fma(r.x, sd.s[12], sd.s[8]);
            pos.y = rf_fma(r.y, sd.s[13], sd.s[9]);
            pos.z = sd.s[10] + RSD_MIN_HEIGHT;
        }
        r = rand4(seed);
        float c0 = rf_sqrt(-2.0f * rf_logf(r.x));
        float c1 = rf_sqrt(-2.0f * rf_logf(r.y));
        float c2 = rf_sqrt(-2.0f * rf_logf(r.z));
        float c3 = r.w;
        r = rand4(seed);
        const float two_pi = 2.0f * RSD_PI_F;
        c0 *= rf_cosf(two_pi * r.x);
        c1 *= rf_cosf(two_pi * r.y);
        c2 *= rf_cosf(two_pi * r.z);
        float sin_th_2, cos_th_2;
        rf_sincosf(RSD_PI_F * c3, &sin_th_2, &cos_th_2);
        float sqxy = rf_sqrt(c0 * c0 + c2 * c2);
        float sqxyz = rf_sqrt((c0 * c0 + c1 * c1) + c2 * c2);
        float sa, ca;
        rf_sincosf(0.5f * rf_acos_unit(rf_div(c1, sqxyz)), &sa, &ca);
        ori.x = (rf_div((c0 * ca) * sin_th_2, sqxyz) - rf_div(rf_div((c0 * c1) * sin_th_2, sqxyz) * sa, sqxy))
                + rf_div((c2 * cos_th_2) * sa, sqxy);
        ori.y = rf_div(sin_th_2, sqxyz) * (sqxy * sa + c1 * ca);
        ori.z = (rf_div((c2 * ca) * sin_th_2, sqxyz) - rf_div((c0 * cos_th_2) * sa, sqxy))
                - rf_div(rf_div(((c1 * c2) * sin_th_2) * sa, sqxy), sqxyz);
        ori.w = cos_th_2 * ca;
        vel = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        tum = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        rcs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.rnd[i] = seed;
        return true;
    }
    return false;
}

// Velocity and tumble update of db_atts (rs.cl:236-261, 876-897) from the
// drag / moment coefficients cd (body frame) and cm.
__device__ __forceinline__ void debris_forces(const DebrisK &a, float4 ori, float4 &vel, float4 &tum, float4 bg, float4 ur,
                                              float4 cd, float4 cm) {
    const F16 &sd = a.sd;
    const float dt = sd.s[11];
    const uint32_t concept = __float_as_uint(sd.s[5]);
    const F16 &ad = a.adm_desc;
    const float Ta = ad.s[15];
    cd = qrot(cd, ori);
    float urn = (ur.x * ur.x + ur.y * ur.y) + ur.z * ur.z;
    float ta_urn = Ta * urn;
    float4 du = make_float4(ta_urn * cd.x + 0.0f, ta_urn * cd.y + 0.0f, ta_urn * cd.z + (-9.8f), ta_urn * cd.w + 0.0f);
    float dwx = rf_radians((ta_urn * ad.s[12]) * cm.x);
    float dwy = rf_radians((ta_urn * ad.s[13]) * cm.y);
    float dwz = rf_radians((ta_urn * ad.s[14]) * cm.z);
    // |v_xy + dv_xy| > 3 |v_bg,xy| (rs.cl:881) on squared norms
    float nvx = vel.x + du.x * dt, nvy = vel.y + du.y * dt;
    if ((concept & 8u) && nvx * nvx + nvy * nvy > 9.0f * (bg.x * bg.x + bg.y * bg.y)) {
        vel.x = bg.x;
        vel.y = bg.y;
        vel.z = vel.z + du.z * dt;
    } else {
        vel.x = vel.x + du.x * dt;
        vel.y = vel.y + du.y * dt;
        vel.z = vel.z + du.z * dt;
        vel.w = vel.w + du.w * 0.0f;
    }
    float sx, cx, sy, cy, sz, cz;
    rf_sincosf(dwx * dt, &sx, &cx);
    rf_sincosf(dwy * dt, &sy, &cy);
    rf_sincosf(dwz * dt, &sz, &cz);
    float4 tq = make_float4((cx * sy) * sz + (sx * cy) * cz, (cx * sy) * cz - (sx * cy) * sz,
                            (cx * cy) * sz + (sx * sy) * cz, (cx * cy) * cz - (sx * sy) * sz);
    tum = norm4(tq);
}

__device__ __forceinline__ bool debris_advance(const DebrisK &a, size_t i, float4 &pos, float4 &ori, float4 &vel, float4 &tum, float4 &rcs, const float4 *itab) {
    const F16 &sd = a.sd;
    const uint32_t concept = __float_as_uint(sd.s[5]);
    if (debris_move(a, i, pos, ori, vel, tum, rcs)) return true;
    float4 c = wind_index(pos, a.les, sd, itab, a.widx);
    float4 bg = sample3d(a.wind, c.x, c.y, c.z);
    float4 ur = make_float4(bg.x - vel.x, bg.y - vel.y, bg.z - vel.z, bg.w - vel.w);
    float4 uh = qrot(norm4(ur), qconj(ori));
    float beta = rf_acos_unit(uh.x);
    float alpha = rf_atan2f(uh.z, uh.y);
    if (alpha < 0.0f) {
        alpha = RSD_PI_F + alpha;
        beta = -beta;
    }
    const F16 &ad = a.adm_desc;
    float cu = rf_fma(beta, ad.s[0], ad.s[4]);
    float cv = rf_fma(alpha, ad.s[1], ad.s[5]);
    float4 cd = sample2d(a.adm_cd, cu, cv);
    float4 cm = sample2d(a.adm_cm, cu, cv);
    debris_forces(a, ori, vel, tum, bg, ur, cd, cm);
    rcs = debris_rcs(pos, ori, a.rcs_re, a.rcs_im, a.rcs_desc, concept);
    return false;
}

// SURR: the surrogate network replaces the ADM and RCS table lookups; every
// wave evaluates it for its 64 lanes at once (BF16 MFMA, weights in LDS).
// Debris scatterers [lo, hi) of one type: advance and/or pulse terms into the
// block's LDS gate sums (cleared, window tables staged by the caller).  SURR:
// the surrogate network in LDS (nw, nb) with per-wave scratch nscr.
template <bool ADV, bool PULSE, bool SURR>
__device__ __forceinline__ void debris_range(const DebrisK &a, size_t lo, size_t hi, double *acc, const float4 *itab,
                                             const uint16_t *nw, const float *nb, float *nscr) {
    const bool use_win = PULSE && a.pk.fold == RSD_FOLD_WINDOW;
    const float2 *wl = win_weights(acc, a.pk);
    const float *grl = win_gates(acc, a.pk);
    for (size_t base = lo; base < hi; base += blockDim.x) {
        const size_t i = base + threadIdx.x;
        const bool live = i < hi;
        float4 p = make_float4(0.0f, 1.0f, 0.0f, 0.0f), o = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        if (live) {
            p = a.pos[i];
            o = a.ori[i];
        }
        float4 rcs;
        bool rcs_known = false, respawned = false;
        if (SURR) {
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f), t = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            float4 bg = v, ur = v;
            if (ADV && live) {
                v = a.vel[i];
                t = a.tum[i];
                respawned = debris_move(a, i, p, o, v, t, rcs);
                if (!respawned) {
                    float4 c = wind_index(p, a.les, a.sd, itab, a.widx);
                    bg = sample3d(a.wind, c.x, c.y, c.z);
                    ur = make_float4(bg.x - v.x, bg.y - v.y, bg.z - v.z, bg.w - v.w);
                }
            }
            const bool need_adv = ADV && live && !respawned;
            const bool need_ref = PULSE && live && a.refresh && !(need_adv && a.reuse_rcs);
            float f[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (need_adv || need_ref) {
                float4 R = radar_rel_quat(p, o);
                float4 uh = need_adv ? qrot(norm4(ur), qconj(o)) : make_float4(1.0f, 0.0f, 0.0f, 0.0f);
                f[0] = R.x; f[1] = R.y; f[2] = R.z; f[3] = R.w;
                f[4] = uh.x; f[5] = uh.y; f[6] = uh.z;
                // still air relative to the body (u_r = 0) has no flow direction:
                // the table path then reads texel 0, the network gets u = 0
#pragma unroll
                for (int k = 4; k < 7; k++) f[k] = f[k] == f[k] ? f[k] : 0.0f;
            }
            float out[MLP_OUT];
            mlp_wave_eval<MlpPadded, RSD_MLP_TILES, RSD_MLP_XR>(nw, nb, f, nscr + (threadIdx.x >> 6) * 64 * RSD_MLP_SCRW, out);
            if (ADV && live) {
                if (need_adv) {
                    debris_forces(a, o, v, t, bg, ur, make_float4(out[0], out[1], out[2], 0.0f),
                                  make_float4(out[3], out[4], out[5], 0.0f));
                    rcs = rcs_from_amps(out, __float_as_uint(a.sd.s[5]));
                }
                a.pos[i] = p;
                a.ori[i] = o;
                a.vel[i] = v;
                a.tum[i] = t;
                rcs_known = true;
            }
            if (PULSE && live) {
                if (need_ref) {
                    rcs = rcs_from_amps(out, a.pulse_concept);
                    rcs_known = true;
                } else if (!rcs_known) {
                    rcs = a.rcs[i];
                }
            }
        } else {
        if (ADV && live) {
            float4 v = a.vel[i], t = a.tum[i];
            respawned = debris_advance(a, i, p, o, v, t, rcs, itab);
            a.pos[i] = p;
            a.ori[i] = o;
            a.vel[i] = v;
            a.tum[i] = t;
            rcs_known = true;
        }
        if (PULSE && live) {
            if (a.refresh && ADV && !respawned && a.reuse_rcs) {
                // the advance just evaluated the same table at the same pose
            } else if (a.refresh) {
                rcs = debris_rcs(p, o, a.prcs_re, a.prcs_im, a.prcs_desc, a.pulse_concept);
                rcs_known = true;
            } else if (!rcs_known) {
                rcs = a.rcs[i];
            }
        }
        }
        if (rcs_known && live) a.rcs[i] = rcs;
        if (PULSE) {
            float4 sa = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float r = 0.0f;
            bool on = live && scat_signal(p, rcs, a.pk, a.pk.angw, sa, r);
            if (use_win) lane_fold(on, r, sa, a.pk, acc, wl, grl);
            else fold_gates(on, r, sa, a.pk, acc);
        }
    }
}

// Debris types of one launch: blocks [first[j], first[j + 1]) take type j.
#define RSD_STEP_DEBRIS 3
struct StepDebris {
    int nt;                             // debris types in this launch
    int first[RSD_STEP_DEBRIS + 1];     // first block of each type; first[nt]: the rest
    DebrisK t[RSD_STEP_DEBRIS];
};

__device__ __forceinline__ int step_type(const StepDebris &d) {
    int j = 0;
    while (j + 1 < d.nt && (int)blockIdx.x >= d.first[j + 1]) j++;
    return j;
}

// Debris launch of one step on the side stream (surrogate-driven types, or
// table types that do not ride in the rain launch): up to RSD_STEP_DEBRIS
// types in one grid, so their blocks share the device instead of queueing
// one type after another.  Three waves per SIMD: the surrogate variant fits
// 168 VGPRs without spilling (two waves at its natural 244; 4,421 against
// 3,887 pulses/s at C5).  Four (128 VGPRs, LDS under 40 KB with the outputs
// exchanged in three rounds) measured no faster.
template <bool ADV, bool PULSE, bool SURR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SURR ? RSD_DEBRIS_SURR_WAVES : 3))) void k_debris(StepDebris d) {
    extern __shared__ double acc[];
    __shared__ __attribute__((aligned(16))) uint16_t nw[SURR ? MLP_PWN : 2];
    __shared__ __attribute__((aligned(16))) float nb[SURR ? MLP_BN : 4];
    __shared__ __attribute__((aligned(16))) float nscr[SURR ? 4 * 64 * RSD_MLP_SCRW : 4];
    const int j = step_type(d);
    const DebrisK &a = d.t[j];
    if (SURR) {
        mlp_stage_padded(nw, a.net.w);
        for (int q = threadIdx.x; q < MLP_BN; q += blockDim.x) nb[q] = a.net.b[q];
    }
    if (PULSE) acc_clear(acc, 4 * a.pk.R);
    else __syncthreads();
    size_t lo, hi;
    const unsigned L = block_range_at(a.begin, a.end, blockIdx.x - (unsigned)d.first[j],
                                      (unsigned)(d.first[j + 1] - d.first[j]), lo, hi);
    if (PULSE && a.pk.fold == RSD_FOLD_WINDOW) stage_window(acc, a.pk);
    // index tables read from global memory (L1 / L2): this kernel's LDS holds
    // the network
    debris_range<ADV, PULSE, SURR>(a, lo, hi, acc, a.widx.t, nw, nb, nscr);
    if (PULSE) acc_flush_row(acc, a.pk, a.pk.row0 + (int)L);
}

// Rain launch of one step.  Debris types driven by the tables ride in the same
// launch (no second stream, no cross-stream events per pulse): blocks
// [first[j], first[j + 1]) take debris type j, the rain takes the blocks after
// first[nt] (a multiple of 8, so the rain blocks' partial rows keep their XCD
// order).

#ifndef RSD_RAIN_WAVES
#define RSD_RAIN_WAVES 4    // waves per SIMD the fused rain kernel is compiled for
#endif
// The time-interpolated variants (TI: eight more texels of the next volume in
// flight) take three waves per SIMD: at four they spill 44 VGPRs to scratch.
template <bool ADV, bool PULSE, bool HILO, bool TI = false, bool FAST = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TI ? 3 : RSD_RAIN_WAVES))) void k_rain(RainK a, StepDebris d) {
    extern __shared__ double acc[];
    __shared__ float4 itab[RSD_ITAB_ROWS];
    const unsigned nd = (unsigned)d.first[d.nt];
    const unsigned nr = gridDim.x - a.fin_blocks;      // end of the rain blocks
    if (blockIdx.x >= nr) {     // the previous pulse's finalize, in the slots the last rain blocks free
        const int j = (int)(blockIdx.x - nr);
        if (j < a.fin.n4r) finalize_column(a.fin, j);
        return;
    }
    if (ADV) itab_stage(itab, a.widx.t);     // riding debris share the rain's wind slot
    if (blockIdx.x < nd) {
        const int j = step_type(d);
        const DebrisK &da = d.t[j];
        if (PULSE) acc_clear(acc, 4 * da.pk.R);
        else __syncthreads();
        if (PULSE && da.pk.fold == RSD_FOLD_WINDOW) stage_window(acc, da.pk);
        size_t lo, hi;
        const unsigned L = block_range_at(da.begin, da.end, blockIdx.x - (unsigned)d.first[j],
                                          (unsigned)(d.first[j + 1] - d.first[j]), lo, hi);
        debris_range<ADV, PULSE, false>(da, lo, hi, acc, itab, nullptr, nullptr, nullptr);
        if (PULSE) acc_flush_row(acc, da.pk, da.pk.row0 + (int)L);
        return;
    }
    if (PULSE) acc_clear(acc, 4 * a.pk.R);
    else __syncthreads();
    size_t lo, hi;
    const unsigned L = block_range_at(0, a.n, blockIdx.x - nd, nr - nd, lo, hi, true);
    const bool use_win = PULSE && a.pk.fold == RSD_FOLD_WINDOW;
    if (PULSE) stage_rain_tables(acc, a);
    if (use_win) stage_window(acc, a.pk);
    rain_range<ADV, PULSE, HILO, TI, FAST>(a, lo, hi, acc, itab, use_win);
    if (PULSE) acc_flush_row(acc, a.pk, a.pk.row0 + (int)L);
}

// ---------------------------------------------------------------------------
// Fixed-position background (concept F): fp_atts (rs.cl:590-648) fused with the
// pulse.  Positions never move; each advance reads the wind and the Cn^2
// volume at the position, sets vel = wind, turns the stored phase by the
// radial wind's Doppler over dt and writes rcs = (a cos phi, a sin phi, cn2,
// phi) with a = 10^cn2 (the V half of rcs carries cn2 and the phase, as the
// reference's does).  The population is built gate by gate, so the scatterers
// of a wave share a gate window like range-sorted rain.
// ---------------------------------------------------------------------------
struct FixedK {
    const float4 *pos;
    float4 *vel, *rcs;
    unsigned long long n;
    Tex3 wind, cpxx;
    F16 les;
    F16 sd;                 // sim_desc of the advance
    WIdx widx;
    PulseK pk;
};

__device__ __forceinline__ void fixed_step(const F16 &sd, float4 p, float4 uvwt, float4 cp, float4 &rcs) {
    const float wav_num = sd.s[4], dt = sd.s[11];
    const float len = rf_len3(p.x, p.y, p.z);
    const float nx = rf_div(p.x, len), ny = rf_div(p.y, len), nz = rf_div(p.z, len);
    const float d = (nx * uvwt.x + ny * uvwt.y) + nz * uvwt.z;
    const float phi = rcs.w - (wav_num * d) * dt;
    const float cn2 = cp.x;
    const float a = rf_powf_pos(10.0f, cn2);
    float sn, cs;
    rf_sincosf(phi, &sn, &cs);
    rcs = make_float4(a * cs, a * sn, cn2, rf_atan2f(sn, cs));
}

template <bool ADV, bool PULSE, bool WIN, bool TRI>
__device__ __forceinline__ void fixed_range(const FixedK &a, size_t lo, size_t hi, double *acc, const float4 *itab) {
    const float2 *wl = win_weights(acc, a.pk);
    const float *grl = win_gates(acc, a.pk);
    WaveWin win;
    win.kb = RSD_WIN_EMPTY;
    wwin_clear(win);
    for (size_t base = lo; base < hi; base += blockDim.x) {
        const size_t i = base + threadIdx.x;
        const bool live = i < hi;
        float4 p = make_float4(0.0f, 1.0f, 0.0f, 0.0f), rcs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (live) {
            p = a.pos[i];
            rcs = a.rcs[i];
        }
        if (ADV && live) {
            const float4 c = wind_index(p, a.les, a.sd, itab, a.widx);
            const float4 uvwt = sample3d(a.wind, c.x, c.y, c.z);
            const float4 cp = sample3d(a.cpxx, c.x, c.y, c.z);
            fixed_step(a.sd, p, uvwt, cp, rcs);
            float4 v = a.vel[i];
            v.x = uvwt.x;
            v.y = uvwt.y;
            v.z = uvwt.z;
            a.vel[i] = v;
            a.rcs[i] = rcs;
        }
        if (PULSE) {
            float4 sa = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float r = 0.0f;
            bool on = live && scat_signal(p, rcs, a.pk, a.pk.angw, sa, r);
            if (WIN) wwin_fold<TRI>(win, on, r, sa, a.pk, acc, wl, grl);
            else fold_gates(on, r, sa, a.pk, acc);
        }
    }
    if (WIN) wwin_flush(win, a.pk, acc);
}

template <bool ADV, bool PULSE>
__global__ __launch_bounds__(256) void k_fixed(FixedK a) {
    extern __shared__ double acc[];
    __shared__ float4 itab[RSD_ITAB_ROWS];
    if (ADV) itab_stage(itab, a.widx.t);
    if (PULSE) acc_clear(acc, 4 * a.pk.R);
    else __syncthreads();
    size_t lo, hi;
    const unsigned L = block_range_at(0, a.n, blockIdx.x, gridDim.x, lo, hi, true);
    const bool use_win = PULSE && a.pk.fold == RSD_FOLD_WINDOW;
    if (use_win) stage_window(acc, a.pk);
    if (use_win && a.pk.tri) fixed_range<ADV, PULSE, true, true>(a, lo, hi, acc, itab);
    else if (use_win) fixed_range<ADV, PULSE, true, false>(a, lo, hi, acc, itab);
    else fixed_range<ADV, PULSE, false, false>(a, lo, hi, acc, itab);
    if (PULSE) acc_flush_row(acc, a.pk, a.pk.row0 + (int)L);
}

// Sum the per-block partial rows of one pulse: one block per output value,
// reading its contiguous column.  Columns are [component][gate]; the pulse
// buffer is [gate][component].
__global__ __launch_bounds__(256) void k_finalize(const double *partials, int rows, int n4r, double *out, double *out2) {
    FinK a;
    a.partials = partials;
    a.out = out;
    a.out2 = out2;
    a.rows = rows;
    a.n4r = n4r;
    finalize_column(a, (int)blockIdx.x);
}

// Pulse-pair moments of one gate over the last M pulses of the IQ ring
// (slots oldest .. newest): lag-0 powers, lag-1 autocorrelation of H and the
// H-V cross-correlation, all in binary64.
struct MomK {
    const double *ring;     // [ring_m][4 R], gate-major (Ih, Qh, Iv, Qv)
    int ring_m, first, m, R;
    double gain2;           // radar gain squared (the IQ are stored before the gain)
    double vfac;            // -lambda / (4 pi PRT)
    double range_start, range_delta;
    float4 *out;            // (Z dB, V m/s, ZDR dB, rho_hv) per gate
};

__global__ __launch_bounds__(256) void k_moments(MomK a) {
    __shared__ double red[6][256];
    const int g = blockIdx.x;
    double p_h = 0.0, p_v = 0.0, r1_re = 0.0, r1_im = 0.0, c_re = 0.0, c_im = 0.0;
    for (int k = threadIdx.x; k < a.m; k += blockDim.x) {
        const double *q = a.ring + (size_t)((a.first + k) % a.ring_m) * 4 * a.R + 4 * g;
        const double hi = q[0], hq = q[1], vi = q[2], vq = q[3];
        p_h += hi * hi + hq * hq;
        p_v += vi * vi + vq * vq;
        c_re += hi * vi + hq * vq;      // H conj(V)
        c_im += hq * vi - hi * vq;
        if (k + 1 < a.m) {
            const double *n = a.ring + (size_t)((a.first + k + 1) % a.ring_m) * 4 * a.R + 4 * g;
            r1_re += hi * n[0] + hq * n[1];  // conj(H_k) H_k+1
            r1_im += hi * n[1] - hq * n[0];
        }
    }
    red[0][threadIdx.x] = p_h; red[1][threadIdx.x] = p_v; red[2][threadIdx.x] = r1_re;
    red[3][threadIdx.x] = r1_im; red[4][threadIdx.x] = c_re; red[5][threadIdx.x] = c_im;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 6; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = (double)a.m;
        const double r0h = red[0][0] / m * a.gain2, r0v = red[1][0] / m * a.gain2;
        const double cxr = red[4][0] / m * a.gain2, cxi = red[5][0] / m * a.gain2;
        const double rg = a.range_start + (double)g * a.range_delta;
        const double z = 10.0 * log10(r0h) + 20.0 * log10(fabs(rg));
        const double v = a.vfac * atan2(red[3][0], red[2][0]);
        const double zdr = 10.0 * log10(r0h / r0v);
        const double rho = sqrt(cxr * cxr + cxi * cxi) / sqrt(r0h * r0v);
        a.out[g] = make_float4((float)z, (float)v, (float)zdr, (float)rho);
    }
}

// Bandwidth probe: copy of n float4 (the reference's io kernel, rs.cl:374-379).
// Each block copies 4 x 256 consecutive float4, all four reads issued before
// the writes.
__global__ __launch_bounds__(256) void k_io(const float4 *__restrict__ in, float4 *__restrict__ out, unsigned long long n) {
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t i = base + (size_t)k * 256;
        v[k] = i < n ? in[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const size_t i = base + (size_t)k * 256;
        if (i < n) out[i] = v[k];
    }
}

// Device evaluation of the binary32 functions the kernels use (self-test).
__global__ void k_math_probe(int fn, const float *a, const float *b, float *out, unsigned long long n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = a[i], y = b[i], s, c;
    float r = 0.0f;
    switch (fn) {
        case 10: r = rf_acos_fast(x); break;
        case 1: r = rf_logf(x); break;
        case 2: rf_sincosf(x, &s, &c); r = s; break;
        case 3: rf_sincosf(x, &s, &c); r = c; break;
        case 4: r = rf_atan2f(x, y); break;
        case 5: r = rf_acosf(x); break;
        case 6: r = rf_sqrt(x); break;
        case 7: r = rf_div(x, y); break;
        case 8: rf_sincos_turns(x, &s, &c); r = s; break;
        case 9: rf_sincos_turns(x, &s, &c); r = c; break;
        case 11: {  // transposed wave sum of a[j] = x + j y (n a multiple of 64)
            float v[32];
#pragma unroll
            for (int j = 0; j < 32; j++) v[j] = rf_fma((float)j, y, x);
            r = wave_transpose_sum(v);
            break;
        }
        default: break;
    }
    out[i] = r;
}

// ---------------------------------------------------------------------------
// Spatial order of the rain streams.
// ---------------------------------------------------------------------------
// Spread the low 10 bits of v to bit positions 0, 3, 6, ... 27.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t cell_of(float x, float o, float inv) {
    float c = __builtin_fminf(__builtin_fmaxf((x - o) * inv, 0.0f), 1023.0f);
    return (uint32_t)c;
}

struct SortK {
    const float4 *pos;
    uint32_t *key, *idx;
    unsigned long long n;
    float ox, oy, oz, ix, iy, iz;   // box origin, 16 / box size
    float r0, ir;                    // first shell range, shells per metre
    uint32_t shells;                 // number of range shells (<= 2^18)
};

// Key = range shell (high 18 bits) | Morton index of the 16^3 box cell (low 12
// bits), and the identity permutation.
__global__ void k_sort_keys(SortK a) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.n) return;
    float4 p = a.pos[j];
    float r = __builtin_sqrtf((p.x * p.x + p.y * p.y) + p.z * p.z);
    float sh = __builtin_fminf(__builtin_fmaxf((r - a.r0) * a.ir, 0.0f), (float)(a.shells - 1));
    uint32_t cx = cell_of(p.x, a.ox, a.ix) >> 6, cy = cell_of(p.y, a.oy, a.iy) >> 6, cz = cell_of(p.z, a.oz, a.iz) >> 6;
    a.key[j] = ((uint32_t)sh << 12) | spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2);
    a.idx[j] = (uint32_t)j;
}

struct PermK {
    const float4 *pos, *vel, *rcs, *plo;
    const uint4 *rnd;
    const uint32_t *id, *perm;
    float4 *pos2, *vel2, *rcs2, *plo2;
    uint4 *rnd2;
    uint32_t *id2;
    unsigned long long n, n_rain;
};

// Gather the permuted streams: stored j <- stored perm[j] for rain, identity
// for the debris range (copied so the buffers can be swapped whole).  A rain
// rcs entry is only meaningful while the scatterer is flagged (sign of pos.w).
__global__ void k_permute(PermK a) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.n) return;
    const bool rain = j < a.n_rain;
    const size_t s = rain ? a.perm[j] : j;
    const float4 p = a.pos[s];
    a.pos2[j] = p;
    a.vel2[j] = a.vel[s];
    if (!rain || __builtin_signbit(p.w)) a.rcs2[j] = a.rcs[s];
    a.rnd2[j] = a.rnd[s];
    if (a.plo) a.plo2[j] = a.plo[s];
    if (rain) a.id2[j] = a.id[s];
}

// Scatter the streams back to the reference order (stored j -> id[j]).
__global__ void k_unpermute(PermK a) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.n) return;
    const bool rain = j < a.n_rain;
    const size_t o = rain ? a.id[j] : j;
    const float4 p = a.pos[j];
    a.pos2[o] = p;
    a.vel2[o] = a.vel[j];
    if (!rain || __builtin_signbit(p.w)) a.rcs2[o] = a.rcs[j];
    a.rnd2[o] = a.rnd[j];
    if (a.plo) a.plo2[o] = a.plo[j];
    if (rain) a.id2[o] = (uint32_t)o;
}

// Low position parts in the reference order (hi/lo mode downloads).
__global__ void k_lo_view(const float4 *plo, const uint32_t *id, unsigned long long n_rain, unsigned long long n,
                          float4 *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (id && i < n_rain) ? id[i] : i;
    out[o] = plo[i];
}

__global__ void k_iota(uint32_t *id, unsigned long long n) {
    size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) id[j] = (uint32_t)j;
}

// Mark every rain scatterer's uploaded rcs as authoritative (sign bit of pos.w).
__global__ void k_flag_rain(float4 *pos, unsigned long long n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float4 p = pos[i];
        p.w = -__builtin_fabsf(p.w);
        pos[i] = p;
    }
}

// Download view of the streams: |pos.w|, reconstructed rain rcs, and the
// rs.cl:1050-1110 signal / auxiliary quantities for the last pulse's beam.
struct ViewK {
    const float4 *pos, *rcs, *aux_in, *vel;
    const uint4 *rnd;
    const uint32_t *id;     // reference index of each stored rain scatterer (NULL: identity)
    float4 *pos_out, *rcs_out, *sig_out, *aux_out, *vel_out;
    uint4 *rnd_out;
    unsigned long long n, n_rain;
    const float4 *drop;
    F4 drop_desc;
    PulseK pk;
    int have_pulse;
    int fixed;              // background of fixed positions: its rcs is the stream's
};

__global__ void k_view(ViewK a) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const size_t o = (a.id && i < a.n_rain) ? a.id[i] : i;
    float4 p = a.pos[i];
    bool stale = __builtin_signbit(p.w);
    float4 rcs;
    if (i < a.n_rain) {
        p.w = __builtin_fabsf(p.w);
        rcs = (stale || a.fixed) ? a.rcs[i] : ellipsoid_rcs(p, a.drop, a.drop_desc);
    } else {
        rcs = a.rcs[i];
    }
    if (a.pos_out) {
        a.pos_out[o] = p;
        a.rcs_out[o] = rcs;
    }
    if (a.vel_out) {
        a.vel_out[o] = a.vel[i];
        a.rnd_out[o] = a.rnd[i];
    }
    if (!a.sig_out) return;
    float4 aux = a.aux_in[o];    // aux stays in the reference order
    float4 sig = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (a.have_pulse) {
        const PulseK &k = a.pk;
        float r = rf_len3(p.x, p.y, p.z);
        const float inv_r = rf_rcp_fast(r);
        float wa = table_lerp(beam_angle(p, k.sd), k.angw, k.angw_desc);
        float atten = inv_r * inv_r;
        float ss, cc;
        if (k.phase_mode == RSD_PHASE_F64) {
            double turns = (double)r * k.two_over_lambda;
            rf_sincos_turns((float)(turns - floor(turns)), &ss, &cc);
        } else {
            rf_sincosf(r * k.sd.s[4], &ss, &cc);
        }
        sig = make_float4((rcs.x * cc - rcs.y * (-ss)) * atten, (rcs.y * cc + rcs.x * (-ss)) * atten,
                          (rcs.z * cc - rcs.w * (-ss)) * atten, (rcs.w * cc + rcs.z * (-ss)) * atten);
        aux.x = r;
        aux.y = aux.y + k.sd.s[15];
        aux.w = wa;
    }
    a.sig_out[o] = sig;
    a.aux_out[o] = aux;
}

// ---------------------------------------------------------------------------
// Scratch of rsd_make_pulse_noalloc: one array per GPU, held until the process
// ends.  Bytes [0, RSD_SCRATCH_GATE_BYTES) hold the binary32 gate ranges, the
// rest the per-block binary64 partial gate sums.  Host code uses the address
// hipGetSymbolAddress returns for the current device (scratch_dev), never the
// array's name.
// ---------------------------------------------------------------------------
#define RSD_SCRATCH_BYTES (16u << 20)
#define RSD_SCRATCH_GATE_BYTES (RSD_MAX_GATES * sizeof(float))
#define RSD_SCRATCH_PARTIALS ((RSD_SCRATCH_BYTES - RSD_SCRATCH_GATE_BYTES) / sizeof(double))
#define RSD_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned char scratch[RSD_SCRATCH_BYTES];
static void *scratch_dev[RSD_SCRATCH_DEVICES];
static pthread_mutex_t scratch_lock = PTHREAD_MUTEX_INITIALIZER;

// Padded copy of a 3-D table (see Tex3): texel (i, j, k) of the padded volume is
// texel (clamp(i - RSD_PAD_LO), clamp(j - RSD_PAD_LO), clamp(k - RSD_PAD_LO)) of
// the table.
// Texel pairs of two volumes of one padded shape: out[2 t] = a[t], out[2 t + 1] = b[t].
__global__ void k_wind_pair(const float4 *__restrict__ a, const float4 *__restrict__ b, float4 *__restrict__ out,
                            unsigned long long n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    out[2 * t] = a[t];
    out[2 * t + 1] = b[t];
}

__global__ void k_pad3d(const float4 *__restrict__ src, float4 *__restrict__ dst, int nx, int ny, int nz) {
    const uint32_t px = (uint32_t)nx + RSD_PAD_LO + RSD_PAD_HI, py = (uint32_t)ny + RSD_PAD_LO + RSD_PAD_HI;
    const uint32_t pz = (uint32_t)nz + RSD_PAD_LO + RSD_PAD_HI;
    const size_t n = (size_t)px * py * pz;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % px) - RSD_PAD_LO, j = (int)((q / px) % py) - RSD_PAD_LO, k = (int)(q / ((size_t)px * py)) - RSD_PAD_LO;
        const int ci = i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i);
        const int cj = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
        const int ck = k < 0 ? 0 : (k > nz - 1 ? nz - 1 : k);
        dst[q] = src[((size_t)ck * ny + cj) * nx + ci];
    }
}

// rs.cl:1189 / 1214: gate k's range as start plus k binary32 additions of delta.
__global__ void k_gate_ranges(float start, float delta, int R, float *out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float r = start;
        for (int k = 0; k < R; k++) {
            out[k] = r;
            r += delta;
        }
    }
}

// Where launch_step takes its scratch and output when no allocation is allowed.
struct StaticPulse {
    double *partials;       // RSD_SCRATCH_PARTIALS doubles
    float *gate_r;          // RSD_MAX_GATES floats
    double *out;            // 4 R doubles of the caller
};

// ---------------------------------------------------------------------------
// Host side of the C-ABI.
// ---------------------------------------------------------------------------
struct rsd_engine {
    int device;
    hipStream_t stream;
    hipStream_t side;       // debris kernels, forked from / joined into stream
    hipEvent_t ev_fork, ev_join;
    float4 *wind_stage[2];      // pinned staging of the stream-ordered wind uploads
    size_t wind_stage_n[2];     // texels of each staging buffer
    int wind_ok[2];             // slot holds an uploaded table
    hipEvent_t ev_wind[2];      // last copy out of each staging buffer
    char err[512];
    size_t n, n_rain;
    size_t origins[RSD_MAX_TYPES], counts[RSD_MAX_TYPES];
    float4 *pos, *vel, *ori, *tum, *rcs, *aux;
    uint4 *rnd;
    float4 *view;           // download staging: 4 x n float4 (pos, rcs, vel, rnd)
    // signal and aux streams of the last pulse (rs.cl:1050-1110 writes them at
    // the pulse): 2 x n float4, reference order, valid while sigaux_saved
    float4 *sigaux;
    int sigaux_saved;
    // spatial order (rain range only): ping-pong streams, reference index,
    // radix-sort keys / permutation and hipCUB scratch
    float4 *pos2, *vel2, *rcs2;
    uint4 *rnd2;
    uint32_t *id, *id2, *skey, *skey2, *sperm, *sperm2;
    int sort_bits;          // key bits of the current sort
    float4 *plo, *plo2;     // hi/lo positions: low parts (and ping-pong copy)
    int pos_hilo;           // 1: positions carried as hi/lo pairs
    // surrogate networks per debris type (k_debris<.., true>)
    MlpDev net[RSD_MAX_TYPES];
    int surr[RSD_MAX_TYPES];
    // batch membership: pulses are held back and launched by the batch
    struct rsd_batch *batch;
    int have_pending_pulse;
    rsd_pulse pending_pulse;
    int batch_wait;         // the engine stream must wait for the batch's last launch
    int batch_rejoin;       // launched on its own stream in the last round: the batch waits for it
    void *sort_tmp;
    size_t sort_tmp_bytes;
    int sort_interval;      // pulses between sorts (0 = never)
    int fold;               // RSD_FOLD_* strategy of the gate sums
    // IQ history for the moment estimator: ring of ring_m pulses of 4 R doubles
    double *ring;
    int ring_m, ring_R, ring_next, ring_count;
    float4 *mom_d;
    int since_sort;
    int permuted;           // stored rain order differs from the reference order
    float4 *wind[2];        // padded volumes (see Tex3)
    int wnx[2], wny[2], wnz[2];
    float4 *raw;            // an unpadded table on its way to a padded slot
    size_t raw_n;
    float4 *cpxx[2];        // (cn2, p, _, _) volume of each LES slot (concept F)
    int cnx[2], cny[2], cnz[2];
    int bg_fixed;           // the background advances as fp_atts (last step's kind)
    F16 les[2];
    float4 *adm_cd[RSD_MAX_TABLES], *adm_cm[RSD_MAX_TABLES];
    int adm_nx[RSD_MAX_TABLES], adm_ny[RSD_MAX_TABLES];
    F16 adm_desc[RSD_MAX_TABLES];
    float4 *rcs_re[RSD_MAX_TABLES], *rcs_im[RSD_MAX_TABLES];
    int rcs_nx[RSD_MAX_TABLES], rcs_ny[RSD_MAX_TABLES];
    F16 rcs_desc[RSD_MAX_TABLES];
    float4 *drop;
    F4 drop_desc;
    // Debris flux field: one inverse CDF per wind slot (the step's wind_slot
    // selects it, as the reference pairs dff_icdf[les_id] with the LES frame,
    // rs.c:3351-3373, 5358), one descriptor; a slot's later uploads go through
    // pinned staging in stream order.
    float *dff[2];
    int dff_n[2];
    float *dff_stage[2];
    hipEvent_t ev_dff[2];
    F16 dff_desc;
    float *rangew;
    float rangew_h[1024];
    int rangew_n;
    F4 rangew_desc;
    float *angw;
    int angw_n;
    F4 angw_desc;
    float *gate_r;
    int gate_R;
    float gate_start, gate_delta;
    double gate_drift;      // max |gate_r[k] - (start + k delta)| of the binary32 sums
    double *partials;
    size_t partial_cap;
    // Rain pulses alternate between partials and partials2 (part_cur: the one the
    // next writes), so that the finalize of pulse k, run by the first blocks of
    // pulse k + 1's launch, reads the other buffer.  fin_on: a rain pulse whose
    // finalize has not been launched yet (fin holds its arguments).
    double *partials2;
    size_t partial_cap2;
    int part_cur;
    int fin_on;
    FinK fin;
    double *pulse_d;         // device address of pulse_h: the finalize writes the
    double *pulse_h;         // pulse straight into pinned host memory
    int have_pending;
    rsd_step pending;
    int have_pulse;
    rsd_pulse last_pulse;
    int block, max_blocks;   // max_blocks 0: rain launches take one resident round
    int cus;                 // compute units of the device
    int rain_per_cu;         // resident fused rain blocks per compute unit (0: not known yet)
    int rain_occ_hilo;       // variant and LDS size rain_per_cu was computed for
    size_t rain_occ_lds;
    hipEvent_t ev0, ev1;
    // Stretched-grid index tables per wind slot, built by
    // rsd_wind_index_table for the slot's LES descriptor and the step's box
    // (rebuilt when either changes) and copied stream-ordered from pinned
    // staging.
    float4 *itab_d[2], *itab_h[2];
    // Time-interpolated wind: the volumes of slot wpair_slot and the other slot
    // as texel pairs (k_wind_pair), rebuilt when either slot's volume changes
    // (wgen counts each slot's uploads).
    float4 *wpair;
    size_t wpair_cap;
    int wpair_ok, wpair_slot;
    unsigned wgen[2], wpair_gen[2];
    hipEvent_t ev_itab[2];
    float itab_key[2][22];
    int itab_ok[2];
    WIdx itab_w[2];
    int prof_on, prof_n;
    int prof_stride, prof_tick;  // time every prof_stride-th rain launch
    hipEvent_t *prof_ev;     // start/stop pairs around the rain kernel
    int *prof_kind;          // bit 0 advance, bit 1 pulse
};

#define RSD_PROF_CAP 16384

#define RSD_TUNE_REPS 3

struct rsd_batch {
    int device, cus;
    hipStream_t stream;
    int nd;
    rsd_engine **eng;
    // descriptors: two pinned host buffers (filled alternately) and one device copy
    char *desc_h[2];
    char *desc_d;
    size_t desc_cap;
    int buf;
    hipEvent_t ev_buf[2];
    hipEvent_t ev_launch;
    // work items
    int2 *items_d;
    int items_cap, n_items, chunk, grid;
    size_t items_sig;
    int chunk_req, grid_req;       // 0 = autotune
    // graphs (one per descriptor buffer)
    int use_graph;
    hipGraphExec_t gexec[2];
    size_t gsig[2];
    // autotuner
    int tuning, tune_idx, tune_rep, tune_pending;
    double tune_sum, tune_best;
    int best_chunk, best_grid;
    hipEvent_t tev0, tev1;
    hipEvent_t t0, t1;
    char err[256];
};

static int flush_pending(rsd_engine *e);
static int flush_finalize(rsd_engine *e);
static void join_batch(rsd_engine *e);
extern "C" int rsd_batch_launch(struct rsd_batch *b);

static int fail(rsd_engine *e, int code, const char *fmt, ...) {
    if (e) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(e->err, sizeof(e->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(e, call)                                                                          \
    do {                                                                                         \
        hipError_t _r = (call);                                                                  \
        if (_r != hipSuccess) return fail((e), RSD_ERR_HIP, "%s: %s", #call, hipGetErrorString(_r)); \
    } while (0)

static void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// Chebyshev nodes of [0, 1]: 0.5 + 0.5 cos((2k + 1) pi / 8), k = 0..3.
static const double itab_cos1 = 0x1.d906bcf328d46p-1, itab_cos3 = 0x1.87de2a6aea964p-2;

// Pieces of one axis: P(x) = s0 log1p(s4 x) on x = (j + f) / ih, interpolated
// at the Chebyshev nodes u0..u3 of f in binary64 (Newton divided differences,
// expanded into monomials of f), coefficients rounded to binary32.
static void itab_axis(double s0, double s4, float ihf, float4 *rows) {
    const double u0 = 0.5 + 0.5 * itab_cos1, u1 = 0.5 + 0.5 * itab_cos3;
    const double u2 = 0.5 - 0.5 * itab_cos3, u3 = 0.5 - 0.5 * itab_cos1;
    const double ih = (double)ihf;
    for (int j = 0; j < RSD_ITAB_N; j++) {
        const double y0 = s0 * log1p(s4 * (((double)j + u0) / ih));
        const double y1 = s0 * log1p(s4 * (((double)j + u1) / ih));
        const double y2 = s0 * log1p(s4 * (((double)j + u2) / ih));
        const double y3 = s0 * log1p(s4 * (((double)j + u3) / ih));
        const double d01 = (y1 - y0) / (u1 - u0), d12 = (y2 - y1) / (u2 - u1), d23 = (y3 - y2) / (u3 - u2);
        const double d012 = (d12 - d01) / (u2 - u0), d123 = (d23 - d12) / (u3 - u1);
        const double A = (d123 - d012) / (u3 - u0), B = d012, C = d01;
        const double c3 = A;
        const double c2 = B - A * ((u0 + u1) + u2);
        const double c1 = (C - B * (u0 + u1)) + A * ((u0 * u1 + u0 * u2) + u1 * u2);
        const double c0 = ((y0 - C * u0) + B * (u0 * u1)) - A * ((u0 * u1) * u2);
        rows[j] = make_float4((float)c0, (float)c1, (float)c2, (float)c3);
    }
}

extern "C" int rsd_wind_index_table(const float *les, const float *sd, float *rows, float *ih, float *jm) {
    if (!les || !sd || !rows || !ih || !jm) return RSD_ERR_ARG;
    // origins as the kernels round them: the box centre in x and y, 0 in z
    const float o[3] = {sd[8] + 0.5f * sd[12], sd[9] + 0.5f * sd[13], 0.0f};
    for (int a = 0; a < 3; a++) {
        const float lo = sd[8 + a], hi = sd[8 + a] + sd[12 + a];
        double X = fmax(fabs((double)lo - (double)o[a]), fabs((double)hi - (double)o[a]));
        X = fmax(X * (1.0 + 0x1p-10), 1.0e-3);
        ih[a] = (float)((double)RSD_ITAB_N / X);
        jm[a] = (float)(RSD_ITAB_N - 1);
        itab_axis((double)les[a], (double)les[4 + a], ih[a], (float4 *)rows + a * RSD_ITAB_N);
    }
    return RSD_OK;
}

extern "C" int rsd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int rsd_device_props(int device_index, int *cus, int *max_wg, size_t *lds) {
    hipDeviceProp_t prop;
    if (device_index < 0 || device_index >= rsd_device_count()) return RSD_ERR_NO_DEVICE;
    if (hipGetDeviceProperties(&prop, device_index) != hipSuccess) return RSD_ERR_HIP;
    if (cus) *cus = prop.multiProcessorCount;
    if (max_wg) *max_wg = prop.maxThreadsPerBlock;
    if (lds) *lds = prop.sharedMemPerBlock;
    return RSD_OK;
}

extern "C" int rsd_open(int device_index, rsd_engine **out) {
    if (!out) return RSD_ERR_ARG;
    *out = NULL;
    int n = rsd_device_count();
    if (n <= 0 || device_index < 0 || device_index >= n) return RSD_ERR_NO_DEVICE;
    rsd_engine *e = (rsd_engine *)calloc(1, sizeof(rsd_engine));
    if (!e) return RSD_ERR_ALLOC;
    e->device = device_index;
    e->block = 256;
    e->max_blocks = 0;
    if (rsd_device_props(device_index, &e->cus, NULL, NULL) != RSD_OK || e->cus < 1) e->cus = 1;
    e->sort_interval = 64;
    e->fold = RSD_FOLD_WINDOW;
    if (hipSetDevice(device_index) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
        hipHostMalloc((void **)&e->pulse_h, 4 * RSD_MAX_GATES * sizeof(double), hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void **)&e->pulse_d, e->pulse_h, 0) != hipSuccess ||
        hipMalloc((void **)&e->itab_d[0], RSD_ITAB_ROWS * sizeof(float4)) != hipSuccess ||
        hipMalloc((void **)&e->itab_d[1], RSD_ITAB_ROWS * sizeof(float4)) != hipSuccess ||
        hipHostMalloc((void **)&e->itab_h[0], RSD_ITAB_ROWS * sizeof(float4), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&e->itab_h[1], RSD_ITAB_ROWS * sizeof(float4), hipHostMallocDefault) != hipSuccess ||
        hipMemsetAsync(e->itab_d[0], 0, RSD_ITAB_ROWS * sizeof(float4), e->stream) != hipSuccess ||
        hipMemsetAsync(e->itab_d[1], 0, RSD_ITAB_ROWS * sizeof(float4), e->stream) != hipSuccess) {
        free(e);
        return RSD_ERR_HIP;
    }
    // the device's scratch address for rsd_make_pulse_noalloc, once per device
    if (device_index < RSD_SCRATCH_DEVICES) {
        pthread_mutex_lock(&scratch_lock);
        hipError_t st = hipSuccess;
        if (!scratch_dev[device_index]) {
            void *ptr = NULL;
            st = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (st == hipSuccess) scratch_dev[device_index] = ptr;
        }
        pthread_mutex_unlock(&scratch_lock);
        if (st != hipSuccess) {
            free(e);
            return RSD_ERR_HIP;
        }
    }
    *out = e;
    return RSD_OK;
}

static void free_state(rsd_engine *e) {
    dfree(e->pos); dfree(e->vel); dfree(e->ori); dfree(e->tum); dfree(e->rcs); dfree(e->aux);
    dfree(e->rnd); dfree(e->view); dfree(e->sigaux);
    e->sigaux = NULL;
    e->sigaux_saved = 0;
    dfree(e->pos2); dfree(e->vel2); dfree(e->rcs2); dfree(e->rnd2);
    dfree(e->id); dfree(e->id2); dfree(e->skey); dfree(e->skey2); dfree(e->sperm); dfree(e->sperm2);
    dfree(e->sort_tmp);
    dfree(e->plo); dfree(e->plo2);
    e->plo = e->plo2 = NULL;
    e->pos = e->vel = e->ori = e->tum = e->rcs = e->aux = e->view = NULL;
    e->pos2 = e->vel2 = e->rcs2 = NULL;
    e->rnd = e->rnd2 = NULL;
    e->id = e->id2 = e->skey = e->skey2 = e->sperm = e->sperm2 = NULL;
    e->sort_tmp = NULL;
    e->sort_tmp_bytes = 0;
    e->permuted = 0;
    e->since_sort = 0;
}

extern "C" void rsd_close(rsd_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->batch) {   // leave the batch: launch what is recorded, then drop this engine
        rsd_batch *b = e->batch;
        if (e->have_pending_pulse) (void)rsd_batch_launch(b);
        (void)hipStreamSynchronize(b->stream);
        int k = 0;
        for (int i = 0; i < b->nd; i++)
            if (b->eng[i] != e) b->eng[k++] = b->eng[i];
        b->nd = k;
        b->items_sig = 0;
        e->batch = NULL;
    }
    (void)hipStreamSynchronize(e->stream);
    free_state(e);
    for (int s = 0; s < 2; s++) {
        dfree(e->wind[s]);
        dfree(e->cpxx[s]);
        if (s == 0) dfree(e->raw);
        if (e->wind_stage[s]) (void)hipHostFree(e->wind_stage[s]);
        if (e->dff_stage[s]) (void)hipHostFree(e->dff_stage[s]);
        if (e->ev_dff[s]) (void)hipEventDestroy(e->ev_dff[s]);
        if (e->ev_wind[s]) (void)hipEventDestroy(e->ev_wind[s]);
        if (e->ev_itab[s]) (void)hipEventDestroy(e->ev_itab[s]);
        dfree(e->itab_d[s]);
        if (e->itab_h[s]) (void)hipHostFree(e->itab_h[s]);
    }
    for (int t = 0; t < RSD_MAX_TABLES; t++) {
        dfree(e->adm_cd[t]); dfree(e->adm_cm[t]); dfree(e->rcs_re[t]); dfree(e->rcs_im[t]);
    }
    dfree(e->drop); dfree(e->dff[0]); dfree(e->dff[1]); dfree(e->rangew); dfree(e->angw); dfree(e->gate_r); dfree(e->partials); dfree(e->partials2);
    dfree(e->ring);
    dfree(e->wpair);
    dfree(e->mom_d);
    if (e->pulse_h) (void)hipHostFree(e->pulse_h);
    if (e->prof_ev) {
        for (int i = 0; i < 2 * RSD_PROF_CAP; i++) (void)hipEventDestroy(e->prof_ev[i]);
        free(e->prof_ev);
        free(e->prof_kind);
    }
    (void)hipEventDestroy(e->ev0);
    (void)hipEventDestroy(e->ev1);
    (void)hipEventDestroy(e->ev_fork);
    (void)hipEventDestroy(e->ev_join);
    (void)hipStreamSynchronize(e->side);
    (void)hipStreamDestroy(e->side);
    (void)hipStreamDestroy(e->stream);
    free(e);
}

extern "C" const char *rsd_last_error(const rsd_engine *e) { return e ? e->err : "no engine"; }

extern "C" int rsd_set_launch(rsd_engine *e, int block_size, int max_blocks) {
    if (!e) return RSD_ERR_ARG;
    if (block_size == 0) block_size = 256;
    if (block_size != 64 && block_size != 128 && block_size != 256) return fail(e, RSD_ERR_ARG, "block size %d not in {64,128,256}", block_size);
    if (max_blocks < 0 || max_blocks > 65536) return fail(e, RSD_ERR_ARG, "max_blocks %d out of range", max_blocks);
    e->block = block_size;
    e->max_blocks = max_blocks;
    return RSD_OK;
}

extern "C" int rsd_alloc_state(rsd_engine *e, size_t num_scats, const size_t *origins, const size_t *counts) {
    if (!e || !origins || !counts || num_scats == 0) return fail(e, RSD_ERR_ARG, "bad population");
    size_t total = 0;
    for (int k = 0; k < RSD_MAX_TYPES; k++) {
        if (counts[k] && origins[k] + counts[k] > num_scats) return fail(e, RSD_ERR_ARG, "type %d range beyond num_scats", k);
        total += counts[k];
    }
    if (total != num_scats || origins[0] != 0) return fail(e, RSD_ERR_ARG, "type counts (%zu) != num_scats (%zu)", total, num_scats);
    HIPCHK(e, hipSetDevice(e->device));
    {
        int st = flush_finalize(e);
        if (st) return st;
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    free_state(e);
    e->n = num_scats;
    e->n_rain = counts[0];
    memcpy(e->origins, origins, sizeof(e->origins));
    memcpy(e->counts, counts, sizeof(e->counts));
    const size_t b = num_scats * sizeof(float4);
    HIPCHK(e, hipMalloc((void **)&e->pos, b));
    HIPCHK(e, hipMalloc((void **)&e->vel, b));
    HIPCHK(e, hipMalloc((void **)&e->ori, b));
    HIPCHK(e, hipMalloc((void **)&e->tum, b));
    HIPCHK(e, hipMalloc((void **)&e->rcs, b));
    HIPCHK(e, hipMalloc((void **)&e->aux, b));
    HIPCHK(e, hipMalloc((void **)&e->rnd, num_scats * sizeof(uint4)));
    HIPCHK(e, hipMemsetAsync(e->rcs, 0, b, e->stream));
    HIPCHK(e, hipMemsetAsync(e->aux, 0, b, e->stream));
    if (e->pos_hilo) {
        HIPCHK(e, hipMalloc((void **)&e->plo, b));
        HIPCHK(e, hipMemsetAsync(e->plo, 0, b, e->stream));
    }
    e->have_pending = 0;
    e->have_pulse = 0;
    return RSD_OK;
}

#define RSD_DEFAULT_BLOCKS 2048

static int blocks_for(const rsd_engine *e, size_t n) {
    size_t b = (n + (size_t)e->block - 1) / (size_t)e->block;
    const size_t cap = e->max_blocks > 0 ? (size_t)e->max_blocks : RSD_DEFAULT_BLOCKS;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

// Blocks of the fused rain kernel that are resident on the device at once
// (occupancy of the launch's variant and LDS size, times the compute units).
// Launches of a multiple of that many equal blocks run in whole rounds (no
// block waits alone for a free slot at the end).
static int rain_resident_blocks(rsd_engine *e, bool hilo, bool fast, bool ti, size_t lds) {
    const int key = (int)hilo | ((int)fast << 1) | ((int)ti << 2);
    if (!e->rain_per_cu || e->rain_occ_hilo != key || e->rain_occ_lds != lds) {
        int nb = 0;
        const void *k;
        if (ti) k = hilo ? (fast ? reinterpret_cast<const void *>(&k_rain<true, true, true, true, true>)
                                 : reinterpret_cast<const void *>(&k_rain<true, true, true, true>))
                         : (fast ? reinterpret_cast<const void *>(&k_rain<true, true, false, true, true>)
                                 : reinterpret_cast<const void *>(&k_rain<true, true, false, true>));
        else k = hilo ? (fast ? reinterpret_cast<const void *>(&k_rain<true, true, true, false, true>)
                              : reinterpret_cast<const void *>(&k_rain<true, true, true>))
                      : (fast ? reinterpret_cast<const void *>(&k_rain<true, true, false, false, true>)
                              : reinterpret_cast<const void *>(&k_rain<true, true, false>));
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, e->block, lds) != hipSuccess || nb < 1) nb = 1;
        e->rain_per_cu = nb;
        e->rain_occ_hilo = key;
        e->rain_occ_lds = lds;
    }
    return e->rain_per_cu * e->cus;
}

// Partial-sum buffer `which` (0: partials, 1: partials2) for rows x 4 R values.
static int ensure_partials(rsd_engine *e, size_t rows, int R, int which = 0) {
    double **buf = which ? &e->partials2 : &e->partials;
    size_t *cap = which ? &e->partial_cap2 : &e->partial_cap;
    size_t need = rows * 4 * (size_t)R;
    if (need > *cap) {
        dfree(*buf);
        *buf = NULL;
        HIPCHK(e, hipMalloc((void **)buf, need * sizeof(double)));
        *cap = need;
    }
    return RSD_OK;
}

// Launch the finalize of the last rain pulse if its launch has not run yet
// (the pulse's output and IQ ring slot are complete once the stream reaches
// this point).
static int flush_finalize(rsd_engine *e) {
    if (!e->fin_on) return RSD_OK;
    e->fin_on = 0;
    hipLaunchKernelGGL(k_finalize, dim3(e->fin.n4r), dim3(256), 0, e->stream, e->fin.partials, e->fin.rows, e->fin.n4r,
                       e->fin.out, e->fin.out2);
    HIPCHK(e, hipGetLastError());
    return RSD_OK;
}

// Largest |start + k delta (binary32 sums) - (start + k delta)| over the gates.
static double gate_drift_of(float start, float delta, int R) {
    float r = start;
    double drift = 0.0;
    for (int k = 0; k < R; k++) {
        double d = fabs((double)r - ((double)start + (double)k * (double)delta));
        if (d > drift) drift = d;
        r += delta;
    }
    return drift;
}

static int ensure_gates(rsd_engine *e, float start, float delta, int R) {
    if (e->gate_r && e->gate_R == R && e->gate_start == start && e->gate_delta == delta) return RSD_OK;
    float *h = (float *)malloc((size_t)R * sizeof(float));
    if (!h) return fail(e, RSD_ERR_ALLOC, "gate table");
    float r = start;
    double drift = 0.0;
    for (int k = 0; k < R; k++) {  // rs.cl:1189 / 1214: repeated binary32 additions
        h[k] = r;
        double d = fabs((double)r - ((double)start + (double)k * (double)delta));
        if (d > drift) drift = d;
        r += delta;
    }
    e->gate_drift = drift;
    dfree(e->gate_r);
    e->gate_r = NULL;
    hipError_t st = hipMalloc((void **)&e->gate_r, (size_t)R * sizeof(float));
    // on the engine's (non-blocking) stream, so the kernels after it see it
    if (st == hipSuccess) st = hipMemcpyAsync(e->gate_r, h, (size_t)R * sizeof(float), hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    free(h);
    if (st != hipSuccess) return fail(e, RSD_ERR_HIP, "gate table: %s", hipGetErrorString(st));
    e->gate_R = R;
    e->gate_start = start;
    e->gate_delta = delta;
    return RSD_OK;
}

// Support of the range weight in metres: outside [lo, hi] every gate weight is 0
// (gates start + k delta, k < R, summed in binary32 with the given drift).
static void range_support(const rsd_engine *e, float start, float delta, int gates, double drift, float *lo, float *hi,
                          int *dense) {
    const int n = e->rangew_n;
    const float *W = e->rangew_h;
    const float xs = e->rangew_desc.s[0], x0 = e->rangew_desc.s[1];
    *dense = 0;
    if (n < 1 || W[0] != 0.0f || W[n - 1] != 0.0f || !(xs > 0.0f)) {
        *dense = 1;
        *lo = -INFINITY;
        *hi = INFINITY;
        return;
    }
    int f = -1, l = -1;
    for (int i = 0; i < n; i++) {
        if (W[i] != 0.0f) {
            if (f < 0) f = i;
            l = i;
        }
    }
    if (f < 0) {  // all-zero table: no gate is ever weighted
        *lo = 1.0f;
        *hi = -1.0f;
        return;
    }
    if (!(delta > 0.0f)) {     // the gate-index bounds below assume increasing ranges
        *dense = 1;
        *lo = -INFINITY;
        *hi = INFINITY;
        return;
    }
    // Margin for the drift of the summed gate ranges and the rounding of the
    // binary32 gate-index bounds computed from r in the kernels.
    const double R = (double)gates;
    const double pad = drift + 1e-3 * delta + 1e-6 * (fabs((double)start) + R * delta) + 1e-3;
    *lo = (float)(((double)(f - 1) - x0) / xs - pad);
    *hi = (float)(((double)(l + 1) - x0) / xs + pad);
}

// Pulse arguments; sp != NULL: gate ranges in the scratch (no allocation).
static int fill_pulse(rsd_engine *e, const rsd_pulse *p, PulseK &k, const StaticPulse *sp = NULL) {
    if (!e->angw || !e->rangew) return fail(e, RSD_ERR_STATE, "weight tables not set");
    double drift;
    if (sp) {
        drift = gate_drift_of(p->range_start, p->range_delta, (int)p->range_count);
    } else {
        int st = ensure_gates(e, p->range_start, p->range_delta, (int)p->range_count);
        if (st) return st;
        drift = e->gate_drift;
    }
    memcpy(k.sd.s, p->sim_desc, sizeof(k.sd.s));
    k.two_over_lambda = p->two_over_lambda;
    k.angw = e->angw;
    k.angw_desc = e->angw_desc;
    k.rangew = e->rangew;
    k.rangew_n = e->rangew_n;
    k.rangew_desc = e->rangew_desc;
    k.gate_r = sp ? sp->gate_r : e->gate_r;
    k.R = (int)p->range_count;
    k.gate_start = p->range_start;
    k.gate_inv_delta = p->range_delta != 0.0f ? 1.0f / p->range_delta : 0.0f;
    range_support(e, p->range_start, p->range_delta, k.R, drift, &k.sup_lo, &k.sup_hi, &k.dense);
    k.tri = e->rangew_n == 3 && e->rangew_h[0] == 0.0f && e->rangew_h[1] == 1.0f && e->rangew_h[2] == 0.0f &&
            e->rangew_desc.s[2] == 2.0f;
    k.tri_s0 = e->rangew_desc.s[0];
    k.tri_s1 = e->rangew_desc.s[1] - 1.0f;
    if (p->range_delta == 0.0f) k.dense = 1;
    k.phase_mode = p->phase_mode;
    k.fold = e->fold;
    // The windows need every scatterer's gates inside RSD_WWIN consecutive gates
    // (with room to move); wider supports take the wave sums.
    // Its LDS tables (weight pairs and gate ranges) must leave room for the
    // debris kernels' static LDS next to the 4 R binary64 gate sums.
    if (k.fold == RSD_FOLD_WINDOW &&
        (k.dense || (k.sup_hi - k.sup_lo) * k.gate_inv_delta >= (float)(RSD_WWIN - 1) ||
         (size_t)k.R * (4 * sizeof(double) + sizeof(float)) + (size_t)k.rangew_n * sizeof(float2) > 96 * 1024))
        k.fold = RSD_FOLD_WAVE32;
    k.partials = sp ? sp->partials : e->partials;
    k.row0 = 0;
    return RSD_OK;
}

static size_t padded_texels(int nx, int ny, int nz) {
    return (size_t)(nx + RSD_PAD_LO + RSD_PAD_HI) * (ny + RSD_PAD_LO + RSD_PAD_HI) * (nz + RSD_PAD_LO + RSD_PAD_HI);
}

static Tex3 tex3_of(const float4 *padded, int nx, int ny, int nz) {
    Tex3 t;
    memset(&t, 0, sizeof(t));
    t.d = padded;
    t.nx = nx; t.ny = ny; t.nz = nz;
    t.px = (uint32_t)(nx + RSD_PAD_LO + RSD_PAD_HI);
    t.pxy = t.px * (uint32_t)(ny + RSD_PAD_LO + RSD_PAD_HI);
    t.fx = (float)nx; t.fy = (float)ny; t.fz = (float)nz;
    return t;
}

// The raw upload buffer, at least `texels` long (a larger one replaces it once
// the pads already queued have read it).
static int ensure_raw(rsd_engine *e, size_t texels) {
    if (e->raw && e->raw_n >= texels) return RSD_OK;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    dfree(e->raw);
    e->raw = NULL;
    e->raw_n = 0;
    HIPCHK(e, hipMalloc((void **)&e->raw, texels * sizeof(float4)));
    e->raw_n = texels;
    return RSD_OK;
}

// Padded copy of the raw buffer's nx x ny x nz table into dst, on the stream.
static int pad_raw(rsd_engine *e, float4 *dst, int nx, int ny, int nz) {
    const size_t n = padded_texels(nx, ny, nz);
    size_t nb = (n + 255) / 256;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL(k_pad3d, dim3((unsigned)nb), dim3(256), 0, e->stream, (const float4 *)e->raw, dst, nx, ny, nz);
    HIPCHK(e, hipGetLastError());
    return RSD_OK;
}

// Time interpolation towards the other LES slot (same shape), weight alpha.
// The texel pairs of slot `slot` (volume k) and the other slot (k + 1),
// rebuilt on the engine's stream after an upload into either; null when they
// cannot be allocated (the launch then reads the two volumes apart).
static const float4 *wind_pair(rsd_engine *e, int slot) {
    const int o = 1 - slot;
    const size_t pt = padded_texels(e->wnx[slot], e->wny[slot], e->wnz[slot]);
    if (!e->wpair || e->wpair_cap < 2 * pt) {
        if (e->wpair) {
            if (hipStreamSynchronize(e->stream) != hipSuccess) return NULL;
            dfree(e->wpair);
        }
        e->wpair = NULL;
        e->wpair_cap = 0;
        e->wpair_ok = 0;
        if (hipMalloc((void **)&e->wpair, 2 * pt * sizeof(float4)) != hipSuccess) {
            e->wpair = NULL;
            return NULL;
        }
        e->wpair_cap = 2 * pt;
    }
    if (!e->wpair_ok || e->wpair_slot != slot || e->wpair_gen[0] != e->wgen[slot] || e->wpair_gen[1] != e->wgen[o]) {
        hipLaunchKernelGGL(k_wind_pair, dim3((unsigned)((pt + 255) / 256)), dim3(256), 0, e->stream,
                           (const float4 *)e->wind[slot], (const float4 *)e->wind[o], e->wpair, (unsigned long long)pt);
        if (hipGetLastError() != hipSuccess) return NULL;
        e->wpair_ok = 1;
        e->wpair_slot = slot;
        e->wpair_gen[0] = e->wgen[slot];
        e->wpair_gen[1] = e->wgen[o];
    }
    return e->wpair;
}

static int blend_slot(rsd_engine *e, int slot, Tex3 *t, float alpha) {
    const int o = 1 - slot;
    if (!e->wind_ok[o] || e->wnx[o] != e->wnx[slot] || e->wny[o] != e->wny[slot] || e->wnz[o] != e->wnz[slot])
        return fail(e, RSD_ERR_STATE, "wind slot %d (the next volume) not set for the time interpolation", o);
    if (!(alpha >= 0.0f && alpha <= 1.0f)) return fail(e, RSD_ERR_ARG, "time interpolation weight %g outside [0, 1]", alpha);
    t->d2 = e->wind[o];
    t->alpha = alpha;
    return RSD_OK;
}

// The index tables of a wind slot for the box of sim_desc sd: rebuilt and
// copied (stream-ordered) when the slot's LES descriptor or the box changed;
// *rebuilt tells the caller (a launch on another stream waits for the copy).
static int wind_itab(rsd_engine *e, int slot, const float *sd, WIdx *w, int *rebuilt) {
    float key[22];
    memcpy(key, e->les[slot].s, 16 * sizeof(float));
    memcpy(key + 16, sd + 8, 3 * sizeof(float));
    memcpy(key + 19, sd + 12, 3 * sizeof(float));
    if (rebuilt) *rebuilt = 0;
    if (!e->itab_ok[slot] || memcmp(key, e->itab_key[slot], sizeof(key)) != 0) {
        if (!e->ev_itab[slot]) HIPCHK(e, hipEventCreateWithFlags(&e->ev_itab[slot], hipEventDisableTiming));
        else HIPCHK(e, hipEventSynchronize(e->ev_itab[slot]));     // the staging's last copy is done
        WIdx &t = e->itab_w[slot];
        int st = rsd_wind_index_table(e->les[slot].s, sd, (float *)e->itab_h[slot], t.ih, t.jm);
        if (st) return fail(e, st, "wind index table");
        t.t = e->itab_d[slot];
        HIPCHK(e, hipMemcpyAsync(e->itab_d[slot], e->itab_h[slot], RSD_ITAB_ROWS * sizeof(float4), hipMemcpyHostToDevice,
                                 e->stream));
        HIPCHK(e, hipEventRecord(e->ev_itab[slot], e->stream));
        memcpy(e->itab_key[slot], key, sizeof(key));
        e->itab_ok[slot] = 1;
        if (rebuilt) *rebuilt = 1;
    }
    *w = e->itab_w[slot];
    w->t = e->itab_d[slot];
    return RSD_OK;
}

static int fill_rain(rsd_engine *e, const rsd_step *s, RainK &a) {
    int slot = s ? s->wind_slot : 0;
    if (s && (!e->wind_ok[slot])) return fail(e, RSD_ERR_STATE, "wind slot %d not set", slot);
    a.widx.t = e->itab_d[slot];     // read only by advances
    if (s) {
        int st = wind_itab(e, slot, s->sim_desc, &a.widx, NULL);
        if (st) return st;
    }
    if (!e->drop) return fail(e, RSD_ERR_STATE, "drop RCS table not set");
    a.pos = e->pos; a.vel = e->vel; a.rcs = e->rcs; a.rnd = e->rnd;
    a.n = e->n_rain;
    a.wind = tex3_of(e->wind[slot], e->wnx[slot], e->wny[slot], e->wnz[slot]);
    if (s && s->wind_blend) {
        int st = blend_slot(e, slot, &a.wind, s->wind_alpha);
        if (st) return st;
        a.wpair = wind_pair(e, slot);
    }
    a.les = e->les[slot];
    if (s) memcpy(a.sd.s, s->sim_desc, sizeof(a.sd.s));
    a.drop = e->drop;
    a.drop_desc = e->drop_desc;
    a.dragged = s ? (s->background_kind == RSD_ADVANCE_DRAGGED) : 1u;
    a.hx = a.sd.s[8] + a.sd.s[12];
    a.hy = a.sd.s[9] + a.sd.s[13];
    a.hz = a.sd.s[10] + a.sd.s[14];
    a.cx = a.sd.s[8] + 0.5f * a.sd.s[12];
    a.cy = a.sd.s[9] + 0.5f * a.sd.s[13];
    return RSD_OK;
}

// The fused advance + pulse launches that the specialised rain kernel (k_rain
// FAST) covers: dragged rain, a stretched wind table, drop and angular tables that fit in LDS and a binary64 phase (or
// hi/lo positions).
static bool rain_fast(const rsd_engine *e, const rsd_step *s, const rsd_pulse *p) {
    if (!s || !p || s->background_kind != RSD_ADVANCE_DRAGGED) return false;
    const F16 &l = e->les[s->wind_slot];
    uint32_t spacing;
    memcpy(&spacing, &l.s[7], sizeof(spacing));
    if (spacing != 7u) return false;
    const int dn = (int)e->drop_desc.s[2] + 1, an = (int)e->angw_desc.s[2] + 1;
    if (dn < 1 || dn > RSD_LDS_DROP_MAX || an < 1 || an > RSD_LDS_ANGW_MAX) return false;
    if (!e->plo && (unsigned long long)e->n_rain * sizeof(float4) > RSD_BUF_MAX_BYTES) return false;
    return e->plo != NULL || p->phase_mode == RSD_PHASE_F64;
}

// Stage the drop RCS and angular weight tables of a rain pulse launch in LDS
// when they fit; returns the launch's dynamic LDS bytes (base: without them).
static size_t set_rain_tables(const rsd_engine *e, RainK &a, int R, size_t base) {
    const int dn = (int)e->drop_desc.s[2] + 1, an = (int)e->angw_desc.s[2] + 1;
    if (dn < 1 || dn > RSD_LDS_DROP_MAX || an < 1 || an > RSD_LDS_ANGW_MAX) return base;
    a.tab_lds = 1;
    a.drop_n = dn;
    a.angw_n = an;
    const size_t off = rain_tab_offset(R, a.pk.fold == RSD_FOLD_WINDOW, e->rangew_n);
    a.tab_off = (unsigned)off;
    return off + (size_t)dn * sizeof(float4) + (size_t)an * sizeof(float);
}

static int fill_debris(rsd_engine *e, const rsd_step *s, const rsd_pulse *p, int type, DebrisK &a) {
    a.pos = e->pos; a.ori = e->ori; a.vel = e->vel; a.tum = e->tum; a.rcs = e->rcs; a.rnd = e->rnd;
    a.begin = e->origins[type];
    a.end = e->origins[type] + e->counts[type];
    a.widx.t = e->itab_d[s ? s->wind_slot : 0];
    a.dff = e->dff[s ? s->wind_slot : 0];
    a.dff_desc = e->dff_desc;
    if (s) {
        int slot = s->wind_slot, ai = s->adm_index[type], ri = s->rcs_index[type];
        if (!e->wind_ok[slot]) return fail(e, RSD_ERR_STATE, "wind slot %d not set", slot);
        if (ai < 0 || ai >= RSD_MAX_TABLES || !e->adm_cd[ai]) return fail(e, RSD_ERR_STATE, "ADM table %d not set", ai);
        if (ri < 0 || ri >= RSD_MAX_TABLES || !e->rcs_re[ri]) return fail(e, RSD_ERR_STATE, "RCS table %d not set", ri);
        a.wind = tex3_of(e->wind[slot], e->wnx[slot], e->wny[slot], e->wnz[slot]);
        if (s->wind_blend) {
            int st = blend_slot(e, slot, &a.wind, s->wind_alpha);
            if (st) return st;
        }
        a.les = e->les[slot];
        memcpy(a.sd.s, s->sim_desc, sizeof(a.sd.s));
        int st = wind_itab(e, slot, s->sim_desc, &a.widx, NULL);
        if (st) return st;
        a.adm_cd.d = e->adm_cd[ai]; a.adm_cm.d = e->adm_cm[ai];
        a.adm_cd.nx = a.adm_cm.nx = e->adm_nx[ai]; a.adm_cd.ny = a.adm_cm.ny = e->adm_ny[ai];
        a.adm_desc = e->adm_desc[ai];
        a.rcs_re.d = e->rcs_re[ri]; a.rcs_im.d = e->rcs_im[ri];
        a.rcs_re.nx = a.rcs_im.nx = e->rcs_nx[ri]; a.rcs_re.ny = a.rcs_im.ny = e->rcs_ny[ri];
        a.rcs_desc = e->rcs_desc[ri];
    }
    if (p) {
        int ri = p->rcs_index[type];
        if (ri < 0 || ri >= RSD_MAX_TABLES || !e->rcs_re[ri]) return fail(e, RSD_ERR_STATE, "RCS table %d not set", ri);
        a.prcs_re.d = e->rcs_re[ri]; a.prcs_im.d = e->rcs_im[ri];
        a.prcs_re.nx = a.prcs_im.nx = e->rcs_nx[ri]; a.prcs_re.ny = a.prcs_im.ny = e->rcs_ny[ri];
        a.prcs_desc = e->rcs_desc[ri];
        a.refresh = p->debris_rcs_refresh;
        uint32_t pc;
        memcpy(&pc, &p->sim_desc[5], sizeof(pc));
        a.pulse_concept = pc;
        if (s) {
            uint32_t sc;
            memcpy(&sc, &s->sim_desc[5], sizeof(sc));
            a.reuse_rcs = (s->rcs_index[type] == ri && sc == pc) ? 1u : 0u;
        }
    }
    return RSD_OK;
}

// Debris populations small next to the rain ride in the rain launch; larger
// ones, and types with a surrogate network, run on the side stream.
// Table-driven debris types ride in the rain launch (up to RSD_STEP_DEBRIS of
// them); surrogate-driven ones run in k_debris<.., true> on the side stream.
static bool debris_rides(const rsd_engine *e, int k) {
    return e->n_rain > 0 && !e->surr[k] && e->counts[k] > 0;
}

// Launch the kernels of one step: advance (s != NULL) and/or pulse (p != NULL).
// sp != NULL: the pulse's partial sums and gate ranges go to the scratch, the
// rain grid shrinks to the scratch's rows, the result goes to sp->out (and the
// engine's pulse buffer) and nothing is allocated.
static int launch_step(rsd_engine *e, const rsd_step *s, const rsd_pulse *p, const StaticPulse *sp = NULL) {
    HIPCHK(e, hipSetDevice(e->device));
    join_batch(e);
    const int R = p ? (int)p->range_count : 0;
    if (p && (R < 1 || R > RSD_MAX_GATES)) return fail(e, RSD_ERR_ARG, "range_count %d out of range", R);
    // A rain pulse (k_rain) runs the previous rain pulse's finalize in its first
    // blocks and leaves its own to the next one; every other launch finalizes
    // what is pending first and its own pulse at once.
    const bool fixed_bg = s ? s->background_kind == RSD_ADVANCE_FIXED : e->bg_fixed != 0;
    // An engine inside a batch (a launch of its own beside the batched ones)
    // finalizes at once, from the second partial buffer: the batch kernel may
    // write the first meanwhile and finalizes the engine's batched pulses.
    const bool defer = p && !sp && e->n_rain && !fixed_bg && !e->batch;
    const int part_buf = defer ? e->part_cur : (e->batch ? 1 : 0);
    if (!defer) {
        int st = flush_finalize(e);
        if (st) return st;
    }
    if (s && e->wind_ok[s->wind_slot]) {
        // the index tables of the step's wind slot, copied (if rebuilt) on the
        // main stream before the side stream forks from it
        WIdx w;
        int st = wind_itab(e, s->wind_slot, s->sim_desc, &w, NULL);
        if (st) return st;
    }
    // Rows of the partial-sum buffer: one per block of every launch, rain
    // first, then each debris type in type order.
    int deb_blocks[RSD_MAX_TYPES] = {0};
    bool ride[RSD_MAX_TYPES] = {false};
    int n_ride = 0, ride_blocks = 0, last_ride = -1;
    for (int k = 1; k < RSD_MAX_TYPES; k++) {
        if (!e->counts[k]) continue;
        deb_blocks[k] = blocks_for(e, e->counts[k]);
        if (n_ride < RSD_STEP_DEBRIS && debris_rides(e, k)) {
            ride[k] = true;
            n_ride++;
            ride_blocks += deb_blocks[k];
            last_ride = k;
        }
    }
    if (last_ride > 0 && ride_blocks % 8) {  // the rain blocks keep their XCD order
        deb_blocks[last_ride] += 8 - ride_blocks % 8;
        ride_blocks += 8 - ride_blocks % 8;
    }
    bool any_debris = false;       // types on the side stream
    for (int k = 1; k < RSD_MAX_TYPES; k++)
        if (e->counts[k] && !ride[k]) any_debris = true;
    // Rain blocks: by default, when no debris kernel shares the device, two
    // resident rounds of equal blocks (next to the riding debris blocks).
    int rain_blocks = e->n_rain ? blocks_for(e, e->n_rain) : 0;
    if (e->n_rain && e->max_blocks == 0 && !any_debris && s && p) {
        const size_t lds_est = (size_t)4 * R * sizeof(double) + (size_t)e->rangew_n * sizeof(float2) +
                               (size_t)R * sizeof(float) + 16 + (size_t)RSD_LDS_DROP_MAX * sizeof(float4) +
                               (size_t)RSD_LDS_ANGW_MAX * sizeof(float);
        const int res = rain_resident_blocks(e, e->plo != NULL, rain_fast(e, s, p), s->wind_blend != 0, lds_est);
        // riders taking a round or more: the rain keeps the default grid
        const int round = ride_blocks < res ? 2 * res - ride_blocks : rain_blocks;
        const size_t need = (e->n_rain + 63) / 64;   // at least a wave per block
        rain_blocks = round < 1 ? 1 : ((size_t)round < need ? round : (int)need);
    }
    if (sp && p) {
        // rows of the scratch at 4 R doubles each; the debris keep theirs
        size_t deb_rows = 0;
        for (int k = 1; k < RSD_MAX_TYPES; k++) deb_rows += (size_t)deb_blocks[k];
        const size_t cap_rows = RSD_SCRATCH_PARTIALS / (4 * (size_t)R);
        if (deb_rows + (e->n_rain ? 1 : 0) > cap_rows)
            return fail(e, RSD_ERR_ARG, "the scratch holds %zu partial rows of %d gates; %zu debris blocks need more "
                        "(at most %u bytes of scratch)", cap_rows, R, deb_rows, (unsigned)RSD_SCRATCH_BYTES);
        if (e->n_rain && (size_t)rain_blocks > cap_rows - deb_rows) rain_blocks = (int)(cap_rows - deb_rows);
    }
    int rows = rain_blocks, row_of[RSD_MAX_TYPES] = {0};
    for (int k = 1; k < RSD_MAX_TYPES; k++) {
        row_of[k] = rows;
        rows += deb_blocks[k];
    }
    PulseK pk;
    memset(&pk, 0, sizeof(pk));
    if (p) {
        int st = sp ? RSD_OK : ensure_partials(e, (size_t)rows, R, part_buf);
        if (st) return st;
        st = fill_pulse(e, p, pk, sp);
        if (st) return st;
        if (!sp && part_buf) pk.partials = e->partials2;
        pk.rows = rows;
        if (sp) {
            hipLaunchKernelGGL(k_gate_ranges, dim3(1), dim3(64), 0, e->stream, p->range_start, p->range_delta, R,
                               sp->gate_r);
            HIPCHK(e, hipGetLastError());
        }
    }
    const size_t lds = p ? (size_t)4 * R * sizeof(double) + (pk.fold == RSD_FOLD_WINDOW ? (size_t)e->rangew_n * sizeof(float2) + (size_t)R * sizeof(float) : 0) : 0;
    if (any_debris) {
        HIPCHK(e, hipEventRecord(e->ev_fork, e->stream));
        HIPCHK(e, hipStreamWaitEvent(e->side, e->ev_fork, 0));
        // side types in launches of up to RSD_STEP_DEBRIS types, surrogate
        // and table types apart (the surrogate kernel runs 256-thread blocks)
        for (int surr = 0; surr < 2; surr++) {
            StepDebris d;
            memset(&d, 0, sizeof(d));
            for (int k = 1; k < RSD_MAX_TYPES; k++) {
                const bool last = k == RSD_MAX_TYPES - 1;
                if (e->counts[k] && !ride[k] && (e->surr[k] ? 1 : 0) == surr) {
                    DebrisK &a = d.t[d.nt];
                    int st = fill_debris(e, s, p, k, a);
                    if (st) return st;
                    a.pk = pk;
                    a.pk.row0 = row_of[k];
                    if (surr) a.net = e->net[k];
                    d.first[d.nt + 1] = d.first[d.nt] + deb_blocks[k];
                    d.nt++;
                }
                if (d.nt == RSD_STEP_DEBRIS || (last && d.nt > 0)) {
                    const dim3 grid(d.first[d.nt]);
                    if (surr) {
                        if (s && p) hipLaunchKernelGGL((k_debris<true, true, true>), grid, dim3(256), lds, e->side, d);
                        else if (s) hipLaunchKernelGGL((k_debris<true, false, true>), grid, dim3(256), 0, e->side, d);
                        else hipLaunchKernelGGL((k_debris<false, true, true>), grid, dim3(256), lds, e->side, d);
                    } else {
                        if (s && p) hipLaunchKernelGGL((k_debris<true, true, false>), grid, dim3(e->block), lds, e->side, d);
                        else if (s) hipLaunchKernelGGL((k_debris<true, false, false>), grid, dim3(e->block), 0, e->side, d);
                        else hipLaunchKernelGGL((k_debris<false, true, false>), grid, dim3(e->block), lds, e->side, d);
                    }
                    HIPCHK(e, hipGetLastError());
                    memset(&d, 0, sizeof(d));
                }
            }
        }
        HIPCHK(e, hipEventRecord(e->ev_join, e->side));
    }
    const bool fixed = s ? s->background_kind == RSD_ADVANCE_FIXED : e->bg_fixed != 0;
    if (e->n_rain && fixed) {
        if (ride_blocks || any_debris) return fail(e, RSD_ERR_STATE, "a fixed-position background takes no debris");
        FixedK a;
        memset(&a, 0, sizeof(a));
        const int slot = s ? s->wind_slot : 0;
        if (s && (!e->wind_ok[slot] || !e->cpxx[slot] || e->cnx[slot] != e->wnx[slot] || e->cny[slot] != e->wny[slot] ||
                  e->cnz[slot] != e->wnz[slot]))
            return fail(e, RSD_ERR_STATE, "wind and Cn2 volume of slot %d not set", slot);
        a.pos = e->pos; a.vel = e->vel; a.rcs = e->rcs;
        a.n = e->n_rain;
        a.wind = tex3_of(e->wind[slot], e->wnx[slot], e->wny[slot], e->wnz[slot]);
        a.cpxx = tex3_of(e->cpxx[slot], e->cnx[slot], e->cny[slot], e->cnz[slot]);
        if (s && s->wind_blend) {
            int st = blend_slot(e, slot, &a.wind, s->wind_alpha);
            if (st) return st;
            if (!e->cpxx[1 - slot] || e->cnx[1 - slot] != e->cnx[slot] || e->cny[1 - slot] != e->cny[slot] ||
                e->cnz[1 - slot] != e->cnz[slot])
                return fail(e, RSD_ERR_STATE, "Cn2 volume of slot %d not set for the time interpolation", 1 - slot);
            a.cpxx.d2 = e->cpxx[1 - slot];
            a.cpxx.alpha = s->wind_alpha;
        }
        a.les = e->les[slot];
        a.widx.t = e->itab_d[slot];
        if (s) {
            memcpy(a.sd.s, s->sim_desc, sizeof(a.sd.s));
            int st = wind_itab(e, slot, s->sim_desc, &a.widx, NULL);
            if (st) return st;
        }
        a.pk = pk;
        a.pk.row0 = 0;
        const dim3 grid(rain_blocks);
        if (s && p) hipLaunchKernelGGL((k_fixed<true, true>), grid, dim3(e->block), lds, e->stream, a);
        else if (s) hipLaunchKernelGGL((k_fixed<true, false>), grid, dim3(e->block), 0, e->stream, a);
        else hipLaunchKernelGGL((k_fixed<false, true>), grid, dim3(e->block), lds, e->stream, a);
        HIPCHK(e, hipGetLastError());
    } else if (e->n_rain) {
        RainK a;
        memset(&a, 0, sizeof(a));
        int st = fill_rain(e, s, a);
        if (st) return st;
        a.pk = pk;
        a.pk.row0 = 0;
        a.plo = e->plo;
        StepDebris d;
        memset(&d, 0, sizeof(d));
        for (int k = 1; k < RSD_MAX_TYPES; k++) {
            if (!ride[k]) continue;
            DebrisK &da = d.t[d.nt];
            st = fill_debris(e, s, p, k, da);
            if (st) return st;
            da.pk = pk;
            da.pk.row0 = row_of[k];
            d.first[d.nt + 1] = d.first[d.nt] + deb_blocks[k];
            d.nt++;
        }
        const size_t lds_rain = p ? set_rain_tables(e, a, R, lds) : lds;
        if (e->fin_on) {        // defer holds here (only rain pulses leave a finalize pending)
            a.fin = e->fin;
            a.fin_blocks = (unsigned)e->fin.n4r;
            e->fin_on = 0;
        }
        const dim3 grid(ride_blocks + a.fin_blocks + rain_blocks);
        const bool timed = e->prof_on && e->prof_n < RSD_PROF_CAP && (e->prof_tick++ % e->prof_stride) == 0;
        // A timed launch carries its events in the dispatch itself
        // (hipExtLaunchKernelGGL): they bracket the kernel's execution, not the
        // queue's processing of separate event packets.
        hipEvent_t t0 = timed ? e->prof_ev[2 * e->prof_n] : nullptr, t1 = timed ? e->prof_ev[2 * e->prof_n + 1] : nullptr;
        const bool ti = a.wind.d2 != nullptr;
        const bool fast = rain_fast(e, s, p) && a.tab_lds && (!ti || a.wpair);
        if (fast) {
            if (e->plo && ti) hipExtLaunchKernelGGL((k_rain<true, true, true, true, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
            else if (ti) hipExtLaunchKernelGGL((k_rain<true, true, false, true, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
            else if (e->plo) hipExtLaunchKernelGGL((k_rain<true, true, true, false, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
            else hipExtLaunchKernelGGL((k_rain<true, true, false, false, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
        } else if (e->plo && ti) {
            if (s && p) hipExtLaunchKernelGGL((k_rain<true, true, true, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
            else if (s) hipExtLaunchKernelGGL((k_rain<true, false, true, true>), grid, dim3(e->block), 0, e->stream, t0, t1, 0, a, d);
            else hipExtLaunchKernelGGL((k_rain<false, true, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
        } else if (ti) {
            if (s && p) hipExtLaunchKernelGGL((k_rain<true, true, false, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
            else if (s) hipExtLaunchKernelGGL((k_rain<true, false, false, true>), grid, dim3(e->block), 0, e->stream, t0, t1, 0, a, d);
            else hipExtLaunchKernelGGL((k_rain<false, true, false>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
        } else if (e->plo) {
            if (s && p) hipExtLaunchKernelGGL((k_rain<true, true, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
            else if (s) hipExtLaunchKernelGGL((k_rain<true, false, true>), grid, dim3(e->block), 0, e->stream, t0, t1, 0, a, d);
            else hipExtLaunchKernelGGL((k_rain<false, true, true>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
        } else {
            if (s && p) hipExtLaunchKernelGGL((k_rain<true, true, false>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
            else if (s) hipExtLaunchKernelGGL((k_rain<true, false, false>), grid, dim3(e->block), 0, e->stream, t0, t1, 0, a, d);
            else hipExtLaunchKernelGGL((k_rain<false, true, false>), grid, dim3(e->block), lds_rain, e->stream, t0, t1, 0, a, d);
        }
        HIPCHK(e, hipGetLastError());
        if (timed) {
            e->prof_kind[e->prof_n] = (s ? 1 : 0) | (p ? 2 : 0);
            e->prof_n++;
        }
    }
    if (any_debris) HIPCHK(e, hipStreamWaitEvent(e->stream, e->ev_join, 0));
    if (p && sp) {
        hipLaunchKernelGGL(k_finalize, dim3(4 * R), dim3(256), 0, e->stream, (const double *)sp->partials, rows, 4 * R,
                           sp->out, e->pulse_d);
        HIPCHK(e, hipGetLastError());
        e->last_pulse = *p;
        e->have_pulse = 1;
        e->sigaux_saved = 0;
    } else if (p) {
        double *slot = NULL;
        if (e->ring && e->ring_R == R) {
            slot = e->ring + (size_t)e->ring_next * 4 * R;
            e->ring_next = (e->ring_next + 1) % e->ring_m;
            if (e->ring_count < e->ring_m) e->ring_count++;
        }
        if (defer) {
            e->fin.partials = pk.partials;
            e->fin.rows = rows;
            e->fin.n4r = 4 * R;
            e->fin.out = e->pulse_d;
            e->fin.out2 = slot;
            e->fin_on = 1;
            e->part_cur ^= 1;
        } else {
            hipLaunchKernelGGL(k_finalize, dim3(4 * R), dim3(256), 0, e->stream, (const double *)pk.partials, rows,
                               4 * R, e->pulse_d, slot);
            HIPCHK(e, hipGetLastError());
        }
        e->last_pulse = *p;
        e->have_pulse = 1;
        e->sigaux_saved = 0;
    }
    return RSD_OK;
}

static int ensure_sort_buffers(rsd_engine *e) {
    if (e->pos2) return RSD_OK;
    const size_t b = e->n * sizeof(float4), bi = e->n_rain * sizeof(uint32_t);
    HIPCHK(e, hipMalloc((void **)&e->pos2, b));
    HIPCHK(e, hipMalloc((void **)&e->vel2, b));
    HIPCHK(e, hipMalloc((void **)&e->rcs2, b));
    HIPCHK(e, hipMalloc((void **)&e->rnd2, e->n * sizeof(uint4)));
    if (e->plo) HIPCHK(e, hipMalloc((void **)&e->plo2, b));
    HIPCHK(e, hipMalloc((void **)&e->id, bi));
    HIPCHK(e, hipMalloc((void **)&e->id2, bi));
    HIPCHK(e, hipMalloc((void **)&e->skey, bi));
    HIPCHK(e, hipMalloc((void **)&e->skey2, bi));
    HIPCHK(e, hipMalloc((void **)&e->sperm, bi));
    HIPCHK(e, hipMalloc((void **)&e->sperm2, bi));
    size_t tmp = 0;
    HIPCHK(e, hipcub::DeviceRadixSort::SortPairs(NULL, tmp, e->skey, e->skey2, e->sperm, e->sperm2, (int)e->n_rain, 0, 30, e->stream));
    HIPCHK(e, hipMalloc(&e->sort_tmp, tmp));
    e->sort_tmp_bytes = tmp;
    return RSD_OK;
}

static void swap_streams(rsd_engine *e) {
    float4 *t;
    t = e->pos; e->pos = e->pos2; e->pos2 = t;
    t = e->vel; e->vel = e->vel2; e->vel2 = t;
    t = e->rcs; e->rcs = e->rcs2; e->rcs2 = t;
    uint4 *r = e->rnd; e->rnd = e->rnd2; e->rnd2 = r;
    uint32_t *u = e->id; e->id = e->id2; e->id2 = u;
    t = e->plo; e->plo = e->plo2; e->plo2 = t;
}

static PermK perm_args(const rsd_engine *e) {
    PermK a;
    a.pos = e->pos; a.vel = e->vel; a.rcs = e->rcs; a.rnd = e->rnd; a.plo = e->plo;
    a.id = e->id; a.perm = e->sperm2;
    a.pos2 = e->pos2; a.vel2 = e->vel2; a.rcs2 = e->rcs2; a.rnd2 = e->rnd2; a.plo2 = e->plo2;
    a.id2 = e->id2;
    a.n = e->n;
    a.n_rain = e->n_rain;
    return a;
}

// Reorder the rain streams by range shell and box cell (sim_desc slots 8-10
// box origin, 12-14 size; shell width = gate spacing / 8).
static int sort_rain(rsd_engine *e, const float *sd, float gate_delta) {
    if (e->n_rain < 2 || e->n_rain > 0x7fffffffu) return RSD_OK;
    if (!(sd[12] > 0.0f && sd[13] > 0.0f && sd[14] > 0.0f)) return RSD_OK;
    int st = ensure_sort_buffers(e);
    if (st) return st;
    const unsigned nb_rain = (unsigned)((e->n_rain + 255) / 256), nb_all = (unsigned)((e->n + 255) / 256);
    if (!e->permuted) {
        hipLaunchKernelGGL(k_iota, dim3(nb_rain), dim3(256), 0, e->stream, e->id, (unsigned long long)e->n_rain);
        HIPCHK(e, hipGetLastError());
    }
    SortK k;
    k.pos = e->pos; k.key = e->skey; k.idx = e->sperm; k.n = e->n_rain;
    k.ox = sd[8]; k.oy = sd[9]; k.oz = sd[10];
    k.ix = 1024.0f / sd[12]; k.iy = 1024.0f / sd[13]; k.iz = 1024.0f / sd[14];
    // nearest and farthest range of the box from the radar at the origin
    double near2 = 0.0, far2 = 0.0;
    for (int a = 0; a < 3; a++) {
        double lo = sd[8 + a], hi = (double)sd[8 + a] + sd[12 + a];
        double dn = lo > 0.0 ? lo : (hi < 0.0 ? -hi : 0.0);
        double df = fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi);
        near2 += dn * dn;
        far2 += df * df;
    }
    const double rn = sqrt(near2), rf = sqrt(far2);
    double width = fabs((double)gate_delta) / 8.0;
    if (!(width > 0.0)) width = (rf - rn) / 4096.0;
    if ((rf - rn) / width > 262143.0) width = (rf - rn) / 262143.0;
    k.r0 = (float)rn;
    k.ir = (float)(1.0 / width);
    k.shells = (uint32_t)((rf - rn) / width) + 1u;
    if (k.shells > 262144u) k.shells = 262144u;
    int bits = 12;
    while ((1u << (bits - 12)) < k.shells) bits++;
    hipLaunchKernelGGL(k_sort_keys, dim3(nb_rain), dim3(256), 0, e->stream, k);
    HIPCHK(e, hipGetLastError());
    size_t tmp = e->sort_tmp_bytes;
    HIPCHK(e, hipcub::DeviceRadixSort::SortPairs(e->sort_tmp, tmp, e->skey, e->skey2, e->sperm, e->sperm2, (int)e->n_rain, 0, bits, e->stream));
    PermK a = perm_args(e);
    hipLaunchKernelGGL(k_permute, dim3(nb_all), dim3(256), 0, e->stream, a);
    HIPCHK(e, hipGetLastError());
    swap_streams(e);
    e->permuted = 1;
    e->since_sort = 0;
    return RSD_OK;
}

// Restore the reference order of the rain streams.
static int unsort_rain(rsd_engine *e) {
    if (!e->permuted) return RSD_OK;
    const unsigned nb_all = (unsigned)((e->n + 255) / 256);
    PermK a = perm_args(e);
    hipLaunchKernelGGL(k_unpermute, dim3(nb_all), dim3(256), 0, e->stream, a);
    HIPCHK(e, hipGetLastError());
    swap_streams(e);
    e->permuted = 0;
    return RSD_OK;
}

extern "C" int rsd_set_position_precision(rsd_engine *e, int hilo) {
    if (!e || (hilo != 0 && hilo != 1)) return fail(e, RSD_ERR_ARG, "position precision %d unknown", hilo);
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    join_batch(e);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->pos_hilo = hilo;
    if (!e->pos) return RSD_OK;
    const size_t b = e->n * sizeof(float4);
    if (hilo && !e->plo) {
        HIPCHK(e, hipMalloc((void **)&e->plo, b));
        HIPCHK(e, hipMemsetAsync(e->plo, 0, b, e->stream));
        if (e->pos2 && !e->plo2) HIPCHK(e, hipMalloc((void **)&e->plo2, b));
    } else if (!hilo && e->plo) {
        dfree(e->plo); dfree(e->plo2);
        e->plo = e->plo2 = NULL;
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

extern "C" int rsd_download_pos_lo(rsd_engine *e, float *lo) {
    if (!e || !lo) return RSD_ERR_ARG;
    if (!e->plo) return fail(e, RSD_ERR_STATE, "positions are not hi/lo");
    int st = flush_pending(e);
    if (st) return st;
    HIPCHK(e, hipSetDevice(e->device));
    join_batch(e);
    const size_t b = e->n * sizeof(float4);
    if (!e->view) HIPCHK(e, hipMalloc((void **)&e->view, 4 * b));
    unsigned nb = (unsigned)((e->n + 255) / 256);
    hipLaunchKernelGGL(k_lo_view, dim3(nb), dim3(256), 0, e->stream, (const float4 *)e->plo,
                       (const uint32_t *)(e->permuted ? e->id : NULL), (unsigned long long)e->n_rain,
                       (unsigned long long)e->n, e->view);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(lo, e->view, b, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

extern "C" int rsd_set_iq_ring(rsd_engine *e, int pulses, int range_count) {
    if (!e) return RSD_ERR_ARG;
    if (pulses < 0 || range_count < 0 || range_count > RSD_MAX_GATES) return fail(e, RSD_ERR_ARG, "bad IQ ring size");
    HIPCHK(e, hipSetDevice(e->device));
    int st0 = flush_pending(e);
    if (st0) return st0;
    join_batch(e);
    st0 = flush_finalize(e);        // its ring slot is in the ring freed below
    if (st0) return st0;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    dfree(e->ring);
    dfree(e->mom_d);
    e->ring = NULL;
    e->mom_d = NULL;
    e->ring_m = e->ring_R = e->ring_next = e->ring_count = 0;
    if (pulses == 0) return RSD_OK;
    if (pulses < 2 || range_count < 1) return fail(e, RSD_ERR_ARG, "IQ ring needs >= 2 pulses and >= 1 gate");
    HIPCHK(e, hipMalloc((void **)&e->ring, (size_t)pulses * 4 * range_count * sizeof(double)));
    HIPCHK(e, hipMalloc((void **)&e->mom_d, (size_t)range_count * sizeof(float4)));
    e->ring_m = pulses;
    e->ring_R = range_count;
    return RSD_OK;
}

extern "C" int rsd_moments(rsd_engine *e, int pulses, const rsd_moment_params *mp, float *out) {
    if (!e || !mp || !out) return RSD_ERR_ARG;
    if (!e->ring) return fail(e, RSD_ERR_STATE, "no IQ ring (rsd_set_iq_ring)");
    int st = flush_pending(e);
    if (st) return st;
    if (pulses < 2 || pulses > e->ring_m) return fail(e, RSD_ERR_ARG, "moments over %d pulses: ring holds %d", pulses, e->ring_m);
    if (e->ring_count < pulses) return fail(e, RSD_ERR_STATE, "only %d pulses recorded", e->ring_count);
    HIPCHK(e, hipSetDevice(e->device));
    join_batch(e);
    st = flush_finalize(e);
    if (st) return st;
    MomK a;
    a.ring = e->ring;
    a.ring_m = e->ring_m;
    a.first = (e->ring_next - pulses + e->ring_m) % e->ring_m;
    a.m = pulses;
    a.R = e->ring_R;
    a.gain2 = mp->gain * mp->gain;
    a.vfac = -mp->lambda / (4.0 * M_PI * mp->prt);
    a.range_start = mp->range_start;
    a.range_delta = mp->range_delta;
    a.out = e->mom_d;
    hipLaunchKernelGGL(k_moments, dim3(e->ring_R), dim3(256), 0, e->stream, a);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(out, e->mom_d, (size_t)e->ring_R * sizeof(float4), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

extern "C" int rsd_set_fold(rsd_engine *e, int mode) {
    if (!e) return RSD_ERR_ARG;
    if (mode < RSD_FOLD_LANE || mode > RSD_FOLD_WINDOW) return fail(e, RSD_ERR_ARG, "fold mode %d unknown", mode);
    e->fold = mode;
    return RSD_OK;
}

extern "C" int rsd_set_sort(rsd_engine *e, int interval) {
    if (!e) return RSD_ERR_ARG;
    if (interval < 0) return fail(e, RSD_ERR_ARG, "sort interval %d < 0", interval);
    e->sort_interval = interval;
    return RSD_OK;
}

extern "C" int rsd_get_sort(const rsd_engine *e) { return e ? e->sort_interval : 0; }

static int save_signal(rsd_engine *e);

// Runs a held-back advance on its own.  The signal and aux streams of the last
// pulse are first taken from the positions that pulse saw (the reference
// computes them in the pulse and keeps them through the advance).
static int flush_pending(rsd_engine *e) {
    if (e->batch && e->have_pending_pulse) {
        int st = rsd_batch_launch(e->batch);
        if (st) return fail(e, st, "batch launch failed");
    }
    if (!e->have_pending) return RSD_OK;
    if (!e->sigaux_saved) {
        int st = save_signal(e);
        if (st) return st;
    }
    e->have_pending = 0;
    return launch_step(e, &e->pending, NULL);
}

extern "C" int rsd_upload_state(rsd_engine *e, const float *pos, const float *vel, const float *ori,
                                const float *tum, const float *aux, const float *rcs, const uint32_t *rnd) {
    if (!e || !e->pos) return fail(e, RSD_ERR_STATE, "state not allocated");
    int st = flush_pending(e);
    if (st) return st;
    HIPCHK(e, hipSetDevice(e->device));
    join_batch(e);
    st = unsort_rain(e);
    if (st) return st;
    const size_t b = e->n * sizeof(float4);
    if (pos) HIPCHK(e, hipMemcpyAsync(e->pos, pos, b, hipMemcpyHostToDevice, e->stream));
    if (pos && e->plo) HIPCHK(e, hipMemsetAsync(e->plo, 0, b, e->stream));
    if (vel) HIPCHK(e, hipMemcpyAsync(e->vel, vel, b, hipMemcpyHostToDevice, e->stream));
    if (ori) HIPCHK(e, hipMemcpyAsync(e->ori, ori, b, hipMemcpyHostToDevice, e->stream));
    if (tum) HIPCHK(e, hipMemcpyAsync(e->tum, tum, b, hipMemcpyHostToDevice, e->stream));
    if (aux) HIPCHK(e, hipMemcpyAsync(e->aux, aux, b, hipMemcpyHostToDevice, e->stream));
    if (rcs) HIPCHK(e, hipMemcpyAsync(e->rcs, rcs, b, hipMemcpyHostToDevice, e->stream));
    if (rnd) HIPCHK(e, hipMemcpyAsync(e->rnd, rnd, e->n * sizeof(uint4), hipMemcpyHostToDevice, e->stream));
    if (pos && e->n_rain) {
        unsigned nb = (unsigned)((e->n_rain + 255) / 256);
        hipLaunchKernelGGL(k_flag_rain, dim3(nb), dim3(256), 0, e->stream, e->pos, (unsigned long long)e->n_rain);
        HIPCHK(e, hipGetLastError());
    }
    if (pos || aux || rcs) e->sigaux_saved = 0;   // recomputed from the uploaded state
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

static void view_common(rsd_engine *e, ViewK &v) {
    memset(&v, 0, sizeof(v));
    v.pos = e->pos; v.rcs = e->rcs; v.aux_in = e->aux; v.vel = e->vel; v.rnd = e->rnd;
    v.id = e->permuted ? e->id : NULL;
    v.n = e->n;
    v.n_rain = e->n_rain;
    v.drop = e->drop;
    v.drop_desc = e->drop_desc;
    v.fixed = e->bg_fixed;
}

// Signal and aux streams of the last pulse from the stored state into sigaux
// (zero signal and the stored aux before any pulse).  Valid until the next
// pulse; the stored positions must be the ones the pulse saw.
static int save_signal(rsd_engine *e) {
    if (e->n_rain && !e->drop && !e->bg_fixed) return fail(e, RSD_ERR_STATE, "drop RCS table not set");
    HIPCHK(e, hipSetDevice(e->device));
    join_batch(e);
    if (!e->sigaux) HIPCHK(e, hipMalloc((void **)&e->sigaux, 2 * e->n * sizeof(float4)));
    ViewK v;
    view_common(e, v);
    v.sig_out = e->sigaux;
    v.aux_out = e->sigaux + e->n;
    if (e->have_pulse) {
        int st = fill_pulse(e, &e->last_pulse, v.pk);
        if (st) return st;
        v.have_pulse = 1;
    }
    unsigned nb = (unsigned)((e->n + 255) / 256);
    hipLaunchKernelGGL(k_view, dim3(nb), dim3(256), 0, e->stream, v);
    HIPCHK(e, hipGetLastError());
    e->sigaux_saved = 1;
    return RSD_OK;
}

extern "C" int rsd_download_state(rsd_engine *e, float *pos, float *vel, float *ori, float *tum,
                                  float *aux, float *rcs, float *sig, uint32_t *rnd) {
    if (!e || !e->pos) return fail(e, RSD_ERR_STATE, "state not allocated");
    int st;
    if ((sig || aux) && !e->sigaux_saved && !(e->batch && e->have_pending_pulse)) {
        st = save_signal(e);    // no advance since the last pulse: its positions are stored
        if (st) return st;
    }
    st = flush_pending(e);
    if (st) return st;
    join_batch(e);
    if ((sig || aux) && !e->sigaux_saved) {
        st = save_signal(e);    // the batch launch above made the pulse
        if (st) return st;
    }
    HIPCHK(e, hipSetDevice(e->device));
    const size_t b = e->n * sizeof(float4);
    const bool perm = e->permuted && (vel || rnd);
    if (pos || rcs || perm) {
        if (!e->view) HIPCHK(e, hipMalloc((void **)&e->view, 4 * b));
        if (e->n_rain && !e->drop && !e->bg_fixed) return fail(e, RSD_ERR_STATE, "drop RCS table not set");
        ViewK v;
        view_common(e, v);
        v.pos_out = e->view; v.rcs_out = e->view + e->n;
        if (perm) {
            v.vel_out = e->view + 2 * e->n;
            v.rnd_out = (uint4 *)(e->view + 3 * e->n);
        }
        unsigned nb = (unsigned)((e->n + 255) / 256);
        hipLaunchKernelGGL(k_view, dim3(nb), dim3(256), 0, e->stream, v);
        HIPCHK(e, hipGetLastError());
        if (pos) HIPCHK(e, hipMemcpyAsync(pos, v.pos_out, b, hipMemcpyDeviceToHost, e->stream));
        if (rcs) HIPCHK(e, hipMemcpyAsync(rcs, v.rcs_out, b, hipMemcpyDeviceToHost, e->stream));
        if (perm && vel) HIPCHK(e, hipMemcpyAsync(vel, v.vel_out, b, hipMemcpyDeviceToHost, e->stream));
        if (perm && rnd) HIPCHK(e, hipMemcpyAsync(rnd, v.rnd_out, e->n * sizeof(uint4), hipMemcpyDeviceToHost, e->stream));
    }
    if (sig) HIPCHK(e, hipMemcpyAsync(sig, e->sigaux, b, hipMemcpyDeviceToHost, e->stream));
    if (aux) HIPCHK(e, hipMemcpyAsync(aux, e->sigaux + e->n, b, hipMemcpyDeviceToHost, e->stream));
    if (vel && !perm) HIPCHK(e, hipMemcpyAsync(vel, e->vel, b, hipMemcpyDeviceToHost, e->stream));
    if (ori) HIPCHK(e, hipMemcpyAsync(ori, e->ori, b, hipMemcpyDeviceToHost, e->stream));
    if (tum) HIPCHK(e, hipMemcpyAsync(tum, e->tum, b, hipMemcpyDeviceToHost, e->stream));
    if (rnd && !perm) HIPCHK(e, hipMemcpyAsync(rnd, e->rnd, e->n * sizeof(uint4), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

static int upload_tex(rsd_engine *e, float4 **dst, const float *src, size_t texels) {
    dfree(*dst);
    *dst = NULL;
    HIPCHK(e, hipMalloc((void **)dst, texels * sizeof(float4)));
    HIPCHK(e, hipMemcpyAsync(*dst, src, texels * sizeof(float4), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

extern "C" int rsd_set_wind(rsd_engine *e, int slot, const float *uvwt, int nx, int ny, int nz, const float *desc) {
    if (!e || slot < 0 || slot > 1 || !uvwt || !desc || nx < 1 || ny < 1 || nz < 1) return fail(e, RSD_ERR_ARG, "bad wind table");
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    const size_t texels = (size_t)nx * ny * nz, pt = padded_texels(nx, ny, nz);
    if (e->wind[slot] && e->wnx[slot] == nx && e->wny[slot] == ny && e->wnz[slot] == nz) {
        // New frame into a slot of the same shape (the LES switch of
        // RS_advance_time): a stream-ordered copy through pinned staging and a
        // pad into the slot, so the queue keeps running; launches already
        // enqueued finish first.
        join_batch(e);
        if (!e->ev_wind[slot]) HIPCHK(e, hipEventCreateWithFlags(&e->ev_wind[slot], hipEventDisableTiming));
        else HIPCHK(e, hipEventSynchronize(e->ev_wind[slot]));   // the staging buffer's last copy is done
        if (e->wind_stage_n[slot] < texels) {
            if (e->wind_stage[slot]) HIPCHK(e, hipHostFree(e->wind_stage[slot]));
            e->wind_stage[slot] = NULL;
            e->wind_stage_n[slot] = 0;
            HIPCHK(e, hipHostMalloc((void **)&e->wind_stage[slot], texels * sizeof(float4), hipHostMallocDefault));
            e->wind_stage_n[slot] = texels;
        }
        st = ensure_raw(e, texels);
        if (st) return st;
        memcpy(e->wind_stage[slot], uvwt, texels * sizeof(float4));
        HIPCHK(e, hipMemcpyAsync(e->raw, e->wind_stage[slot], texels * sizeof(float4), hipMemcpyHostToDevice, e->stream));
        HIPCHK(e, hipEventRecord(e->ev_wind[slot], e->stream));
        st = pad_raw(e, e->wind[slot], nx, ny, nz);
        if (st) return st;
    } else {
        join_batch(e);
        HIPCHK(e, hipStreamSynchronize(e->stream));
        dfree(e->wind[slot]);
        e->wind[slot] = NULL;
        HIPCHK(e, hipMalloc((void **)&e->wind[slot], pt * sizeof(float4)));
        st = ensure_raw(e, texels);
        if (st) return st;
        HIPCHK(e, hipMemcpyAsync(e->raw, uvwt, texels * sizeof(float4), hipMemcpyHostToDevice, e->stream));
        st = pad_raw(e, e->wind[slot], nx, ny, nz);
        if (st) return st;
        HIPCHK(e, hipStreamSynchronize(e->stream));
        // the other slot takes the next frame of the same shape: allocate it now
        // so that the switch needs no allocation (and no device-wide wait)
        const int o = 1 - slot;
        if (!e->wind[o] || e->wnx[o] != nx || e->wny[o] != ny || e->wnz[o] != nz) {
            dfree(e->wind[o]);
            e->wind[o] = NULL;
            HIPCHK(e, hipMalloc((void **)&e->wind[o], pt * sizeof(float4)));
            e->wnx[o] = nx; e->wny[o] = ny; e->wnz[o] = nz;
            e->wind_ok[o] = 0;
            if (!e->wind_stage[o]) {
                HIPCHK(e, hipHostMalloc((void **)&e->wind_stage[o], texels * sizeof(float4), hipHostMallocDefault));
                e->wind_stage_n[o] = texels;
            }
        }
    }
    e->wnx[slot] = nx; e->wny[slot] = ny; e->wnz[slot] = nz;
    e->wind_ok[slot] = 1;
    e->wgen[slot]++;
    memcpy(e->les[slot].s, desc, sizeof(e->les[slot].s));
    return RSD_OK;
}

extern "C" int rsd_set_cpxx(rsd_engine *e, int slot, const float *cpxx, int nx, int ny, int nz) {
    if (!e || !cpxx || slot < 0 || slot > 1 || nx < 1 || ny < 1 || nz < 1) return fail(e, RSD_ERR_ARG, "bad Cn2 volume");
    int st = flush_pending(e);
    if (st) return st;
    HIPCHK(e, hipSetDevice(e->device));
    join_batch(e);
    HIPCHK(e, hipStreamSynchronize(e->stream));      // kernels of the slot's previous volume are done
    const size_t texels = (size_t)nx * ny * nz;
    if (!e->cpxx[slot] || e->cnx[slot] != nx || e->cny[slot] != ny || e->cnz[slot] != nz) {
        dfree(e->cpxx[slot]);
        e->cpxx[slot] = NULL;
        HIPCHK(e, hipMalloc((void **)&e->cpxx[slot], padded_texels(nx, ny, nz) * sizeof(float4)));
    }
    st = ensure_raw(e, texels);
    if (st) return st;
    HIPCHK(e, hipMemcpyAsync(e->raw, cpxx, texels * sizeof(float4), hipMemcpyHostToDevice, e->stream));
    st = pad_raw(e, e->cpxx[slot], nx, ny, nz);
    if (st) return st;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->cnx[slot] = nx;
    e->cny[slot] = ny;
    e->cnz[slot] = nz;
    return RSD_OK;
}

extern "C" int rsd_set_adm(rsd_engine *e, int t, const float *cd, const float *cm, int nx, int ny, const float *desc) {
    if (!e || t < 0 || t >= RSD_MAX_TABLES || !cd || !cm || !desc || nx < 1 || ny < 1) return fail(e, RSD_ERR_ARG, "bad ADM table");
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    st = upload_tex(e, &e->adm_cd[t], cd, (size_t)nx * ny);
    if (st) return st;
    st = upload_tex(e, &e->adm_cm[t], cm, (size_t)nx * ny);
    if (st) return st;
    e->adm_nx[t] = nx; e->adm_ny[t] = ny;
    memcpy(e->adm_desc[t].s, desc, sizeof(e->adm_desc[t].s));
    return RSD_OK;
}

extern "C" int rsd_set_rcs(rsd_engine *e, int t, const float *re, const float *im, int nx, int ny, const float *desc) {
    if (!e || t < 0 || t >= RSD_MAX_TABLES || !re || !im || !desc || nx < 1 || ny < 1) return fail(e, RSD_ERR_ARG, "bad RCS table");
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    st = upload_tex(e, &e->rcs_re[t], re, (size_t)nx * ny);
    if (st) return st;
    st = upload_tex(e, &e->rcs_im[t], im, (size_t)nx * ny);
    if (st) return st;
    e->rcs_nx[t] = nx; e->rcs_ny[t] = ny;
    memcpy(e->rcs_desc[t].s, desc, sizeof(e->rcs_desc[t].s));
    return RSD_OK;
}

extern "C" int rsd_set_drop_rcs(rsd_engine *e, const float *table, int n, const float *desc4) {
    if (!e || !table || !desc4 || n < 1) return fail(e, RSD_ERR_ARG, "bad drop RCS table");
    if (!(desc4[2] <= (float)(n - 1))) return fail(e, RSD_ERR_ARG, "drop RCS max index %g beyond %d entries", desc4[2], n);
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    st = upload_tex(e, &e->drop, table, (size_t)n);
    if (st) return st;
    memcpy(e->drop_desc.s, desc4, sizeof(e->drop_desc.s));
    return RSD_OK;
}

static int check_dff(rsd_engine *e, int slot, const float *icdf, int count, const float *desc16) {
    if (!e || slot < -1 || slot > 1 || count < 0 || (count > 0 && (!icdf || !desc16)))
        return fail(e, RSD_ERR_ARG, "bad debris flux field");
    if (count > 0 && !(desc16[15] >= 0.0f && desc16[15] <= (float)(count - 1)))
        return fail(e, RSD_ERR_ARG, "flux field max index %g beyond %d entries", desc16[15], count);
    if (count > 0 && !(desc16[13] > 0.0f)) return fail(e, RSD_ERR_ARG, "flux field row length %g", desc16[13]);
    return RSD_OK;
}

// Both slots (re)allocated and filled with one table, after the queue drains.
static int dff_both(rsd_engine *e, const float *icdf, int count, const float *desc16) {
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (int k = 0; k < 2; k++) {
        if (e->ev_dff[k]) HIPCHK(e, hipEventSynchronize(e->ev_dff[k]));
        dfree(e->dff[k]);
        e->dff[k] = NULL;
        e->dff_n[k] = 0;
        if (e->dff_stage[k]) HIPCHK(e, hipHostFree(e->dff_stage[k]));      // sized for the old count
        e->dff_stage[k] = NULL;
    }
    memset(&e->dff_desc, 0, sizeof(e->dff_desc));
    if (count == 0) return RSD_OK;
    for (int k = 0; k < 2; k++) {
        HIPCHK(e, hipMalloc((void **)&e->dff[k], (size_t)count * sizeof(float)));
        HIPCHK(e, hipMemcpyAsync(e->dff[k], icdf, (size_t)count * sizeof(float), hipMemcpyHostToDevice, e->stream));
        e->dff_n[k] = count;
    }
    HIPCHK(e, hipStreamSynchronize(e->stream));
    memcpy(e->dff_desc.s, desc16, sizeof(e->dff_desc.s));
    return RSD_OK;
}

extern "C" int rsd_set_dff(rsd_engine *e, const float *icdf, int count, const float *desc16) {
    int st = check_dff(e, -1, icdf, count, desc16);
    if (st) return st;
    HIPCHK(e, hipSetDevice(e->device));
    st = flush_pending(e);
    if (st) return st;
    join_batch(e);
    return dff_both(e, icdf, count, desc16);
}

extern "C" int rsd_set_dff_slot(rsd_engine *e, int slot, const float *icdf, int count, const float *desc16) {
    int st = check_dff(e, slot, icdf, count, desc16);
    if (st) return st;
    if (slot < 0) return fail(e, RSD_ERR_ARG, "wind slot %d", slot);
    HIPCHK(e, hipSetDevice(e->device));
    st = flush_pending(e);
    if (st) return st;
    join_batch(e);
    // first field, a removal or a new size: both slots at once, as the first
    // upload of the reference fills both buffers
    if (count == 0 || !e->dff[0] || !e->dff[1] || e->dff_n[0] != count || e->dff_n[1] != count)
        return dff_both(e, icdf, count, desc16);
    // a new field for one slot (the LES frame loaded into it): stream-ordered,
    // so launches already enqueued read the old table and the queue keeps running
    if (!e->ev_dff[slot]) HIPCHK(e, hipEventCreateWithFlags(&e->ev_dff[slot], hipEventDisableTiming));
    else HIPCHK(e, hipEventSynchronize(e->ev_dff[slot]));     // the staging buffer's last copy is done
    if (!e->dff_stage[slot]) HIPCHK(e, hipHostMalloc((void **)&e->dff_stage[slot], (size_t)count * sizeof(float), hipHostMallocDefault));
    memcpy(e->dff_stage[slot], icdf, (size_t)count * sizeof(float));
    HIPCHK(e, hipMemcpyAsync(e->dff[slot], e->dff_stage[slot], (size_t)count * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipEventRecord(e->ev_dff[slot], e->stream));
    memcpy(e->dff_desc.s, desc16, sizeof(e->dff_desc.s));
    return RSD_OK;
}

static int upload_floats(rsd_engine *e, float **dst, const float *src, int n) {
    dfree(*dst);
    *dst = NULL;
    HIPCHK(e, hipMalloc((void **)dst, (size_t)n * sizeof(float)));
    HIPCHK(e, hipMemcpyAsync(*dst, src, (size_t)n * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

extern "C" int rsd_set_range_weight(rsd_engine *e, const float *w, int n, const float *desc4) {
    if (!e || !w || !desc4 || n < 1 || n > 1024) return fail(e, RSD_ERR_ARG, "bad range weight table");
    if (!(desc4[2] <= (float)(n - 1))) return fail(e, RSD_ERR_ARG, "range weight max index beyond table");
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    st = upload_floats(e, &e->rangew, w, n);
    if (st) return st;
    memcpy(e->rangew_h, w, (size_t)n * sizeof(float));
    e->rangew_n = n;
    memcpy(e->rangew_desc.s, desc4, sizeof(e->rangew_desc.s));
    return RSD_OK;
}

extern "C" int rsd_set_angular_weight(rsd_engine *e, const float *w, int n, const float *desc4) {
    if (!e || !w || !desc4 || n < 1) return fail(e, RSD_ERR_ARG, "bad angular weight table");
    if (!(desc4[2] <= (float)(n - 1))) return fail(e, RSD_ERR_ARG, "angular weight max index beyond table");
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    st = upload_floats(e, &e->angw, w, n);
    if (st) return st;
    e->angw_n = n;
    memcpy(e->angw_desc.s, desc4, sizeof(e->angw_desc.s));
    return RSD_OK;
}

extern "C" int rsd_advance(rsd_engine *e, const rsd_step *step) {
    if (!e || !step) return fail(e, RSD_ERR_ARG, "null step");
    if (!e->pos) return fail(e, RSD_ERR_STATE, "state not allocated");
    int st = flush_pending(e);
    if (st) return st;
    e->pending = *step;
    e->have_pending = 1;
    e->bg_fixed = step->background_kind == RSD_ADVANCE_FIXED;
    return RSD_OK;
}

extern "C" int rsd_make_pulse(rsd_engine *e, const rsd_pulse *p) {
    if (!e || !p) return fail(e, RSD_ERR_ARG, "null pulse");
    if (!e->pos) return fail(e, RSD_ERR_STATE, "state not allocated");
    if (e->batch) {
        if (e->have_pending_pulse) {
            int st = rsd_batch_launch(e->batch);
            if (st) return st;
        }
        if (p->range_count < 1 || p->range_count > RSD_MAX_GATES) return fail(e, RSD_ERR_ARG, "range_count out of range");
        e->pending_pulse = *p;
        e->have_pending_pulse = 1;
        return RSD_OK;
    }
    join_batch(e);
    if (e->sort_interval > 0 && e->n_rain > 1 && !e->bg_fixed && (!e->permuted || e->since_sort >= e->sort_interval)) {
        HIPCHK(e, hipSetDevice(e->device));
        int st = sort_rain(e, e->have_pending ? e->pending.sim_desc : p->sim_desc, p->range_delta);
        if (st) return st;
    }
    e->since_sort++;
    if (e->have_pending) {
        e->have_pending = 0;
        return launch_step(e, &e->pending, p);
    }
    return launch_step(e, NULL, p);
}

extern "C" int rsd_make_pulse_noalloc(rsd_engine *e, const rsd_pulse *p, double *out) {
    if (!e || !p || !out) return fail(e, RSD_ERR_ARG, "null engine, pulse or output");
    if (!e->pos) return fail(e, RSD_ERR_STATE, "state not allocated");
    if (e->batch) return fail(e, RSD_ERR_STATE, "engine is in a batch");
    if (p->range_count < 1 || p->range_count > RSD_MAX_GATES)
        return fail(e, RSD_ERR_ARG, "range_count %u out of range (1..%d)", p->range_count, RSD_MAX_GATES);
    if (e->device >= RSD_SCRATCH_DEVICES || !scratch_dev[e->device])
        return fail(e, RSD_ERR_STATE, "no scratch address for device %d", e->device);
    StaticPulse sp;
    sp.gate_r = (float *)scratch_dev[e->device];
    sp.partials = (double *)((unsigned char *)scratch_dev[e->device] + RSD_SCRATCH_GATE_BYTES);
    sp.out = out;
    e->since_sort++;
    if (e->have_pending) {
        e->have_pending = 0;
        return launch_step(e, &e->pending, p, &sp);
    }
    return launch_step(e, NULL, p, &sp);
}

extern "C" int rsd_read_pulse(rsd_engine *e, double *out, uint32_t R) {
    if (!e || !out) return fail(e, RSD_ERR_ARG, "null output");
    if (e->batch && e->have_pending_pulse) {
        int st = rsd_batch_launch(e->batch);
        if (st) return st;
    }
    join_batch(e);
    if (!e->have_pulse) return fail(e, RSD_ERR_STATE, "no pulse made yet");
    if (R != e->last_pulse.range_count) return fail(e, RSD_ERR_ARG, "range_count %u != %u of the last pulse", R, e->last_pulse.range_count);
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_finalize(e);
    if (st) return st;
    HIPCHK(e, hipStreamSynchronize(e->stream));   // the finalize wrote pulse_h itself
    memcpy(out, e->pulse_h, (size_t)4 * R * sizeof(double));
    return RSD_OK;
}

extern "C" int rsd_sync(rsd_engine *e) {
    if (!e) return RSD_ERR_ARG;
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    join_batch(e);
    st = flush_finalize(e);
    if (st) return st;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return RSD_OK;
}

extern "C" int rsd_math_probe(int device, int fn, const float *a, const float *b, float *out, size_t n) {
    if (!a || !b || !out || n == 0) return RSD_ERR_ARG;
    if (fn == 11 && n % 64 != 0) return RSD_ERR_ARG;   // whole waves only
    if (device < 0 || device >= rsd_device_count()) return RSD_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return RSD_ERR_HIP;
    float *d = NULL;
    if (hipMalloc((void **)&d, 3 * n * sizeof(float)) != hipSuccess) return RSD_ERR_ALLOC;
    int rc = RSD_OK;
    if (hipMemcpy(d, a, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, b, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = RSD_ERR_HIP;
    if (rc == RSD_OK) {
        hipLaunchKernelGGL(k_math_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, (const float *)d,
                           (const float *)(d + n), d + 2 * n, (unsigned long long)n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = RSD_ERR_HIP;
    }
    if (rc == RSD_OK && hipMemcpy(out, d + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) rc = RSD_ERR_HIP;
    (void)hipFree(d);
    return rc;
}

extern "C" int rsd_io_test(rsd_engine *e, size_t n, int reps, double *gbps) {
    if (!e || n == 0 || reps < 1 || reps > 1024 || !gbps) return fail(e, RSD_ERR_ARG, "bad io test arguments");
    if ((n + 1023) / 1024 > 0x7fffffffull) return fail(e, RSD_ERR_ARG, "io test size %zu too large", n);
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    float4 *buf = NULL;
    HIPCHK(e, hipMalloc((void **)&buf, 2 * n * sizeof(float4)));
    hipEvent_t ev[2];
    hipError_t r = hipEventCreate(&ev[0]);
    if (r == hipSuccess) r = hipEventCreate(&ev[1]);
    if (r == hipSuccess) r = hipMemsetAsync(buf, 0, 2 * n * sizeof(float4), e->stream);
    const unsigned blocks = (unsigned)((n + 1023) / 1024);
    // every dispatch carries its own events (hipExtLaunchKernelGGL): the rate is
    // that of the copies themselves, not of the queue between them
    double total_ms = 0.0;
    for (int k = 0; k < reps + 1 && r == hipSuccess; k++) {
        hipExtLaunchKernelGGL(k_io, dim3(blocks), dim3(256), 0, e->stream, ev[0], ev[1], 0, (const float4 *)buf, buf + n,
                              (unsigned long long)n);
        r = hipGetLastError();
        if (r == hipSuccess) r = hipEventSynchronize(ev[1]);
        float ms = 0.0f;
        if (r == hipSuccess) r = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (k > 0) total_ms += ms;      // the first copy warms up
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    (void)hipFree(buf);
    if (r != hipSuccess) return fail(e, RSD_ERR_HIP, "io test: %s", hipGetErrorString(r));
    *gbps = 2.0 * (double)n * sizeof(float4) * reps / (total_ms * 1e-3) / 1e9;
    return RSD_OK;
}

extern "C" int rsd_profile(rsd_engine *e, int enable) {
    if (!e) return RSD_ERR_ARG;
    HIPCHK(e, hipSetDevice(e->device));
    if (enable && !e->prof_ev) {
        e->prof_ev = (hipEvent_t *)calloc(2 * RSD_PROF_CAP, sizeof(hipEvent_t));
        e->prof_kind = (int *)calloc(RSD_PROF_CAP, sizeof(int));
        if (!e->prof_ev || !e->prof_kind) return fail(e, RSD_ERR_ALLOC, "profile events");
        // timing only: no system-scope release when the dispatch completes (that
        // release writes the L2's dirty lines back and would lengthen the timed
        // launch beyond an untimed one)
        for (int i = 0; i < 2 * RSD_PROF_CAP; i++)
            HIPCHK(e, hipEventCreateWithFlags(&e->prof_ev[i], hipEventDisableSystemFence));
    }
    if (enable < 0) return fail(e, RSD_ERR_ARG, "profile stride %d", enable);
    e->prof_on = enable ? 1 : 0;
    e->prof_stride = enable > 0 ? enable : 1;
    e->prof_tick = 0;
    e->prof_n = 0;
    return RSD_OK;
}

extern "C" int rsd_profile_read(rsd_engine *e, int kind, double *total_ms, int *count) {
    if (!e || !total_ms || !count) return RSD_ERR_ARG;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    double t = 0.0;
    int c = 0;
    for (int i = 0; i < e->prof_n; i++) {
        if (kind >= 0 && e->prof_kind[i] != kind) continue;
        float ms = 0.0f;
        HIPCHK(e, hipEventElapsedTime(&ms, e->prof_ev[2 * i], e->prof_ev[2 * i + 1]));
        t += ms;
        c++;
    }
    *total_ms = t;
    *count = c;
    return RSD_OK;
}

extern "C" int rsd_timer_start(rsd_engine *e) {
    if (!e) return RSD_ERR_ARG;
    HIPCHK(e, hipEventRecord(e->ev0, e->stream));
    return RSD_OK;
}

extern "C" int rsd_timer_stop(rsd_engine *e, double *ms) {
    if (!e || !ms) return RSD_ERR_ARG;
    int st = flush_pending(e);
    if (st) return st;
    HIPCHK(e, hipEventRecord(e->ev1, e->stream));
    HIPCHK(e, hipEventSynchronize(e->ev1));
    float t = 0.0f;
    HIPCHK(e, hipEventElapsedTime(&t, e->ev0, e->ev1));
    *ms = (double)t;
    return RSD_OK;
}

// ---------------------------------------------------------------------------
// Batches: many engines (independent rain domains) advanced and pulsed by one
// persistent launch (k_rain_batch) and one finalize, optionally replayed from
// a captured HIP graph, with launch shape chosen by an online autotuner.
// ---------------------------------------------------------------------------

static const int TUNE_CHUNK[] = {1024, 2048, 4096};
static const int TUNE_GRID[] = {2, 4, 8};     // blocks per compute unit
#define RSD_TUNE_N 9

static void join_batch(rsd_engine *e) {
    if (e->batch && e->batch_wait) {
        (void)hipStreamWaitEvent(e->stream, e->batch->ev_launch, 0);
        e->batch_wait = 0;
    }
}

static int bfail(rsd_batch *b, int code, const char *msg) {
    snprintf(b->err, sizeof(b->err), "%s", msg);
    return code;
}

extern "C" const char *rsd_batch_last_error(const rsd_batch *b) { return b ? b->err : "no batch"; }

extern "C" int rsd_batch_create(rsd_engine *const *engines, int count, rsd_batch **out) {
    if (!engines || count < 1 || !out) return RSD_ERR_ARG;
    *out = NULL;
    const int dev = engines[0]->device;
    for (int i = 0; i < count; i++) {
        if (!engines[i] || engines[i]->device != dev || engines[i]->batch) return RSD_ERR_ARG;
        for (int j = 0; j < i; j++)
            if (engines[j] == engines[i]) return RSD_ERR_ARG;
    }
    if (hipSetDevice(dev) != hipSuccess) return RSD_ERR_HIP;
    for (int i = 0; i < count; i++)
        if (flush_finalize(engines[i]) != RSD_OK) return RSD_ERR_HIP;
    rsd_batch *b = (rsd_batch *)calloc(1, sizeof(rsd_batch));
    if (!b) return RSD_ERR_ALLOC;
    b->eng = (rsd_engine **)calloc((size_t)count, sizeof(rsd_engine *));
    if (!b->eng) {
        free(b);
        return RSD_ERR_ALLOC;
    }
    memcpy(b->eng, engines, (size_t)count * sizeof(rsd_engine *));
    b->nd = count;
    b->device = dev;
    b->use_graph = 1;
    b->tuning = 1;
    b->tune_best = 1e30;
    hipDeviceProp_t prop;
    if (hipSetDevice(dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_buf[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_buf[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_launch, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->tev0) != hipSuccess || hipEventCreate(&b->tev1) != hipSuccess ||
        hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess) {
        free(b->eng);
        free(b);
        return RSD_ERR_HIP;
    }
    b->cus = prop.multiProcessorCount;
    (void)hipEventRecord(b->ev_buf[0], b->stream);
    (void)hipEventRecord(b->ev_buf[1], b->stream);
    (void)hipEventRecord(b->ev_launch, b->stream);
    for (int i = 0; i < count; i++) {
        rsd_engine *e = engines[i];
        // wait for whatever the engine's stream still holds (its last pulse's
        // finalize included), then stagger the domains' range reorders over the
        // sort interval
        (void)hipEventRecord(e->ev_fork, e->stream);
        (void)hipStreamWaitEvent(b->stream, e->ev_fork, 0);
        if (e->sort_interval > 0 && e->permuted) e->since_sort = (int)((long)i * e->sort_interval / count);
        e->batch = b;
    }
    *out = b;
    return RSD_OK;
}

extern "C" void rsd_batch_destroy(rsd_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (int i = 0; i < b->nd; i++) {
        rsd_engine *e = b->eng[i];
        if (e->have_pending_pulse) (void)rsd_batch_launch(b);
    }
    (void)hipStreamSynchronize(b->stream);
    for (int i = 0; i < b->nd; i++) {
        b->eng[i]->batch = NULL;
        b->eng[i]->batch_wait = 0;
    }
    for (int k = 0; k < 2; k++) {
        if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
        if (b->desc_h[k]) (void)hipHostFree(b->desc_h[k]);
        (void)hipEventDestroy(b->ev_buf[k]);
    }
    dfree(b->desc_d);
    dfree(b->items_d);
    (void)hipEventDestroy(b->ev_launch);
    (void)hipEventDestroy(b->tev0);
    (void)hipEventDestroy(b->tev1);
    (void)hipEventDestroy(b->t0);
    (void)hipEventDestroy(b->t1);
    (void)hipStreamDestroy(b->stream);
    free(b->eng);
    free(b);
}

extern "C" int rsd_batch_set_launch(rsd_batch *b, int chunk, int blocks_per_cu) {
    if (!b) return RSD_ERR_ARG;
    if (chunk != 0 && (chunk < 256 || chunk % 256 != 0 || chunk > (1 << 20))) return bfail(b, RSD_ERR_ARG, "chunk must be a multiple of 256");
    if (blocks_per_cu < 0 || blocks_per_cu > 64) return bfail(b, RSD_ERR_ARG, "blocks per CU out of range");
    b->chunk_req = chunk;
    b->grid_req = blocks_per_cu;
    b->tuning = (chunk == 0 || blocks_per_cu == 0);
    b->tune_idx = b->tune_rep = b->tune_pending = 0;
    b->tune_sum = 0.0;
    b->tune_best = 1e30;
    return RSD_OK;
}

extern "C" int rsd_batch_set_graph(rsd_batch *b, int on) {
    if (!b) return RSD_ERR_ARG;
    b->use_graph = on ? 1 : 0;
    return RSD_OK;
}

extern "C" int rsd_batch_get_launch(rsd_batch *b, int *chunk, int *blocks_per_cu, int *tuning) {
    if (!b) return RSD_ERR_ARG;
    if (chunk) *chunk = b->tuning ? 0 : b->best_chunk;
    if (blocks_per_cu) *blocks_per_cu = b->tuning ? 0 : b->best_grid;
    if (tuning) *tuning = b->tuning;
    return RSD_OK;
}

// Launch shape of this launch (and autotuner bookkeeping of the previous one).
static void batch_shape(rsd_batch *b, int *chunk, int *bpc) {
    if (!b->tuning) {
        *chunk = b->chunk_req ? b->chunk_req : b->best_chunk;
        *bpc = b->grid_req ? b->grid_req : b->best_grid;
        return;
    }
    if (b->tune_pending) {   // time of the previous tuning launch
        float ms = 0.0f;
        (void)hipEventSynchronize(b->tev1);
        (void)hipEventElapsedTime(&ms, b->tev0, b->tev1);
        if (b->tune_rep > 0) b->tune_sum += ms;   // the first launch of a shape warms it up
        b->tune_rep++;
        b->tune_pending = 0;
        if (b->tune_rep == RSD_TUNE_REPS) {
            const double avg = b->tune_sum / (RSD_TUNE_REPS - 1);
            const int c = b->chunk_req ? b->chunk_req : TUNE_CHUNK[b->tune_idx / 3];
            const int g = b->grid_req ? b->grid_req : TUNE_GRID[b->tune_idx % 3];
            if (avg < b->tune_best) {
                b->tune_best = avg;
                b->best_chunk = c;
                b->best_grid = g;
            }
            b->tune_idx++;
            b->tune_rep = 0;
            b->tune_sum = 0.0;
            if (b->tune_idx == RSD_TUNE_N || (b->chunk_req && b->tune_idx % 3 == 0 && b->tune_idx >= 3) ||
                (b->grid_req && b->tune_idx >= 3)) {
                b->tuning = 0;
                *chunk = b->chunk_req ? b->chunk_req : b->best_chunk;
                *bpc = b->grid_req ? b->grid_req : b->best_grid;
                return;
            }
        }
    }
    int idx = b->tune_idx;
    if (b->grid_req) idx *= 3;            // only the chunk varies
    *chunk = b->chunk_req ? b->chunk_req : TUNE_CHUNK[(idx / 3) % 3];
    *bpc = b->grid_req ? b->grid_req : TUNE_GRID[idx % 3];
}

extern "C" int rsd_batch_launch(rsd_batch *b) {
    if (!b) return RSD_ERR_ARG;
    if (hipSetDevice(b->device) != hipSuccess) return bfail(b, RSD_ERR_HIP, "hipSetDevice");
    // Engines with a pending pulse: rain-only, binary32 positions go through the
    // batched launch; anything else is launched on its own stream.
    rsd_engine *sel[1024];
    int ns = 0, any_adv = -1, mixed = 0;
    for (int i = 0; i < b->nd; i++) {
        rsd_engine *e = b->eng[i];
        if (!e->have_pending_pulse) continue;
        const int batched = (e->n == e->n_rain) && !e->plo && e->n_rain > 0 && ns < 1024;
        int adv = e->have_pending ? 1 : 0;
        if (batched && (any_adv < 0 || any_adv == adv)) {
            any_adv = adv;
            sel[ns++] = e;
        } else {
            mixed = 1;
            e->have_pending_pulse = 0;
            int st;
            if (e->have_pending) {
                e->have_pending = 0;
                st = launch_step(e, &e->pending, &e->pending_pulse);
            } else {
                st = launch_step(e, NULL, &e->pending_pulse);
            }
            if (st) return bfail(b, st, e->err);
        }
    }
    (void)mixed;
    if (ns == 0) return RSD_OK;
    // A domain due for its range reorder sits the batched launch out: the sort
    // and the domain's own fused launch run on its stream beside the batch
    // kernel, whose work items for it find no scatterers and whose finalize
    // leaves its pulse alone (the item list, and so the graph, stay as they
    // are); the next round's batch waits for that stream.  Only those streams
    // have new work the batch must wait for (engine transfers are synchronous
    // and fallback launches are not in the batch).
    unsigned char sat_out[1024];
    for (int k = 0; k < ns; k++) {
        rsd_engine *e = sel[k];
        sat_out[k] = 0;
        if (e->batch_rejoin) {
            if (hipEventRecord(e->ev_fork, e->stream) != hipSuccess ||
                hipStreamWaitEvent(b->stream, e->ev_fork, 0) != hipSuccess)
                return bfail(b, RSD_ERR_HIP, "stream join");
            e->batch_rejoin = 0;
        }
        if (e->sort_interval > 0 && e->n_rain > 1 && (!e->permuted || e->since_sort >= e->sort_interval)) {
            join_batch(e);
            int st = sort_rain(e, e->have_pending ? e->pending.sim_desc : e->pending_pulse.sim_desc,
                               e->pending_pulse.range_delta);
            if (st) return bfail(b, st, e->err);
            sat_out[k] = 1;
            e->have_pending_pulse = 0;
            const bool adv = e->have_pending != 0;
            e->have_pending = 0;
            st = launch_step(e, adv ? &e->pending : NULL, &e->pending_pulse);
            if (st) return bfail(b, st, e->err);
            e->batch_rejoin = 1;
        }
        e->since_sort++;
    }
    int chunk, bpc;
    batch_shape(b, &chunk, &bpc);
    // work items and partial buffers
    size_t sig = (size_t)chunk * 1315423911u;
    int n_items = 0;
    for (int k = 0; k < ns; k++) {
        const int rows = (int)((sel[k]->n_rain + chunk - 1) / chunk);
        n_items += rows;
        sig = sig * 31u + (size_t)sel[k]->n_rain + (size_t)(uintptr_t)sel[k];
    }
    if (sig != b->items_sig || n_items != b->n_items) {
        if (n_items > b->items_cap) {
            dfree(b->items_d);
            b->items_d = NULL;
            if (hipMalloc((void **)&b->items_d, (size_t)n_items * sizeof(int2)) != hipSuccess) return bfail(b, RSD_ERR_ALLOC, "items");
            b->items_cap = n_items;
        }
        int2 *h = (int2 *)malloc((size_t)n_items * sizeof(int2));
        if (!h) return bfail(b, RSD_ERR_ALLOC, "items");
        int t = 0;
        for (int k = 0; k < ns; k++) {
            const int rows = (int)((sel[k]->n_rain + chunk - 1) / chunk);
            for (int c = 0; c < rows; c++) h[t++] = make_int2(k, c);
        }
        hipError_t r = hipMemcpyAsync(b->items_d, h, (size_t)n_items * sizeof(int2), hipMemcpyHostToDevice, b->stream);
        if (r == hipSuccess) r = hipStreamSynchronize(b->stream);
        free(h);
        if (r != hipSuccess) return bfail(b, RSD_ERR_HIP, "items upload");
        b->items_sig = sig;
        b->n_items = n_items;
    }
    int grid = bpc * b->cus;
    if (grid > n_items) grid = n_items;
    if (grid >= 8) grid &= ~7;
    // descriptors
    const size_t rk_bytes = (size_t)ns * sizeof(RainK), fk_bytes = (size_t)ns * sizeof(FinK);
    const size_t desc_bytes = rk_bytes + fk_bytes + (size_t)(ns + 1) * sizeof(int);
    if (desc_bytes > b->desc_cap) {
        (void)hipStreamSynchronize(b->stream);
        for (int k2 = 0; k2 < 2; k2++) {
            if (b->desc_h[k2]) (void)hipHostFree(b->desc_h[k2]);
            b->desc_h[k2] = NULL;
            if (b->gexec[k2]) (void)hipGraphExecDestroy(b->gexec[k2]);
            b->gexec[k2] = NULL;
            b->gsig[k2] = 0;
        }
        dfree(b->desc_d);
        b->desc_d = NULL;
        if (hipHostMalloc((void **)&b->desc_h[0], desc_bytes, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&b->desc_h[1], desc_bytes, hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void **)&b->desc_d, desc_bytes) != hipSuccess)
            return bfail(b, RSD_ERR_ALLOC, "descriptors");
        b->desc_cap = desc_bytes;
    }
    const int buf = b->buf;
    if (hipEventSynchronize(b->ev_buf[buf]) != hipSuccess) return bfail(b, RSD_ERR_HIP, "descriptor buffer wait");
    RainK *rk = (RainK *)b->desc_h[buf];
    FinK *fk = (FinK *)(b->desc_h[buf] + rk_bytes);
    int *first = (int *)(b->desc_h[buf] + rk_bytes + fk_bytes);
    size_t lds = 0;
    int total4r = 0;
    bool fast = any_adv == 1;
    for (int k = 0; k < ns; k++) {
        rsd_engine *e = sel[k];
        const rsd_pulse *p = &e->pending_pulse;
        const int R = (int)p->range_count;
        const int rows = (int)((e->n_rain + chunk - 1) / chunk);
        if (fast && !(rain_fast(e, &e->pending, p) && !e->plo)) fast = false;
        int st = ensure_partials(e, (size_t)rows, R);
        if (st) return bfail(b, st, e->err);
        if (any_adv && e->wind_ok[e->pending.wind_slot]) {
            // index tables copied on the engine's stream: wait for a rebuilt
            // one here, the batch runs on its own stream
            WIdx w;
            int rebuilt = 0;
            st = wind_itab(e, e->pending.wind_slot, e->pending.sim_desc, &w, &rebuilt);
            if (st) return bfail(b, st, e->err);
            if (rebuilt && hipEventSynchronize(e->ev_itab[e->pending.wind_slot]) != hipSuccess)
                return bfail(b, RSD_ERR_HIP, "index table copy");
        }
        RainK a;
        memset(&a, 0, sizeof(a));
        st = fill_rain(e, any_adv ? &e->pending : NULL, a);
        if (st) return bfail(b, st, e->err);
        if (a.wind.d2) return bfail(b, RSD_ERR_STATE, "batched domains take no time-interpolated wind");
        st = fill_pulse(e, p, a.pk);
        if (st) return bfail(b, st, e->err);
        a.pk.row0 = 0;
        a.pk.rows = rows;
        a.plo = NULL;
        if (sat_out[k]) a.n = 0;
        rk[k] = a;
        FinK f;
        f.partials = e->partials;
        f.rows = rows;
        f.n4r = 4 * R;
        f.out = sat_out[k] ? NULL : e->pulse_d;
        f.out2 = NULL;
        if (!sat_out[k] && e->ring && e->ring_R == R) {
            f.out2 = e->ring + (size_t)e->ring_next * 4 * R;
            e->ring_next = (e->ring_next + 1) % e->ring_m;
            if (e->ring_count < e->ring_m) e->ring_count++;
        }
        fk[k] = f;
        first[k] = total4r;
        total4r += 4 * R;
        size_t l = (size_t)4 * R * sizeof(double) + (a.pk.fold == RSD_FOLD_WINDOW ? (size_t)e->rangew_n * sizeof(float2) + (size_t)R * sizeof(float) : 0);
        l = set_rain_tables(e, rk[k], R, l);
        if (!rk[k].tab_lds) fast = false;
        if (l > lds) lds = l;
    }
    first[ns] = total4r;
    const RainK *rk_d = (const RainK *)b->desc_d;
    const FinK *fk_d = (const FinK *)(b->desc_d + rk_bytes);
    const int *first_d = (const int *)(b->desc_d + rk_bytes + fk_bytes);
    const bool timed = b->tuning;
    // everything the graph depends on besides descriptor contents
    const size_t gsig = (((((size_t)n_items * 131u + (size_t)grid) * 131u + lds) * 131u + (size_t)chunk) * 131u +
                         (size_t)total4r) * 8u + (size_t)fast * 4u + (size_t)any_adv * 2u + 1u + (desc_bytes << 20);
    auto enqueue = [&](hipStream_t st) -> hipError_t {
        hipError_t r = hipMemcpyAsync(b->desc_d, b->desc_h[buf], desc_bytes, hipMemcpyHostToDevice, st);
        if (r != hipSuccess) return r;
        if (fast)
            hipLaunchKernelGGL((k_rain_batch<true, true, true>), dim3(grid), dim3(256), lds, st, rk_d,
                               (const int2 *)b->items_d, n_items, chunk);
        else if (any_adv)
            hipLaunchKernelGGL((k_rain_batch<true, true>), dim3(grid), dim3(256), lds, st, rk_d,
                               (const int2 *)b->items_d, n_items, chunk);
        else
            hipLaunchKernelGGL((k_rain_batch<false, true>), dim3(grid), dim3(256), lds, st, rk_d,
                               (const int2 *)b->items_d, n_items, chunk);
        r = hipGetLastError();
        if (r != hipSuccess) return r;
        hipLaunchKernelGGL(k_finalize_batch, dim3(total4r), dim3(256), 0, st, fk_d, first_d, ns);
        return hipGetLastError();
    };
    if (timed) (void)hipEventRecord(b->tev0, b->stream);
    if (b->use_graph && !timed) {
        if (!b->gexec[buf] || b->gsig[buf] != gsig) {
            if (b->gexec[buf]) (void)hipGraphExecDestroy(b->gexec[buf]);
            b->gexec[buf] = NULL;
            hipGraph_t g = NULL;
            if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return bfail(b, RSD_ERR_HIP, "begin capture");
            hipError_t r = enqueue(b->stream);
            hipError_t r2 = hipStreamEndCapture(b->stream, &g);
            if (r != hipSuccess || r2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(b, RSD_ERR_HIP, "capture");
            }
            r = hipGraphInstantiate(&b->gexec[buf], g, NULL, NULL, 0);
            (void)hipGraphDestroy(g);
            if (r != hipSuccess) return bfail(b, RSD_ERR_HIP, "graph instantiate");
            b->gsig[buf] = gsig;
        }
        if (hipGraphLaunch(b->gexec[buf], b->stream) != hipSuccess) return bfail(b, RSD_ERR_HIP, "graph launch");
    } else {
        if (enqueue(b->stream) != hipSuccess) return bfail(b, RSD_ERR_HIP, "batch launch");
    }
    if (timed) {
        (void)hipEventRecord(b->tev1, b->stream);
        b->tune_pending = 1;
    }
    if (hipEventRecord(b->ev_buf[buf], b->stream) != hipSuccess || hipEventRecord(b->ev_launch, b->stream) != hipSuccess)
        return bfail(b, RSD_ERR_HIP, "event");
    b->buf ^= 1;
    for (int k = 0; k < ns; k++) {
        rsd_engine *e = sel[k];
        e->have_pending = 0;
        e->have_pending_pulse = 0;
        e->last_pulse = e->pending_pulse;
        e->have_pulse = 1;
        e->sigaux_saved = 0;
        e->batch_wait = 1;
    }
    return RSD_OK;
}

extern "C" int rsd_batch_sync(rsd_batch *b) {
    if (!b) return RSD_ERR_ARG;
    for (int i = 0; i < b->nd; i++)
        if (b->eng[i]->have_pending_pulse) {
            int st = rsd_batch_launch(b);
            if (st) return st;
            break;
        }
    if (hipSetDevice(b->device) != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess)
        return bfail(b, RSD_ERR_HIP, "sync");
    for (int i = 0; i < b->nd; i++) {
        int st = rsd_sync(b->eng[i]);
        if (st) return bfail(b, st, b->eng[i]->err);
    }
    return RSD_OK;
}

extern "C" int rsd_batch_timer_start(rsd_batch *b) {
    if (!b) return RSD_ERR_ARG;
    (void)hipSetDevice(b->device);
    return hipEventRecord(b->t0, b->stream) == hipSuccess ? RSD_OK : RSD_ERR_HIP;
}

extern "C" int rsd_batch_timer_stop(rsd_batch *b, double *ms) {
    if (!b || !ms) return RSD_ERR_ARG;
    for (int i = 0; i < b->nd; i++)
        if (b->eng[i]->have_pending_pulse) {
            int st = rsd_batch_launch(b);
            if (st) return st;
            break;
        }
    float f = 0.0f;
    if (hipEventRecord(b->t1, b->stream) != hipSuccess || hipEventSynchronize(b->t1) != hipSuccess ||
        hipEventElapsedTime(&f, b->t0, b->t1) != hipSuccess)
        return RSD_ERR_HIP;
    *ms = f;
    return RSD_OK;
}

// ---------------------------------------------------------------------------
// Training samples of the surrogate network from the table path.
// ---------------------------------------------------------------------------
struct SampK {
    Tex2 cd, cm, rr, ri;
    F16 ad, rd;
    unsigned long long n, seed;
    float *x, *t;           // [n][7], [n][12]
};

__device__ __forceinline__ float samp_uniform(uint64_t &s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((float)(z >> 40) + 0.5f) * 0x1.0p-24f;
}

__global__ void k_debris_samples(SampK a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    uint64_t s = a.seed * 0xD1B54A32D192ED03ull + (uint64_t)i * 0x9E3779B97F4A7C15ull;
    float g[8];
    for (int k = 0; k < 8; k += 2) {     // Box-Muller pairs
        const float u1 = samp_uniform(s), u2 = samp_uniform(s);
        const float rr = __builtin_sqrtf(-2.0f * __logf(u1));
        float sn, cs;
        __sincosf(6.2831853f * u2, &sn, &cs);
        g[k] = rr * cs;
        g[k + 1] = rr * sn;
    }
    float qn = __builtin_sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3]);
    float4 R = make_float4(g[0] / qn, g[1] / qn, g[2] / qn, g[3] / qn);
    if (R.w < 0.0f) R = make_float4(-R.x, -R.y, -R.z, -R.w);
    float un = __builtin_sqrtf(g[4] * g[4] + g[5] * g[5] + g[6] * g[6]);
    float4 uh = make_float4(g[4] / un, g[5] / un, g[6] / un, 0.0f);
    float t[12];
    table_targets(R, uh, a.cd, a.cm, a.ad, a.rr, a.ri, a.rd, t);
    float *x = a.x + i * MLP_IN;
    x[0] = R.x; x[1] = R.y; x[2] = R.z; x[3] = R.w;
    x[4] = uh.x; x[5] = uh.y; x[6] = uh.z;
    for (int k = 0; k < 12; k++) a.t[i * 12 + k] = t[k];
}

extern "C" int rsd_debris_samples(rsd_engine *e, int adm_table, int rcs_table, size_t n, uint64_t seed, float *x,
                                  float *t) {
    if (!e || !x || !t || n == 0) return RSD_ERR_ARG;
    if (adm_table < 0 || adm_table >= RSD_MAX_TABLES || !e->adm_cd[adm_table]) return fail(e, RSD_ERR_STATE, "ADM table %d not set", adm_table);
    if (rcs_table < 0 || rcs_table >= RSD_MAX_TABLES || !e->rcs_re[rcs_table]) return fail(e, RSD_ERR_STATE, "RCS table %d not set", rcs_table);
    HIPCHK(e, hipSetDevice(e->device));
    SampK a;
    a.cd.d = e->adm_cd[adm_table]; a.cm.d = e->adm_cm[adm_table];
    a.cd.nx = a.cm.nx = e->adm_nx[adm_table]; a.cd.ny = a.cm.ny = e->adm_ny[adm_table];
    a.ad = e->adm_desc[adm_table];
    a.rr.d = e->rcs_re[rcs_table]; a.ri.d = e->rcs_im[rcs_table];
    a.rr.nx = a.ri.nx = e->rcs_nx[rcs_table]; a.rr.ny = a.ri.ny = e->rcs_ny[rcs_table];
    a.rd = e->rcs_desc[rcs_table];
    a.n = n;
    a.seed = seed;
    HIPCHK(e, hipMalloc((void **)&a.x, n * MLP_IN * sizeof(float)));
    hipError_t r = hipMalloc((void **)&a.t, n * 12 * sizeof(float));
    if (r == hipSuccess) {
        hipLaunchKernelGGL(k_debris_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, a);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipMemcpyAsync(x, a.x, n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost, e->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(t, a.t, n * 12 * sizeof(float), hipMemcpyDeviceToHost, e->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    dfree(a.x);
    dfree(a.t);
    HIPCHK(e, r);
    return RSD_OK;
}

// Training samples at the population's own states: for every body of one
// debris type, the network inputs the debris kernel would form (R from the
// position and orientation, u from the velocity and the wind at the position)
// and the table path's targets at them.  A body in still air relative to the
// wind (u undefined) gets u = (1, 0, 0) and the targets there.
struct StateSampK {
    const float4 *pos, *ori, *vel;
    unsigned long long begin, n;
    Tex3 wind;
    F16 les, sd;
    WIdx widx;
    Tex2 cd, cm, rr, ri;
    F16 ad, rd;
    float *x, *t;
};

__global__ void k_debris_state_samples(StateSampK a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float4 p = a.pos[a.begin + i], o = a.ori[a.begin + i], v = a.vel[a.begin + i];
    const float4 c = wind_index(p, a.les, a.sd, a.widx.t, a.widx);
    const float4 bg = sample3d(a.wind, c.x, c.y, c.z);
    const float4 ur = make_float4(bg.x - v.x, bg.y - v.y, bg.z - v.z, bg.w - v.w);
    float4 uh = qrot(norm4(ur), qconj(o));
    if (!(uh.x == uh.x && uh.y == uh.y && uh.z == uh.z)) uh = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    const float4 R = radar_rel_quat(p, o);
    float t[12];
    table_targets(R, uh, a.cd, a.cm, a.ad, a.rr, a.ri, a.rd, t);
    float *x = a.x + i * MLP_IN;
    x[0] = R.x; x[1] = R.y; x[2] = R.z; x[3] = R.w;
    x[4] = uh.x; x[5] = uh.y; x[6] = uh.z;
    for (int k = 0; k < 12; k++) a.t[i * 12 + k] = t[k];
}

extern "C" int rsd_debris_state_samples(rsd_engine *e, int type, int adm_table, int rcs_table, int wind_slot,
                                        const float *sim_desc, size_t cap, float *x, float *t, size_t *n_out) {
    if (!e || !x || !t || !n_out || !sim_desc || type < 1 || type >= RSD_MAX_TYPES) return RSD_ERR_ARG;
    if (adm_table < 0 || adm_table >= RSD_MAX_TABLES || !e->adm_cd[adm_table]) return fail(e, RSD_ERR_STATE, "ADM table %d not set", adm_table);
    if (rcs_table < 0 || rcs_table >= RSD_MAX_TABLES || !e->rcs_re[rcs_table]) return fail(e, RSD_ERR_STATE, "RCS table %d not set", rcs_table);
    if (wind_slot < 0 || wind_slot > 1 || !e->wind_ok[wind_slot]) return fail(e, RSD_ERR_STATE, "wind slot %d not set", wind_slot);
    if (!e->pos || !e->ori || !e->vel) return fail(e, RSD_ERR_STATE, "state not allocated");
    HIPCHK(e, hipSetDevice(e->device));
    int st = flush_pending(e);
    if (st) return st;
    const size_t n = e->counts[type] < cap ? e->counts[type] : cap;
    *n_out = n;
    if (n == 0) return RSD_OK;
    StateSampK a;
    memset(&a, 0, sizeof(a));
    a.pos = e->pos; a.ori = e->ori; a.vel = e->vel;
    a.begin = e->origins[type];
    a.n = n;
    a.wind = tex3_of(e->wind[wind_slot], e->wnx[wind_slot], e->wny[wind_slot], e->wnz[wind_slot]);
    a.les = e->les[wind_slot];
    memcpy(a.sd.s, sim_desc, sizeof(a.sd.s));
    st = wind_itab(e, wind_slot, sim_desc, &a.widx, NULL);
    if (st) return st;
    a.cd.d = e->adm_cd[adm_table]; a.cm.d = e->adm_cm[adm_table];
    a.cd.nx = a.cm.nx = e->adm_nx[adm_table]; a.cd.ny = a.cm.ny = e->adm_ny[adm_table];
    a.ad = e->adm_desc[adm_table];
    a.rr.d = e->rcs_re[rcs_table]; a.ri.d = e->rcs_im[rcs_table];
    a.rr.nx = a.ri.nx = e->rcs_nx[rcs_table]; a.rr.ny = a.ri.ny = e->rcs_ny[rcs_table];
    a.rd = e->rcs_desc[rcs_table];
    HIPCHK(e, hipMalloc((void **)&a.x, n * MLP_IN * sizeof(float)));
    hipError_t r = hipMalloc((void **)&a.t, n * 12 * sizeof(float));
    if (r == hipSuccess) {
        hipLaunchKernelGGL(k_debris_state_samples, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, a);
        r = hipGetLastError();
    }
    if (r == hipSuccess) r = hipMemcpyAsync(x, a.x, n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost, e->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(t, a.t, n * 12 * sizeof(float), hipMemcpyDeviceToHost, e->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    dfree(a.x);
    dfree(a.t);
    HIPCHK(e, r);
    return RSD_OK;
}

extern "C" int rsd_set_debris_surrogate(rsd_engine *e, int type, const rsd_mlp *m) {
    if (!e || type < 1 || type >= RSD_MAX_TYPES) return fail(e, RSD_ERR_ARG, "debris type %d out of range", type);
    int st = flush_pending(e);
    if (st) return st;
    if (!m) {
        e->surr[type] = 0;
        return RSD_OK;
    }
    int dev = -1;
    MlpDev v;
    if (rsd_mlp_device_view(m, &v, &dev) != RSD_OK || dev != e->device) return fail(e, RSD_ERR_ARG, "network on another device");
    e->net[type] = v;
    e->surr[type] = 1;
    return RSD_OK;
}
