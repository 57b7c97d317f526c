/*
 * rs_oracle.c -- CPU oracle for SimRadar's per-pulse scatterer engine.
 *
 * TEST INFRASTRUCTURE.  Built by oracle/Makefile into oracle/_build/liboracle.so;
 * only tests/, the smoke check and bench.py's cpu_baseline leg load it.
 *
 * Documented choices where the reference leaves the result to the OpenCL device
 * (rs.cl compiled by the driver; no device here can run it):
 *   - elementary functions: ofmath.h (one fixed algorithm each);
 *   - OpenCL length / normalize / dot: plain left-to-right sums of squares, a
 *     correctly rounded sqrt and a correctly rounded division per component;
 *   - a*b + c written without fma in rs.cl is evaluated unfused; fma() in rs.cl is
 *     a fused multiply-add here too;
 *   - linear image filtering (rs.cl:8): texel pair i0 = floor(u - 0.5), i1 = i0 + 1,
 *     both clamped to the edge, weight a = (u - 0.5) - i0 in binary32, blended by
 *     lerp(p, q, a) = fma(a, q, fma(-a, p, p)), x first, then y, then z;
 *   - pown(r, -2) = 1 / (r * r);
 *   - per-gate sums in binary64 (the reference's pass-1/pass-2 order depends on
 *     the device's work-group shape).
 */
#include "rs_oracle.h"
#include "ofmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MIN_HEIGHT 2.0f
#define ORC_PI_F 0x1.921fb6p+1f
#define ORC_PI_2_F 0x1.921fb6p+0f
#define ORC_SQRT1_2_F 0x1.6a09e6p-1f

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

void orc_rand4(ou4 *s, of4 *r) {
    /* rs.cl:35-44: s <- (s * 16807) & 0x7FFFFFFF in uint32, r = float(s) / 2147483647 */
    const float n = 1.0f / 2147483647.0f;
    s->x = (s->x * 16807u) & 0x7FFFFFFFu;
    s->y = (s->y * 16807u) & 0x7FFFFFFFu;
    s->z = (s->z * 16807u) & 0x7FFFFFFFu;
    s->w = (s->w * 16807u) & 0x7FFFFFFFu;
    r->x = (float)s->x * n;
    r->y = (float)s->y * n;
    r->z = (float)s->z * n;
    r->w = (float)s->w * n;
}

/* Stretched-grid index of rs.cl:130-151, c = s8 + copysign(s0 log1p(s4 |r|), r),
 * as the engine documents its evaluation (include/rsd.h rsd_wind_index_table):
 * P(x) = s0 log1p(s4 x) from ORC_ITAB_N cubic pieces per axis over [0, X], X the
 * box's extent from the axis origin times 1 + 2^-10, each piece the binary64
 * cubic through P at the four Chebyshev nodes u_k = 1/2 + cos((2k + 1) pi / 8) / 2
 * of its local coordinate f (x = (j + f) / ih), coefficients in binary32. */
#define ORC_ITAB_N 256

static void itab_piece_axis(double s0, double s4, float ihf, float *rows) {
    const double k1 = 0x1.d906bcf328d46p-1, k3 = 0x1.87de2a6aea964p-2;   /* cos(pi/8), cos(3 pi/8) */
    const double u[4] = {0.5 + 0.5 * k1, 0.5 + 0.5 * k3, 0.5 - 0.5 * k3, 0.5 - 0.5 * k1};
    const double ih = (double)ihf;
    for (int j = 0; j < ORC_ITAB_N; j++) {
        double y[4];
        for (int k = 0; k < 4; k++) y[k] = s0 * log1p(s4 * (((double)j + u[k]) / ih));
        /* Newton divided differences, then the Newton form expanded in powers of f */
        const double d01 = (y[1] - y[0]) / (u[1] - u[0]);
        const double d12 = (y[2] - y[1]) / (u[2] - u[1]);
        const double d23 = (y[3] - y[2]) / (u[3] - u[2]);
        const double d012 = (d12 - d01) / (u[2] - u[0]);
        const double d123 = (d23 - d12) / (u[3] - u[1]);
        const double d0123 = (d123 - d012) / (u[3] - u[0]);
        float *r = rows + 4 * j;
        r[0] = (float)(((y[0] - d01 * u[0]) + d012 * (u[0] * u[1])) - d0123 * ((u[0] * u[1]) * u[2]));
        r[1] = (float)((d01 - d012 * (u[0] + u[1])) + d0123 * ((u[0] * u[1] + u[0] * u[2]) + u[1] * u[2]));
        r[2] = (float)(d012 - d0123 * ((u[0] + u[1]) + u[2]));
        r[3] = (float)d0123;
    }
}

void orc_wind_index_table(const float *d, const float *sd, float *rows, float *ih, float *jm) {
    const float org[3] = {sd[8] + 0.5f * sd[12], sd[9] + 0.5f * sd[13], 0.0f};
    for (int a = 0; a < 3; a++) {
        const float lo = sd[8 + a];
        const float hi = sd[8 + a] + sd[12 + a];
        double X = fabs((double)lo - (double)org[a]);
        const double Xh = fabs((double)hi - (double)org[a]);
        if (Xh > X) X = Xh;
        X = X * (1.0 + 0x1p-10);
        if (X < 1.0e-3) X = 1.0e-3;
        ih[a] = (float)((double)ORC_ITAB_N / X);
        jm[a] = (float)(ORC_ITAB_N - 1);
        itab_piece_axis((double)d[a], (double)d[4 + a], ih[a], rows + 4 * ORC_ITAB_N * a);
    }
}

static inline float itab_eval(float r, const float *T, float ih, float jm, float base) {
    const float t = fabsf(r) * ih;
    const float jf = fminf(floorf(t), jm);
    const float f = t - jf;
    const float *c = T + 4 * (int)jf;
    const float p = fmaf(fmaf(fmaf(c[3], f, c[2]), f, c[1]), f, c[0]);
    return copysignf(p, r) + base;
}

/* The tables of the last (descriptor, box) seen by this thread. */
static __thread float itab_key[22];
static __thread int itab_have;
static __thread float itab_rows[3 * ORC_ITAB_N * 4], itab_ih[3], itab_jm[3];

of4 orc_wind_index(of4 pos, const float *d, const float *sd) {
    of4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t spacing = fbits(d[7]);
    if (spacing == 7u) {
        /* rs.cl:130-151: relative to the box centre in x, y and to 0 in z */
        float key[22];
        memcpy(key, d, 16 * sizeof(float));
        memcpy(key + 16, sd + 8, 3 * sizeof(float));
        memcpy(key + 19, sd + 12, 3 * sizeof(float));
        if (!itab_have || memcmp(key, itab_key, sizeof(key)) != 0) {
            orc_wind_index_table(d, sd, itab_rows, itab_ih, itab_jm);
            memcpy(itab_key, key, sizeof(key));
            itab_have = 1;
        }
        float cx = sd[8] + 0.5f * sd[12];
        float cy = sd[9] + 0.5f * sd[13];
        c.x = itab_eval(pos.x - cx, itab_rows, itab_ih[0], itab_jm[0], d[8]);
        c.y = itab_eval(pos.y - cy, itab_rows + 4 * ORC_ITAB_N, itab_ih[1], itab_jm[1], d[9]);
        c.z = itab_eval(pos.z - 0.0f, itab_rows + 8 * ORC_ITAB_N, itab_ih[2], itab_jm[2], d[10]);
    } else if (spacing == 0u) {
        /* rs.cl:173 */
        c.x = fmaf(pos.x, d[0], d[4]);
        c.y = fmaf(pos.y, d[1], d[5]);
        c.z = fmaf(pos.z, d[2], d[6]);
    }
    return c;
}

static inline void tex_axis(float u, int n, int *i0, int *i1, float *a) {
    float uu = u - 0.5f;
    float fl = floorf(uu);
    *a = uu - fl;
    float fc = fminf(fmaxf(fl, -1.0f), (float)n);
    int k = (int)fc;
    *i0 = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
    *i1 = k + 1 < 0 ? 0 : (k + 1 > n - 1 ? n - 1 : k + 1);
}

static inline float lerpf_(float p, float q, float t) { return fmaf(t, q, fmaf(-t, p, p)); }
static inline of4 lerp4(of4 p, of4 q, float t) {
    of4 o = {lerpf_(p.x, q.x, t), lerpf_(p.y, q.y, t), lerpf_(p.z, q.z, t), lerpf_(p.w, q.w, t)};
    return o;
}

of4 orc_sample3d(const orc_tex *t, float u, float v, float w) {
    int i0, i1, j0, j1, k0, k1;
    float a, b, c;
    tex_axis(u, t->nx, &i0, &i1, &a);
    tex_axis(v, t->ny, &j0, &j1, &b);
    tex_axis(w, t->nz, &k0, &k1, &c);
    const of4 *D = t->data;
    size_t nx = (size_t)t->nx, nxy = (size_t)t->nx * (size_t)t->ny;
    of4 x00 = lerp4(D[k0 * nxy + j0 * nx + i0], D[k0 * nxy + j0 * nx + i1], a);
    of4 x10 = lerp4(D[k0 * nxy + j1 * nx + i0], D[k0 * nxy + j1 * nx + i1], a);
    of4 x01 = lerp4(D[k1 * nxy + j0 * nx + i0], D[k1 * nxy + j0 * nx + i1], a);
    of4 x11 = lerp4(D[k1 * nxy + j1 * nx + i0], D[k1 * nxy + j1 * nx + i1], a);
    of4 y0 = lerp4(x00, x10, b);
    of4 y1 = lerp4(x01, x11, b);
    return lerp4(y0, y1, c);
}

/* Time interpolation between LES volumes: the wind (and Cn^2 volume) of the
 * advances blend towards the next volume with weight orc_blend_alpha,
 * lerp(sample of the current volume, sample of the next, alpha) per channel,
 * when a next volume is set. */
static const orc_tex *orc_blend_wind = NULL, *orc_blend_cpxx = NULL;
static float orc_blend_alpha = 0.0f;

void orc_set_blend(const orc_tex *wind_next, const orc_tex *cpxx_next, float alpha) {
    orc_blend_wind = wind_next;
    orc_blend_cpxx = cpxx_next;
    orc_blend_alpha = alpha;
}

static of4 wind_at(const orc_tex *t, of4 c) {
    of4 a = orc_sample3d(t, c.x, c.y, c.z);
    if (!orc_blend_wind) return a;
    return lerp4(a, orc_sample3d(orc_blend_wind, c.x, c.y, c.z), orc_blend_alpha);
}

static of4 cpxx_at(const orc_tex *t, of4 c) {
    of4 a = orc_sample3d(t, c.x, c.y, c.z);
    if (!orc_blend_wind || !orc_blend_cpxx) return a;
    return lerp4(a, orc_sample3d(orc_blend_cpxx, c.x, c.y, c.z), orc_blend_alpha);
}

of4 orc_sample2d(const orc_tex *t, float u, float v) {
    int i0, i1, j0, j1;
    float a, b;
    tex_axis(u, t->nx, &i0, &i1, &a);
    tex_axis(v, t->ny, &j0, &j1, &b);
    const of4 *D = t->data;
    size_t nx = (size_t)t->nx;
    of4 y0 = lerp4(D[j0 * nx + i0], D[j0 * nx + i1], a);
    of4 y1 = lerp4(D[j1 * nx + i0], D[j1 * nx + i1], a);
    return lerp4(y0, y1, b);
}

of4 orc_ellipsoid_rcs(of4 pos, const of4 *table, const float *td) {
    /* rs.cl:268-280: truncating index, H = alpha_x, V blended by cos^2(elevation) */
    float fidx = of_clampf(fmaf(pos.w, td[0], td[1]), 0.0f, td[2]);
    uint32_t idx = (uint32_t)fidx;
    of4 xz = table[idx];
    float beta = of_atan2f(pos.z, of_len2(pos.x, pos.y));
    float cb = of_cosf(beta);
    of4 r;
    r.x = xz.x;
    r.y = xz.y;
    r.z = xz.x + ((xz.z - xz.x) * cb) * cb;
    r.w = xz.y + ((xz.w - xz.y) * cb) * cb;
    return r;
}

/* rs.cl:54-85 */
static inline of4 qmul(of4 l, of4 r) {
    of4 o;
    o.x = ((l.w * r.x - l.z * r.y) + l.y * r.z) + l.x * r.w;
    o.y = ((l.w * r.y + l.z * r.x) + l.y * r.w) - l.x * r.z;
    o.z = ((l.w * r.z + l.z * r.w) - l.y * r.x) + l.x * r.y;
    o.w = ((l.w * r.w - l.z * r.z) - l.y * r.y) - l.x * r.x;
    return o;
}
static inline of4 qconj(of4 q) { of4 o = {-q.x, -q.y, -q.z, q.w}; return o; }
static inline of4 qrot(of4 v, of4 q) { return qmul(qmul(q, v), qconj(q)); }
static inline of4 norm4(of4 q) {
    float n = of_len4(q.x, q.y, q.z, q.w);
    of4 o = {q.x / n, q.y / n, q.z / n, q.w / n};
    return o;
}

void orc_quat2abg(of4 q, float *alpha, float *beta, float *gamma) {
    /* rs.cl:317-332 (matlab/quat2abg.m): 3-2-3 Euler angles, lumped into gamma
     * within about 1 degree of the poles */
    float beta_arg = ((q.w * q.w + q.z * q.z) - q.y * q.y) - q.x * q.x;
    if (beta_arg > 0.999847f || beta_arg < -0.999847f) {
        *alpha = 0.0f;
        *beta = 0.0f;
        *gamma = (of_signf(q.z) * of_acos_unit(q.w)) * 2.0f;
    } else {
        *alpha = of_atan2f(q.y * q.z - q.w * q.x, q.x * q.z + q.w * q.y);
        *beta = of_acosf(((q.w * q.w + q.z * q.z) - q.y * q.y) - q.x * q.x);
        *gamma = of_atan2f(q.y * q.z + q.w * q.x, q.w * q.y - q.x * q.z);
    }
}

of4 orc_qmul(of4 l, of4 r) { return qmul(l, r); }

of4 orc_debris_rcs(of4 pos, of4 ori, const orc_tex *rr, const orc_tex *ri, const float *d, const float *sd) {
    uint32_t concept = fbits(sd[5]);
    float el = of_atan2f(pos.z, of_len2(pos.x, pos.y));
    float az = of_atan2f(pos.x, pos.y);
    float se, ce, sa, ca;
    of_sincosf(0.5f * (el + ORC_PI_2_F), &se, &ce);
    of_sincosf(0.5f * az, &sa, &ca);
    of4 oc;
    oc.x = (-se * ca + ce * sa) * ORC_SQRT1_2_F;
    oc.y = (ce * ca + se * sa) * ORC_SQRT1_2_F;
    oc.z = (se * ca + ce * sa) * ORC_SQRT1_2_F;
    oc.w = (ce * ca + (-se) * sa) * ORC_SQRT1_2_F;
    of4 R = qmul(ori, oc);
    of4 q = {R.x, R.z, -R.y, R.w};
    float alpha, beta, gamma;
    orc_quat2abg(q, &alpha, &beta, &gamma);
    float u = fmaf(alpha, d[0], d[4]);
    float v = fmaf(beta, d[1], d[5]);
    of4 re = orc_sample2d(rr, u, v);
    of4 im = orc_sample2d(ri, u, v);
    float sg, cg;
    of_sincosf(gamma, &sg, &cg);
    of4 ss;
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

static inline int is_outside(of4 pos, const float *sd) {
    float ex = sd[8] + sd[12], ey = sd[9] + sd[13], ez = sd[10] + sd[14];
    return (pos.x <= sd[8]) | (pos.y <= sd[9]) | (pos.z <= sd[10]) |
           (pos.x >= ex) | (pos.y >= ey) | (pos.z >= ez);
}

void orc_el_atts(of4 *p, of4 *v, of4 *x, ou4 *y, size_t begin, size_t end,
                 const orc_tex *wind, const float *les_desc,
                 const of4 *drop, const float *dd, const float *sd) {
    const float dt = sd[11];
    const uint32_t concept = fbits(sd[5]);
    const float rho_air = 1.225f;
    const float rho_over_mu_air = 6.7308e4f;
    for (size_t i = begin; i < end; i++) {
        of4 pos = p[i], vel = v[i], rcs = x[i];
        ou4 seed = y[i];
        pos.x = fmaf(vel.x, dt, pos.x);
        pos.y = fmaf(vel.y, dt, pos.y);
        pos.z = fmaf(vel.z, dt, pos.z);
        pos.w = fmaf(vel.w, 0.0f, pos.w);
        const float aom = 0.003006012f / pos.w;
        if (is_outside(pos, sd)) {
            of4 r;
            orc_rand4(&seed, &r);
            pos.x = fmaf(r.x, sd[12], sd[8]);
            pos.y = fmaf(r.y, sd[13], sd[9]);
            pos.z = fmaf(r.z, sd[14], sd[10]);
            vel.x = vel.y = vel.z = vel.w = 0.0f;
        } else {
            of4 c = orc_wind_index(pos, les_desc, sd);
            of4 bg = wind_at(wind, c);
            of4 dv = {bg.x - vel.x, bg.y - vel.y, bg.z - vel.z, bg.w - vel.w};
            float dva = of_len3(dv.x, dv.y, dv.z);
            if (dva > 1.0e-3f) {
                float re = (rho_over_mu_air * (2.0f * pos.w)) * dva;
                float cd = (24.0f / re + 6.0f / (1.0f + sqrtf(re))) + 0.4f;
                float k = (((0.5f * rho_air) * cd) * aom) * dva;
                of4 du = {k * dv.x + 0.0f, k * dv.y + 0.0f, k * dv.z + (-9.8f), k * dv.w + 0.0f};
                vel.x = vel.x + du.x * dt;
                vel.y = vel.y + du.y * dt;
                vel.z = vel.z + du.z * dt;
                vel.w = vel.w + du.w * 0.0f;
                /* |v| > max(1, 3 |v_bg|) (rs.cl:716), evaluated on squared norms */
                float v2 = ((vel.x * vel.x + vel.y * vel.y) + vel.z * vel.z) + vel.w * vel.w;
                float b2 = ((bg.x * bg.x + bg.y * bg.y) + bg.z * bg.z) + bg.w * bg.w;
                if ((concept & 8u) && v2 > fmaxf(1.0f, 9.0f * b2)) {
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
            rcs = orc_ellipsoid_rcs(pos, drop, dd);
        }
        p[i] = pos;
        v[i] = vel;
        x[i] = rcs;
        y[i] = seed;
    }
}

void orc_bg_atts(of4 *p, of4 *v, of4 *x, ou4 *y, size_t begin, size_t end,
                 const orc_tex *wind, const float *les_desc,
                 const of4 *drop, const float *dd, const float *sd) {
    const float dt = sd[11];
    for (size_t i = begin; i < end; i++) {
        of4 pos = p[i], vel = v[i];
        pos.x = pos.x + vel.x * dt;
        pos.y = pos.y + vel.y * dt;
        pos.z = pos.z + vel.z * dt;
        pos.w = pos.w + vel.w * 0.0f;
        if (is_outside(pos, sd)) {
            ou4 seed = y[i];
            of4 r;
            orc_rand4(&seed, &r);
            pos.x = r.x * sd[12] + sd[8];
            pos.y = r.y * sd[13] + sd[9];
            pos.z = r.z * sd[14] + sd[10];
            vel.x = vel.y = vel.z = vel.w = 0.0f;
            p[i] = pos;
            v[i] = vel;
            y[i] = seed;
            continue;
        }
        of4 c = orc_wind_index(pos, les_desc, sd);
        vel = wind_at(wind, c);
        p[i] = pos;
        v[i] = vel;
        x[i] = orc_ellipsoid_rcs(pos, drop, dd);
    }
}

/* rs.cl:590-648 -- fp_atts over [begin, end): positions fixed, vel = wind at
 * the position, phase turned by the radial wind over dt, rcs = (a cos phi,
 * a sin phi, cn2, phi) with a = 10^cn2 from the (cn2, p, _, _) volume. */
void orc_fp_atts(of4 *p, of4 *v, of4 *x, size_t begin, size_t end, const orc_tex *wind, const orc_tex *cpxx,
                 const float *les_desc, const float *sd) {
    const float wav_num = sd[4], dt = sd[11];
    for (size_t i = begin; i < end; i++) {
        const of4 pos = p[i];
        of4 vel = v[i], rcs = x[i];
        of4 c = orc_wind_index(pos, les_desc, sd);
        of4 uvwt = wind_at(wind, c);
        of4 cp = cpxx_at(cpxx, c);
        float len = of_len3(pos.x, pos.y, pos.z);
        float nx = pos.x / len, ny = pos.y / len, nz = pos.z / len;
        float d = (nx * uvwt.x + ny * uvwt.y) + nz * uvwt.z;
        float phi = rcs.w - (wav_num * d) * dt;
        float cn2 = cp.x;
        float a = of_powf_pos(10.0f, cn2);
        float s = of_sinf(phi), cc = of_cosf(phi);
        rcs.x = a * cc;
        rcs.y = a * s;
        rcs.z = cn2;
        rcs.w = of_atan2f(s, cc);
        vel.x = uvwt.x;
        vel.y = uvwt.y;
        vel.z = uvwt.z;
        v[i] = vel;
        x[i] = rcs;
    }
}

/* Debris flux field (concept C): a copy of the inverse CDF and its
 * descriptor (rs.c:3392-3404 layout); count 0 switches the placement off. */
static float *orc_dff_tab = NULL;
static float orc_dff_desc[16];

void orc_set_dff(const float *icdf, int count, const float *desc16) {
    free(orc_dff_tab);
    orc_dff_tab = NULL;
    memset(orc_dff_desc, 0, sizeof(orc_dff_desc));
    if (count <= 0) return;
    orc_dff_tab = (float *)malloc((size_t)count * sizeof(float));
    memcpy(orc_dff_tab, icdf, (size_t)count * sizeof(float));
    memcpy(orc_dff_desc, desc16, sizeof(orc_dff_desc));
}

/* Inverse CDF from a cell PDF, rs.c:3247-3320 (CDF summed in binary32, kept in
 * binary64; a run of empty cells is stepped back over).  0 or -1 as the
 * reference's two error returns. */
int orc_dff_icdf_from_pdf(const float *pdf, int pdf_count, float *tab, int n) {
    const int cdf_count = pdf_count + 1;
    double *cdf = (double *)malloc((size_t)cdf_count * sizeof(double));
    float cumsum = 0.0f;
    int k;
    for (k = 0; k < pdf_count; k++) {
        cdf[k] = cumsum;
        cumsum += pdf[k];
        if (cumsum > 1.0f) {
            if (k < pdf_count - 2) { free(cdf); return -1; }
            if (k < pdf_count - 1) cumsum = 1.0f;
        }
    }
    cdf[pdf_count] = 1.0;
    for (k = 0; k < n; k++) {
        float x = (float)k / (float)(n - 1);
        int b = 0, e;
        while (b < cdf_count && cdf[b] <= x) b++;
        b = b > 0 ? b - 1 : 0;
        e = b + 1 < pdf_count ? b + 1 : pdf_count;
        if (cdf[b] == cdf[e]) {
            while (b > 0 && cdf[b] >= x) b--;
            e = b + 1 < pdf_count ? b + 1 : pdf_count;
        }
        if (b == e) { tab[k] = (float)b; continue; }
        if (!(cdf[b] <= x && x <= cdf[e])) { free(cdf); return -1; }
        float a = (float)((x - cdf[b]) / (cdf[e] - cdf[b]));
        tab[k] = (float)b + a * ((float)e - (float)b);
    }
    free(cdf);
    return 0;
}

/* Respawn placement under the flux field, rs.cl:795-836. */
static void orc_dff_place(of4 r, const float *sd, of4 *pos) {
    const float *dd = orc_dff_desc;
    const float sf = dd[15];
    float i0, i1;
    const float f0 = of_clampf(fmaf(r.x, sf, 0.0f), 0.0f, sf);
    const float f1 = of_clampf(fmaf(r.x, sf, 1.0f), 0.0f, sf);
    const float dec = of_fractf(f0, &i0);
    of_fractf(f1, &i1);
    const float c = of_mixf(orc_dff_tab[(uint32_t)i0], orc_dff_tab[(uint32_t)i1], dec);
    float y = floorf(c / dd[13]);
    float x = fmaf(y, -dd[13], c);
    y = y + r.y;
    if (dd[14] == 1.0f) {
        const float rx = x - fmaf(0.5f, dd[8], 0.5f), ry = y - fmaf(0.5f, dd[9], 0.5f);
        x = fmaf(of_powf_pos(dd[4], fabsf(rx)), -dd[0], dd[0]);
        y = fmaf(of_powf_pos(dd[5], fabsf(ry)), -dd[1], dd[1]);
        x = copysignf(x, rx) + fmaf(0.5f, sd[12], sd[8]);
        y = copysignf(y, ry) + fmaf(0.5f, sd[13], sd[9]);
    } else if (dd[14] == 2.0f) {
        x = fmaf(x, sd[12] * dd[0], sd[8]);
        y = fmaf(y, sd[13] * dd[1], sd[9]);
    } else {
        x = fmaf(x, dd[0], sd[8]);
        y = fmaf(y, dd[1], sd[9]);
    }
    pos->x = x;
    pos->y = y;
    pos->z = MIN_HEIGHT;
}

/* Test entry: the placement of one respawn draw r under the flux field set by
 * orc_set_dff (pos.z = MIN_HEIGHT); returns -1 when no field is set. */
int orc_dff_place_probe(of4 r, const float *sd, of4 *pos) {
    if (!orc_dff_tab) return -1;
    orc_dff_place(r, sd, pos);
    return 0;
}

void orc_db_atts(of4 *p, of4 *o, of4 *v, of4 *t, of4 *x, ou4 *y, size_t begin, size_t end,
                 const orc_tex *wind, const float *les_desc,
                 const orc_tex *adm_cd, const orc_tex *adm_cm, const float *ad,
                 const orc_tex *rcs_real, const orc_tex *rcs_imag, const float *rd,
                 const float *sd) {
    const float dt = sd[11];
    const uint32_t concept = fbits(sd[5]);
    for (size_t i = begin; i < end; i++) {
        of4 pos = p[i], ori = o[i], vel = v[i], tum = t[i];
        ori = norm4(qmul(ori, tum));
        pos.x = pos.x + vel.x * dt;
        pos.y = pos.y + vel.y * dt;
        pos.z = pos.z + vel.z * dt;
        pos.w = pos.w + vel.w * 0.0f;
        if (is_outside(pos, sd)) {
            /* rs.cl:790-873: flux-field or uniform placement */
            ou4 seed = y[i];
            of4 r;
            orc_rand4(&seed, &r);
            if ((concept & 128u) && orc_dff_tab) {
                orc_dff_place(r, sd, &pos);
            } else {
                pos.x = fmaf(r.x, sd[12], sd[8]);
                pos.y = fmaf(r.y, sd[13], sd[9]);
                pos.z = sd[10] + MIN_HEIGHT;
            }
            orc_rand4(&seed, &r);
            float c0 = sqrtf(-2.0f * of_logf(r.x));
            float c1 = sqrtf(-2.0f * of_logf(r.y));
            float c2 = sqrtf(-2.0f * of_logf(r.z));
            float c3 = r.w;
            orc_rand4(&seed, &r);
            const float two_pi = 2.0f * ORC_PI_F;
            c0 *= of_cosf(two_pi * r.x);
            c1 *= of_cosf(two_pi * r.y);
            c2 *= of_cosf(two_pi * r.z);
            float sin_th_2, cos_th_2;
            of_sincosf(ORC_PI_F * c3, &sin_th_2, &cos_th_2);
            float sqxy = sqrtf(c0 * c0 + c2 * c2);
            float sqxyz = sqrtf((c0 * c0 + c1 * c1) + c2 * c2);
            float sa, ca;
            of_sincosf(0.5f * of_acos_unit(c1 / sqxyz), &sa, &ca);
            ori.x = (((c0 * ca) * sin_th_2) / sqxyz - ((((c0 * c1) * sin_th_2) / sqxyz) * sa) / sqxy)
                    + ((c2 * cos_th_2) * sa) / sqxy;
            ori.y = (sin_th_2 / sqxyz) * (sqxy * sa + c1 * ca);
            ori.z = (((c2 * ca) * sin_th_2) / sqxyz - ((c0 * cos_th_2) * sa) / sqxy)
                    - ((((c1 * c2) * sin_th_2) * sa) / sqxy) / sqxyz;
            ori.w = cos_th_2 * ca;
            vel.x = vel.y = vel.z = vel.w = 0.0f;
            tum.x = tum.y = tum.z = 0.0f;
            tum.w = 1.0f;
            of4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
            p[i] = pos;
            o[i] = ori;
            v[i] = vel;
            t[i] = tum;
            x[i] = zero;
            y[i] = seed;
            continue;
        }
        of4 c = orc_wind_index(pos, les_desc, sd);
        of4 bg = wind_at(wind, c);
        /* rs.cl:195-261 compute_dudt_dwdt */
        of4 ur = {bg.x - vel.x, bg.y - vel.y, bg.z - vel.z, bg.w - vel.w};
        of4 uh = qrot(norm4(ur), qconj(ori));
        float beta = of_acos_unit(uh.x);
        float alpha = of_atan2f(uh.z, uh.y);
        if (alpha < 0.0f) {
            alpha = ORC_PI_F + alpha;
            beta = -beta;
        }
        float cu = fmaf(beta, ad[0], ad[4]);
        float cv = fmaf(alpha, ad[1], ad[5]);
        of4 cd = orc_sample2d(adm_cd, cu, cv);
        of4 cm = orc_sample2d(adm_cm, cu, cv);
        const float Ta = ad[15];
        cd = qrot(cd, ori);
        float urn = (ur.x * ur.x + ur.y * ur.y) + ur.z * ur.z;
        float ta_urn = Ta * urn;
        of4 du = {ta_urn * cd.x + 0.0f, ta_urn * cd.y + 0.0f, ta_urn * cd.z + (-9.8f), ta_urn * cd.w + 0.0f};
        of4 dw;
        dw.x = of_radiansf((ta_urn * ad[12]) * cm.x);
        dw.y = of_radiansf((ta_urn * ad[13]) * cm.y);
        dw.z = of_radiansf((ta_urn * ad[14]) * cm.z);
        dw.w = of_radiansf((ta_urn * 0.0f) * cm.w);
        /* |v_xy + dv_xy| > 3 |v_bg,xy| (rs.cl:881), evaluated on squared norms */
        float nx_ = vel.x + du.x * dt, ny_ = vel.y + du.y * dt;
        if ((concept & 8u) && nx_ * nx_ + ny_ * ny_ > 9.0f * (bg.x * bg.x + bg.y * bg.y)) {
            vel.x = bg.x;
            vel.y = bg.y;
            vel.z = vel.z + du.z * dt;
        } else {
            vel.x = vel.x + du.x * dt;
            vel.y = vel.y + du.y * dt;
            vel.z = vel.z + du.z * dt;
            vel.w = vel.w + du.w * 0.0f;
        }
        float ax = dw.x * dt, ay = dw.y * dt, az = dw.z * dt;
        float sx, cx, sy, cy, sz, cz;
        of_sincosf(ax, &sx, &cx);
        of_sincosf(ay, &sy, &cy);
        of_sincosf(az, &sz, &cz);
        of4 tq;
        tq.x = (cx * sy) * sz + (sx * cy) * cz;
        tq.y = (cx * sy) * cz - (sx * cy) * sz;
        tq.z = (cx * cy) * sz + (sx * sy) * cz;
        tq.w = (cx * cy) * cz - (sx * sy) * sz;
        tum = norm4(tq);
        of4 rcs = orc_debris_rcs(pos, ori, rcs_real, rcs_imag, rd, sd);
        p[i] = pos;
        o[i] = ori;
        v[i] = vel;
        t[i] = tum;
        x[i] = rcs;
    }
}

void orc_db_rcs(const of4 *p, const of4 *o, of4 *x, size_t begin, size_t end,
                const orc_tex *rr, const orc_tex *ri, const float *rd, const float *sd) {
    for (size_t i = begin; i < end; i++) x[i] = orc_debris_rcs(p[i], o[i], rr, ri, rd, sd);
}

/* Angular weight (rs.cl:1065-1093).  The angle between the beam b and the
 * position p, acos(b . p / |p|) in rs.cl:1066, is evaluated in the
 * well-conditioned form atan2(|b x p|, b . p): near the beam axis the cosine form
 * turns one ulp of the cosine into 3.5e-4 rad, a 1% change of the main-lobe
 * weight. */
static inline float ang_weight(of4 pos, const float *w, const float *d4, const float *sd, float *range) {
    float r = of_len3(pos.x, pos.y, pos.z);
    float cx = sd[1] * pos.z - sd[2] * pos.y;
    float cy = sd[2] * pos.x - sd[0] * pos.z;
    float cz = sd[0] * pos.y - sd[1] * pos.x;
    float angle = of_atan2f(of_len3(cx, cy, cz), (sd[0] * pos.x + sd[1] * pos.y) + sd[2] * pos.z);
    float x0 = of_clampf(fmaf(angle, d4[0], d4[1]), 0.0f, d4[2]);
    float x1 = of_clampf(fmaf(angle, d4[0], d4[1] + 1.0f), 0.0f, d4[2]);
    float ip;
    float dec = of_fractf(x0, &ip);
    uint32_t i0 = (uint32_t)ip;
    uint32_t i1 = (uint32_t)floorf(x1);
    *range = r;
    return of_mixf(w[i0], w[i1], dec);
}

void orc_scat_sig_aux(of4 *s, of4 *a, const of4 *p, const of4 *x, size_t n,
                      const float *w, const float *d4, const float *sd) {
    for (size_t i = 0; i < n; i++) {
        of4 aux = a[i];
        float r;
        float wa = ang_weight(p[i], w, d4, sd, &r);
        aux.x = r;
        aux.y = aux.y + sd[15];
        aux.w = wa;
        float atten = 1.0f / (r * r);
        float phase = r * sd[4];
        float ss, cc;
        of_sincosf(phase, &ss, &cc);
        of4 b = x[i];
        of4 sig;
        sig.x = (b.x * cc - b.y * (-ss)) * atten;
        sig.y = (b.y * cc + b.x * (-ss)) * atten;
        sig.z = (b.z * cc - b.w * (-ss)) * atten;
        sig.w = (b.w * cc + b.z * (-ss)) * atten;
        s[i] = sig;
        a[i] = aux;
    }
}

void orc_gate_ranges(float start, float delta, int R, float *rk) {
    float r = start;
    for (int k = 0; k < R; k++) {
        rk[k] = r;
        r += delta;
    }
}

static inline float range_weight(float dr, const float *W, const float *d4) {
    float x0 = of_clampf(fmaf(dr, d4[0], d4[1]), 0.0f, d4[2]);
    float x1 = of_clampf(fmaf(dr, d4[0], d4[1] + 1.0f), 0.0f, d4[2]);
    float ip;
    float dec = of_fractf(x0, &ip);
    return of_mixf(W[(uint32_t)ip], W[(uint32_t)floorf(x1)], dec);
}

static inline double range_weight_d(double dr, const float *W, const float *d4) {
    double x0 = fmin(fmax(dr * (double)d4[0] + (double)d4[1], 0.0), (double)d4[2]);
    double fl = floor(x0);
    double dec = x0 - fl;
    uint32_t i0 = (uint32_t)fl;
    uint32_t i1 = (double)i0 + 1.0 > (double)d4[2] ? i0 : i0 + 1;
    return (double)W[i0] + ((double)W[i1] - (double)W[i0]) * dec;
}

static inline double ang_weight_d(double x, double y, double z, const float *w, const float *d4,
                                  const float *sd, double *range) {
    double r = sqrt(x * x + y * y + z * z);
    double bx = sd[0], by = sd[1], bz = sd[2];
    double cx = by * z - bz * y, cy = bz * x - bx * z, cz = bx * y - by * x;
    double angle = atan2(sqrt(cx * cx + cy * cy + cz * cz), bx * x + by * y + bz * z);
    *range = r;
    return range_weight_d(angle, w, d4);
}

void orc_scatterer_signal(const of4 *p, const of4 *x, const float *w, const float *d4,
                          const float *sd, double kd, int mode,
                          double *sig4, float *range_out, float *wa_out) {
    of4 b = *x;
    float r;
    float wa = ang_weight(*p, w, d4, sd, &r);
    float atten = 1.0f / (r * r);
    float cc, ss;
    if (mode == ORC_PHASE_F32) {
        of_sincosf(r * sd[4], &ss, &cc);
    } else {
        double ph = (double)r * kd;
        cc = (float)cos(ph);
        ss = (float)sin(ph);
    }
    float sx = (b.x * cc - b.y * (-ss)) * atten;
    float sy = (b.y * cc + b.x * (-ss)) * atten;
    float sz = (b.z * cc - b.w * (-ss)) * atten;
    float sw = (b.w * cc + b.z * (-ss)) * atten;
    sig4[0] = sx * wa;
    sig4[1] = sy * wa;
    sig4[2] = sz * wa;
    sig4[3] = sw * wa;
    *range_out = r;
    *wa_out = wa;
}

void orc_make_pulse(double *out, const of4 *p, const double *pos64, const of4 *x, size_t n,
                    const float *angw, const float *angw_d4,
                    const float *rangew, const float *rangew_d4,
                    float range_start, float range_delta, int R,
                    const float *sd, double kd, int mode) {
    float *rk = (float *)malloc((size_t)R * sizeof(float));
    orc_gate_ranges(range_start, range_delta, R, rk);
    memset(out, 0, (size_t)R * 4 * sizeof(double));
    for (size_t i = 0; i < n; i++) {
        if (mode == ORC_FULL_F64) {
            double r;
            double wa = ang_weight_d(pos64[3 * i], pos64[3 * i + 1], pos64[3 * i + 2], angw, angw_d4, sd, &r);
            if (wa == 0.0) continue;
            double atten = 1.0 / (r * r);
            double ph = r * kd;
            double cc = cos(ph), ss = sin(ph);
            of4 b = x[i];
            double s0 = ((double)b.x * cc + (double)b.y * ss) * atten * wa;
            double s1 = ((double)b.y * cc - (double)b.x * ss) * atten * wa;
            double s2 = ((double)b.z * cc + (double)b.w * ss) * atten * wa;
            double s3 = ((double)b.w * cc - (double)b.z * ss) * atten * wa;
            for (int k = 0; k < R; k++) {
                double wr = range_weight_d(r - (double)range_start - (double)k * (double)range_delta, rangew, rangew_d4);
                if (wr == 0.0) continue;
                out[4 * k + 0] += wr * s0;
                out[4 * k + 1] += wr * s1;
                out[4 * k + 2] += wr * s2;
                out[4 * k + 3] += wr * s3;
            }
            continue;
        }
        double sig[4];
        float r, wa;
        orc_scatterer_signal(&p[i], &x[i], angw, angw_d4, sd, kd, mode, sig, &r, &wa);
        float sa0 = (float)sig[0], sa1 = (float)sig[1], sa2 = (float)sig[2], sa3 = (float)sig[3];
        for (int k = 0; k < R; k++) {
            float wr = range_weight(r - rk[k], rangew, rangew_d4);
            out[4 * k + 0] += (double)(wr * sa0);
            out[4 * k + 1] += (double)(wr * sa1);
            out[4 * k + 2] += (double)(wr * sa2);
            out[4 * k + 3] += (double)(wr * sa3);
        }
    }
    free(rk);
}

float orc_read_table(const float *table, float index_last, float index) {
    float fl = floorf(index);
    float alpha = index - fl;
    if (index <= 0.0f) return table[0];
    if (fl >= index_last) return table[(unsigned int)index_last];
    unsigned int i = (unsigned int)fl;
    return table[i] + alpha * (table[i + 1] - table[i]);
}

void orc_populate(of4 *pos, of4 *vel, of4 *ori, of4 *tum, of4 *aux, of4 *rcs, ou4 *rnd,
                  const size_t *counts, int num_types, const float *sd,
                  unsigned int seed, int dsd_count, const float *dsd_r, const float *dsd_cdf,
                  int uniform_dsd) {
    const float ox = sd[8], oy = sd[9], oz = sd[10];
    const float sx = sd[12], sy = sd[13];
    const float rmax = (float)RAND_MAX;
    srand(seed);
    size_t i = 0;
    for (int k = 0; k < num_types; k++) {
        for (size_t n = 0; n < counts[k]; n++, i++) {
            pos[i].x = (float)rand() / rmax * sx + ox;
            pos[i].y = (float)rand() / rmax * sy + oy;
            pos[i].z = (float)rand() / rmax * 10.0f + oz;
            pos[i].w = 0.0f;
            aux[i].x = 0.0f;
            aux[i].y = (float)rand() / rmax;
            aux[i].z = 0.0f;
            aux[i].w = 1.0f;
            vel[i].x = vel[i].y = vel[i].z = vel[i].w = 0.0f;
            ori[i].x = 0.5f;
            ori[i].y = -0.5f;
            ori[i].z = -0.5f;
            ori[i].w = 0.5f;
            tum[i].x = tum[i].y = tum[i].z = 0.0f;
            tum[i].w = 1.0f;
            rcs[i].x = 1.0f;
            rcs[i].y = 0.0f;
            rcs[i].z = 1.0f;
            rcs[i].w = 0.0f;
            rnd[i].x = (uint32_t)rand();
            rnd[i].y = (uint32_t)rand();
            rnd[i].z = (uint32_t)rand();
            rnd[i].w = (uint32_t)rand();
        }
    }
    srand(seed + seed);
    if (dsd_count <= 0) return;
    for (size_t n = 0; n < counts[0]; n++) {
        float a = (float)rand() / rmax;
        int bin = 0;
        if (uniform_dsd) {
            bin = (int)(a * (float)dsd_count);
            /* a == 1.0 when rand() >= 2^31 - 64; the reference would read past the bins */
            if (bin > dsd_count - 1) bin = dsd_count - 1;
        } else {
            int k = dsd_count;
            while (k > 0) {
                k--;
                if (a >= dsd_cdf[k]) {
                    bin = k;
                    break;
                }
            }
        }
        pos[n].w = dsd_r[bin];
        aux[n].z = ((float)bin + 0.5f) / (float)dsd_count;
    }
}

void orc_drop_rcs_table(of4 *table, float *desc4, float wave_number, float drop_scale,
                        int transparent, int uniform_dsd, int dsd_count,
                        const float *dsd_r, const float *dsd_pdf) {
    const double k_0 = wave_number * 0.5f;
    const double epsilon_0 = 8.85418782e-12f;
    const double er[4] = {78.669, 18.2257, 78.669, 18.2257};
    const double sc = k_0 * k_0 / (4.0f * M_PI * epsilon_0) * 1.0e-9f * drop_scale;
    const int n = 96;
    if (transparent) {
        memset(table, 0, (size_t)n * sizeof(of4));
    } else {
        for (int i = 0; i < n; i++) {
            double d = 0.5 + (double)i * 0.1;
            double d2 = d * d, d3 = d2 * d, d4 = d3 * d;
            double rba = 1.0048 + (0.0057e-1 * d) - (2.628e-2 * d2) + (3.682e-3 * d3) - (1.677e-4 * d4);
            double rab = 1.0f / rba;
            double fsq = rab * rab - 1.0;
            double f = sqrt(fsq);
            double lz = (1.0 + fsq) / fsq * (1.0 - atan(f) / f);
            double lx = (1.0 - lz) * 0.5;
            double vol = M_PI * d3 / 6.0;
            double a0 = vol * epsilon_0, a2 = vol * epsilon_0;
            double nu0 = a0 * er[0] - 0.0 * er[1];
            double nu1 = 0.0 * er[0] + a0 * er[1];
            double nu2 = a2 * er[2] - 0.0 * er[3];
            double nu3 = 0.0 * er[2] + a2 * er[3];
            double de0 = lx * er[0] + 1.0, de1 = lx * er[1];
            double de2 = lz * er[2] + 1.0, de3 = lz * er[3];
            double bm01 = de0 * de0 + de1 * de1;
            double bm23 = de2 * de2 + de3 * de3;
            double al0 = (nu0 * de0 + nu1 * de1) / bm01;
            double al1 = (nu1 * de0 - nu0 * de1) / bm01;
            double al2 = (nu2 * de2 + nu3 * de3) / bm23;
            double al3 = (nu3 * de2 - nu2 * de3) / bm23;
            table[i].x = (float)(sc * al0);
            table[i].y = (float)(sc * al1);
            table[i].z = (float)(sc * al2);
            table[i].w = (float)(sc * al3);
        }
        if (uniform_dsd) {
            const float pr = 1.0f / (float)dsd_count;
            of4 *copy = (of4 *)malloc((size_t)n * sizeof(of4));
            memcpy(copy, table, (size_t)n * sizeof(of4));
            memset(table, 0, (size_t)n * sizeof(of4));
            for (int i = 0; i < dsd_count; i++) {
                int k = (int)(dsd_r[i] * 20000.0f) - 5;
                float s = sqrtf(dsd_pdf[i] / pr);
                table[k].x = copy[k].x * s;
                table[k].y = copy[k].y * s;
                table[k].z = copy[k].z * s;
                table[k].w = copy[k].w * s;
            }
            free(copy);
        }
    }
    orc_table1d_desc(desc4, 0.25e-3f, 0.05e-3f, n);
    desc4[3] = drop_scale;
}

void orc_table1d_desc(float *d4, float start, float delta, int n) {
    d4[0] = 1.0f / delta;
    d4[1] = -start * d4[0];
    d4[2] = (float)n - 1.0f;
    d4[3] = 0.0f;
}

void orc_angular_weight_standard(float *w, float *d4, float bw_rad) {
    const int n = 32;
    float b = 1.27f * M_PI / bw_rad;
    float delta = 1.0f / 360.0f * M_PI;
    for (int i = 0; i < n; i++) {
        float a = (float)i * delta;
        float c = b * sinf(a);
        if (i == 0) w[i] = 1.0f;
        else if (i == n - 1) w[i] = 0.0f;
        else w[i] = 8.0f * jn(2, c) / (c * c);
    }
    orc_table1d_desc(d4, 0.0f, delta, n);
}

float orc_radar_gain(float gain_dbi, float watt) {
    return powf(10.0f, 0.1f * gain_dbi) * sqrtf(watt) / (4.0f * M_PI) * 1.0e6f;
}

void orc_el_atts_f64(double *p3, double *v3, const float *radius, size_t n,
                     const orc_tex *wind, const float *ld, const float *sd, int bounded) {
    const double dt = sd[11];
    const double cx = (double)sd[8] + 0.5 * (double)sd[12];
    const double cy = (double)sd[9] + 0.5 * (double)sd[13];
    const size_t nx = (size_t)wind->nx, ny = (size_t)wind->ny, nz = (size_t)wind->nz;
    for (size_t i = 0; i < n; i++) {
        double *p = p3 + 3 * i, *v = v3 + 3 * i;
        double w = radius[i];
        p[0] += v[0] * dt;
        p[1] += v[1] * dt;
        p[2] += v[2] * dt;
        double rel[3] = {p[0] - cx, p[1] - cy, p[2]};
        double u[3], bg[4] = {0, 0, 0, 0};
        int i0[3], i1[3];
        double a[3];
        size_t dims[3] = {nx, ny, nz};
        for (int c = 0; c < 3; c++) {
            if (fbits(ld[7]) == 7u) {
                u[c] = copysign((double)ld[c], rel[c]) * log1p((double)ld[4 + c] * fabs(rel[c])) + (double)ld[8 + c];
            } else {
                u[c] = p[c] * (double)ld[c] + (double)ld[4 + c];
            }
            double uu = u[c] - 0.5, fl = floor(uu);
            a[c] = uu - fl;
            long k = (long)fl;
            i0[c] = k < 0 ? 0 : (k > (long)dims[c] - 1 ? (int)dims[c] - 1 : (int)k);
            i1[c] = k + 1 < 0 ? 0 : (k + 1 > (long)dims[c] - 1 ? (int)dims[c] - 1 : (int)(k + 1));
        }
        for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    double wgt = (dx ? a[0] : 1.0 - a[0]) * (dy ? a[1] : 1.0 - a[1]) * (dz ? a[2] : 1.0 - a[2]);
                    size_t idx = (size_t)(dz ? i1[2] : i0[2]) * nx * ny + (size_t)(dy ? i1[1] : i0[1]) * nx +
                                 (size_t)(dx ? i1[0] : i0[0]);
                    of4 T = wind->data[idx];
                    bg[0] += wgt * T.x;
                    bg[1] += wgt * T.y;
                    bg[2] += wgt * T.z;
                    bg[3] += wgt * T.w;
                }
        double dv[3] = {bg[0] - v[0], bg[1] - v[1], bg[2] - v[2]};
        double dva = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        if (dva > 1.0e-3) {
            double re = 6.7308e4 * (2.0 * w) * dva;
            double cd = 24.0 / re + 6.0 / (1.0 + sqrt(re)) + 0.4;
            double k = 0.5 * 1.225 * cd * (0.003006012 / w) * dva;
            v[0] += k * dv[0] * dt;
            v[1] += k * dv[1] * dt;
            v[2] += (k * dv[2] - 9.8) * dt;
            if (bounded) {
                double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                double bn = sqrt(bg[0] * bg[0] + bg[1] * bg[1] + bg[2] * bg[2] + bg[3] * bg[3]);
                if (vn > fmax(1.0, 3.0 * bn)) {
                    v[0] = bg[0];
                    v[1] = bg[1];
                    v[2] = bg[2] - 9.8 * dt;
                }
            }
        } else {
            v[2] += -9.8 * dt;
        }
    }
}

void orc_les_desc(float *d, int stretched, uint32_t nx, uint32_t ny, uint32_t nz,
                  float ax, float ay, float az, float rx, float ry, float rz, float tr) {
    memset(d, 0, 16 * sizeof(float));
    uint32_t spacing = stretched ? 7u : 0u;
    memcpy(&d[7], &spacing, 4);
    if (stretched) {
        /* coordinate = copysign(1 / ln r, p) log1p((r - 1) / a |p|) + (n - 1) / 2; z offset 0 */
        d[0] = 1.0f / log(rx); d[4] = (rx - 1.0f) / ax; d[8] = 0.5f * (float)(nx - 1);
        d[1] = 1.0f / log(ry); d[5] = (ry - 1.0f) / ay; d[9] = 0.5f * (float)(ny - 1);
        d[2] = 1.0f / log(rz); d[6] = (rz - 1.0f) / az; d[10] = 0.0f;
    } else {
        d[0] = 1.0f / rx; d[4] = (float)(nx - 1) * 0.5f; d[8] = (float)nx - 1.0f;
        d[1] = 1.0f / ry; d[5] = (float)(ny - 1) * 0.5f; d[9] = (float)ny - 1.0f;
        d[2] = 1.0f / rz; d[6] = 0.0f; d[10] = (float)nz - 1.0f;
    }
    d[3] = tr;
}
