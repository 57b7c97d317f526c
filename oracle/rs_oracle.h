/*
 * rs_oracle.h -- CPU oracle for SimRadar's per-pulse scatterer engine.
 *
 * TEST INFRASTRUCTURE.  Only tests/, the smoke check and bench.py's cpu_baseline
 * leg may build, link or call this library, and only as the checker / CPU
 * baseline -- never as the thing measured or shipped.
 *
 * The oracle restates the reference's per-scatterer algorithms (rs.cl) and the
 * host steps that feed them (rs.c) on plain arrays laid out as the reference lays
 * them out: one float4 stream per attribute, float16 descriptors with the slot
 * meanings of rs_enum.h.  Each function cites the reference lines it follows.
 */
#ifndef RS_ORACLE_H
#define RS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

typedef struct { float x, y, z, w; } of4;
typedef struct { uint32_t x, y, z, w; } ou4;

/* A float4 texture: width x height (x depth) texels, x fastest. */
typedef struct { const of4 *data; int nx, ny, nz; } orc_tex;

/* Precision of the phase / accumulation in orc_make_pulse. */
enum {
    ORC_PHASE_F32 = 0,   /* phase = r * (4 pi / lambda) in binary32, as rs.cl:1096-1100 */
    ORC_PHASE_F64 = 1,   /* binary32 range, phase from a binary64 product (engine precision) */
    ORC_FULL_F64  = 2    /* positions, range, weights and phase all in binary64 */
};

/* rs.cl:35-44 -- four 31-bit multiplicative LCG lanes, one step each. */
void orc_rand4(ou4 *seed, of4 *r);

/* rs.cl:123-176 -- table coordinates of a position (stretched axes through the
 * index tables below). */
of4 orc_wind_index(of4 pos, const float *les_desc, const float *sim_desc);
/* The stretched-grid index tables for a descriptor and the box of sim_desc:
 * rows[3 * 256 * 4] (x, y, z pieces), ih[3], jm[3]. */
void orc_wind_index_table(const float *les_desc, const float *sim_desc, float *rows, float *ih, float *jm);

/* Time interpolation of the advances' wind (and Cn^2) towards the next LES
 * volume with weight alpha; wind_next NULL switches it off (the reference's
 * piecewise-constant volumes, rs.c:5189-5203). */
void orc_set_blend(const orc_tex *wind_next, const orc_tex *cpxx_next, float alpha);

/* OpenCL linear filter, unnormalised coordinates, clamp-to-edge (rs.cl:8). */
of4 orc_sample3d(const orc_tex *t, float u, float v, float w);
of4 orc_sample2d(const orc_tex *t, float u, float v);

/* rs.cl:268-280 */
of4 orc_ellipsoid_rcs(of4 pos, const of4 *table, const float *table_desc4);
/* rs.cl:317-332 -- 3-2-3 Euler angles of a relative quaternion. */
void orc_quat2abg(of4 q, float *alpha, float *beta, float *gamma);
/* rs.cl:54-60 */
of4 orc_qmul(of4 l, of4 r);
/* rs.cl:282-359 */
of4 orc_debris_rcs(of4 pos, of4 ori, const orc_tex *rcs_real, const orc_tex *rcs_imag,
                   const float *rcs_desc, const float *sim_desc);

/* rs.cl:653-739 -- el_atts over [begin, end). */
void orc_el_atts(of4 *p, of4 *v, of4 *x, ou4 *y, size_t begin, size_t end,
                 const orc_tex *wind, const float *les_desc,
                 const of4 *drop_rcs, const float *drop_desc4, const float *sim_desc);
/* rs.cl:533-580 -- bg_atts over [begin, end). */
void orc_bg_atts(of4 *p, of4 *v, of4 *x, ou4 *y, size_t begin, size_t end,
                 const orc_tex *wind, const float *les_desc,
                 const of4 *drop_rcs, const float *drop_desc4, const float *sim_desc);
/* rs.cl:590-648 -- fp_atts over [begin, end) (concept F): cpxx is the
 * (cn2, p, _, _) volume sampled like the wind. */
void orc_fp_atts(of4 *p, of4 *v, of4 *x, size_t begin, size_t end, const orc_tex *wind, const orc_tex *cpxx,
                 const float *les_desc, const float *sim_desc);
/* rs.cl:744-907 -- db_atts over [begin, end) (debris flux field concept not covered). */
void orc_set_dff(const float *icdf, int count, const float *desc16);
int orc_dff_icdf_from_pdf(const float *pdf, int pdf_count, float *tab, int n);
/* Respawn placement of one draw r (rs.cl:795-836) under the field of
 * orc_set_dff; -1 when none is set. */
int orc_dff_place_probe(of4 r, const float *sim_desc, of4 *pos);
void orc_db_atts(of4 *p, of4 *o, of4 *v, of4 *t, of4 *x, ou4 *y, size_t begin, size_t end,
                 const orc_tex *wind, const float *les_desc,
                 const orc_tex *adm_cd, const orc_tex *adm_cm, const float *adm_desc,
                 const orc_tex *rcs_real, const orc_tex *rcs_imag, const float *rcs_desc,
                 const float *sim_desc);
/* rs.cl:909-919 */
void orc_db_rcs(const of4 *p, const of4 *o, of4 *x, size_t begin, size_t end,
                const orc_tex *rcs_real, const orc_tex *rcs_imag, const float *rcs_desc,
                const float *sim_desc);

/* rs.cl:1050-1110 -- binary32 signal and auxiliary streams. */
void orc_scat_sig_aux(of4 *s, of4 *a, const of4 *p, const of4 *x, size_t n,
                      const float *angw, const float *angw_desc4, const float *sim_desc);

/* Gate ranges as rs.cl:1189/1214 form them: start, then repeated binary32 adds. */
void orc_gate_ranges(float range_start, float range_delta, int R, float *r_k);

/*
 * Per-gate complex sums (rs.cl:1050-1110 + 1125-1301 + 1304-1441): for every
 * scatterer and every gate, range weight x angular weight x signal, summed in
 * binary64 (brute force over all R gates).  out[4*k + c] = (Ih, Qh, Iv, Qv) of
 * gate k before the radar gain.  wave_number_d is 4 pi / lambda in binary64
 * (used by ORC_PHASE_F64 / ORC_FULL_F64); sim_desc[4] is the binary32 one.
 * pos64 (x,y,z per scatterer) is read only by ORC_FULL_F64 and may be NULL.
 */
void orc_make_pulse(double *out, const of4 *p, const double *pos64, const of4 *x, size_t n,
                    const float *angw, const float *angw_desc4,
                    const float *rangew, const float *rangew_desc4,
                    float range_start, float range_delta, int R,
                    const float *sim_desc, double wave_number_d, int mode);

/* Single-scatterer contributions for tests: returns the signal (Ih,Qh,Iv,Qv) x
 * angular weight and the binary32 range of one scatterer (engine precision). */
void orc_scatterer_signal(const of4 *p, const of4 *x, const float *angw, const float *angw_desc4,
                          const float *sim_desc, double wave_number_d, int mode,
                          double *sig4, float *range_out, float *wa_out);

/* rs.c:358-374 -- linear interpolation used by the CPU loop of the pulse test. */
float orc_read_table(const float *table, float index_last, float index);

/* rs.c:4568-4799 -- population of the scatterer streams with glibc rand().
 * counts[k] scatterers of type k, contiguous, rain first.  box = sim_desc slots
 * 8-10 (origin) and 12-14 (size).  dsd_* may be NULL when dsd_count == 0. */
void orc_populate(of4 *pos, of4 *vel, of4 *ori, of4 *tum, of4 *aux, of4 *rcs, ou4 *rnd,
                  const size_t *counts, int num_types, const float *sim_desc,
                  unsigned int seed, int dsd_count, const float *dsd_r, const float *dsd_cdf,
                  int uniform_dsd);

/* rs.c:5922-6029 -- Rayleigh oblate-spheroid table (96 rows) and its desc. */
void orc_drop_rcs_table(of4 *table96, float *desc4, float wave_number, float drop_scale,
                        int transparent, int uniform_dsd, int dsd_count,
                        const float *dsd_r, const float *dsd_pdf);

/* rs.c:2826-2851 -- 32-entry standard antenna pattern and desc. */
void orc_angular_weight_standard(float *w32, float *desc4, float beamwidth_rad);

/* rs.c:2472-2484 / 2553-2559 -- generic 1-D table description. */
void orc_table1d_desc(float *desc4, float index_start, float index_delta, int n);

/* rs.c:5084 -- radar gain applied to the merged pulse. */
float orc_radar_gain(float gain_dbi, float tx_power_watt);

/* FP64 restatement of el_atts for the precision study (no respawn).  Positions
 * and velocities in binary64, wind sampled with binary64 weights. */
void orc_el_atts_f64(double *p3, double *v3, const float *radius, size_t n,
                     const orc_tex *wind, const float *les_desc, const float *sim_desc, int bounded);

/* rs.c:3029-3070 -- LES descriptor of a stretched or uniform grid. */
void orc_les_desc(float *desc16, int stretched, uint32_t nx, uint32_t ny, uint32_t nz,
                  float ax, float ay, float az, float rx, float ry, float rz, float tr);

#endif
