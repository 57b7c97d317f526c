/*
 * rs.h -- the RS_* C API of SimRadar's per-pulse scatterer engine, served by
 * the gfx950 engine (include/rsd.h) instead of OpenCL.
 *
 * Every prototype below keeps the reference's name, argument meaning and
 * calling pattern (rs.h:366-500; private setters rs_priv.h:109-123), so a
 * caller written for the reference (simradar.c:812-1218, simple_ppi.c:24-98)
 * compiles against it.  RSHandle keeps the fields such callers read
 * (params, sim_tic, sim_desc, num_scats, counts, num_types, num_workers,
 * random_seed, sim_concept, verb, status, scat_*, pulse, pulse_tmp); the
 * OpenCL objects of RSWorker are replaced by one opaque engine pointer.
 *
 * Additions (no counterpart in the reference) are marked "addition".
 */
#ifndef RS_H
#define RS_H

#include "rs_types.h"
#include "rs_enum.h"
#include "rs_providers.h"
#include "rsd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RS_INDENT "    "

typedef struct _rs_pulse_params {
    unsigned int num_scats;
    unsigned int user_max_groups;
    unsigned int user_max_work_items;
    unsigned int cl_pass_2_method;
    unsigned int range_count;
    float        range_start;
    float        range_delta;
    unsigned int entry_counts[2];
    unsigned int group_counts[2];
    size_t       global[2];
    size_t       local[2];
    size_t       local_mem_size[2];
} RSMakePulseParams;

typedef struct _rs_table {
    float    x0;
    float    xm;
    float    dx;
    uint32_t reserved;
    float    *data;
} RSTable;

typedef struct _rs_table2d {
    float     xs, xo, xm;
    uint32_t  x_;
    float     ys, yo, ym;
    uint32_t  y_;
    cl_float4 *data;
} RSTable2D;

typedef struct _rs_table3d {
    float     xs, xo, xm;
    uint32_t  x_;
    float     ys, yo, ym;
    uint32_t  y_;
    float     zs, zo, zm;
    uint32_t  z_;
    float     tr;
    float     reserved1;
    float     reserved2;
    uint32_t  spacing;
    cl_float4 *uvwt;
    cl_float4 *cpxx;
} RSTable3D;

typedef union {
    cl_float __attribute__((aligned(16))) s[4];
    __extension__ struct { cl_float r, a, e, w; };
    __extension__ struct { cl_float s0, s1, s2, s3; };
} RSPolar;

typedef struct _rs_box {
    RSPolar origin;
    RSPolar size;
} RSBox;

typedef uint32_t RSTable1DDescription;
typedef uint32_t RSTable3DDescription;
typedef uint32_t RSTable3DStaggeredDescription;
typedef uint32_t RSTableDescription;
typedef uint32_t RSTableStaggeredDescription;
typedef uint32_t RSSimulationDescription;
typedef uint32_t RSDropSizeDistribution;
typedef uint32_t RSTableSpacing;
typedef uint32_t RSSimulationConcept;

/* One GPU worker (rs.h:150-266 with the OpenCL objects folded into `engine`). */
typedef struct _rs_worker {
    char              name;
    char              verb;
    rsd_engine        *engine;
    cl_uint           num_cus;
    size_t            num_scats;
    size_t            origins[RS_MAX_DEBRIS_TYPES];
    size_t            counts[RS_MAX_DEBRIS_TYPES];
    RSMakePulseParams make_pulse_params;
    cl_float4         range_weight_desc;
    cl_float4         angular_weight_desc;
    cl_float4         rcs_ellipsoid_desc;
    cl_float16        adm_desc[RS_MAX_ADM_TABLES];
    cl_uint           adm_count;
    cl_float16        rcs_desc[RS_MAX_RCS_TABLES];
    cl_uint           rcs_count;
    cl_float16        les_desc;
    unsigned int      les_id;
    cl_float16        dff_desc;
    float             *dff_icdf[2];   /* host copies of the flux field of each LES slot */
    int               dff_count[2];
    size_t            mem_size;
    size_t            mem_usage;
} RSWorker;

typedef struct _rs_handle RSHandle;

struct _rs_handle {
    char                verb;
    char                method;
    RSParams            params;
    unsigned int        random_seed;
    char                status;
    RSfloat             sim_tic;
    RSfloat             sim_toc;
    cl_float16          sim_desc;
    RSSimulationConcept sim_concept;
    uint32_t            vel_idx;
    uint32_t            vel_count;
    uint32_t            adm_idx;
    uint32_t            rcs_idx;
    LESTable            vel_desc;
    ADMTable            adm_desc[RS_MAX_DEBRIS_TYPES];
    RCSTable            rcs_desc[RS_MAX_DEBRIS_TYPES];
    size_t              num_scats;
    size_t              num_types;
    size_t              counts[RS_MAX_DEBRIS_TYPES];
    cl_uint4            *scat_uid;
    cl_float4           *scat_pos;
    cl_float4           *scat_vel;
    cl_float4           *scat_ori;
    cl_float4           *scat_tum;
    cl_float4           *scat_aux;
    cl_float4           *scat_rcs;
    cl_float4           *scat_sig;
    cl_uint4            *scat_rnd;
    cl_float4           *pulse;
    cl_float4           *pulse_tmp[RS_MAX_GPU_DEVICE];
    size_t              mem_size;
    cl_uint             num_devs;
    cl_uint             num_workers;
    cl_uint             preferred_multiple;
    cl_uint             num_cus[RS_MAX_GPU_DEVICE];
    cl_uint             vendors[RS_MAX_GPU_DEVICE];
    char                has_vbo_from_gl;
    RSWorker            workers[RS_MAX_GPU_DEVICE];
    size_t              offset[RS_MAX_GPU_DEVICE];
    ssize_t             num_anchors;
    ssize_t             num_anchor_lines;
    cl_uint4            draw_mode;
    cl_float4           *anchor_pos;
    cl_float4           *anchor_lines;
    char                dsd_name;
    RSfloat             dsd_n0;
    RSfloat             dsd_lambda;
    RSfloat             dsd_mu;
    int                 dsd_count;
    RSfloat             dsd_nd_sum;
    size_t              *dsd_pop;
    RSfloat             *dsd_pdf;
    RSfloat             *dsd_cdf;
    RSfloat             *dsd_r;
    char                summary[2048];
    LESHandle           L;
    OBJHandle           O;
    POSHandle           P;
    /* additions */
    size_t              population_override;    /* RS_set_population */
    float               range_cap;              /* RS_set_range_cap (reference: 10 km) */
    uint32_t            phase_mode;             /* RSD_PHASE_F64 (default) or RSD_PHASE_F32 */
    double              *pulse_acc;             /* binary64 gate sums of the last pulse, R x 4 */
    int                 last_error;             /* RSD_* code of the last failed engine call */
    char                owns_pattern;
    uint32_t            les_blend;              /* RS_set_les_time_interpolation */
    float               wind_alpha;             /* time-interpolation weight of the last advance */
};

unsigned int RS_gpu_count(void);
char *RS_version_string(void);
int RS_indent_copy(char *dst, char *src, const int width);

RSHandle *RS_init_with_path(const char *bundle_path, RSMethod method, const uint8_t gpu_mask,
                            cl_context_properties sharegroup, const char verb);
RSHandle *RS_init_for_selected_gpu(const uint8_t gpu_mask, const char verb);
RSHandle *RS_init_for_cpu_verbose(const char verb);
RSHandle *RS_init_verbose(const char verb);
RSHandle *RS_init(void);
void RS_free(RSHandle *H);

RSMakePulseParams RS_make_pulse_params(const cl_uint count, const cl_uint group_size_multiple, const cl_uint user_group_counts,
                                       const cl_uint max_local_mem_size, const float range_start, const float range_delta,
                                       const unsigned int range_count);

void RS_set_concept(RSHandle *H, RSSimulationConcept c);
char *RS_simulation_concept_string(RSHandle *H);
char *RS_simulation_concept_bulleted_string(RSHandle *H);

void RS_set_prt(RSHandle *H, const RSfloat prt);
void RS_set_lambda(RSHandle *H, const RSfloat lambda);
void RS_set_density(RSHandle *H, const RSfloat density);
void RS_set_antenna_params(RSHandle *H, RSfloat beamwidth_deg, RSfloat gain_dbi);
void RS_set_tx_params(RSHandle *H, RSfloat pulsewidth, RSfloat tx_power_watt);
void RS_set_sampling_spacing(RSHandle *H, const RSfloat range, const RSfloat azimuth, const RSfloat elevation);
void RS_set_scan_box(RSHandle *H, RSBox box);
void RS_set_scan_extent(RSHandle *H,
                        RSfloat range_start, RSfloat range_end, RSfloat range_gate,
                        RSfloat azimuth_start, RSfloat azimuth_end, RSfloat azimuth_gate,
                        RSfloat elevation_start, RSfloat elevation_end, RSfloat elevation_gate);
void RS_set_beam_pos(RSHandle *H, RSfloat az_deg, RSfloat el_deg);
void RS_set_verbosity(RSHandle *H, const char verb);
void RS_set_debris_count(RSHandle *H, const int debris_id, const size_t count);
size_t RS_get_debris_count(RSHandle *H, const int debris_id);
size_t RS_get_worker_debris_count(RSHandle *H, const int debris_id, const int worker_id);
size_t RS_get_all_worker_debris_counts(RSHandle *H, const int debris_id, size_t counts[]);
RSVolume RS_get_domain(RSHandle *H);

void RS_set_dsd(RSHandle *H, const float *nd, const float *diameters, const int count, const char name);
void RS_set_dsd_to_mp(RSHandle *H);
void RS_set_dsd_to_mp_with_sizes(RSHandle *H, const float *ds, const int count);

void RS_set_range_weight(RSHandle *H, const float *weights, const float table_index_start, const float table_index_delta, unsigned int table_size);
void RS_set_range_weight_to_triangle(RSHandle *H, float pulse_width_m);
void RS_set_angular_weight(RSHandle *H, const float *weights, const float table_index_start, const float table_index_delta, unsigned int table_size);
void RS_set_angular_weight_to_standard(RSHandle *H, float beamwidth_rad);
void RS_set_angular_weight_to_double_cone(RSHandle *H, float beamwidth_rad);

void RS_set_vel_data_to_config(RSHandle *H, LESConfig c);
void RS_set_vel_data_to_LES_table(RSHandle *H, const LESTable *table);
void RS_set_vel_data_to_uniform(RSHandle *H, cl_float4 velocity);
void RS_set_vel_data_to_cube27(RSHandle *H);
void RS_set_vel_data_to_cube125(RSHandle *H);
void RS_clear_vel_data(RSHandle *H);

void RS_set_debris_flux_field_by_pdf(RSHandle *H, RSTable2D *map, const float *pdf);
void RS_set_debris_flux_field_by_icdf(RSHandle *H, RSTable2D *map, const float *icdf);
void RS_set_debris_flux_field_to_center_cell_of_3x3(RSHandle *H);
void RS_set_debris_flux_field_to_checker_board(RSHandle *H, const int);
void RS_set_debris_flux_field_to_checker_board_stretched(RSHandle *H, const LESTable *table);
void RS_set_debris_flux_field_from_LES(RSHandle *H, const LESTable *leslie);

void RS_set_scan_pattern(RSHandle *H, const POSPattern *scan_pattern);
void RS_set_scan_pattern_with_string(RSHandle *H, const char *scan_string);

void RS_set_random_seed(RSHandle *H, const unsigned int seed);
void RS_add_debris(RSHandle *H, OBJConfig type, const size_t count);

void RS_io_test(RSHandle *H);

void RS_populate(RSHandle *H);

void RS_upload(RSHandle *H);
void RS_download(RSHandle *H);
void RS_download_position_only(RSHandle *H);
void RS_download_orientation_only(RSHandle *H);
void RS_download_pulse_only(RSHandle *H);

void RS_compute_rcs_ellipsoids(RSHandle *H);

void RS_advance_time(RSHandle *H);
void RS_advance_beam(RSHandle *H);
void RS_make_pulse(RSHandle *H);

RSTable RS_table_init(size_t numel);
void RS_table_free(RSTable T);
RSTable2D RS_table2d_init(size_t numel);
void RS_table2d_free(RSTable2D T);
RSTable3D RS_table3d_init(size_t numel);
void RS_table3d_free(RSTable3D T);

void RS_show_radar_params(RSHandle *H);
void RS_show_scat_pos(RSHandle *H);
void RS_show_scat_sig(RSHandle *H);
void RS_show_scat_att(RSHandle *H);
void RS_show_pulse(RSHandle *H);

RSBox RS_suggest_scan_domain(RSHandle *H);
void RS_revise_debris_counts_to_gpu_preference(RSHandle *H);
char *RS_simulation_description(RSHandle *H);

/* Private setters (rs_priv.h:109-123), used to inject tables. */
void RS_set_vel_data(RSHandle *H, const RSTable3D table);
void RS_set_adm_data(RSHandle *H, const RSTable2D table_cd, const RSTable2D table_cm);
void RS_set_rcs_data(RSHandle *H, const RSTable2D table_real, const RSTable2D table_imag);
void RS_set_adm_data_to_ADM_table(RSHandle *H, const ADMTable *table);
void RS_set_adm_data_to_unity(RSHandle *H);
void RS_clear_adm_data(RSHandle *H);
void RS_set_rcs_data_to_RCS_table(RSHandle *H, const RCSTable *table);
void RS_set_rcs_data_to_unity(RSHandle *H);
void RS_clear_rcs_data(RSHandle *H);
void RS_set_rcs_ellipsoid_table(RSHandle *H, const cl_float4 *weights, const float table_index_start,
                                const float table_index_delta, unsigned int table_size);
void RS_revise_population(RSHandle *H);
void RS_update_origins_offsets(RSHandle *H);
void RS_merge_pulse_tmp(RSHandle *H);

/* addition: a handle without a GPU (table, geometry and population logic only;
 * online calls report an error).  For callers and tests on machines without a
 * device. */
RSHandle *RS_init_host_only(const char verb);
/* addition: exact background population, bypassing the rounding of
 * RS_revise_population (rs.c:2172-2178) to device-preferred multiples. */
void RS_set_population(RSHandle *H, size_t rain_count);
/* addition: maximum padded range of the scan extent (default 10 km, rs.c:1561) */
void RS_set_range_cap(RSHandle *H, float r_max);
/* addition: phase precision of the pulse (RSD_PHASE_F64 default). */
void RS_set_phase_precision(RSHandle *H, uint32_t phase_mode);
/* addition: binary64 gate sums of the last pulse (R x 4, before the gain);
 * returns the number of gates. */
int RS_get_pulse_f64(RSHandle *H, double *out);
/* addition: radar moments (Z, V, ZDR, rho_hv) over the last `pulses` pulses,
 * computed on the device from its IQ history (no counterpart in the reference) */
void RS_set_moment_pulses(RSHandle *H, int pulses);
/* addition: one pulse round of many handles whose engines form the batch b
 * (include/rsd.h rsd_batch_create): make_pulse of all, one device launch,
 * advance_time + advance_beam of all; returns RSD_OK or the launch error */
int RS_batch_round(RSHandle *const *handles, int count, rsd_batch *b);
int RS_get_moments(RSHandle *H, int pulses, float *out);
/* addition: descriptor accessors (which: 0 sim, 1 les, 2 range weight,
 * 3 angular weight, 4 drop rcs, 16+t adm table t, 32+t rcs table t);
 * returns the number of floats written (4 or 16). */
int RS_get_descriptor(RSHandle *H, int which, float *out16);
/* addition: radar gain applied by RS_merge_pulse_tmp (rs.c:5084). */
float RS_radar_gain(RSHandle *H);
/* addition: engine of worker 0, for timing and launch control. */
rsd_engine *RS_engine(RSHandle *H);
/* addition: plain-data view of the handle for bindings that cannot read the
 * struct directly. */
typedef struct {
    uint64_t num_scats;
    uint64_t num_types;
    uint64_t counts[RS_MAX_DEBRIS_TYPES];
    uint64_t origins[RS_MAX_DEBRIS_TYPES];
    float    sim_tic;
    float    sim_toc;
    uint32_t status;
    uint32_t sim_concept;
    uint32_t random_seed;
    uint32_t vel_idx;
    uint32_t les_id;
    uint32_t range_count;
    int32_t  last_error;
    int32_t  dsd_count;
    uint32_t adm_count;
    uint32_t rcs_count;
    uint32_t les_blend;     /* 1: LES volumes interpolated in time */
    float    wind_alpha;    /* weight of the next volume in the last advance */
} RSInfo;
void RS_get_info(RSHandle *H, RSInfo *out);
/* addition: on = 1 interpolates the wind linearly in time between LES volume k
 * (at k tp) and volume k + 1, both resident in the two slots, with weight
 * (sim_tic - k tp) / tp; on = 0 (default) keeps the reference's volume switch
 * every tp (rs.c:5189-5203). */
void RS_set_les_time_interpolation(RSHandle *H, int on);
/* addition: host copies of the streams (which: 0 pos, 1 vel, 2 ori, 3 tum,
 * 4 aux, 5 rcs, 6 sig, 7 rnd, 8 pulse, 9 anchors); NULL before RS_populate. */
void *RS_get_host_stream(RSHandle *H, int which);
/* addition: RSParams of the handle. */
void RS_get_params(RSHandle *H, RSParams *out);
/* addition: RS_suggest_scan_domain / RS_set_scan_box with the box as eight
 * floats (origin r, a, e, w, size r, a, e, w) for bindings that cannot pass a
 * 16-byte-aligned struct by value. */
void RS_suggest_scan_domain_values(RSHandle *H, float *box8);
void RS_set_scan_box_values(RSHandle *H, const float *box8);
/* addition: host LES volume `frame` of the handle's provider (float4 texels,
 * x fastest) and its dimensions; NULL when there is none. */
const float *RS_get_les_frame(RSHandle *H, int frame, uint32_t *dims3);
/* addition: the (cn2, p, _, _) volume of an LES frame, laid out as
 * RS_get_les_frame (made for a synthetic frame on the first request). */
const float *RS_get_les_cpxx(RSHandle *H, int frame, uint32_t *dims3);
/* addition: float4 texels of the t-th debris object's tables as uploaded
 * (which: 0 ADM cd, 1 ADM cm, 2 RCS real, 3 RCS imag); dims2 = (nx, ny).
 * out4 may be NULL to query the size; returns 0 on success. */
int RS_get_debris_table(RSHandle *H, int t, int which, float *out4, uint32_t *dims2);
/* addition: DSD bins (radius, pdf, cdf; count floats each); returns count. */
int RS_get_dsd(RSHandle *H, float *r, float *pdf, float *cdf, int cap);
/* addition: the debris flux field of LES slot `slot` (0 or 1) as uploaded:
 * inverse CDF (up to cap entries; icdf may be NULL to query) and the shared
 * descriptor (desc16 may be NULL); returns the entry count, 0 when none. */
int RS_get_debris_flux_field(RSHandle *H, int slot, float *icdf, int cap, float *desc16);

#ifdef __cplusplus
}
#endif

#endif
