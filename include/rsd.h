/*
 * rsd.h -- C-ABI of the gfx950 scatterer engine (the device layer under RS_*).
 *
 * The reference drives OpenCL directly from rs.c.  Each group of entry points
 * below replaces one of those call clusters; the RS_* host runtime
 * (include/rs.h) is written against this interface only.
 *
 *   rsd_open / rsd_close            RS_worker_init / RS_worker_free (rs.c:383-620)
 *   rsd_alloc_state                 RS_worker_malloc buffer creation (rs.c:623-944)
 *   rsd_upload_state                RS_upload (rs.c:5124-5175)
 *   rsd_download_state              RS_download, RS_download_position_only,
 *                                   RS_download_orientation_only (rs.c:4961-5064)
 *   rsd_set_wind                    RS_set_vel_data image upload (rs.c:2854-2977)
 *   rsd_set_adm                     RS_set_adm_data (rs.c:3642-3762)
 *   rsd_set_rcs                     RS_set_rcs_data (rs.c:3852-3972)
 *   rsd_set_drop_rcs                RS_set_rcs_ellipsoid_table (rs.c:2472-2545)
 *   rsd_set_dff                     RS_set_debris_flux_field_by_icdf (rs.c:3331-3404)
 *   rsd_set_dff_slot                its per-buffer write, dff_icdf[les_id] (rs.c:3364-3366)
 *   rsd_set_range_weight            RS_set_range_weight (rs.c:2548-2618)
 *   rsd_set_angular_weight          RS_set_angular_weight (rs.c:2626-2698)
 *   rsd_advance                     enqueue block of RS_advance_time (rs.c:5327-5388)
 *   rsd_make_pulse                  enqueue block of RS_make_pulse (rs.c:5526-5584)
 *   rsd_read_pulse                  RS_download_pulse_only read (rs.c:5096-5121)
 *
 * Conventions: plain pointers and sizes only; every function returns RSD_OK (0)
 * or a negative RSD_ERR_* code, never exits; rsd_last_error() gives the text of
 * the most recent failure on that engine.  Streams are the reference's float4 /
 * uint4 layouts (16 bytes per scatterer per attribute), descriptors are the
 * reference's float16 / float4 slot layouts (rs_enum.h).  One engine drives one
 * GPU; calls on one engine must not overlap.
 */
#ifndef RSD_H
#define RSD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSD_OK              0
#define RSD_ERR_NO_DEVICE  -1
#define RSD_ERR_ALLOC      -2
#define RSD_ERR_ARG        -3
#define RSD_ERR_HIP        -4
#define RSD_ERR_STATE      -5

#define RSD_MAX_TYPES       8
#define RSD_MAX_TABLES      8
#define RSD_MAX_GATES       4096

typedef struct rsd_engine rsd_engine;

/* Advance kind of the background (type 0) scatterers. */
#define RSD_ADVANCE_DRAGGED     1u   /* el_atts (concept D, rs.cl:653-739) */
#define RSD_ADVANCE_BACKGROUND  0u   /* bg_atts (rs.cl:533-580) */
#define RSD_ADVANCE_FIXED       2u   /* fp_atts (concept F, rs.cl:590-648): fixed positions,
                                        rcs from the Cn^2 volume with an accumulated phase */

/* Precision of the phase term of the pulse. */
#define RSD_PHASE_F64           0u   /* 4 pi r / lambda reduced mod 2 pi in binary64 */
#define RSD_PHASE_F32           1u   /* r * (4 pi / lambda) in binary32, as rs.cl:1096 */

typedef struct {
    float    sim_desc[16];           /* reference sim_desc at the RS_advance_time call */
    int32_t  wind_slot;              /* active LES slot (0 or 1) */
    int32_t  adm_index[RSD_MAX_TYPES];   /* ADM table of each debris type (entry 0 unused) */
    int32_t  rcs_index[RSD_MAX_TYPES];   /* RCS table of each debris type (entry 0 unused) */
    uint32_t background_kind;        /* RSD_ADVANCE_DRAGGED, RSD_ADVANCE_BACKGROUND or RSD_ADVANCE_FIXED */
    uint32_t wind_blend;             /* 1: wind = lerp(slot wind_slot, the other slot, wind_alpha); 0: the slot alone */
    float    wind_alpha;             /* time-interpolation weight in [0, 1] (wind_blend only) */
} rsd_step;

typedef struct {
    float    sim_desc[16];           /* reference sim_desc at the RS_make_pulse call (beam) */
    double   two_over_lambda;        /* 2 / lambda in binary64: phase in turns per metre */
    float    range_start;            /* first gate range (m) */
    float    range_delta;            /* gate spacing (m) */
    uint32_t range_count;            /* number of gates R */
    int32_t  rcs_index[RSD_MAX_TYPES];
    uint32_t debris_rcs_refresh;     /* 1 when the beam moved since the last pulse (rs.c:5531) */
    uint32_t phase_mode;             /* RSD_PHASE_F64 or RSD_PHASE_F32 */
} rsd_pulse;

int  rsd_device_count(void);
/* Compute units, maximum work-group size and LDS bytes per work-group of a device. */
int  rsd_device_props(int device_index, int *compute_units, int *max_work_group, size_t *local_mem);
int  rsd_open(int device_index, rsd_engine **out);
void rsd_close(rsd_engine *e);
const char *rsd_last_error(const rsd_engine *e);

/* Scatterer population: counts[k] scatterers of type k at origins[k] (type 0 =
 * rain at origin 0, debris types contiguous after it), num_scats in total. */
int rsd_alloc_state(rsd_engine *e, size_t num_scats, const size_t *origins, const size_t *counts);

/* Full-stream transfers (num_scats elements each; any pointer may be NULL on
 * download).  sig and aux on download are those of the most recent pulse: its
 * beam and the positions it saw, also after an advance since (rs.cl:1050-1110
 * writes them in the pulse and the advance leaves them).  Before any pulse sig
 * is zero and aux is the stored stream. */
int rsd_upload_state(rsd_engine *e, const float *pos, const float *vel, const float *ori,
                     const float *tum, const float *aux, const float *rcs, const uint32_t *rnd);
int rsd_download_state(rsd_engine *e, float *pos, float *vel, float *ori, float *tum,
                       float *aux, float *rcs, float *sig, uint32_t *rnd);

/* Tables.  Texels are float4, x fastest.  Descriptors use the reference slots. */
int rsd_set_wind(rsd_engine *e, int slot, const float *uvwt, int nx, int ny, int nz, const float *les_desc16);
/* Index tables of a stretched LES descriptor (slot 7 = 7) for the box of
 * sim_desc16 (slots 8-10 origin, 12-14 size), as every advance evaluates the
 * index c = s8 + copysign(s0 log1p(s4 |r|), r) of rs.cl:130-151: per axis
 * (x, y, z) 256 cubic pieces of P(x) = s0 log1p(s4 x) over x in [0, X], X the
 * box's extent from the axis origin (box centre in x and y, 0 in z) times
 * 1 + 2^-10; rows[4 (256 a + j) + 0..3] = binary32 monomial coefficients of
 * piece j in f, x = (j + f) / ih[a], interpolating binary64 P at the Chebyshev
 * nodes 0.5 + 0.5 cos((2k + 1) pi / 8); jm[a] = 255.  The engine builds them
 * with this function on the host (no device needed) whenever a wind slot's
 * descriptor or the box changes.  No counterpart in the reference (which
 * evaluates log1p per scatterer). */
int rsd_wind_index_table(const float *les_desc16, const float *sim_desc16, float *rows, float *ih, float *jm);
/* The (cn2, p, _, _) volume of the LES slot (the second image RS_set_vel_data
 * uploads, rs.c:2925), read with the slot's wind descriptor by the concept-F
 * advance; same shape as the slot's wind table.  Waits for the device. */
int rsd_set_cpxx(rsd_engine *e, int slot, const float *cpxx, int nx, int ny, int nz);
int rsd_set_adm(rsd_engine *e, int table, const float *cd, const float *cm, int nx, int ny, const float *adm_desc16);
int rsd_set_rcs(rsd_engine *e, int table, const float *re, const float *im, int nx, int ny, const float *rcs_desc16);
int rsd_set_drop_rcs(rsd_engine *e, const float *table4, int n, const float *desc4);
int rsd_set_range_weight(rsd_engine *e, const float *w, int n, const float *desc4);
int rsd_set_angular_weight(rsd_engine *e, const float *w, int n, const float *desc4);
/* Debris flux field (concept C): inverse CDF of `count` entries and its
 * descriptor (RS_set_debris_flux_field_by_icdf, rs.c:3331-3404, read by
 * db_atts at rs.cl:795-836); count 0 removes it.  Used on debris respawn
 * while the step's concept has bit 7 set.  The engine keeps one table per
 * wind slot and a step reads its wind_slot's (the reference's dff_icdf[les_id],
 * rs.c:3351-3373, 5358); rsd_set_dff sets both. */
int rsd_set_dff(rsd_engine *e, const float *icdf, int count, const float *desc16);
/* The table of one wind slot (0 or 1), e.g. the field of the LES frame just
 * loaded into that slot: written in stream order through pinned staging, so
 * steps already enqueued read the previous table.  The first table, a removal
 * (count 0) or a change of size sets both slots, as rsd_set_dff.  Replaces the
 * per-buffer write of RS_set_debris_flux_field_by_icdf (rs.c:3364-3366). */
int rsd_set_dff_slot(rsd_engine *e, int slot, const float *icdf, int count, const float *desc16);

/* One PRT of every scatterer type.  May be held back and fused with the next
 * rsd_make_pulse; any transfer or rsd_sync performs it first. */
int rsd_advance(rsd_engine *e, const rsd_step *step);

/* Per-gate complex sums of all scatterers (before the radar gain) into the
 * engine's pulse buffer.  Enqueued; rsd_read_pulse waits for it. */
int rsd_make_pulse(rsd_engine *e, const rsd_pulse *pulse);

/* rsd_make_pulse without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same advance and
 * pulse as rsd_make_pulse (a held-back advance is fused in), with its per-block
 * partial sums and gate ranges in a 16 MiB array of each GPU that the library
 * holds until the process ends.  The result, out[4 k + c] = (Ih, Qh, Iv, Qv) of
 * gate k before the radar gain, goes to `out` (device memory, or host memory
 * the device can write) and to the engine's pulse buffer (rsd_read_pulse).
 * Never reorders the streams and does not append to the IQ ring.  The rain
 * grid shrinks to the array's rows (4 range_count doubles each); a population
 * whose debris blocks alone need more rows is rejected with RSD_ERR_ARG.
 * Engines in a batch are rejected.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race. */
int rsd_make_pulse_noalloc(rsd_engine *e, const rsd_pulse *pulse, double *out);

/* Blocking read of the last pulse: out[4 k + c] = (Ih, Qh, Iv, Qv) of gate k. */
int rsd_read_pulse(rsd_engine *e, double *out, uint32_t range_count);

int rsd_sync(rsd_engine *e);

/* Timing of everything enqueued between the two calls, in milliseconds,
 * measured with HIP events on the engine's stream. */
int rsd_timer_start(rsd_engine *e);
int rsd_timer_stop(rsd_engine *e, double *ms);

/* Per-launch HIP-event timing of the background (rain) kernel on the engine's
 * stream: rsd_profile(e, n) clears and starts recording every n-th launch
 * (n = 1: all; 0 stops); rsd_profile_read sums the durations of the recorded
 * launches of one kind (1 advance only, 2 pulse only, 3 fused advance + pulse,
 * -1 all). */
int rsd_profile(rsd_engine *e, int enable);
int rsd_profile_read(rsd_engine *e, int kind, double *total_ms, int *count);

/* Bandwidth probe (RS_io_test, rs.c:4389-4427): reps copies of n float4; returns
 * the achieved read + write rate in GB/s. */
int rsd_io_test(rsd_engine *e, size_t n, int reps, double *gbps);

/* Launch shape of the scatterer kernels.  block_size 64, 128 or 256 (0: 256);
 * max_blocks caps the grid of every scatterer launch (1..65536).  max_blocks 0
 * (the default): a fused rain launch with no debris kernel beside it takes two
 * resident rounds of equal blocks (occupancy x compute units, less the blocks
 * of the riding debris types); other launches take at most 2048 blocks. */
int rsd_set_launch(rsd_engine *e, int block_size, int max_blocks);

/* Spatial order of the rain streams: every `interval` pulses (default 64,
 * 0 = never) the engine reorders the stored rain scatterers by range shell
 * (1/8 gate) and box cell.  Invisible to callers: downloads return the
 * reference order and uploads restore it.  No counterpart in the reference
 * (its streams stay in index order); it lets each lane sum its scatterers'
 * gate terms in registers and keeps wind-table gathers local. */
int rsd_set_sort(rsd_engine *e, int interval);
/* Pulses between reorders currently in effect (0: never). */
int rsd_get_sort(const rsd_engine *e);

/* Strategy of the per-gate sums inside a work-group (results agree to
 * floating-point reassociation; all end in binary64 LDS and block partials):
 *   RSD_FOLD_LANE    every lane adds its own terms to the LDS gate sums;
 *   RSD_FOLD_WAVE64  where a wave's scatterers fall into at most 16 gates, the
 *   RSD_FOLD_WAVE32  wave sums each gate's terms (binary64 / binary32 partials
 *                    of 64 terms) and one lane adds them;
 *   RSD_FOLD_WINDOW  (default) terms are summed in registers (binary32) over a
 *                    window of 8 consecutive gates and added to LDS once per
 *                    launch: for rain the window is shared by the 64 lanes of a
 *                    wave (streams in range order) and reduced in one transposed
 *                    wave pass; for debris each lane keeps its own window.
 *                    Taken when the range weight spans under 6 gate spacings
 *                    and its table ends in zeros, else WAVE32.
 * No counterpart in the reference (one work-item per scatterer, dense gate loop,
 * rs.cl:1182-1217). */
#define RSD_FOLD_LANE    0
#define RSD_FOLD_WAVE64  1
#define RSD_FOLD_WAVE32  2
#define RSD_FOLD_WINDOW  3
int rsd_set_fold(rsd_engine *e, int mode);

/* Position precision (no counterpart in the reference, which keeps binary32
 * positions): hilo = 1 carries every position as a pair of binary32 values
 * hi + lo (hi in the pos stream, lo in a stream of its own, reset to 0 on
 * upload and on respawn); the advance adds v dt to the pair with TwoProduct /
 * TwoSum and the pulse forms the range for the phase in binary64 from hi + lo.
 * The physics (wind lookup, drag, box test) reads hi.  Downloads return hi in
 * pos; rsd_download_pos_lo returns lo in the reference order. */
int rsd_set_position_precision(rsd_engine *e, int hilo);
int rsd_download_pos_lo(rsd_engine *e, float *lo);

/* Radar moments (no counterpart in the reference, which has no estimator; an
 * unused zdr() exists at rs.c:376-378).  rsd_set_iq_ring keeps the last
 * `pulses` pulses of `range_count` gates on the device (0 frees it); every
 * rsd_make_pulse with that gate count appends to it.  rsd_moments computes, per
 * gate over the most recent `pulses` pulses (2 <= pulses <= ring size), with
 * H = Ih + j Qh and V = Iv + j Qv scaled by `gain`:
 *   R0h = <|H_k|^2>, R0v = <|V_k|^2>, R1 = <conj(H_k) H_k+1>, C = <H_k conj(V_k)>
 *   Z   = 10 log10(R0h) + 20 log10(r_g)   (dB; range-corrected power, radar constant 0)
 *   V   = -lambda / (4 pi PRT) arg(R1)    (m/s, positive away from the radar)
 *   ZDR = 10 log10(R0h / R0v)             (dB)
 *   rho = |C| / sqrt(R0h R0v)
 * with r_g = range_start + g range_delta; out[4 g + (0..3)] = (Z, V, ZDR, rho). */
typedef struct {
    double lambda;          /* wavelength (m) */
    double prt;             /* pulse repetition time (s) */
    double gain;            /* amplitude gain applied to the IQ (RS radar gain) */
    double range_start;     /* range of gate 0 (m) */
    double range_delta;     /* gate spacing (m) */
} rsd_moment_params;
int rsd_set_iq_ring(rsd_engine *e, int pulses, int range_count);
int rsd_moments(rsd_engine *e, int pulses, const rsd_moment_params *params, float *out);

/* Batches (no counterpart in the reference, which runs one domain per handle):
 * many engines on one GPU, each an independent rain domain of its own size,
 * tables and pulse parameters, advanced and pulsed by ONE persistent kernel
 * launch plus one finalize per pulse round.  While an engine belongs to a
 * batch, rsd_make_pulse only records the pulse (and the held-back advance);
 * rsd_batch_launch launches every recorded pulse (rsd_read_pulse,
 * downloads and rsd_sync launch it too).  Rain-only engines with binary32
 * positions share the persistent launch; others are launched on their own
 * streams.  The launch is captured once as a HIP graph (descriptor upload,
 * persistent kernel, finalize) and replayed every round; the chunk size and
 * blocks per compute unit are picked by an online autotuner over the first
 * rounds (0 = tune) -- results differ only by reassociation of the gate sums.
 * Destroy the batch before closing its engines. */
typedef struct rsd_batch rsd_batch;
int  rsd_batch_create(rsd_engine *const *engines, int count, rsd_batch **out);
void rsd_batch_destroy(rsd_batch *b);
int  rsd_batch_launch(rsd_batch *b);
int  rsd_batch_sync(rsd_batch *b);
int  rsd_batch_set_launch(rsd_batch *b, int chunk, int blocks_per_cu);
int  rsd_batch_get_launch(rsd_batch *b, int *chunk, int *blocks_per_cu, int *tuning);
int  rsd_batch_set_graph(rsd_batch *b, int on);
int  rsd_batch_timer_start(rsd_batch *b);
int  rsd_batch_timer_stop(rsd_batch *b, double *ms);
const char *rsd_batch_last_error(const rsd_batch *b);

/* Debris surrogate network (no counterpart in the reference; it replaces the
 * ADM + RCS table lookups of rs.cl:195-359).  7 inputs (radar-relative
 * orientation quaternion R = ori (x) o_conj(el, az), body-frame unit air-flow
 * direction u) -> 4 hidden layers of 64 (ReLU) -> 12 outputs (cd xyz, cm xyz,
 * gamma-rotated hh, vv, hv as re, im), each output scaled and shifted.  BF16
 * weights and activations, FP32 accumulation, on MFMA; FP32 master weights.
 * Parameters, flat (RSD_MLP_PARAMS floats): W0[64][7], b0[64], W1[64][64], b1,
 * W2, b2, W3, b3, W4[12][64], b4[12] (row-major [out][in]).
 * rsd_mlp_gradient: loss = mean over the n x 12 outputs of (y - t)^2 (outputs
 * before scaling, t normalised), gradient in the same flat layout; n a
 * multiple of 256.  rsd_mlp_set_data uploads a training set (x n x 7, t n x 12
 * normalised); rsd_mlp_train runs Adam (cosine-decayed lr) over it. */
#define RSD_MLP_PARAMS 13772
typedef struct rsd_mlp rsd_mlp;
int  rsd_mlp_create(int device_index, uint64_t seed, rsd_mlp **out);
void rsd_mlp_destroy(rsd_mlp *m);
int  rsd_mlp_set_params(rsd_mlp *m, const float *params);
int  rsd_mlp_get_params(rsd_mlp *m, float *params);
int  rsd_mlp_set_output_scaling(rsd_mlp *m, const float *scale12, const float *shift12);
int  rsd_mlp_forward(rsd_mlp *m, const float *x, float *y, size_t n);
int  rsd_mlp_gradient(rsd_mlp *m, const float *x, const float *t, size_t n, float *grad, double *loss);
int  rsd_mlp_set_data(rsd_mlp *m, const float *x, const float *t, size_t n);
int  rsd_mlp_train(rsd_mlp *m, int epochs, int batch, float lr, double *last_loss);
int  rsd_mlp_sync(rsd_mlp *m);
const char *rsd_mlp_last_error(const rsd_mlp *m);

/* Training samples from the table path: n random orientations R (uniform on
 * SO(3), R.w >= 0) and flow directions u (uniform on the sphere) with the
 * network's 12 targets read from ADM table adm_table and RCS table rcs_table
 * exactly as the debris kernels read them; x[n][7] = (R, u), t[n][12]. */
int rsd_debris_samples(rsd_engine *e, int adm_table, int rcs_table, size_t n, uint64_t seed, float *x, float *t);
/* Training samples at the population's current states: for the first
 * min(count, cap) bodies of debris type `type`, x[n][7] = the inputs the debris
 * kernel forms (R from position and orientation, u from velocity and the wind
 * of wind_slot at the position, for the box of sim_desc16) and t[n][12] = the
 * targets read from ADM table adm_table and RCS table rcs_table there; *n_out =
 * the rows written.  No counterpart in the reference. */
int rsd_debris_state_samples(rsd_engine *e, int type, int adm_table, int rcs_table, int wind_slot,
                             const float *sim_desc16, size_t cap, float *x, float *t, size_t *n_out);
/* Use network m (NULL: the tables) for debris type `type` (>= 1) of engine e:
 * the debris kernel evaluates it once per scatterer and pulse in place of the
 * ADM and RCS lookups.  m must stay alive while it is in use. */
int rsd_set_debris_surrogate(rsd_engine *e, int type, const rsd_mlp *m);

/* Self-test: evaluate one of the device's binary32 functions on n inputs
 * (fn: 1 log(a), 2 sin(a), 3 cos(a), 4 atan2(a, b), 5 acos(a), 6 sqrt(a),
 * 7 a / b, 8 sin(2 pi a), 9 cos(2 pi a), 10 the pulse path's fast acos(a),
 * 11 the gate window's transposed wave sum of a + j b, j = 0..31 (n a multiple
 * of 64); other values return 0).  No counterpart in the reference; it lets a
 * caller check the device arithmetic the advection uses. */
int rsd_math_probe(int device_index, int fn, const float *a, const float *b, float *out, size_t n);

#ifdef __cplusplus
}
#endif

#endif
