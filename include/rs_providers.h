/*
 * rs_providers.h -- table providers and scan patterns used by the RS_* API.
 *
 * Interfaces follow the reference's les.h:47-84, adm.h:13-71, rcs.h:20-51,
 * obj.h:13-47, pos.h:22-66 and log.h, because RS_* signatures and the
 * RSHandle fields callers read are written in these types.  The reference's
 * providers read external data files (LES Fortran records, .adm, .rcs) that are
 * not part of the source tree.  These providers read the same files when they
 * are found where the reference looks for them (the given path, or
 * <cwd>/Contents/Resources/..., $HOME/Documents/tables, $HOME/Downloads/tables,
 * $HOME/tables, and for LES also $HOME/{Desktop,Documents,Downloads}/les:
 * les/<config>/fort.10_2 + LES_mean_1_6_fnum<k>.dat, adm/<name>.adm,
 * rcs/<name>.rcs); otherwise every configuration is generated in memory from
 * documented analytic fields and a seeded generator, so a run is reproducible
 * on any machine:
 *
 *   LES "suctvort": 176 x 176 x 80 stretched grid (a = 2 m, r = 1.0212 in x/y,
 *     1.05 in z; les.c:159-171), Rankine vortex at the grid centre (core 50 m,
 *     40 m/s) plus i.i.d. N(0, 2^2) m/s turbulence per component and node; frame
 *     f uses noise stream f; t = 0; tp = 2 s, tr = 50 s (les.c:139, 170).
 *   ADM "plate"/"square_plate" and RCS "leaf"/"plate"/"woodboard"/"brick": 1-degree
 *     grids of smooth trigonometric fits plus seeded N(0, sigma^2) noise.
 *   OBJ dimensions and densities as obj.c:102-148 (ADM_dimension_set).
 */
#ifndef RS_PROVIDERS_H
#define RS_PROVIDERS_H

#include "rs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- logging (log.h) ---- */
char *now(void);
char *commaint(long long num);
char *commafloat(float num);
void rsprint(const char *format, ...);

/* ---- LES (les.h) ---- */
#define LESConfigNull                 ""
#define LESConfigFlat                 "flat"
#define LESConfigTwoCell              "twocell"
#define LESConfigSuctionVortices      "suctvort"
#define LESConfigSuctionVorticesLarge "suctvort_large"

typedef void *LESHandle;
typedef char *LESConfig;
typedef float LESFloat4[4];

typedef struct _les_value {
    float *a, *x, *y, *z, *u, *v, *w, *p, *t;
} LESValue;

typedef struct _les_table {
    uint32_t  nx, ny, nz, nn, nt, nc;
    bool      is_stretched;
    float     tr;
    float     tp;
    float     ax, ay, az;
    float     rx, ry, rz;
    LESValue  data;
    LESFloat4 *uvwt;
    LESFloat4 *cpxx;
    float     *flux;
} LESTable;

LESHandle LES_init_with_config_path(const LESConfig config, const char *path);
LESHandle LES_init(void);
void LES_free(LESHandle);
LESTable *LES_get_frame(const LESHandle, const int n);
/* The frame's (cn2, p, 0, 0) volume, made on the first request for a synthetic
 * frame (the synthetic Cn^2 profile is documented in rs_providers.c). */
const LESFloat4 *LES_get_cpxx(const LESHandle, LESTable *t);
float LES_get_table_period(const LESHandle);
size_t LES_get_table_count(const LESHandle);
/* Synthetic generator parameters (seed of the turbulence, number of frames). */
LESHandle LES_init_synthetic(uint32_t nx, uint32_t ny, uint32_t nz, float a_xy, float r_xy, float a_z, float r_z,
                             float core_m, float vmax, float sigma, uint64_t seed, uint32_t frames);

/* ---- ADM (adm.h) ---- */
#define ADMConfigModelPlate  "plate"
#define ADMConfigSquarePlate "square_plate"
#define ADMConfigRoofTile    "roof_tile"
typedef void *ADMHandle;
typedef char *ADMConfig;

typedef struct _adm_base {
    float x, y, z, rho, mass, Ta, inv_inln_x, inv_inln_y, inv_inln_z;
} ADMBase;

typedef struct _adm_data {
    float *b, *a, *cdx, *cdy, *cdz, *cmx, *cmy, *cmz;
} ADMData;

typedef struct _adm_table {
    uint32_t nb;
    uint32_t na;
    uint32_t nn;
    ADMBase  phys;
    ADMData  data;
    char     name[1024];
    char     path[1024];
} ADMTable;

void ADM_compute_properties(ADMBase *phys);
void ADM_dimension_set(ADMTable *T, const float x, const float y, const float z, const float r);
/* nb x na grid (beta fastest) of a synthetic air-drag model; seed selects the noise. */
ADMTable *ADM_table_synthetic(const char *name, uint32_t nb, uint32_t na, float noise, uint64_t seed);
void ADM_table_free(ADMTable *T);
/* One .adm file (adm.c:152-225 layout); NULL if it cannot be read. */
ADMTable *ADM_table_read(const char *path);
/* adm.h handles: NULL when no folder with adm/square_plate.adm is found. */
ADMHandle ADM_init_with_path(const char *path);
ADMHandle ADM_init(void);
ADMTable *ADM_get_table(const ADMHandle, const ADMConfig);
void ADM_free(ADMHandle);

/* ---- RCS (rcs.h) ---- */
#define RCSConfigLeaf      "leaf"
#define RCSConfigWoodBoard "woodboard"
#define RCSConfigBrick     "brick"
#define RCSConfigPlate     "plate"
typedef void *RCSHandle;
typedef char *RCSConfig;

typedef struct _rcs_data {
    float *a, *b, *hh_real, *vv_real, *hv_real, *hh_imag, *vv_imag, *hv_imag;
} RCSData;

typedef struct _rcs_table {
    uint32_t na;
    uint32_t nb;
    uint32_t nn;
    RCSData  data;
    char     name[1024];
    char     path[1024];
    float    lambda;
} RCSTable;

/* na x nb grid (alpha fastest) of synthetic complex hh/vv/hv amplitudes. */
RCSTable *RCS_table_synthetic(const char *name, uint32_t na, uint32_t nb, float scale, float noise, uint64_t seed);
void RCS_table_free(RCSTable *T);
/* One .rcs file (rcs.c:181-246 layout); NULL if it cannot be read. */
RCSTable *RCS_table_read(const char *path);
/* rcs.h handles: NULL when no folder with rcs/leaf.rcs is found. */
RCSHandle RCS_init_with_path(const char *path);
RCSHandle RCS_init(void);
RCSTable *RCS_get_table(const RCSHandle, const RCSConfig);
void RCS_free(RCSHandle);

/* ---- OBJ (obj.h) ---- */
typedef int OBJConfig;
enum OBJConfig {
    OBJConfigUnknown      = 0,
    OBJConfigLeaf         = 1,
    OBJConfigLeafBig      = 2,
    OBJConfigWoodboard2x4 = 3,
    OBJConfigWoodboard4x8 = 4,
    OBJConfigMetalSheet   = 5,
    OBJConfigBrick        = 6,
    OBJConfigCount
};

typedef struct _obj_table {
    ADMTable *adm_table;
    RCSTable *rcs_table;
} OBJTable;

typedef void *OBJHandle;
OBJHandle OBJ_init(void);
/* Tables of a type come from files when the ADM / RCS folders are found from
 * path as adm.c / rcs.c search them, else from the synthetic generators. */
OBJHandle OBJ_init_with_path(const char *path);
void OBJ_free(OBJHandle);
OBJTable *OBJ_get_table(const OBJHandle, OBJConfig type);
/* The t-th table handed out by OBJ_get_table (NULL beyond the count). */
OBJTable *OBJ_table_at(const OBJHandle, int t);

/* ---- scan patterns (pos.h) ---- */
#define POS_MAX_PATTERN_COUNT 10000
#define POS_MAX_SWEEP_COUNT   50
typedef void *POSHandle;

typedef struct pos_pos {
    float    az;
    float    el;
    uint32_t index;
    uint32_t count;
} POSPosition;

typedef struct pos_sweep {
    float azStart, azEnd, azDelta, elStart, elEnd, elDelta;
} POSSweep;

typedef struct pos_pattern {
    char        mode;
    char        reserved1, reserved2, reserved3;
    uint32_t    index;
    uint32_t    count;
    uint32_t    sweepIndex;
    uint32_t    sweepCount;
    POSPosition positions[POS_MAX_PATTERN_COUNT];
    POSSweep    sweeps[POS_MAX_SWEEP_COUNT];
    float       az;
    float       el;
    uint32_t    tic;
} POSPattern;

POSPattern *POS_init(void);
POSPattern *POS_init_with_string(const char *);
void POS_free(POSPattern *);
int POS_get_next_angles(POSPattern *scan);
int POS_parse_from_string(POSPattern *scan, const char *string);
bool POS_is_ppi(const POSPattern *scan);
bool POS_is_rhi(const POSPattern *scan);
bool POS_is_dbs(const POSPattern *scan);
bool POS_is_empty(const POSPattern *scan);

/* ---- seeded generator shared by the synthetic providers ---- */
uint64_t rs_splitmix64(uint64_t *state);
double rs_uniform(uint64_t *state);
double rs_gauss(uint64_t *state);

#ifdef __cplusplus
}
#endif

#endif
