/*
 * rs_types.h -- value types and constants of the RS_* API.
 *
 * Same names, sizes and meanings as the reference's rs_types.h / rs_const.h /
 * rs_enum.h, because callers compile against them (rs.h:11-23).  The reference
 * pulls the cl_* value types from <CL/cl.h> (rs.h:20-25); this build has no
 * OpenCL runtime, so the value types are defined here with the same 16-byte
 * alignment and member names (.s[], .x .y .z .w, .s0 ... .sF, .lo / .hi).
 */
#ifndef RS_TYPES_H
#define RS_TYPES_H

#include <math.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <sys/types.h>

#if !defined(__OPENCL_CL_H) && !defined(RS_HAVE_CL_TYPES)
#define RS_HAVE_CL_TYPES 1
typedef float cl_float;
typedef double cl_double;
typedef uint32_t cl_uint;
typedef int32_t cl_int;
typedef uint64_t cl_ulong;
typedef intptr_t cl_context_properties;

typedef union {
    cl_float __attribute__((aligned(8))) s[2];
    __extension__ struct { cl_float x, y; };
    __extension__ struct { cl_float s0, s1; };
} cl_float2;

typedef union {
    cl_float __attribute__((aligned(16))) s[4];
    __extension__ struct { cl_float x, y, z, w; };
    __extension__ struct { cl_float s0, s1, s2, s3; };
    __extension__ struct { cl_float2 lo, hi; };
} cl_float4;

typedef union {
    cl_float __attribute__((aligned(64))) s[16];
    __extension__ struct { cl_float x, y, z, w, __spacer4, __spacer5, __spacer6, __spacer7,
                                 __spacer8, __spacer9, sa, sb, sc, sd, se, sf; };
    __extension__ struct { cl_float s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, sA, sB, sC, sD, sE, sF; };
} cl_float16;

typedef union {
    cl_uint __attribute__((aligned(16))) s[4];
    __extension__ struct { cl_uint x, y, z, w; };
    __extension__ struct { cl_uint s0, s1, s2, s3; };
} cl_uint4;

typedef union {
    cl_double __attribute__((aligned(32))) s[4];
    __extension__ struct { cl_double x, y, z, w; };
    __extension__ struct { cl_double s0, s1, s2, s3; };
} cl_double4;
#endif

#define RSfloat float

#if !defined(TRUE) && !defined(FALSE)
#define TRUE ((bool)1)
#define FALSE ((bool)0)
#endif
#ifndef MAX
#define MAX(X, Y) ((X) > (Y) ? (X) : (Y))
#endif
#ifndef MIN
#define MIN(X, Y) ((X) > (Y) ? (Y) : (X))
#endif
#define DTIME(T_begin, T_end) ((double)(T_end.tv_sec - T_begin.tv_sec) + 1.0e-6 * (double)(T_end.tv_usec - T_begin.tv_usec))

/* rs_const.h */
#define RS_VERSION_STRING       "2.0.0-mi355x"
#define RS_DOMAIN_PAD           2.0
#define RS_MAX_STR              4096
#define RS_MAX_GPU_PLATFORM     40
#define RS_MAX_GPU_DEVICE       8
#define RS_ALIGN_SIZE           128
#define RS_MAX_GATES            512
#define RS_CL_GROUP_ITEMS       64
#define RS_MAX_DEBRIS_TYPES     8
#define RS_MAX_ADM_TABLES       RS_MAX_DEBRIS_TYPES
#define RS_MAX_RCS_TABLES       RS_MAX_DEBRIS_TYPES
#define RS_MAX_NUM_SCATS        120000000
#define RS_BODY_PER_CELL        100.0f
#define RS_PARAMS_LAMBDA        0.1f
#define RS_PARAMS_PRT           1.0e-3f
#define RS_PARAMS_TAU           0.2e-6f
#define RS_PARAMS_PULSEWIDTH    RS_PARAMS_TAU
#define RS_PARAMS_BEAMWIDTH     1.0f
#define RS_PARAMS_GATEWIDTH     30.0f

enum RSStatus {
    RSStatusNull                       = 0,
    RSStatusPopulationDefined          = 1 << 1,
    RSStatusOriginsOffsetsDefined      = 1 << 2,
    RSStatusWorkersAllocated           = 1 << 3,
    RSStatusDomainPopulated            = 1 << 4,
    RSStatusScattererSignalNeedsUpdate = 1 << 5,
    RSStatusDebrisRCSNeedsUpdate       = 1 << 6
};

enum RS_CL_PASS_2 {
    RS_CL_PASS_2_UNIVERSAL,
    RS_CL_PASS_2_IN_RANGE,
    RS_CL_PASS_2_IN_LOCAL
};

typedef char RSMethod;
enum RSMethod {
    RS_METHOD_CPU,
    RS_METHOD_GPU
};

typedef struct _rs_point { RSfloat x, y, z; } RSPoint;
typedef struct _rs_volume { RSPoint origin; RSPoint size; } RSVolume;

typedef struct _rs_params {
    RSfloat  c;
    RSfloat  prt;
    RSfloat  loss;
    RSfloat  lambda;
    RSfloat  tx_power_watt;
    RSfloat  antenna_gain_dbi;
    RSfloat  antenna_bw_deg;
    RSfloat  tau;
    RSfloat  range_start;
    RSfloat  range_end;
    RSfloat  range_delta;
    RSfloat  azimuth_start_deg;
    RSfloat  azimuth_end_deg;
    RSfloat  azimuth_delta_deg;
    RSfloat  elevation_start_deg;
    RSfloat  elevation_end_deg;
    RSfloat  elevation_delta_deg;
    RSfloat  domain_pad_factor;
    RSfloat  body_per_cell;
    RSfloat  prf;
    RSfloat  va;
    RSfloat  fn;
    RSfloat  antenna_bw_rad;
    RSfloat  dr;
    uint32_t range_count;
} RSParams;

#endif
