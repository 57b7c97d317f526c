/*
 * rs_host.c -- the RS_* host runtime (plain C) over the gfx950 engine.
 *
 * Same call semantics as the reference's rs.c: setters fill RSParams and the
 * float16 descriptors, geometry and population are computed on the host with
 * the same binary32 / binary64 mix, and the per-pulse calls hand one step of
 * work to the device layer (include/rsd.h).  Where the reference exits the
 * process on a bad state, this runtime prints the same ERROR line and returns
 * (the caller sees H->last_error).  Citations name the reference function each
 * block follows.
 */
#include "rs.h"

#include <unistd.h>

#define RS_SQRT1_2 0.70710678118654752440

static float bits_to_float(uint32_t u) { float f; memcpy(&f, &u, sizeof(f)); return f; }

static int engine_ok(RSHandle *H, int rc, const char *what) {
    if (rc == RSD_OK) return 1;
    H->last_error = rc;
    rsprint("ERROR: %s failed (%d): %s", what, rc, rsd_last_error(H->workers[0].engine));
    return 0;
}

static int has_engine(const RSHandle *H) { return H->workers[0].engine != NULL; }

/* rs.c:946-952 */
static void RS_update_computed_properties(RSHandle *H) {
    H->params.prf = 1.0f / H->params.prt;
    H->params.va = 0.25f * H->params.lambda * H->params.prf;
    H->params.fn = 0.5 / H->params.prf;
    H->params.antenna_bw_rad = H->params.antenna_bw_deg / 180.0f * M_PI;
    H->params.dr = 0.5f * H->params.c * H->params.tau;
}

unsigned int RS_gpu_count(void) { return (unsigned int)rsd_device_count(); }

char *RS_version_string(void) {
    static char s[32];
    snprintf(s, sizeof(s), "%s", RS_VERSION_STRING);
    return s;
}

int RS_indent_copy(char *dst, char *src, const int width) {
    int k = 0;
    char *e, *s = src;
    if (width == 0) return sprintf(dst, "%s", src);
    char indent[width + 1];
    memset(indent, ' ', (size_t)width);
    indent[width] = '\0';
    do {
        e = strchr(s, '\n');
        if (e) {
            *e = '\0';
            k += sprintf(dst + k, "%s%s\n", indent, s);
            s = e + 1;
        }
    } while (e != NULL);
    k += sprintf(dst + k, "%s%s", indent, s);
    return k;
}

/* ---------------------------------------------------------------- life cycle */

static RSHandle *alloc_handle(RSMethod method) {
    RSHandle *H = NULL;
    if (posix_memalign((void **)&H, RS_ALIGN_SIZE, sizeof(RSHandle))) {
        rsprint("ERROR: Unable to initialize RS Framework.");
        return NULL;
    }
    memset(H, 0, sizeof(RSHandle));
    /* rs.c:1013-1027 */
    H->sim_tic = 0.0f;
    H->status = RSStatusNull;
    H->params.c = 3.0e8f;
    H->params.tau = 0.2e-6f;
    H->params.body_per_cell = RS_BODY_PER_CELL;
    H->params.domain_pad_factor = RS_DOMAIN_PAD;
    H->num_workers = 1;
    H->num_types = 1;
    H->method = method;
    H->random_seed = 19760520;
    H->phase_mode = RSD_PHASE_F64;
    H->range_cap = 10.0e3f;
    for (int i = 0; i < RS_MAX_GPU_DEVICE; i++) H->workers[i].name = (char)i;
    return H;
}

static void set_defaults(RSHandle *H, char verb) {
    /* rs.c:1141-1156 */
    H->verb = verb > 1 ? verb : 0;
    RS_set_prt(H, RS_PARAMS_PRT);
    RS_set_lambda(H, RS_PARAMS_LAMBDA);
    RS_set_antenna_params(H, RS_PARAMS_BEAMWIDTH, 50.0f);
    RS_set_tx_params(H, RS_PARAMS_TAU, 50.0e3f);
    RS_set_beam_pos(H, 5.0f, 1.0f);
    RS_set_sampling_spacing(H, RS_PARAMS_GATEWIDTH, RS_PARAMS_BEAMWIDTH, RS_PARAMS_BEAMWIDTH);
    H->verb = verb;
}

RSHandle *RS_init_with_path(const char *bundle_path, RSMethod method, const uint8_t gpu_mask,
                            cl_context_properties sharegroup, const char verb) {
    (void)bundle_path;
    (void)sharegroup;
    if (method == RS_METHOD_CPU) return RS_init_host_only(verb);
    RSHandle *H = alloc_handle(method);
    if (!H) return NULL;
    int ndev = rsd_device_count();
    int dev = -1;
    for (int k = 0; k < ndev && k < 8; k++) {
        if (gpu_mask & (1u << k)) {
            dev = k;
            break;
        }
    }
    if (ndev <= 0 || dev < 0) {
        rsprint("ERROR: No GPU devices found.");
        free(H);
        return NULL;
    }
    H->num_devs = (cl_uint)ndev;
    /* This runtime drives one GPU per handle; the reference's multi-worker split
     * (rs.c:2259-2358) is a host sum of per-device pulses. */
    H->num_workers = 1;
    int cus = 0, max_wg = 0;
    rsd_device_props(dev, &cus, &max_wg, NULL);
    H->num_cus[0] = (cl_uint)cus;
    H->vendors[0] = RS_GPU_VENDOR_AMD;
    H->preferred_multiple = H->num_cus[0] * 16; /* rs.c:1048-1052 (AMD) */
    H->workers[0].num_cus = (cl_uint)cus;
    if (rsd_open(dev, &H->workers[0].engine) != RSD_OK) {
        rsprint("ERROR: Unable to open GPU %d.", dev);
        free(H);
        return NULL;
    }
    if (verb) rsprint("GPU %d: %d compute units, max work-group %d", dev, cus, max_wg);
    set_defaults(H, verb);
    return H;
}

RSHandle *RS_init_host_only(const char verb) {
    RSHandle *H = alloc_handle(RS_METHOD_CPU);
    if (!H) return NULL;
    H->num_devs = 0;
    H->num_workers = 1;
    H->num_cus[0] = 256; /* MI355X */
    H->workers[0].num_cus = 256;
    H->vendors[0] = RS_GPU_VENDOR_AMD;
    H->preferred_multiple = H->num_cus[0] * 16;
    set_defaults(H, verb);
    return H;
}

RSHandle *RS_init_for_selected_gpu(const uint8_t gpu_mask, const char verb) {
    return RS_init_with_path(".", RS_METHOD_GPU, gpu_mask, 0, verb);
}
RSHandle *RS_init_for_cpu_verbose(const char verb) { return RS_init_with_path(".", RS_METHOD_CPU, 0, 0, verb); }
RSHandle *RS_init_verbose(const char verb) { return RS_init_with_path(".", RS_METHOD_GPU, 0xFF, 0, verb); }
RSHandle *RS_init(void) { return RS_init_with_path(".", RS_METHOD_GPU, 0xFF, 0, 0); }

static void free_scat_memory(RSHandle *H) {
    free(H->scat_uid); free(H->scat_pos); free(H->scat_vel); free(H->scat_ori); free(H->scat_tum);
    free(H->scat_aux); free(H->scat_rcs); free(H->scat_sig); free(H->scat_rnd); free(H->pulse);
    free(H->pulse_acc);
    H->scat_uid = NULL; H->scat_pos = H->scat_vel = H->scat_ori = H->scat_tum = NULL;
    H->scat_aux = H->scat_rcs = H->scat_sig = H->pulse = NULL;
    H->scat_rnd = NULL;
    H->pulse_acc = NULL;
    for (int i = 0; i < RS_MAX_GPU_DEVICE; i++) {
        free(H->pulse_tmp[i]);
        H->pulse_tmp[i] = NULL;
    }
}

void RS_free(RSHandle *H) {
    if (!H) return;
    char v = H->verb;
    LES_free(H->L);
    if (H->O) OBJ_free(H->O);
    if (H->P && H->owns_pattern) POS_free((POSPattern *)H->P);
    for (int i = 0; i < RS_MAX_GPU_DEVICE; i++) {
        if (H->workers[i].engine) rsd_close(H->workers[i].engine);
        free(H->workers[i].dff_icdf[0]);
        free(H->workers[i].dff_icdf[1]);
    }
    free_scat_memory(H);
    free(H->anchor_pos);
    free(H->anchor_lines);
    free(H->dsd_r); free(H->dsd_pdf); free(H->dsd_cdf); free(H->dsd_pop);
    free(H);
    if (v) rsprint("Resources released.");
}

/* rs.c:1284-1385 -- launch geometry of the reference's two-pass reduction.  The
 * gfx950 engine does not use it; it is computed for callers that read it. */
RSMakePulseParams RS_make_pulse_params(const cl_uint count, const cl_uint group_size_multiple, const cl_uint user_max_groups,
                                       const cl_uint max_local_mem_size, const float range_start, const float range_delta,
                                       const unsigned int range_count) {
    RSMakePulseParams param;
    memset(&param, 0, sizeof(param));
    param.num_scats = count;
    param.user_max_groups = user_max_groups;
    param.user_max_work_items = group_size_multiple;
    param.range_start = range_start;
    param.range_delta = range_delta;
    param.range_count = MAX(1, range_count);
    if (user_max_groups > 1024) {
        rsprint("WARNING: user_max_groups > 1024 is capped at 1024.");
        param.user_max_groups = 1024;
    }
    unsigned int work_items = count > param.user_max_work_items * 2 ? param.user_max_work_items : count / 2;
    unsigned int group_count = count <= work_items * 2 ? 1 : (count + work_items * 2 - 1) / (work_items * 2);
    if (group_count > param.user_max_groups) group_count = param.user_max_groups;
    param.entry_counts[0] = count;
    param.group_counts[0] = group_count;
    param.global[0] = (size_t)group_count * work_items;
    param.local[0] = work_items;
    param.local_mem_size[0] = (size_t)range_count * work_items * sizeof(cl_float4);
    while (param.local_mem_size[0] > max_local_mem_size) {
        if (range_count % 2 == 0) {
            work_items /= 2;
            group_count *= 2;
            if (group_count > param.user_max_groups) group_count = param.user_max_groups;
            param.group_counts[0] = group_count;
            param.global[0] = (size_t)group_count * work_items;
            param.local[0] = work_items;
            param.local_mem_size[0] = (size_t)range_count * work_items * sizeof(cl_float4);
        } else {
            rsprint("ERROR: Could not resolve local memory size limits.");
            param.cl_pass_2_method = 0xFFFFFFFFu;
            return param;
        }
    }
    unsigned int work_count = group_count * param.range_count;
    param.entry_counts[1] = work_count;
    work_items = work_count / (param.range_count * 2);
    if (work_items < 1) work_items = 1;
    if (param.local[0] % param.range_count == 0 && group_size_multiple >= work_items) {
        param.cl_pass_2_method = RS_CL_PASS_2_UNIVERSAL;
        param.group_counts[1] = 1;
        param.global[1] = work_items;
        param.local[1] = work_items;
        param.local_mem_size[1] = work_items * sizeof(cl_float4);
    } else if (group_count >= 2 * param.range_count && group_size_multiple >= work_items) {
        param.cl_pass_2_method = RS_CL_PASS_2_IN_LOCAL;
        param.group_counts[1] = 1;
        param.global[1] = work_items;
        param.local[1] = work_items;
        param.local_mem_size[1] = work_items * sizeof(cl_float4);
    } else {
        param.cl_pass_2_method = RS_CL_PASS_2_IN_RANGE;
        param.group_counts[1] = 1;
        param.global[1] = param.range_count;
        param.local[1] = 1;
        param.local_mem_size[1] = sizeof(cl_float4);
    }
    return param;
}

/* ----------------------------------------------------------------- setters */

void RS_set_concept(RSHandle *H, RSSimulationConcept c) { H->sim_concept = c; }

char *RS_simulation_concept_string(RSHandle *H) {
    static char s[64];
    snprintf(s, sizeof(s), "Concepts used: %s%s%s%s%s%s%s",
             H->sim_concept & RSSimulationConceptBoundedParticleVelocity ? "B" : "",
             H->sim_concept & RSSimulationConceptDebrisFluxFromVelocity ? "C" : "",
             H->sim_concept & RSSimulationConceptDraggedBackground ? "D" : "",
             H->sim_concept & RSSimulationConceptFixedScattererPosition ? "F" : "",
             H->sim_concept & RSSimulationConceptTransparentBackground ? "T" : "",
             H->sim_concept & RSSimulationConceptUniformDSDScaledRCS ? "U" : "",
             H->sim_concept & RSSimulationConceptVerticallyPointingRadar ? "V" : "");
    return s;
}

char *RS_simulation_concept_bulleted_string(RSHandle *H) {
    static char s[1024];
    static const struct { uint32_t bit; const char *text; } items[] = {
        {RSSimulationConceptBoundedParticleVelocity, "o B - Bounded Particle Velocity\n"},
        {RSSimulationConceptDebrisFluxFromVelocity, "o C - Debris Concentration from Velocity\n"},
        {RSSimulationConceptDraggedBackground, "o D - Dragged Meteorological Scatterers\n"},
        {RSSimulationConceptFixedScattererPosition, "o F - Fixed Scatterer Positions\n"},
        {RSSimulationConceptTransparentBackground, "o T - Transparent Meteorological Scatterers\n"},
        {RSSimulationConceptUniformDSDScaledRCS, "o U - Uniform DSD with Scaled RCS\n"},
        {RSSimulationConceptVerticallyPointingRadar, "o V - Vertically Pointing Radar\n"},
    };
    snprintf(s, sizeof(s), "Concepts used:\n");
    for (size_t i = 0; i < sizeof(items) / sizeof(items[0]); i++) {
        if (H->sim_concept & items[i].bit) {
            strncat(s, RS_INDENT, sizeof(s) - strlen(s) - 1);
            strncat(s, items[i].text, sizeof(s) - strlen(s) - 1);
        }
    }
    return s;
}

void RS_set_prt(RSHandle *H, const RSfloat prt) {
    H->params.prt = prt;
    H->sim_desc.s[RSSimulationDescriptionPRT] = H->params.prt;
    RS_update_computed_properties(H);
}

void RS_set_lambda(RSHandle *H, const RSfloat lambda) {
    H->params.lambda = lambda;
    H->sim_desc.s[RSSimulationDescriptionWaveNumber] = 4.0f * M_PI / H->params.lambda;
    RS_update_computed_properties(H);
}

void RS_set_density(RSHandle *H, const RSfloat density) {
    if (H->status & RSStatusDomainPopulated) {
        rsprint("Simulation domain has been populated. Density cannot be changed.");
        return;
    }
    H->params.body_per_cell = density;
    RS_update_computed_properties(H);
}

void RS_set_antenna_params(RSHandle *H, RSfloat beamwidth_deg, RSfloat gain_dbi) {
    if (H->status & RSStatusDomainPopulated) {
        rsprint("Simulation domain has been populated. Radar antenna parameters cannot be changed.");
        return;
    }
    H->params.antenna_bw_deg = beamwidth_deg;
    H->params.antenna_gain_dbi = gain_dbi;
    RS_update_computed_properties(H);
    RS_set_angular_weight_to_standard(H, H->params.antenna_bw_rad);
}

void RS_set_tx_params(RSHandle *H, RSfloat pulsewidth, RSfloat tx_power_watt) {
    if (H->status & RSStatusDomainPopulated) {
        rsprint("Simulation domain has been populated. Radar parameters cannot be changed.");
        return;
    }
    H->params.tau = pulsewidth;
    H->params.tx_power_watt = tx_power_watt;
    RS_update_computed_properties(H);
    RS_set_range_weight_to_triangle(H, H->params.dr);
}

void RS_set_sampling_spacing(RSHandle *H, const RSfloat range, const RSfloat azimuth, const RSfloat elevation) {
    H->params.range_delta = range;
    H->params.azimuth_delta_deg = azimuth;
    H->params.elevation_delta_deg = elevation;
}

void RS_set_scan_box(RSHandle *H, RSBox box) {
    if (H->params.range_delta == 0.0f) H->params.range_delta = RS_PARAMS_GATEWIDTH;
    if (H->params.azimuth_delta_deg == 0.0f) H->params.azimuth_delta_deg = RS_PARAMS_BEAMWIDTH;
    if (H->params.elevation_delta_deg == 0.0f) H->params.elevation_delta_deg = RS_PARAMS_BEAMWIDTH;
    RS_set_scan_extent(H,
                       box.origin.r, box.origin.r + box.size.r, H->params.range_delta,
                       box.origin.a, box.origin.a + box.size.a, H->params.azimuth_delta_deg,
                       box.origin.e, box.origin.e + box.size.e, H->params.elevation_delta_deg);
}

static void anchor_add(RSHandle *H, int ii, RSfloat r, RSfloat el, RSfloat az, RSfloat *xmin, RSfloat *xmax,
                       RSfloat *ymin, RSfloat *ymax, RSfloat *zmin, RSfloat *zmax) {
    cl_float4 *a = &H->anchor_pos[ii];
    a->x = r * cos(el) * sin(az);
    a->y = r * cos(el) * cos(az);
    a->z = r * sin(el);
    a->w = 1.0f;
    *xmin = a->x < *xmin ? a->x : *xmin;
    *xmax = a->x > *xmax ? a->x : *xmax;
    *ymin = a->y < *ymin ? a->y : *ymin;
    *ymax = a->y > *ymax ? a->y : *ymax;
    *zmin = a->z < *zmin ? a->z : *zmin;
    *zmax = a->z > *zmax ? a->z : *zmax;
}

/* rs.c:1535-2073 -- work domain from the user scan extent: padded range /
 * azimuth / elevation, anchor points on the near and far range shells, and the
 * bounding box that becomes sim_desc slots 8-10 / 12-14. */
void RS_set_scan_extent(RSHandle *H,
                        RSfloat range_start, RSfloat range_end, RSfloat range_delta,
                        RSfloat azimuth_start, RSfloat azimuth_end, RSfloat azimuth_delta,
                        RSfloat elevation_start, RSfloat elevation_end, RSfloat elevation_delta) {
    if (H->status & RSStatusDomainPopulated) {
        rsprint("Simulation domain has been populated. Scan box cannot be changed.");
        return;
    }
    H->params.range_start = range_start;
    H->params.range_end = range_end;
    H->params.range_delta = range_delta;
    H->params.azimuth_start_deg = azimuth_start;
    H->params.azimuth_end_deg = azimuth_end;
    H->params.azimuth_delta_deg = azimuth_delta;
    H->params.elevation_start_deg = elevation_start;
    H->params.elevation_end_deg = elevation_end;
    H->params.elevation_delta_deg = elevation_delta;

    const bool is_full_sweep = fabs(azimuth_start - 0.0f) < 0.01f && fabs(azimuth_end - 360.0f) < 0.01f;
    const RSParams *P = &H->params;
    const RSfloat r_lo = floor(MAX(P->range_delta, P->range_start - P->domain_pad_factor * P->dr) / P->range_delta) * P->range_delta;
    const RSfloat r_hi = ceil(MIN(H->range_cap, P->range_end + P->domain_pad_factor * P->dr) / P->range_delta) * P->range_delta;
    const RSfloat az_lo = is_full_sweep ? 0.0f : floor(P->azimuth_start_deg - P->domain_pad_factor * P->antenna_bw_deg) / P->antenna_bw_deg * P->antenna_bw_deg;
    const RSfloat az_hi = is_full_sweep ? 360.0f : ceil(P->azimuth_end_deg + P->domain_pad_factor * P->antenna_bw_deg) / P->antenna_bw_deg * P->antenna_bw_deg;
    const RSfloat el_lo = floor(MAX(0.0f, P->elevation_start_deg - P->domain_pad_factor * P->antenna_bw_deg) / P->antenna_bw_deg) * P->antenna_bw_deg;
    const RSfloat el_hi = ceil(MIN(90.0f, P->elevation_end_deg + P->domain_pad_factor * P->antenna_bw_deg) / P->antenna_bw_deg) * P->antenna_bw_deg;
    const RSfloat tiny = 1.0e-5f;
    int nr = 0, naz = 0, nel = 0;
    RSfloat xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY, zmin = INFINITY, zmax = -INFINITY;
    RSfloat r, az, el;
    int ii = 0;

    r = floor(P->range_start / P->range_delta) * P->range_delta;
    while (r <= ceil(P->range_end / P->range_delta) * P->range_delta) {
        r += P->range_delta;
        nr++;
    }
    H->params.range_count = MIN(RS_MAX_GATES, nr);

    if (H->sim_concept & RSSimulationConceptVerticallyPointingRadar) {
        /* rs.c:1602-1690: a profiler's raster of zenith offsets (xx, yy) in
         * degrees, spacing min(elevation, azimuth delta), out to 90 - el_lo;
         * each point gives az = atan2(yy, xx) (radians), el = 90 - |(xx, yy)| */
        const RSfloat delta = MIN(elevation_delta, azimuth_delta);
        const RSfloat edge = el_lo;
        RSfloat xx, yy = edge - 90.0f;
        while (yy < (90.0f - edge) + tiny) {
            xx = el_lo - 90.0f;
            while (xx < (90.0f - edge) + tiny) {
                xx += delta;
                ii++;
            }
            yy += delta;
        }
        H->num_anchors = 2 * ii + 1;
        free(H->anchor_pos);
        H->anchor_pos = (cl_float4 *)malloc((size_t)H->num_anchors * sizeof(cl_float4));
        if (!H->anchor_pos) {
            rsprint("ERROR: Unable to allocate memory for anchors.");
            return;
        }
        ii = 0;
        yy = edge - 90.0f;
        while (yy < (90.0f - edge) + tiny) {
            xx = edge - 90.0f;
            while (xx < (90.0f - edge) + tiny) {
                az = atan2f(yy, xx);
                el = 90.0f - sqrtf(yy * yy + xx * xx);
                el = el / 180.0f * M_PI;
                anchor_add(H, ii++, r_lo, el, az, &xmin, &xmax, &ymin, &ymax, &zmin, &zmax);
                anchor_add(H, ii++, r_hi, el, az, &xmin, &xmax, &ymin, &ymax, &zmin, &zmax);
                xx += delta;
            }
            yy += delta;
        }
        naz = 360;
        nel = H->P ? (int)((POSPattern *)H->P)->count : 1;
    } else {
    az = az_lo;
    while (az <= az_hi + tiny) {
        az += P->azimuth_delta_deg;
        naz++;
    }
    el = el_lo;
    while (el <= el_hi) {
        el += P->elevation_delta_deg;
        nel++;
    }
    if (naz == 0 || nel == 0) {
        rsprint("NEL = %d and/or NAZ = %d resulted in a zero volumne.", nel, naz);
        return;
    }
    H->num_anchors = 2 * naz * nel + 1;
    free(H->anchor_pos);
    H->anchor_pos = (cl_float4 *)malloc((size_t)H->num_anchors * sizeof(cl_float4));
    if (!H->anchor_pos) {
        rsprint("ERROR: Unable to allocate memory for anchors.");
        return;
    }
    el = el_lo / 180.0f * M_PI;
    while (el <= el_hi / 180.0f * M_PI + tiny && ii < H->num_anchors - 1) {
        az = az_lo / 180.0f * M_PI;
        while ((is_full_sweep && az < (az_hi - azimuth_delta) / 180.0f * M_PI) ||
               (!is_full_sweep && az <= az_hi / 180.0f * M_PI + tiny && ii < H->num_anchors - 1)) {
            anchor_add(H, ii++, r_lo, el, az, &xmin, &xmax, &ymin, &ymax, &zmin, &zmax);
            anchor_add(H, ii++, r_hi, el, az, &xmin, &xmax, &ymin, &ymax, &zmin, &zmax);
            az += azimuth_delta / 180.0f * M_PI;
        }
        el += elevation_delta / 180.0f * M_PI;
    }
    }
    H->anchor_pos[ii].x = 0.0f;
    H->anchor_pos[ii].y = 0.0f;
    H->anchor_pos[ii].z = 0.0f;
    H->anchor_pos[ii].w = 5.0f;

    H->sim_desc.s[RSSimulationDescriptionBoundSizeX] = xmax - xmin;
    H->sim_desc.s[RSSimulationDescriptionBoundSizeY] = ymax - ymin;
    H->sim_desc.s[RSSimulationDescriptionBoundSizeZ] = zmax - zmin;
    H->sim_desc.s[RSSimulationDescriptionBoundOriginX] = xmin;
    H->sim_desc.s[RSSimulationDescriptionBoundOriginY] = ymin;
    H->sim_desc.s[RSSimulationDescriptionBoundOriginZ] = zmin;

    snprintf(H->summary, sizeof(H->summary),
             "User domain @\n  R:[  %6.2f ~  %6.2f ] km\n  E:[  %6.2f ~  %6.2f ] deg\n  A:[ %+7.2f ~ %+7.2f ] deg\n"
             "Work domain @\n  R:[  %6.2f ~  %6.2f ] km  (%d gates)\n  E:[  %6.2f ~  %6.2f ] deg  (%d rays)\n  A:[ %+7.2f ~ %+7.2f ] deg  (%d rays)\n"
             "==\n  X:[ %7.2f ~ %7.2f ]  (%.2f) m\n  Y:[ %7.2f ~ %7.2f ]  (%.2f) m\n  Z:[ %7.2f ~ %7.2f ]  (%.2f) m\n",
             1.0e-3 * P->range_start, 1.0e-3 * P->range_end, P->elevation_start_deg, P->elevation_end_deg,
             P->azimuth_start_deg, P->azimuth_end_deg,
             1.0e-3 * r_lo, 1.0e-3 * r_hi, P->range_count, el_lo, el_hi, nel, az_lo, az_hi, naz,
             xmin, xmax, xmax - xmin, ymin, ymax, ymax - ymin, zmin, zmax, zmax - zmin);
    if (H->verb) {
        rsprint("Work domain @ R:[ %5.2f ~ %5.2f ] km   E:[ %5.2f ~ %5.2f ] deg   A:[ %+7.2f ~ %+7.2f ] deg  (%d gates)",
                1.0e-3 * r_lo, 1.0e-3 * r_hi, el_lo, el_hi, az_lo, az_hi, P->range_count);
    }
    RS_revise_population(H);
}

void RS_set_beam_pos(RSHandle *H, RSfloat az_deg, RSfloat el_deg) {
    /* rs.c:2076-2084 */
    H->sim_desc.s[RSSimulationDescriptionBeamUnitX] = cosf(el_deg / 180.0f * M_PI) * sinf(az_deg / 180.0f * M_PI);
    H->sim_desc.s[RSSimulationDescriptionBeamUnitY] = cosf(el_deg / 180.0f * M_PI) * cosf(az_deg / 180.0f * M_PI);
    H->sim_desc.s[RSSimulationDescriptionBeamUnitZ] = sinf(el_deg / 180.0f * M_PI);
    H->status |= RSStatusDebrisRCSNeedsUpdate;
    H->status |= RSStatusScattererSignalNeedsUpdate;
}

void RS_set_verbosity(RSHandle *H, const char verb) { H->verb = verb; }

void RS_set_debris_count(RSHandle *H, const int debris_id, const size_t count) {
    if (debris_id <= 0 || debris_id >= RS_MAX_DEBRIS_TYPES) {
        rsprint("RS_set_debris_count() cannot have debris = %d.", debris_id);
        return;
    }
    if (H->status & RSStatusDomainPopulated) {
        rsprint("ERROR: Debris counts cannot change after RS_populate() in this engine.");
        return;
    }
    H->counts[debris_id] = count;
    H->num_types = 1;
    for (int i = 1; i < RS_MAX_DEBRIS_TYPES; i++) {
        if (H->counts[i] > 0) H->num_types++;
    }
    if (H->status & RSStatusPopulationDefined) RS_revise_population(H);
}

/* addition: the reference caps the padded scan extent at 10 km (rs.c:1561);
 * domains beyond it (a box at 10 km range) lift the cap before RS_set_scan_box. */
void RS_set_range_cap(RSHandle *H, float r_max) {
    H->range_cap = r_max > 0.0f ? r_max : 10.0e3f;
}

void RS_set_population(RSHandle *H, size_t rain_count) {
    if (H->status & RSStatusDomainPopulated) {
        rsprint("Simulation domain has been populated. Population cannot be changed.");
        return;
    }
    H->population_override = rain_count;
    if (H->status & RSStatusPopulationDefined) RS_revise_population(H);
}

void RS_set_phase_precision(RSHandle *H, uint32_t phase_mode) {
    H->phase_mode = phase_mode == RSD_PHASE_F32 ? RSD_PHASE_F32 : RSD_PHASE_F64;
}

/* rs.c:2134-2211 */
void RS_revise_population(RSHandle *H) {
    if (H->sim_concept & RSSimulationConceptVerticallyPointingRadar) {
        /* rs.c:2143-2157: one scatterer per anchor of a plane and gate */
        const size_t per_plane = H->num_anchors > 0 ? (H->num_anchors - 1) / 2 : 0;
        H->counts[0] = H->population_override ? H->population_override : H->params.range_count * per_plane;
        H->num_scats = H->counts[0];
        for (int k = 1; k < RS_MAX_DEBRIS_TYPES; k++) H->num_scats += H->counts[k];
        if (H->verb) rsprint("Anchors / Plane = %s", commaint((long long)per_plane));
        H->status |= RSStatusPopulationDefined;
        return;
    }
    const size_t max_wg = 1024;
    const size_t mul = (size_t)H->num_cus[0] * H->num_workers * max_wg * 2;
    RSfloat r = H->params.range_start;
    RSfloat svol = (H->params.antenna_bw_rad * r) * (H->params.antenna_bw_rad * r) * (H->params.c * H->params.tau * 0.5f);
    RSfloat nvol = H->sim_desc.s[RSSimulationDescriptionBoundSizeX] * H->sim_desc.s[RSSimulationDescriptionBoundSizeY] *
                   H->sim_desc.s[RSSimulationDescriptionBoundSizeZ] / svol;
    H->num_scats = (size_t)(H->params.body_per_cell * nvol);
    size_t preferred_n = H->num_scats;
    if (H->num_scats > 50000) {
        preferred_n = (H->num_scats / mul) * mul;
        if (preferred_n < H->params.body_per_cell * 9 / 10) preferred_n += mul;
    }
    H->counts[0] = H->population_override ? H->population_override : preferred_n;
    H->num_scats = H->counts[0];
    for (int k = 1; k < RS_MAX_DEBRIS_TYPES; k++) H->num_scats += H->counts[k];
    if (H->verb) {
        rsprint("nvol = %s x volumes of %s m^3", commafloat(nvol), commafloat(svol));
        rsprint("Population %s (%s background)", commaint((long long)H->num_scats), commaint((long long)H->counts[0]));
    }
    if (H->num_scats > 20000000 && !H->population_override) {
        rsprint("ERROR: Too many scatterers (%s).", commaint((long long)H->num_scats));
        return;
    }
    H->status |= RSStatusPopulationDefined;
}

void RS_revise_debris_counts_to_gpu_preference(RSHandle *H) {
    for (int i = 0; i < RS_MAX_DEBRIS_TYPES; i++) {
        if (H->counts[i]) {
            H->counts[i] = ((H->counts[i] + H->preferred_multiple - 1) / H->preferred_multiple) * H->preferred_multiple;
        }
    }
}

size_t RS_get_debris_count(RSHandle *H, const int id) { return (id >= 0 && id < RS_MAX_DEBRIS_TYPES) ? H->counts[id] : 0; }
size_t RS_get_worker_debris_count(RSHandle *H, const int id, const int w) {
    return (id >= 0 && id < RS_MAX_DEBRIS_TYPES && w >= 0 && w < (int)H->num_workers) ? H->workers[w].counts[id] : 0;
}
size_t RS_get_all_worker_debris_counts(RSHandle *H, const int id, size_t counts[]) {
    for (cl_uint i = 0; i < H->num_workers; i++) counts[i] = H->workers[i].counts[id];
    return H->counts[id];
}

RSVolume RS_get_domain(RSHandle *H) {
    RSVolume v;
    v.size.x = H->sim_desc.s[RSSimulationDescriptionBoundSizeX];
    v.size.y = H->sim_desc.s[RSSimulationDescriptionBoundSizeY];
    v.size.z = H->sim_desc.s[RSSimulationDescriptionBoundSizeZ];
    v.origin.x = H->sim_desc.s[RSSimulationDescriptionBoundOriginX];
    v.origin.y = H->sim_desc.s[RSSimulationDescriptionBoundOriginY];
    v.origin.z = H->sim_desc.s[RSSimulationDescriptionBoundOriginZ];
    return v;
}

/* rs.c:2259-2358 for one worker: type k occupies [origins[k], origins[k] + counts[k]),
 * debris types packed from the end, background first. */
void RS_update_origins_offsets(RSHandle *H) {
    size_t count = H->num_scats;
    H->offset[0] = 0;
    H->workers[0].num_scats = H->num_scats;
    for (int k = RS_MAX_DEBRIS_TYPES - 1; k > 0; k--) count -= H->counts[k];
    if (H->counts[0] != count) {
        rsprint("ERROR: Inconsistent debris counts.");
        return;
    }
    for (int k = 0; k < RS_MAX_DEBRIS_TYPES; k++) H->workers[0].counts[k] = H->counts[k];
    size_t origin = H->workers[0].num_scats;
    H->workers[0].origins[0] = 0;
    for (int k = RS_MAX_DEBRIS_TYPES - 1; k > 0; k--) {
        if (H->workers[0].counts[k] == 0) continue;
        origin -= H->workers[0].counts[k];
        H->workers[0].origins[k] = origin;
    }
    H->status |= RSStatusOriginsOffsetsDefined;
}

/* --------------------------------------------------------------------- DSD */

void RS_set_dsd(RSHandle *H, const float *nd, const float *diameters, const int count, const char name) {
    if (H->status & RSStatusDomainPopulated) {
        rsprint("Simulation domain has been populated. DSD cannot be changed.");
        return;
    }
    if (count <= 0) {
        rsprint("DSD bin count cannot be 0.");
        return;
    }
    H->dsd_name = name;
    H->dsd_count = count;
    free(H->dsd_r); free(H->dsd_pdf); free(H->dsd_cdf); free(H->dsd_pop);
    H->dsd_r = (RSfloat *)calloc((size_t)count, sizeof(RSfloat));
    H->dsd_pdf = (RSfloat *)calloc((size_t)count, sizeof(RSfloat));
    H->dsd_cdf = (RSfloat *)calloc((size_t)count, sizeof(RSfloat));
    H->dsd_pop = (size_t *)calloc((size_t)count, sizeof(size_t));
    if (!H->dsd_r || !H->dsd_pdf || !H->dsd_cdf || !H->dsd_pop) {
        rsprint("ERROR: Unable to allocate memory for DSD parameterization.");
        return;
    }
    H->dsd_nd_sum = 0.0f;
    for (int i = 0; i < count; i++) H->dsd_nd_sum += nd[i];
    RSfloat lo = 0.0f;
    for (int i = 0; i < count; i++) {
        RSfloat pdf = nd[i] / H->dsd_nd_sum;
        H->dsd_r[i] = 0.5f * diameters[i];
        H->dsd_pdf[i] = pdf;
        H->dsd_cdf[i] = lo;
        lo += pdf;
    }
    if (H->verb) rsprint("Drop concentration ~ %s drops / m^3", commaint((long long)H->dsd_nd_sum));
}

void RS_set_dsd_to_mp_with_sizes(RSHandle *H, const float *ds, const int count) {
    H->dsd_n0 = 8000.0f;
    H->dsd_lambda = 2.3f * 1000.0f;
    RSfloat *n = (RSfloat *)malloc((size_t)count * sizeof(RSfloat));
    if (!n) return;
    for (int i = 0; i < count; i++) n[i] = H->dsd_n0 * exp(-H->dsd_lambda * ds[i]);
    RS_set_dsd(H, n, ds, count, RSDropSizeDistributionMarshallPalmer);
    free(n);
}

void RS_set_dsd_to_mp(RSHandle *H) {
    const float ds[] = {0.001f, 0.002f, 0.003f, 0.004f, 0.005f};
    RS_set_dsd_to_mp_with_sizes(H, ds, 5);
}

/* ------------------------------------------------------------ weight tables */

static void table1d_desc(cl_float4 *d, float start, float delta, unsigned int n) {
    d->s[RSTable1DDescriptionScale] = 1.0f / delta;
    d->s[RSTable1DDescriptionOrigin] = -start * d->s[RSTable1DDescriptionScale];
    d->s[RSTable1DDescriptionMaximum] = (float)n - 1.0f;
}

void RS_set_rcs_ellipsoid_table(RSHandle *H, const cl_float4 *weights, const float start, const float delta, unsigned int n) {
    RSWorker *C = &H->workers[0];
    table1d_desc(&C->rcs_ellipsoid_desc, start, delta, n);
    C->rcs_ellipsoid_desc.s[RSTable1DDescriptionUserConstant] = H->sim_desc.s[RSSimulationDescriptionDropConcentrationScale];
    if (has_engine(H)) engine_ok(H, rsd_set_drop_rcs(C->engine, (const float *)weights, (int)n, C->rcs_ellipsoid_desc.s), "drop RCS table upload");
}

void RS_set_range_weight(RSHandle *H, const float *weights, const float start, const float delta, unsigned int n) {
    RSWorker *C = &H->workers[0];
    cl_float4 d = C->range_weight_desc;
    table1d_desc(&d, start, delta, n);
    C->range_weight_desc = d;
    if (has_engine(H)) engine_ok(H, rsd_set_range_weight(C->engine, weights, (int)n, d.s), "range weight upload");
}

void RS_set_range_weight_to_triangle(RSHandle *H, float pulse_width_m) {
    const float w[3] = {0.0f, 1.0f, 0.0f};
    RS_set_range_weight(H, w, -pulse_width_m, pulse_width_m, 3);
}

void RS_set_angular_weight(RSHandle *H, const float *weights, const float start, const float delta, unsigned int n) {
    RSWorker *C = &H->workers[0];
    table1d_desc(&C->angular_weight_desc, start, delta, n);
    if (has_engine(H)) engine_ok(H, rsd_set_angular_weight(C->engine, weights, (int)n, C->angular_weight_desc.s), "angular weight upload");
}

/* rs.c:2826-2851: one-way pattern 8 J2(c) / c^2, c = 1.27 pi / bw sin(theta), 0.5-degree steps */
void RS_set_angular_weight_to_standard(RSHandle *H, float beamwidth_rad) {
    const unsigned int n = 32;
    float w[32];
    float b = 1.27f * M_PI / beamwidth_rad;
    float delta = 1.0f / 360.0f * M_PI;
    for (unsigned int i = 0; i < n; i++) {
        float a = (float)i * delta;
        float c = b * sinf(a);
        if (i == 0) w[i] = 1.0f;
        else if (i == n - 1) w[i] = 0.0f;
        else w[i] = 8.0f * jn(2, c) / (c * c);
    }
    RS_set_angular_weight(H, w, 0.0f, delta, n);
}

void RS_set_angular_weight_to_double_cone(RSHandle *H, float beamwidth_rad) {
    const float w[] = {1.0f, 0.0f, 0.8f, 0.0f};
    RS_set_angular_weight(H, w, 0.0f, 2.0f * beamwidth_rad, 4);
}

/* --------------------------------------------------------------- wind data */

/* rs.c:2854-2977 -- descriptor slots 0 / 4 / 8 per axis, 3 = refresh time,
 * 7 = spacing bits; upload into the active slot. */
void RS_set_vel_data(RSHandle *H, const RSTable3D table) {
    RSWorker *C = &H->workers[0];
    cl_float16 d = C->les_desc;
    d.s[RSTable3DStaggeredDescriptionFormat] = bits_to_float(table.spacing);
    d.s[0] = table.xs; d.s[4] = table.xo; d.s[8] = table.xm;
    d.s[1] = table.ys; d.s[5] = table.yo; d.s[9] = table.ym;
    d.s[2] = table.zs; d.s[6] = table.zo; d.s[10] = table.zm;
    d.s[RSTable3DDescriptionRefreshTime] = table.tr;
    C->les_desc = d;
    if (has_engine(H) && table.uvwt) {
        engine_ok(H, rsd_set_wind(C->engine, (int)C->les_id, (const float *)table.uvwt, (int)table.x_, (int)table.y_,
                                  (int)table.z_, d.s), "wind table upload");
    }
}

/* The (cn2, p, 0, 0) volume of an LES frame into the active wind slot (the
 * second image of RS_set_vel_data, rs.c:2925); only fp_atts (concept F) reads
 * it, so it is uploaded under concept F only. */
static void upload_cpxx(RSHandle *H, const LESTable *leslie) {
    if (!has_engine(H) || !H->L || !(H->sim_concept & RSSimulationConceptFixedScattererPosition)) return;
    const LESFloat4 *c = LES_get_cpxx(H->L, (LESTable *)leslie);
    if (!c) {
        rsprint("ERROR: No Cn^2 volume for the LES frame.");
        return;
    }
    RSWorker *C = &H->workers[0];
    engine_ok(H, rsd_set_cpxx(C->engine, (int)C->les_id, (const float *)c, (int)leslie->nx, (int)leslie->ny,
                              (int)leslie->nz), "Cn2 volume upload");
}

void RS_set_vel_data_to_LES_table(RSHandle *H, const LESTable *leslie) {
    if (!leslie) {
        rsprint("ERROR: LES input data cannot be NULL.");
        return;
    }
    RSTable3D t;
    memset(&t, 0, sizeof(t));
    /* rs.c:3040-3070 */
    if (leslie->is_stretched) {
        t.spacing = RSTableSpacingStretchedX | RSTableSpacingStretchedY | RSTableSpacingStretchedZ;
        t.x_ = leslie->nx; t.xm = 0.5f * (float)(leslie->nx - 1); t.xs = 1.0f / log(leslie->rx); t.xo = (leslie->rx - 1.0f) / leslie->ax;
        t.y_ = leslie->ny; t.ym = 0.5f * (float)(leslie->ny - 1); t.ys = 1.0f / log(leslie->ry); t.yo = (leslie->ry - 1.0f) / leslie->ay;
        t.z_ = leslie->nz; t.zm = 0.0f;                            t.zs = 1.0f / log(leslie->rz); t.zo = (leslie->rz - 1.0f) / leslie->az;
    } else {
        t.x_ = leslie->nx; t.xm = (float)leslie->nx - 1.0f; t.xs = 1.0f / leslie->rx; t.xo = (float)(leslie->nx - 1) * 0.5f;
        t.y_ = leslie->ny; t.ym = (float)leslie->ny - 1.0f; t.ys = 1.0f / leslie->ry; t.yo = (float)(leslie->ny - 1) * 0.5f;
        t.z_ = leslie->nz; t.zm = (float)leslie->nz - 1.0f; t.zs = 1.0f / leslie->rz; t.zo = 0.0f;
    }
    t.tr = leslie->tr;
    t.uvwt = (cl_float4 *)leslie->uvwt;
    t.cpxx = (cl_float4 *)leslie->cpxx;
    /* rs.c:3106-3111: concept C rebuilds the debris flux field from every frame */
    if (H->sim_concept & RSSimulationConceptDebrisFluxFromVelocity) RS_set_debris_flux_field_from_LES(H, leslie);
    RS_set_vel_data(H, t);
    upload_cpxx(H, leslie);
    H->vel_desc = *leslie;
    memset(&H->vel_desc.data, 0, sizeof(LESValue));
    H->vel_desc.uvwt = NULL;
    H->vel_desc.cpxx = NULL;
    H->vel_desc.flux = NULL;
}

/* Time interpolation: LES volume vel_idx into the slot that is not active
 * (the next volume), then vel_idx moves on. */
static void les_load_next(RSHandle *H) {
    RSWorker *C = &H->workers[0];
    const uint32_t cur = C->les_id;
    C->les_id = 1 - cur;
    RS_set_vel_data_to_LES_table(H, LES_get_frame(H->L, (int)H->vel_idx));
    C->les_id = cur;
    H->vel_idx = H->vel_idx == H->vel_count - 1 ? 0 : H->vel_idx + 1;
}

void RS_set_vel_data_to_config(RSHandle *H, LESConfig c) {
    if (H->L) LES_free(H->L);
    H->L = LES_init_with_config_path(c, NULL);
    if (!H->L) return;
    H->vel_idx = 0;
    H->vel_count = (uint32_t)LES_get_table_count(H->L);
    RS_set_vel_data_to_LES_table(H, LES_get_frame(H->L, 0));
    H->vel_idx = 1;
    if (H->les_blend) les_load_next(H);
}

void RS_set_les_time_interpolation(RSHandle *H, int on) {
    if (H->status & RSStatusDomainPopulated) {
        rsprint("Simulation domain has been populated. The time interpolation cannot be changed.");
        return;
    }
    if ((on != 0) == (H->les_blend != 0)) return;
    H->les_blend = on != 0;
    if (H->les_blend && H->L) les_load_next(H);
    if (!H->les_blend && H->L) H->vel_idx = H->vel_idx == 0 ? H->vel_count - 1 : H->vel_idx - 1;
}

static void set_small_wind(RSHandle *H, RSTable3D *t, int n, const cl_float4 *uvwt) {
    t->uvwt = (cl_float4 *)uvwt;
    t->tr = 1000.0f;
    t->spacing = RSTableSpacingUniform;
    (void)n;
    RS_set_vel_data(H, *t);
}

/* rs.c:3128-3156 */
void RS_set_vel_data_to_uniform(RSHandle *H, cl_float4 velocity) {
    RSVolume domain = RS_get_domain(H);
    RSTable3D t;
    memset(&t, 0, sizeof(t));
    t.x_ = 1; t.xm = 0.0f; t.xs = 1.0f / domain.size.x; t.xo = 0.0f;
    t.y_ = 1; t.ym = 0.0f; t.ys = 1.0f / domain.size.y; t.yo = 0.0f;
    t.z_ = 1; t.zm = 0.0f; t.zs = 1.0f / domain.size.z; t.zo = 0.0f;
    cl_float4 v = velocity;
    v.w = 0.0f;
    set_small_wind(H, &t, 1, &v);
}

/* rs.c:3159-3193 */
void RS_set_vel_data_to_cube27(RSHandle *H) {
    RSVolume domain = RS_get_domain(H);
    RSTable3D t;
    memset(&t, 0, sizeof(t));
    t.x_ = 3; t.xm = 2.0f; t.xs = 3.0f / domain.size.x; t.xo = 1.0f;
    t.y_ = 3; t.ym = 2.0f; t.ys = 3.0f / domain.size.y; t.yo = 1.0f;
    t.z_ = 3; t.zm = 2.0f; t.zs = 3.0f / domain.size.z; t.zo = 0.0f;
    cl_float4 d[27];
    const float v = 10.0f;
    for (int i = 0; i < 27; i++) {
        d[i].x = (float)(i % 3) * v - v;
        d[i].y = (float)((i % 9) / 3) * v - v;
        d[i].z = (float)(i / 9) * v - v;
        d[i].w = 0.0f;
    }
    set_small_wind(H, &t, 27, d);
}

/* rs.c:3196-3234 */
void RS_set_vel_data_to_cube125(RSHandle *H) {
    RSVolume domain = RS_get_domain(H);
    RSTable3D t;
    memset(&t, 0, sizeof(t));
    t.x_ = 5; t.xm = 4.0f; t.xs = 5.0f / domain.size.x; t.xo = -domain.origin.x * t.xs;
    t.y_ = 5; t.ym = 4.0f; t.ys = 5.0f / domain.size.y; t.yo = -domain.origin.y * t.ys;
    t.z_ = 5; t.zm = 4.0f; t.zs = 5.0f / domain.size.z; t.zo = -domain.origin.z * t.zs;
    cl_float4 d[125];
    const float v = 0.5f;
    for (int i = 0; i < 125; i++) {
        d[i].x = (float)(i % 5) * v - 2.0f * v;
        d[i].y = (float)((i % 25) / 5) * v - 2.0f * v;
        d[i].z = (float)(i / 25) * v - 2.0f * v;
        d[i].w = 0.0f;
    }
    set_small_wind(H, &t, 125, d);
}

void RS_clear_vel_data(RSHandle *H) { (void)H; }

/* ------------------------------------------ Debris flux field (concept C) */

/* Inverse-CDF table of n entries from a cell PDF (rs.c:3247-3320): the CDF
 * accumulates the PDF in binary32 and is kept in binary64; entry k is the
 * fractional cell index at which the CDF reaches k / (n - 1), linear inside the
 * cell, and a run of empty cells is stepped back over to the last cell that
 * holds flux.  Returns 0, or -1 for a PDF whose sum passes 1 before its last
 * two cells or a lookup that leaves its bracket (no table is produced). */
int rs_dff_icdf_from_pdf(const float *pdf, int pdf_count, float *tab, int n) {
    if (!pdf || !tab || pdf_count < 1 || n < 2) return -1;
    const int cdf_count = pdf_count + 1;
    double *cdf = (double *)malloc((size_t)cdf_count * sizeof(double));
    if (!cdf) return -1;
    float cumsum = 0.0f;
    for (int k = 0; k < pdf_count; k++) {
        cdf[k] = cumsum;
        cumsum += pdf[k];
        if (cumsum > 1.0f) {
            if (k < pdf_count - 2) {
                rsprint("ERROR: Bad PDF, cumulative sum %.8f > 1 at cell %d of %d.", cumsum, k, pdf_count);
                free(cdf);
                return -1;
            }
            if (k < pdf_count - 1) {
                rsprint("WARNING: PDF cumulative sum %.8f at cell %d of %d clamped to 1.", cumsum, k, pdf_count);
                cumsum = 1.0f;
            }
        }
    }
    cdf[pdf_count] = 1.0;
    for (int k = 0; k < n; k++) {
        const float x = (float)k / (float)(n - 1);
        int lo = 0;
        while (lo < cdf_count && cdf[lo] <= x) lo++;
        lo = lo > 0 ? lo - 1 : 0;
        int hi = lo + 1 < pdf_count ? lo + 1 : pdf_count;
        if (cdf[lo] == cdf[hi]) {
            while (lo > 0 && cdf[lo] >= x) lo--;
            hi = lo + 1 < pdf_count ? lo + 1 : pdf_count;
        }
        const float vl = (float)lo, vh = (float)hi;
        if (lo == hi) {
            tab[k] = vl;
            continue;
        }
        if (!(cdf[lo] <= x && x <= cdf[hi])) {
            rsprint("ERROR: Inverse CDF lookup left its bracket (%d, %d) at %.4f.", lo, hi, x);
            free(cdf);
            return -1;
        }
        const float w = (float)((x - cdf[lo]) / (cdf[hi] - cdf[lo]));
        tab[k] = vl + w * (vh - vl);
    }
    free(cdf);
    return 0;
}

void RS_set_debris_flux_field_by_pdf(RSHandle *H, RSTable2D *map, const float *pdf) {
    const int n = 2048;     /* inverse-CDF entries (rs.c:3250) */
    float *tab = (float *)malloc((size_t)n * sizeof(float));
    if (!tab) return;
    if (rs_dff_icdf_from_pdf(pdf, (int)map->x_, tab, n) == 0) {
        map->x_ = (uint32_t)n;
        RS_set_debris_flux_field_by_icdf(H, map, tab);
    }
    free(tab);
}

/* rs.c:3331-3404 -- descriptor: scales 0-1, stretch ratios or origins 4-5,
 * last cell indices 8-9, row length 13, grid kind 14 (0 uniform, 1 stretched,
 * 2 box-filling test board), last table index 15. */
void RS_set_debris_flux_field_by_icdf(RSHandle *H, RSTable2D *map, const float *icdf) {
    RSWorker *C = &H->workers[0];
    cl_float16 d;
    memset(&d, 0, sizeof(d));
    d.s[0] = map->xs;
    d.s[1] = map->ys;
    d.s[2] = -1.0f;
    d.s[4] = map->xo;
    d.s[5] = map->yo;
    d.s[6] = -1.0f;
    d.s[8] = map->xm;
    d.s[9] = map->ym;
    d.s[10] = -1.0f;
    d.s[13] = map->xm + 1.0f;
    d.s[14] = (float)map->y_;
    d.s[15] = (float)map->x_ - 1.0f;
    C->dff_desc = d;
    /* the slot of the LES frame being set (dff_icdf[les_id], rs.c:3364-3366);
     * the first field, or one of another size, fills both (rs.c:3342-3348) */
    const int n = (int)map->x_;
    const int slot = (int)(C->les_id & 1u);
    const int both = !C->dff_icdf[0] || !C->dff_icdf[1] || C->dff_count[0] != n || C->dff_count[1] != n;
    for (int k = 0; k < 2; k++) {
        if (!both && k != slot) continue;
        float *h = (float *)realloc(C->dff_icdf[k], (size_t)(n > 0 ? n : 1) * sizeof(float));
        if (!h) {
            rsprint("ERROR: Unable to keep the debris flux field.");
            return;
        }
        memcpy(h, icdf, (size_t)n * sizeof(float));
        C->dff_icdf[k] = h;
        C->dff_count[k] = n;
    }
    if (has_engine(H)) engine_ok(H, rsd_set_dff_slot(C->engine, slot, icdf, n, d.s), "debris flux field upload");
}

int RS_get_debris_flux_field(RSHandle *H, int slot, float *icdf, int cap, float *desc16) {
    if (!H || slot < 0 || slot > 1) return 0;
    const RSWorker *C = &H->workers[0];
    const int n = C->dff_icdf[slot] ? C->dff_count[slot] : 0;
    if (icdf && n > 0) memcpy(icdf, C->dff_icdf[slot], (size_t)(n < cap ? n : cap) * sizeof(float));
    if (desc16) memcpy(desc16, C->dff_desc.s, 16 * sizeof(float));
    return n;
}

/* rs.c:3431-3459: all flux in the middle cell (4) of a 3 x 3 field. */
void RS_set_debris_flux_field_to_center_cell_of_3x3(RSHandle *H) {
    const float icdf[2] = {4.0f, 5.0f};
    RSTable2D map;
    memset(&map, 0, sizeof(map));
    map.x_ = 2;
    map.xs = 1.0f / 3.0f;
    map.xo = -1.0f;
    map.xm = 2.0f;
    map.ys = 1.0f / 3.0f;
    map.yo = -1.0f;
    map.ym = 2.0f;
    RS_set_debris_flux_field_by_icdf(H, &map, icdf);
}

/* rs.c:3462-3504: c x c board, flux in the even cells, stretched over the box. */
void RS_set_debris_flux_field_to_checker_board(RSHandle *H, const int c) {
    const int count = c * c;
    if (c < 1) return;
    float *pdf = (float *)calloc((size_t)count, sizeof(float));
    if (!pdf) return;
    for (int k = 0; k < count; k += 2) pdf[k] = 100.0f;
    float sum = 0.0f;
    for (int k = 0; k < count; k++) sum += pdf[k];
    for (int k = 0; k < count; k++) pdf[k] /= sum;
    RSTable2D map;
    memset(&map, 0, sizeof(map));
    map.x_ = (uint32_t)count;
    map.y_ = 2;
    map.xs = 1.0f / (float)c;
    map.ys = 1.0f / (float)c;
    map.xo = -1.0f;
    map.yo = -1.0f;
    map.xm = (float)(c - 1);
    map.ym = (float)(c - 1);
    RS_set_debris_flux_field_by_pdf(H, &map, pdf);
    free(pdf);
}

/* Grid mapping shared by the LES-derived fields (rs.c:3529-3555, 3600-3623):
 * stretched grids pass a / (1 - r) and the ratio (times q), uniform grids the
 * spacing and the half extent. */
static void dff_les_map(RSTable2D *map, const LESTable *l, float q, int cx, int cy) {
    if (l->is_stretched) {
        map->xs = l->ax / (1.0f - l->rx);
        map->ys = l->ay / (1.0f - l->ry);
        map->xo = q * l->rx;
        map->yo = q * l->ry;
    } else {
        map->xs = l->rx;
        map->ys = l->ry;
        map->xo = (float)(l->nx - 1) * 0.5f * l->rx;
        map->yo = (float)(l->ny - 1) * 0.5f * l->ry;
    }
    map->xm = (float)(cx - 1);
    map->ym = (float)(cy - 1);
}

/* rs.c:3507-3562: 9 x 9 board whose even cells carry the area of the stretched
 * LES cells at that distance from the centre (ratio scaled by q = 1.46). */
void RS_set_debris_flux_field_to_checker_board_stretched(RSHandle *H, const LESTable *l) {
    const int c = 9;
    const float q = 1.46f;
    const int count = c * c;
    const float m = 0.5f * (float)(c - 1);
    float pdf[81];
    memset(pdf, 0, sizeof(pdf));
    for (int k = 0; k < count; k += 2) {
        if (l->is_stretched) {
            const int iy = k / c, ix = k % c;
            const float dx = l->ax * powf(q * l->rx, fabsf((float)ix - m));
            const float dy = l->ay * powf(q * l->ry, fabsf((float)iy - m));
            pdf[k] = dx * dy;
        } else {
            pdf[k] = 1.0f;
        }
    }
    float sum = 0.0f;
    for (int k = 0; k < count; k++) sum += pdf[k];
    for (int k = 0; k < count; k++) pdf[k] /= sum;
    RSTable2D map;
    memset(&map, 0, sizeof(map));
    map.x_ = (uint32_t)count;
    map.y_ = 1;
    dff_les_map(&map, l, q, c, c);
    RS_set_debris_flux_field_by_pdf(H, &map, pdf);
}

/* rs.c:3565-3626: flux of the second LES level, |u|^2 per cell times the cell
 * area on a stretched grid, normalised to a PDF (also left in l->flux when the
 * table has that buffer).  The velocity comes from the separate u, v, w arrays
 * or, for frames that only hold the interleaved texels, from uvwt, which holds
 * the same values (les.c:395-398). */
void RS_set_debris_flux_field_from_LES(RSHandle *H, const LESTable *l) {
    const int count = (int)(l->nx * l->ny);
    const float mx = 0.5f * (float)(l->nx - 1), my = 0.5f * (float)(l->ny - 1);
    const int split = l->data.u && l->data.v && l->data.w;
    if (count < 1 || l->nz < 2 || (!split && !l->uvwt)) return;
    float *flux = l->flux ? l->flux : (float *)malloc((size_t)count * sizeof(float));
    if (!flux) return;
    float sum = 0.0f;
    for (int k = 0; k < count; k++) {
        const int i = count + k;
        const float u = split ? l->data.u[i] : l->uvwt[i][0];
        const float vv = split ? l->data.v[i] : l->uvwt[i][1];
        const float w = split ? l->data.w[i] : l->uvwt[i][2];
        float v = u * u + vv * vv + w * w;
        if (l->is_stretched) {
            const int iy = k / (int)l->nx, ix = k % (int)l->nx;
            const float dx = l->ax * powf(l->rx, fabsf((float)ix - mx));
            const float dy = l->ay * powf(l->ry, fabsf((float)iy - my));
            v *= dx * dy;
        }
        flux[k] = v;
        sum += v;
    }
    for (int k = 0; k < count; k++) flux[k] /= sum;
    RSTable2D map;
    memset(&map, 0, sizeof(map));
    map.x_ = (uint32_t)count;
    map.y_ = l->is_stretched ? 1u : 0u;
    dff_les_map(&map, l, 1.0f, (int)l->nx, (int)l->ny);
    RS_set_debris_flux_field_by_pdf(H, &map, flux);
    if (flux != l->flux) free(flux);
}

void RS_set_scan_pattern(RSHandle *H, const POSPattern *scan_pattern) {
    /* rs.c:3628-3633: the pattern is borrowed */
    if (H->P && H->owns_pattern) POS_free((POSPattern *)H->P);
    H->P = (POSHandle)scan_pattern;
    H->owns_pattern = 0;
}

void RS_set_scan_pattern_with_string(RSHandle *H, const char *scan_string) {
    POSPattern *p = POS_init_with_string(scan_string);
    RS_set_scan_pattern(H, p);
    H->owns_pattern = 1;
}

/* ------------------------------------------------------------ ADM and RCS */

/* rs.c:3642-3762 -- descriptor: scales 0-1, origins 4-5, maxima 8-9, inverse
 * inertia lengths 12-14 and Tachikawa number 15 of the table's body. */
void RS_set_adm_data(RSHandle *H, const RSTable2D cd, const RSTable2D cm) {
    RSWorker *C = &H->workers[0];
    const cl_uint t = C->adm_count;
    if (t >= RS_MAX_ADM_TABLES) {
        rsprint("ERROR: Too many ADM tables.");
        return;
    }
    if (cm.x_ * cm.y_ != cd.x_ * cd.y_) return;
    cl_float16 d;
    memset(&d, 0, sizeof(d));
    d.s[RSTable3DDescriptionScaleX] = cd.xs;
    d.s[RSTable3DDescriptionScaleY] = cd.ys;
    d.s[RSTable3DDescriptionOriginX] = cd.xo;
    d.s[RSTable3DDescriptionOriginY] = cd.yo;
    d.s[RSTable3DDescriptionMaximumX] = cd.xm;
    d.s[RSTable3DDescriptionMaximumY] = cd.ym;
    d.s[RSTable3DDescriptionRecipInLnX] = H->adm_desc[t].phys.inv_inln_x;
    d.s[RSTable3DDescriptionRecipInLnY] = H->adm_desc[t].phys.inv_inln_y;
    d.s[RSTable3DDescriptionRecipInLnZ] = H->adm_desc[t].phys.inv_inln_z;
    d.s[RSTable3DDescriptionTachikawa] = H->adm_desc[t].phys.Ta;
    C->adm_desc[t] = d;
    if (has_engine(H) && !engine_ok(H, rsd_set_adm(C->engine, (int)t, (const float *)cd.data, (const float *)cm.data,
                                                   (int)cd.x_, (int)cd.y_, d.s), "ADM table upload")) return;
    C->adm_count++;
}

void RS_set_adm_data_to_ADM_table(RSHandle *H, const ADMTable *adam) {
    RSTable2D cd = RS_table2d_init(adam->nn), cm = RS_table2d_init(adam->nn);
    if (!cd.data || !cm.data) {
        RS_table2d_free(cd);
        RS_table2d_free(cm);
        return;
    }
    /* rs.c:3778-3783: x = beta in [-pi, pi], y = alpha in [0, pi], offsets +0.5 */
    cd.x_ = adam->nb; cd.xm = (float)(cd.x_ - 1); cd.xs = (float)(adam->nb - 1) / (2.0f * M_PI); cd.xo = -(-M_PI) * cd.xs + 0.5f;
    cd.y_ = adam->na; cd.ym = (float)(cd.y_ - 1); cd.ys = (float)(adam->na - 1) / M_PI;           cd.yo = 0.5f;
    cm.x_ = cd.x_; cm.xm = cd.xm; cm.xs = cd.xs; cm.xo = cd.xo;
    cm.y_ = cd.y_; cm.ym = cd.ym; cm.ys = cd.ys; cm.yo = cd.yo;
    for (uint32_t i = 0; i < adam->nn; i++) {
        cd.data[i].x = adam->data.cdx[i]; cd.data[i].y = adam->data.cdy[i]; cd.data[i].z = adam->data.cdz[i]; cd.data[i].w = 0.0f;
        cm.data[i].x = adam->data.cmx[i]; cm.data[i].y = adam->data.cmy[i]; cm.data[i].z = adam->data.cmz[i]; cm.data[i].w = 0.0f;
    }
    const cl_uint t = H->workers[0].adm_count;
    if (t < RS_MAX_ADM_TABLES) {
        H->adm_desc[t] = *adam;
        memset(&H->adm_desc[t].data, 0, sizeof(ADMData));
    }
    RS_set_adm_data(H, cd, cm);
    RS_table2d_free(cd);
    RS_table2d_free(cm);
}

void RS_set_adm_data_to_unity(RSHandle *H) {
    RSTable2D t = RS_table2d_init(9);
    if (!t.data) return;
    t.x_ = 3; t.xm = 2.0f; t.xs = 3.0f / (2.0f * M_PI); t.xo = -(-M_PI) * t.xs;
    t.y_ = 3; t.ym = 2.0f; t.ys = 3.0f / M_PI;          t.yo = 0.0f;
    for (int i = 0; i < 9; i++) {
        t.data[i].x = 1.0f; t.data[i].y = 1.0f; t.data[i].z = 1.0f; t.data[i].w = 0.0f;
    }
    RS_set_adm_data(H, t, t);
    RS_table2d_free(t);
}

void RS_clear_adm_data(RSHandle *H) { H->workers[0].adm_count = 0; }

/* rs.c:3852-3972 */
void RS_set_rcs_data(RSHandle *H, const RSTable2D real, const RSTable2D imag) {
    RSWorker *C = &H->workers[0];
    const cl_uint t = C->rcs_count;
    if (t >= RS_MAX_RCS_TABLES) {
        rsprint("ERROR: Too many RCS tables.");
        return;
    }
    if (imag.x_ * imag.y_ != real.x_ * real.y_) {
        rsprint("WARNING. RS_set_rcs_data() received inconsistent real & imag dimensions");
        return;
    }
    cl_float16 d;
    memset(&d, 0, sizeof(d));
    d.s[RSTable3DDescriptionScaleX] = real.xs;
    d.s[RSTable3DDescriptionScaleY] = real.ys;
    d.s[RSTable3DDescriptionOriginX] = real.xo;
    d.s[RSTable3DDescriptionOriginY] = real.yo;
    d.s[RSTable3DDescriptionMaximumX] = real.xm;
    d.s[RSTable3DDescriptionMaximumY] = real.ym;
    C->rcs_desc[t] = d;
    if (has_engine(H) && !engine_ok(H, rsd_set_rcs(C->engine, (int)t, (const float *)real.data, (const float *)imag.data,
                                                   (int)real.x_, (int)real.y_, d.s), "RCS table upload")) return;
    C->rcs_count++;
}

void RS_set_rcs_data_to_RCS_table(RSHandle *H, const RCSTable *rosie) {
    RSTable2D re = RS_table2d_init(rosie->nn), im = RS_table2d_init(rosie->nn);
    if (!re.data || !im.data) {
        RS_table2d_free(re);
        RS_table2d_free(im);
        return;
    }
    /* rs.c:3987-3992: x = alpha in [-pi, pi], y = beta in [0, pi], no +0.5 */
    re.x_ = rosie->na; re.xm = (float)(re.x_ - 1); re.xs = (float)(rosie->na - 1) / (2.0f * M_PI); re.xo = -(-M_PI) * re.xs;
    re.y_ = rosie->nb; re.ym = (float)(re.y_ - 1); re.ys = (float)(rosie->nb - 1) / M_PI;           re.yo = 0.0f;
    im.x_ = re.x_; im.xm = re.xm; im.xs = re.xs; im.xo = re.xo;
    im.y_ = re.y_; im.ym = re.ym; im.ys = re.ys; im.yo = re.yo;
    for (uint32_t i = 0; i < rosie->nn; i++) {
        re.data[i].x = rosie->data.hh_real[i]; re.data[i].y = rosie->data.vv_real[i]; re.data[i].z = rosie->data.hv_real[i]; re.data[i].w = 0.0f;
        im.data[i].x = rosie->data.hh_imag[i]; im.data[i].y = rosie->data.vv_imag[i]; im.data[i].z = rosie->data.hv_imag[i]; im.data[i].w = 0.0f;
    }
    const cl_uint t = H->workers[0].rcs_count;
    if (t < RS_MAX_RCS_TABLES) {
        H->rcs_desc[t] = *rosie;
        memset(&H->rcs_desc[t].data, 0, sizeof(RCSData));
    }
    RS_set_rcs_data(H, re, im);
    RS_table2d_free(re);
    RS_table2d_free(im);
}

void RS_set_rcs_data_to_unity(RSHandle *H) {
    RSTable2D re = RS_table2d_init(9), im = RS_table2d_init(9);
    if (!re.data || !im.data) {
        RS_table2d_free(re);
        RS_table2d_free(im);
        return;
    }
    re.x_ = 3; re.xm = 2.0f; re.xs = 3.0f / (2.0f * M_PI); re.xo = -(-M_PI) * re.xs;
    re.y_ = 3; re.ym = 2.0f; re.ys = 3.0f / M_PI;          re.yo = 0.0f;
    im.x_ = re.x_; im.xm = re.xm; im.xs = re.xs; im.xo = re.xo;
    im.y_ = re.y_; im.ym = re.ym; im.ys = re.ys; im.yo = re.yo;
    for (int i = 0; i < 9; i++) {
        re.data[i].x = 1.0f; re.data[i].y = 1.0f; re.data[i].z = 1.0f; re.data[i].w = 0.0f;
        im.data[i].x = 0.0f; im.data[i].y = 0.0f; im.data[i].z = 0.0f; im.data[i].w = 0.0f;
    }
    RS_set_rcs_data(H, re, im);
    RS_table2d_free(re);
    RS_table2d_free(im);
}

void RS_clear_rcs_data(RSHandle *H) { H->workers[0].rcs_count = 0; }

void RS_set_random_seed(RSHandle *H, const unsigned int seed) {
    if (H->verb) rsprint("Random number generator set to use seed %s", commaint(seed));
    H->random_seed = seed;
}

/* rs.c:4069-4120 */
void RS_add_debris(RSHandle *H, OBJConfig type, const size_t count) {
    int k = 1;
    if (!H->O) {
        H->O = OBJ_init();
        if (!H->O) {
            rsprint("ERROR: OBJ_init() failed.");
            return;
        }
    }
    while (k < RS_MAX_DEBRIS_TYPES && H->counts[k] > 0) k++;
    if (k == RS_MAX_DEBRIS_TYPES || H->workers[0].adm_count == RS_MAX_ADM_TABLES || H->workers[0].rcs_count == RS_MAX_RCS_TABLES) {
        rsprint("Unable to add more debris type.");
        return;
    }
    OBJTable *obj = OBJ_get_table(H->O, type);
    if (!obj) {
        rsprint("ERROR: Unknown debris type %d.", type);
        return;
    }
    RS_set_adm_data_to_ADM_table(H, obj->adm_table);
    RS_set_rcs_data_to_RCS_table(H, obj->rcs_table);
    H->counts[k] = count;
    H->num_types++;
    if (k != (int)H->workers[0].adm_count || k != (int)H->workers[0].rcs_count) {
        rsprint("WARNING: Inconsistent k = %d vs adm_count = %d vs rcs_count = %d.", k, H->workers[0].adm_count, H->workers[0].rcs_count);
    }
    if (H->status & RSStatusPopulationDefined) RS_revise_population(H);
}

void RS_io_test(RSHandle *H) {
    if (!has_engine(H)) {
        rsprint("ERROR: RS_io_test() needs a GPU.");
        return;
    }
    size_t n = H->num_scats ? H->num_scats : ((size_t)1 << 22);
    double gbps = 0.0;
    if (engine_ok(H, rsd_io_test(H->workers[0].engine, n, 20, &gbps), "io test")) {
        rsprint("io test: %s float4 copied 20 times, %.1f GB/s (read + write)", commaint((long long)n), gbps);
    }
}

/* ------------------------------------------------------------- drop RCS */

/* rs.c:5922-6029 -- Rayleigh polarisabilities of oblate spheroids, binary64 */
void RS_compute_rcs_ellipsoids(RSHandle *H) {
    const double k_0 = H->sim_desc.s[RSSimulationDescriptionWaveNumber] * 0.5f;
    const double epsilon_0 = 8.85418782e-12f;
    const double er[4] = {78.669, 18.2257, 78.669, 18.2257};
    const double sc = k_0 * k_0 / (4.0f * M_PI * epsilon_0) * 1.0e-9f * H->sim_desc.s[RSSimulationDescriptionDropConcentrationScale];
    const int n = 96;
    cl_float4 table[96];
    memset(table, 0, sizeof(table));
    if (!(H->sim_concept & RSSimulationConceptTransparentBackground)) {
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
            /* alpha = V eps0 (eps_r - 1) / (1 + L (eps_r - 1)), complex, per axis */
            double num_re = vol * epsilon_0 * er[0], num_im = vol * epsilon_0 * er[1];
            double dx_re = lx * er[0] + 1.0, dx_im = lx * er[1];
            double dz_re = lz * er[2] + 1.0, dz_im = lz * er[3];
            double mx = dx_re * dx_re + dx_im * dx_im, mz = dz_re * dz_re + dz_im * dz_im;
            table[i].s0 = (cl_float)(sc * ((num_re * dx_re + num_im * dx_im) / mx));
            table[i].s1 = (cl_float)(sc * ((num_im * dx_re - num_re * dx_im) / mx));
            table[i].s2 = (cl_float)(sc * ((num_re * dz_re + num_im * dz_im) / mz));
            table[i].s3 = (cl_float)(sc * ((num_im * dz_re - num_re * dz_im) / mz));
        }
        if ((H->sim_concept & RSSimulationConceptUniformDSDScaledRCS) && H->dsd_count) {
            const float p = 1.0f / (float)H->dsd_count;
            cl_float4 copy[96];
            memcpy(copy, table, sizeof(table));
            memset(table, 0, sizeof(table));
            for (int i = 0; i < H->dsd_count; i++) {
                int k = (int)(H->dsd_r[i] * 20000.0f) - 5;
                if (k < 0 || k >= n) continue;
                float s = sqrtf(H->dsd_pdf[i] / p);
                table[k].s0 = copy[k].s0 * s;
                table[k].s1 = copy[k].s1 * s;
                table[k].s2 = copy[k].s2 * s;
                table[k].s3 = copy[k].s3 * s;
            }
        }
    }
    RS_set_rcs_ellipsoid_table(H, table, 0.25e-3f, 0.05e-3f, (unsigned int)n);
}

/* ---------------------------------------------------------------- populate */

static int alloc_host_streams(RSHandle *H) {
    const size_t n = H->num_scats;
    free_scat_memory(H);
    int bad = 0;
    bad |= posix_memalign((void **)&H->scat_uid, RS_ALIGN_SIZE, n * sizeof(cl_uint4));
    bad |= posix_memalign((void **)&H->scat_pos, RS_ALIGN_SIZE, n * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->scat_vel, RS_ALIGN_SIZE, n * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->scat_ori, RS_ALIGN_SIZE, n * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->scat_tum, RS_ALIGN_SIZE, n * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->scat_aux, RS_ALIGN_SIZE, n * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->scat_rcs, RS_ALIGN_SIZE, n * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->scat_sig, RS_ALIGN_SIZE, n * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->scat_rnd, RS_ALIGN_SIZE, n * sizeof(cl_uint4));
    bad |= posix_memalign((void **)&H->pulse, RS_ALIGN_SIZE, MAX(1u, H->params.range_count) * sizeof(cl_float4));
    bad |= posix_memalign((void **)&H->pulse_tmp[0], RS_ALIGN_SIZE, MAX(1u, H->params.range_count) * sizeof(cl_float4));
    H->pulse_acc = (double *)calloc(4 * (size_t)MAX(1u, H->params.range_count), sizeof(double));
    if (bad || !H->pulse_acc) {
        rsprint("ERROR: Unable to allocate memory for scatterer streams.");
        return 0;
    }
    memset(H->scat_aux, 0, n * sizeof(cl_float4));
    memset(H->scat_sig, 0, n * sizeof(cl_float4));
    memset(H->pulse, 0, H->params.range_count * sizeof(cl_float4));
    H->mem_size = n * (8 * sizeof(cl_float4) + 2 * sizeof(cl_uint4)) + 2 * H->params.range_count * sizeof(cl_float4);
    return 1;
}

/* rs.c:4652-4799 -- glibc rand() stream: per type, per scatterer x, y, z (z within
 * 10 m of the floor), age, four seeds; then srand(2 seed) for the DSD bins. */
static void populate_streams(RSHandle *H) {
    RSVolume domain = RS_get_domain(H);
    uint32_t uid = 0;
    srand(H->random_seed);
    for (size_t k = 0; k < H->num_types; k++) {
        size_t i = H->offset[0] + H->workers[0].origins[k];
        for (size_t n = 0; n < H->workers[0].counts[k]; n++, i++) {
            H->scat_uid[i].s0 = uid++;
            H->scat_uid[i].s1 = (cl_uint)n;
            H->scat_uid[i].s2 = (cl_uint)k;
            H->scat_uid[i].s3 = 0;
            H->scat_pos[i].x = (float)rand() / RAND_MAX * domain.size.x + domain.origin.x;
            H->scat_pos[i].y = (float)rand() / RAND_MAX * domain.size.y + domain.origin.y;
            H->scat_pos[i].z = (float)rand() / RAND_MAX * 10.0f + domain.origin.z;
            H->scat_pos[i].w = 0.0f;
            H->scat_aux[i].s0 = 0.0f;
            H->scat_aux[i].s1 = (float)rand() / RAND_MAX;
            H->scat_aux[i].s2 = 0.0f;
            H->scat_aux[i].s3 = 1.0f;
            H->scat_vel[i].x = H->scat_vel[i].y = H->scat_vel[i].z = H->scat_vel[i].w = 0.0f;
            H->scat_ori[i].x = 0.5f;
            H->scat_ori[i].y = -0.5f;
            H->scat_ori[i].z = -0.5f;
            H->scat_ori[i].w = 0.5f;
            H->scat_tum[i].x = H->scat_tum[i].y = H->scat_tum[i].z = 0.0f;
            H->scat_tum[i].w = 1.0f;
            H->scat_rcs[i].s0 = 1.0f;
            H->scat_rcs[i].s1 = 0.0f;
            H->scat_rcs[i].s2 = 1.0f;
            H->scat_rcs[i].s3 = 0.0f;
            H->scat_rnd[i].s0 = (cl_uint)rand();
            H->scat_rnd[i].s1 = (cl_uint)rand();
            H->scat_rnd[i].s2 = (cl_uint)rand();
            H->scat_rnd[i].s3 = (cl_uint)rand();
        }
    }
    float vol = H->sim_desc.s[RSSimulationDescriptionBoundSizeX] * H->sim_desc.s[RSSimulationDescriptionBoundSizeY] *
                H->sim_desc.s[RSSimulationDescriptionBoundSizeZ];
    srand(H->random_seed + H->random_seed);
    if (H->dsd_name == RSDropSizeDistributionUndefined) {
        rsprint("WARNING: No drop size distribution set; background drops have zero radius.");
        return;
    }
    float drops_per_scat = (vol * H->dsd_nd_sum) / H->counts[0];
    H->sim_desc.s[RSSimulationDescriptionDropConcentrationScale] = sqrt(drops_per_scat);
    size_t i = H->offset[0] + H->workers[0].origins[0];
    for (size_t n = 0; n < H->workers[0].counts[0]; n++, i++) {
        float a = (float)rand() / RAND_MAX;
        int bin = 0;
        if (H->sim_concept & RSSimulationConceptUniformDSDScaledRCS) {
            bin = (int)(a * (float)H->dsd_count);
            /* a == 1.0 (rand() >= 2^31 - 64) would index one past the bins */
            if (bin > H->dsd_count - 1) bin = H->dsd_count - 1;
        } else {
            int k = H->dsd_count;
            while (k > 0) {
                k--;
                if (a >= H->dsd_cdf[k]) {
                    bin = k;
                    break;
                }
            }
        }
        H->dsd_pop[bin]++;
        H->scat_pos[i].w = H->dsd_r[bin];
        H->scat_aux[i].s2 = ((float)bin + 0.5f) / (float)H->dsd_count;
    }
}

/* rs.c:4574-4647: concept F.  The low anchors (even indices) give one plane of
 * directions scaled to range_start; the plane is repeated at every gate range
 * start + n delta.  Each scatterer: random age, rcs (1e-10, 0, 1e-10, phase
 * uniform in [0, 2 pi)), four random seeds.  Debris are not emulated. */
static int populate_fixed(RSHandle *H) {
    const size_t per_plane = H->num_anchors > 0 ? (H->num_anchors - 1) / 2 : 0;
    if (H->num_types > 1) {
        rsprint("ERROR: Debris are not emulated with fixed scatterer positions (concept F).");
        return 0;
    }
    if (per_plane == 0 || H->counts[0] != H->params.range_count * per_plane) {
        rsprint("ERROR: Concept F needs %u gates x %zu anchors = %zu scatterers, not %zu.", H->params.range_count,
                per_plane, (size_t)H->params.range_count * per_plane, H->counts[0]);
        return 0;
    }
    srand(H->random_seed);
    const cl_float4 *A = H->anchor_pos;
    const RSfloat r_lo = sqrtf(A[0].x * A[0].x + A[0].y * A[0].y + A[0].z * A[0].z);
    const RSfloat rs = H->params.range_start, dr = H->params.range_delta;
    size_t i = 0;
    for (size_t k = 0; k + 1 < (size_t)H->num_anchors; k += 2, i++) {
        H->scat_pos[i].x = A[k].x / r_lo * rs;
        H->scat_pos[i].y = A[k].y / r_lo * rs;
        H->scat_pos[i].z = A[k].z / r_lo * rs;
        H->scat_pos[i].w = 0.0f;
    }
    uint32_t uid = 0;
    i = 0;
    for (uint32_t n = 0; n < H->params.range_count; n++) {
        for (size_t k = 0; k < per_plane; k++, i++) {
            H->scat_uid[i].s0 = uid++;
            H->scat_uid[i].s1 = (cl_uint)H->num_scats;
            H->scat_uid[i].s2 = (cl_uint)k;
            H->scat_uid[i].s3 = 0;
            H->scat_pos[i].x = H->scat_pos[k].x / rs * (rs + (float)n * dr);
            H->scat_pos[i].y = H->scat_pos[k].y / rs * (rs + (float)n * dr);
            H->scat_pos[i].z = H->scat_pos[k].z / rs * (rs + (float)n * dr);
            H->scat_pos[i].w = 0.0f;
            H->scat_aux[i].s0 = 0.0f;
            H->scat_aux[i].s1 = (float)rand() / RAND_MAX;
            H->scat_aux[i].s2 = 0.0f;
            H->scat_aux[i].s3 = 1.0f;
            H->scat_vel[i].x = H->scat_vel[i].y = H->scat_vel[i].z = H->scat_vel[i].w = 0.0f;
            H->scat_ori[i].x = H->scat_ori[i].y = H->scat_ori[i].z = 0.0f;
            H->scat_ori[i].w = 1.0f;
            H->scat_tum[i].x = H->scat_tum[i].y = H->scat_tum[i].z = 0.0f;
            H->scat_tum[i].w = 1.0f;
            H->scat_rcs[i].s0 = 1.0e-10f;
            H->scat_rcs[i].s1 = 0.0f;
            H->scat_rcs[i].s2 = 1.0e-10f;
            H->scat_rcs[i].s3 = (float)rand() / RAND_MAX * 2.0 * M_PI;
            H->scat_rnd[i].s0 = (cl_uint)rand();
            H->scat_rnd[i].s1 = (cl_uint)rand();
            H->scat_rnd[i].s2 = (cl_uint)rand();
            H->scat_rnd[i].s3 = (cl_uint)rand();
        }
    }
    return 1;
}

void RS_populate(RSHandle *H) {
    if (H->num_scats > RS_MAX_NUM_SCATS) {
        rsprint("ERROR: Too many scatterers (%s).", commaint((long long)H->num_scats));
        return;
    }
    if (H->L == NULL) RS_set_vel_data_to_config(H, LESConfigSuctionVortices);
    if (H->P == NULL) {
        RS_set_scan_pattern(H, POS_init());
        H->owns_pattern = 1;
    }
    if (H->num_anchors == 0) {
        RSBox box = RS_suggest_scan_domain(H);
        RS_set_scan_box(H, box);
    }
    if (H->workers[0].adm_count == 0) RS_set_adm_data_to_unity(H);
    if (H->workers[0].rcs_count == 0) RS_set_rcs_data_to_unity(H);
    if (H->status & RSStatusDomainPopulated) {
        rsprint("ERROR: Simulation domain has been populated.");
        return;
    }
    if (!(H->status & RSStatusPopulationDefined)) {
        rsprint("ERROR: Population not defined.");
        return;
    }
    if (!alloc_host_streams(H)) return;
    RS_update_origins_offsets(H);
    if (H->sim_concept & RSSimulationConceptFixedScattererPosition) {
        if (!populate_fixed(H)) return;
        /* the frames in the wind slots (concept F may have been set after them) */
        if (H->L && H->vel_count) {
            const uint32_t k = H->vel_count, back = H->les_blend ? 2u : 1u;
            upload_cpxx(H, LES_get_frame(H->L, (int)((H->vel_idx + 2 * k - back) % k)));
            if (H->les_blend) {
                RSWorker *C = &H->workers[0];
                C->les_id = 1 - C->les_id;
                upload_cpxx(H, LES_get_frame(H->L, (int)((H->vel_idx + k - 1) % k)));
                C->les_id = 1 - C->les_id;
            }
        }
    } else {
        populate_streams(H);
    }
    H->sim_tic = 0.0f;
    H->sim_toc = H->vel_desc.tp;
    H->sim_desc.s[RSSimulationDescriptionBeamUnitX] = 0.0f;
    H->sim_desc.s[RSSimulationDescriptionBeamUnitY] = 1.0f;
    H->sim_desc.s[RSSimulationDescriptionBeamUnitZ] = 0.0f;
    H->sim_desc.s[RSSimulationDescriptionTotalParticles] = (float)H->num_scats;
    H->sim_desc.s[RSSimulationDescriptionConcept] = bits_to_float(H->sim_concept);
    H->workers[0].range_weight_desc.s[RSTable1DDescriptionUserConstant] = H->sim_desc.s[RSSimulationDescriptionWaveNumber];
    H->workers[0].rcs_ellipsoid_desc.s[RSTable1DDescriptionUserConstant] = H->sim_desc.s[RSSimulationDescriptionDropConcentrationScale];
    RS_compute_rcs_ellipsoids(H);
    H->workers[0].make_pulse_params = RS_make_pulse_params((cl_uint)H->num_scats, RS_CL_GROUP_ITEMS, 1024, 65536,
                                                           H->params.range_start, H->params.range_delta, H->params.range_count);
    if (has_engine(H)) {
        if (!engine_ok(H, rsd_alloc_state(H->workers[0].engine, H->num_scats, H->workers[0].origins, H->workers[0].counts),
                       "state allocation")) return;
        H->status |= RSStatusWorkersAllocated;
        RS_upload(H);
    }
    H->status |= RSStatusDomainPopulated;
    if (!has_engine(H)) return;
    /* rs.c:4941-4955: one zero-length step and a pulse */
    RS_advance_beam(H);
    H->sim_desc.s[RSSimulationDescriptionPRT] = 0.0f;
    RS_advance_time(H);
    RS_make_pulse(H);
    H->sim_desc.s[RSSimulationDescriptionPRT] = H->params.prt;
    H->sim_tic -= H->params.prt;
    H->sim_desc.s[RSSimulationDescriptionSimTic] = H->sim_tic;
}

/* ---------------------------------------------------------------- transfers */

void RS_upload(RSHandle *H) {
    if (H->num_scats == 0 || !has_engine(H)) return;
    engine_ok(H, rsd_upload_state(H->workers[0].engine, (const float *)H->scat_pos, (const float *)H->scat_vel,
                                  (const float *)H->scat_ori, (const float *)H->scat_tum, (const float *)H->scat_aux,
                                  (const float *)H->scat_rcs, (const uint32_t *)H->scat_rnd), "upload");
}

void RS_merge_pulse_tmp(RSHandle *H) {
    memcpy(H->pulse, H->pulse_tmp[0], H->params.range_count * sizeof(cl_float4));
    float g = RS_radar_gain(H);
    for (uint32_t k = 0; k < H->params.range_count; k++) {
        H->pulse[k].s0 *= g;
        H->pulse[k].s1 *= g;
        H->pulse[k].s2 *= g;
        H->pulse[k].s3 *= g;
    }
}

float RS_radar_gain(RSHandle *H) {
    return powf(10.0f, 0.1f * H->params.antenna_gain_dbi) * sqrtf(H->params.tx_power_watt) / (4.0f * M_PI) * 1.0e6f;
}

void RS_download_pulse_only(RSHandle *H) {
    if (!has_engine(H) || !H->pulse_acc) return;
    const uint32_t R = H->params.range_count;
    if (!engine_ok(H, rsd_read_pulse(H->workers[0].engine, H->pulse_acc, R), "pulse download")) return;
    for (uint32_t k = 0; k < R; k++) {
        for (int c = 0; c < 4; c++) H->pulse_tmp[0][k].s[c] = (float)H->pulse_acc[4 * k + c];
    }
    RS_merge_pulse_tmp(H);
}

void RS_download(RSHandle *H) {
    if (!has_engine(H) || !H->scat_pos) return;
    if (!engine_ok(H, rsd_download_state(H->workers[0].engine, (float *)H->scat_pos, (float *)H->scat_vel, (float *)H->scat_ori,
                                         (float *)H->scat_tum, (float *)H->scat_aux, (float *)H->scat_rcs, (float *)H->scat_sig,
                                         (uint32_t *)H->scat_rnd), "download")) return;
    RS_download_pulse_only(H);
}

void RS_download_position_only(RSHandle *H) {
    if (!has_engine(H) || !H->scat_pos) return;
    engine_ok(H, rsd_download_state(H->workers[0].engine, (float *)H->scat_pos, NULL, NULL, NULL, NULL, NULL, NULL, NULL),
              "position download");
}

void RS_download_orientation_only(RSHandle *H) {
    if (!has_engine(H) || !H->scat_ori) return;
    engine_ok(H, rsd_download_state(H->workers[0].engine, NULL, NULL, (float *)H->scat_ori, NULL, NULL, NULL, NULL, NULL),
              "orientation download");
}

/* ------------------------------------------------------------- time / pulse */

static void fill_table_indices(RSHandle *H, int32_t *adm, int32_t *rcs) {
    /* rs.c:5357-5373: type k uses table index k - 1, cycling over the tables */
    int r = 0, a = 0;
    const RSWorker *C = &H->workers[0];
    for (int k = 1; k < RS_MAX_DEBRIS_TYPES; k++) {
        adm[k] = a;
        rcs[k] = r;
        r = (C->rcs_count == 0 || r == (int)C->rcs_count - 1) ? 0 : r + 1;
        a = (C->adm_count == 0 || a == (int)C->adm_count - 1) ? 0 : a + 1;
    }
    adm[0] = rcs[0] = 0;
}

void RS_advance_time(RSHandle *H) {
    if (!(H->status & RSStatusDomainPopulated)) {
        rsprint("ERROR: Simulation domain not yet populated.");
        return;
    }
    /* rs.c:5189-5203: switch to the next LES volume every tp seconds */
    if (H->L && H->sim_tic >= H->sim_toc) {
        H->sim_toc += H->vel_desc.tp;
        H->workers[0].les_id = H->workers[0].les_id == 1 ? 0 : 1;
        if (H->les_blend) {
            /* the slot now active already holds this volume; the next one
             * goes to the other slot */
            les_load_next(H);
        } else {
            RS_set_vel_data_to_LES_table(H, LES_get_frame(H->L, (int)H->vel_idx));
            H->vel_idx = H->vel_idx == H->vel_count - 1 ? 0 : H->vel_idx + 1;
        }
    }
    H->wind_alpha = 0.0f;
    if (H->les_blend && H->L) {
        const float tp = H->vel_desc.tp;
        H->wind_alpha = fminf(fmaxf((H->sim_tic - (H->sim_toc - tp)) / tp, 0.0f), 1.0f);
    }
    if (has_engine(H)) {
        rsd_step s;
        memset(&s, 0, sizeof(s));
        memcpy(s.sim_desc, H->sim_desc.s, sizeof(s.sim_desc));
        s.wind_slot = (int32_t)H->workers[0].les_id;
        fill_table_indices(H, s.adm_index, s.rcs_index);
        /* rs.c:5337-5350: el_atts under D, else fp_atts under F, else bg_atts */
        s.background_kind = (H->sim_concept & RSSimulationConceptDraggedBackground) ? RSD_ADVANCE_DRAGGED
                            : (H->sim_concept & RSSimulationConceptFixedScattererPosition) ? RSD_ADVANCE_FIXED
                                                                                           : RSD_ADVANCE_BACKGROUND;
        s.wind_blend = H->les_blend && H->L ? 1u : 0u;
        s.wind_alpha = H->wind_alpha;
        engine_ok(H, rsd_advance(H->workers[0].engine, &s), "advance");
    }
    H->sim_tic += H->params.prt;
    H->sim_desc.s[RSSimulationDescriptionSimTic] = H->sim_tic;
    H->status |= RSStatusScattererSignalNeedsUpdate;
}

void RS_advance_beam(RSHandle *H) {
    POSPattern *scan = (POSPattern *)H->P;
    if (!scan) return;
    POS_get_next_angles(scan);
    RS_set_beam_pos(H, scan->az, scan->el);
}

void RS_make_pulse(RSHandle *H) {
    if (!(H->status & RSStatusDomainPopulated)) {
        rsprint("ERROR: Simulation domain not populated.");
        return;
    }
    if (has_engine(H)) {
        rsd_pulse p;
        memset(&p, 0, sizeof(p));
        memcpy(p.sim_desc, H->sim_desc.s, sizeof(p.sim_desc));
        p.two_over_lambda = 2.0 / (double)H->params.lambda;
        p.range_start = H->params.range_start;
        p.range_delta = H->params.range_delta;
        p.range_count = H->params.range_count;
        int32_t adm[RSD_MAX_TYPES];
        fill_table_indices(H, adm, p.rcs_index);
        p.debris_rcs_refresh = (H->status & RSStatusDebrisRCSNeedsUpdate) ? 1u : 0u;
        p.phase_mode = H->phase_mode;
        engine_ok(H, rsd_make_pulse(H->workers[0].engine, &p), "make pulse");
    }
    H->status &= ~RSStatusDebrisRCSNeedsUpdate;
    H->status &= ~RSStatusScattererSignalNeedsUpdate;
}

/* addition: one pulse round of a batch of handles whose engines form the
 * batch b (rsd_batch_create): RS_make_pulse of every handle (recorded by its
 * engine), one rsd_batch_launch, then RS_advance_time and RS_advance_beam of
 * every handle -- the per-pulse loop of simradar.c:1011-1074 over many domains
 * with one device launch. */
int RS_batch_round(RSHandle *const *handles, int count, rsd_batch *b) {
    for (int i = 0; i < count; i++) RS_make_pulse(handles[i]);
    int st = rsd_batch_launch(b);
    if (st != RSD_OK) {
        rsprint("ERROR: batch launch failed (%d): %s", st, rsd_batch_last_error(b));
        for (int i = 0; i < count; i++) handles[i]->last_error = st;
        return st;
    }
    for (int i = 0; i < count; i++) {
        RS_advance_time(handles[i]);
        RS_advance_beam(handles[i]);
    }
    return RSD_OK;
}

int RS_get_pulse_f64(RSHandle *H, double *out) {
    if (!has_engine(H)) return 0;
    const uint32_t R = H->params.range_count;
    if (rsd_read_pulse(H->workers[0].engine, out, R) != RSD_OK) return 0;
    return (int)R;
}

/* Moment estimator over the device IQ history (rsd_set_iq_ring / rsd_moments):
 * RS_set_moment_pulses keeps the last `pulses` pulses of the current gate count;
 * RS_get_moments fills out[4 g + (Z dB, V m/s, ZDR dB, rho_hv)] and returns R. */
void RS_set_moment_pulses(RSHandle *H, int pulses) {
    if (!has_engine(H)) return;
    engine_ok(H, rsd_set_iq_ring(H->workers[0].engine, pulses, (int)H->params.range_count), "IQ ring");
}

int RS_get_moments(RSHandle *H, int pulses, float *out) {
    if (!has_engine(H)) return 0;
    rsd_moment_params mp;
    mp.lambda = H->params.lambda;
    mp.prt = H->params.prt;
    mp.gain = RS_radar_gain(H);
    mp.range_start = H->params.range_start;
    mp.range_delta = H->params.range_delta;
    if (!engine_ok(H, rsd_moments(H->workers[0].engine, pulses, &mp, out), "moments")) return 0;
    return (int)H->params.range_count;
}

int RS_get_descriptor(RSHandle *H, int which, float *out) {
    const RSWorker *C = &H->workers[0];
    if (which == 0) { memcpy(out, H->sim_desc.s, 64); return 16; }
    if (which == 1) { memcpy(out, C->les_desc.s, 64); return 16; }
    if (which == 2) { memcpy(out, C->range_weight_desc.s, 16); return 4; }
    if (which == 3) { memcpy(out, C->angular_weight_desc.s, 16); return 4; }
    if (which == 4) { memcpy(out, C->rcs_ellipsoid_desc.s, 16); return 4; }
    if (which >= 16 && which < 16 + RS_MAX_ADM_TABLES) { memcpy(out, C->adm_desc[which - 16].s, 64); return 16; }
    if (which >= 32 && which < 32 + RS_MAX_RCS_TABLES) { memcpy(out, C->rcs_desc[which - 32].s, 64); return 16; }
    return 0;
}

rsd_engine *RS_engine(RSHandle *H) { return H ? H->workers[0].engine : NULL; }

void RS_get_info(RSHandle *H, RSInfo *o) {
    memset(o, 0, sizeof(*o));
    o->num_scats = H->num_scats;
    o->num_types = H->num_types;
    for (int k = 0; k < RS_MAX_DEBRIS_TYPES; k++) {
        o->counts[k] = H->counts[k];
        o->origins[k] = H->workers[0].origins[k];
    }
    o->sim_tic = H->sim_tic;
    o->sim_toc = H->sim_toc;
    o->status = (uint32_t)(unsigned char)H->status;
    o->sim_concept = H->sim_concept;
    o->random_seed = H->random_seed;
    o->vel_idx = H->vel_idx;
    o->les_id = H->workers[0].les_id;
    o->range_count = H->params.range_count;
    o->last_error = H->last_error;
    o->dsd_count = H->dsd_count;
    o->adm_count = H->workers[0].adm_count;
    o->rcs_count = H->workers[0].rcs_count;
    o->les_blend = H->les_blend;
    o->wind_alpha = H->wind_alpha;
}

void *RS_get_host_stream(RSHandle *H, int which) {
    switch (which) {
        case 0: return H->scat_pos;
        case 1: return H->scat_vel;
        case 2: return H->scat_ori;
        case 3: return H->scat_tum;
        case 4: return H->scat_aux;
        case 5: return H->scat_rcs;
        case 6: return H->scat_sig;
        case 7: return H->scat_rnd;
        case 8: return H->pulse;
        case 9: return H->anchor_pos;
        default: return NULL;
    }
}

void RS_get_params(RSHandle *H, RSParams *out) { *out = H->params; }

int RS_get_dsd(RSHandle *H, float *r, float *pdf, float *cdf, int cap) {
    int n = MIN(cap, H->dsd_count);
    for (int i = 0; i < n; i++) {
        if (r) r[i] = H->dsd_r[i];
        if (pdf) pdf[i] = H->dsd_pdf[i];
        if (cdf) cdf[i] = H->dsd_cdf[i];
    }
    return H->dsd_count;
}

/* ---------------------------------------------------------------- tables */

RSTable RS_table_init(size_t numel) {
    RSTable t = {0.0f, 1.0f, 1.0f, 0, NULL};
    if (posix_memalign((void **)&t.data, RS_ALIGN_SIZE, MAX(1, numel) * sizeof(float))) t.data = NULL;
    return t;
}
void RS_table_free(RSTable T) { free(T.data); }

RSTable2D RS_table2d_init(size_t numel) {
    RSTable2D t;
    memset(&t, 0, sizeof(t));
    if (posix_memalign((void **)&t.data, RS_ALIGN_SIZE, MAX(1, numel) * sizeof(cl_float4))) t.data = NULL;
    return t;
}
void RS_table2d_free(RSTable2D T) { free(T.data); }

RSTable3D RS_table3d_init(size_t numel) {
    RSTable3D t;
    memset(&t, 0, sizeof(t));
    if (posix_memalign((void **)&t.uvwt, RS_ALIGN_SIZE, MAX(1, numel) * sizeof(cl_float4))) t.uvwt = NULL;
    if (posix_memalign((void **)&t.cpxx, RS_ALIGN_SIZE, MAX(1, numel) * sizeof(cl_float4))) t.cpxx = NULL;
    return t;
}
void RS_table3d_free(RSTable3D T) {
    free(T.uvwt);
    free(T.cpxx);
}

/* ------------------------------------------------------------------ show */

void RS_show_radar_params(RSHandle *H) {
    const RSParams *P = &H->params;
    printf(" Wavelength = %.4f m\n PRT = %.3f ms\n Beamwidth = %.2f deg\n Gain = %.1f dBi\n Pulse width = %.2f us\n"
           " Range resolution = %.1f m\n Gates = %u x %.1f m from %.1f m\n",
           P->lambda, 1.0e3f * P->prt, P->antenna_bw_deg, P->antenna_gain_dbi, 1.0e6f * P->tau, P->dr,
           P->range_count, P->range_delta, P->range_start);
}

static void show_stream(const char *name, const cl_float4 *v, size_t n) {
    if (!v) return;
    printf(" %s:\n", name);
    for (size_t i = 0; i < MIN(n, (size_t)5); i++) printf("  %8zu : %+12.4f %+12.4f %+12.4f %+12.4f\n", i, v[i].x, v[i].y, v[i].z, v[i].w);
}
void RS_show_scat_pos(RSHandle *H) { show_stream("pos", H->scat_pos, H->num_scats); }
void RS_show_scat_sig(RSHandle *H) { show_stream("sig", H->scat_sig, H->num_scats); }
void RS_show_scat_att(RSHandle *H) {
    show_stream("pos", H->scat_pos, H->num_scats);
    show_stream("vel", H->scat_vel, H->num_scats);
    show_stream("rcs", H->scat_rcs, H->num_scats);
}
void RS_show_pulse(RSHandle *H) {
    if (!H->pulse) return;
    for (uint32_t k = 0; k < H->params.range_count; k++)
        printf(" %3u : %+12.4e %+12.4e %+12.4e %+12.4e\n", k, H->pulse[k].s0, H->pulse[k].s1, H->pulse[k].s2, H->pulse[k].s3);
}

/* rs.c:5796-5919 (PPI branch) -- box fitted inside the LES domain */
RSBox RS_suggest_scan_domain(RSHandle *H) {
    RSBox box;
    memset(&box, 0, sizeof(box));
    if (H->P == NULL) {
        H->P = POS_init();
        H->owns_pattern = 1;
    }
    if (H->L == NULL) RS_set_vel_data_to_config(H, LESConfigSuctionVortices);
    float w, h;
    if (H->vel_desc.is_stretched) {
        w = H->vel_desc.ax * (1.0f - powf(H->vel_desc.rx, 0.5f * (float)(H->vel_desc.nx - 3))) / (1.0f - H->vel_desc.rx);
        h = H->vel_desc.az * (1.0f - powf(H->vel_desc.rz, (float)(H->vel_desc.nz - 1))) / (1.0f - H->vel_desc.rz);
    } else {
        w = 0.5f * H->vel_desc.nx * H->vel_desc.rx;
        h = H->vel_desc.nz * H->vel_desc.rz;
    }
    if (H->sim_concept & RSSimulationConceptVerticallyPointingRadar) {
        /* rs.c:5830-5866: elevations of the pattern, ranges 100 m .. 2 km */
        const POSPattern *scan = (const POSPattern *)H->P;
        float emin = 90.0f, emax = 0.0f;
        const float rmin = 100.0f;
        const float rmax = ceilf((2000.0f - rmin) / (2.0f * H->params.range_delta)) * 2.0f * H->params.range_delta;
        for (uint32_t j = 0; j < scan->count; j++) {
            emin = MIN(emin, scan->positions[j].el);
            emax = MAX(emax, scan->positions[j].el);
        }
        if (rmax * cos(emin / 180.0f * M_PI) > w)
            rsprint("WARNING. Elevation %.2f at %.2f m exceeds the LES domain width = %.2f.", emin, rmax, w);
        if (rmax * sin(emax / 180.0f * M_PI) > h)
            rsprint("WARNING. Elevation %.2f at %.2f m exceeds the LES domain height = %.2f.", emax, rmax, h);
        const float nr = ceilf((rmax - rmin) / H->params.dr * 0.5f) * 2.0f;
        box.origin.a = 0.0f;
        box.size.a = 360.0f;
        box.origin.e = emin;
        box.size.e = emax - emin;
        box.origin.r = rmax - (nr + RS_DOMAIN_PAD - 1.0f) * H->params.dr;
        box.size.r = floorf(nr - RS_DOMAIN_PAD - 1.0f) * H->params.dr;
        return box;
    }
    (void)h;
    const int nbeams = 16;
    float na = 0.5f * (float)nbeams + RS_DOMAIN_PAD + 0.5f;
    float ne = 18.0f;
    float rmax = w / sinf(na * H->params.antenna_bw_rad);
    float rmin = (rmax - 2.0f * w) / cosf(na * H->params.antenna_bw_rad) / cosf(ne * H->params.antenna_bw_rad);
    float nr = (rmax - rmin) / H->params.dr - 2.0f * RS_DOMAIN_PAD - 1.0f;
    nr = ceilf(nr * 0.5f) * 2.0f;
    if (rmax - rmin < 8.0f * H->params.dr) rsprint("ERROR: Range resolution of the radar is too coarse!");
    box.origin.a = ceilf(-0.5f * (float)nbeams) * H->params.antenna_bw_rad * 180.0f / M_PI;
    box.size.a = nbeams * H->params.antenna_bw_deg;
    box.origin.r = rmax - (nr + 2.0f * RS_DOMAIN_PAD - 1.0f) * H->params.dr;
    box.size.r = floorf(nr - 2.0f * RS_DOMAIN_PAD - 1.0f) * H->params.dr;
    box.origin.e = 0.0f;
    box.size.e = ne * H->params.antenna_bw_deg;
    if (H->verb) {
        rsprint("Best fit with R:[ %5.2f ~ %5.2f ] km   E:[ %5.2f ~ %5.2f ] deg   A:[ %6.2f ~ %6.2f ] deg",
                1.0e-3f * box.origin.r, 1.0e-3f * (box.origin.r + box.size.r), box.origin.e, box.origin.e + box.size.e,
                box.origin.a, box.origin.a + box.size.a);
    }
    return box;
}

const float *RS_get_les_frame(RSHandle *H, int frame, uint32_t *dims3) {
    if (!H->L) return NULL;
    LESTable *t = LES_get_frame(H->L, frame);
    if (!t) return NULL;
    dims3[0] = t->nx;
    dims3[1] = t->ny;
    dims3[2] = t->nz;
    return (const float *)t->uvwt;
}

const float *RS_get_les_cpxx(RSHandle *H, int frame, uint32_t *dims3) {
    if (!H->L) return NULL;
    LESTable *t = LES_get_frame(H->L, frame);
    if (!t) return NULL;
    const LESFloat4 *c = LES_get_cpxx(H->L, t);
    if (!c) return NULL;
    dims3[0] = t->nx;
    dims3[1] = t->ny;
    dims3[2] = t->nz;
    return (const float *)c;
}

int RS_get_debris_table(RSHandle *H, int t, int which, float *out4, uint32_t *dims2) {
    OBJTable *o = H->O ? OBJ_table_at(H->O, t) : NULL;
    if (!o) return -1;
    if (which < 2) {
        const ADMTable *a = o->adm_table;
        dims2[0] = a->nb;
        dims2[1] = a->na;
        if (!out4) return 0;
        const float *px = which == 0 ? a->data.cdx : a->data.cmx;
        const float *py = which == 0 ? a->data.cdy : a->data.cmy;
        const float *pz = which == 0 ? a->data.cdz : a->data.cmz;
        for (uint32_t i = 0; i < a->nn; i++) {
            out4[4 * i] = px[i]; out4[4 * i + 1] = py[i]; out4[4 * i + 2] = pz[i]; out4[4 * i + 3] = 0.0f;
        }
    } else {
        const RCSTable *r = o->rcs_table;
        dims2[0] = r->na;
        dims2[1] = r->nb;
        if (!out4) return 0;
        const float *px = which == 2 ? r->data.hh_real : r->data.hh_imag;
        const float *py = which == 2 ? r->data.vv_real : r->data.vv_imag;
        const float *pz = which == 2 ? r->data.hv_real : r->data.hv_imag;
        for (uint32_t i = 0; i < r->nn; i++) {
            out4[4 * i] = px[i]; out4[4 * i + 1] = py[i]; out4[4 * i + 2] = pz[i]; out4[4 * i + 3] = 0.0f;
        }
    }
    return 0;
}

void RS_suggest_scan_domain_values(RSHandle *H, float *box8) {
    RSBox b = RS_suggest_scan_domain(H);
    memcpy(box8, b.origin.s, 4 * sizeof(float));
    memcpy(box8 + 4, b.size.s, 4 * sizeof(float));
}

void RS_set_scan_box_values(RSHandle *H, const float *box8) {
    RSBox b;
    memcpy(b.origin.s, box8, 4 * sizeof(float));
    memcpy(b.size.s, box8 + 4, 4 * sizeof(float));
    RS_set_scan_box(H, b);
}

char *RS_simulation_description(RSHandle *H) {
    size_t n = strlen(H->summary);
    if (n > 0 && H->summary[n - 1] == '\n') H->summary[n - 1] = 0;
    return H->summary;
}
