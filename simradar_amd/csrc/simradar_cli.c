/*
 * simradar_cli.c -- command-line driver of the engine: the pulse loop of
 * simradar.c:1011-1074 over the RS_* API, writing the reference's .iq file
 * format (iq.h:20-41; simradar.c:340-360): a 1024-byte file header (RSParams,
 * debris counts, scan mode and sweep, seed) followed, per pulse, by a 32-byte
 * pulse header (time, elevation, azimuth) and range_count x (Ih, Qh, Iv, Qv)
 * binary32.
 *
 * Options follow simradar.c:489-517 (same letters and meanings) where the
 * engine has the feature:
 *   -p/--pulses N   -S/--sweep PATTERN   -c/--concept LETTERS (B D U T)
 *   -d/--debris TYPE,COUNT (repeatable)   -s/--seed N   -t/--prt S
 *   -l/--lambda M   -w/--pulsewidth S   -b/--beamwidth DEG   -D/--density X
 *   -g/--mp-dsd D1:D2:... (mm)   -L/--les CONFIG   -W/--warm-up N
 *   -o/--output   -O/--out-dir DIR   -q/--quiet   -v/--verbose   -N/--no-run
 * Additions: --population N (exact rain count), --gate M (range spacing,
 * default 30 m as simradar.c:909), --moments M (print Z / V / ZDR / rho_hv of
 * the last M pulses per gate), --phase32 (binary32 phase as rs.cl:1096),
 * --hilo (binary32 hi/lo positions), --range-cap M.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/time.h>
#include <time.h>

#include "rs.h"

typedef union {
    char raw[1024];
    struct {
        RSParams params;
        uint32_t counts[RS_MAX_DEBRIS_TYPES];
        char scan_mode[16];
        float scan_start;
        float scan_end;
        float scan_delta;
        uint32_t simulation_seed;
    };
} IQFileHeader;

typedef union {
    char raw[32];
    struct {
        float time;
        float el_deg;
        float az_deg;
    };
} IQPulseHeader;

static double now_s(void) {
    struct timeval t;
    gettimeofday(&t, NULL);
    return (double)t.tv_sec + 1e-6 * (double)t.tv_usec;
}

static void usage(const char *prog) {
    printf("usage: %s [-p pulses] [-S sweep] [-c concept] [-d type,count] [-s seed] [-o] [-O dir] ...\n"
           "  -p, --pulses N          pulses to emulate (default 1000)\n"
           "  -S, --sweep P:el,a0:a1:da | R:az,e0:e1:de | D:...   scan pattern (default P:3.0,-12:12:0.01)\n"
           "  -c, --concept LETTERS   B bounded velocity, D dragged background, U uniform DSD, T transparent\n"
           "  -d, --debris TYPE,COUNT debris group (repeatable)\n"
           "  -s, --seed N            random seed\n"
           "  -t, --prt S  -l, --lambda M  -w, --pulsewidth S  -b, --beamwidth DEG  -D, --density X\n"
           "  -g, --mp-dsd D1:D2:...  Marshall-Palmer DSD at these diameters (mm)\n"
           "  -L, --les CONFIG        LES configuration (suctvort or suctvort:cx,cy,seed)\n"
           "  -W, --warm-up N         advances at PRT 1/60 s before the pulses\n"
           "  -o, --output            write the .iq file   -O, --out-dir DIR\n"
           "  --population N  --gate M  --moments M  --phase32  --hilo  --range-cap M\n"
           "  -q, --quiet  -v, --verbose  -N, --no-run  -h, --help\n",
           prog);
}

int main(int argc, char **argv) {
    int pulses = -1, warm = -1, seed = -1, verb = 0, output = 0, no_run = 0, quiet = 0, moments = 0, phase32 = 0,
        hilo = 0;
    float prt = -1.0f, lambda = -1.0f, pw = -1.0f, bw = -1.0f, density = -1.0f, gate = 30.0f, range_cap = -1.0f;
    long population = -1;
    uint32_t concept = 0;
    char sweep[256] = "", les[128] = "", outdir[512] = ".";
    int dtype[RS_MAX_DEBRIS_TYPES], dcount[RS_MAX_DEBRIS_TYPES], ngroups = 0;
    float dsd[32];
    int ndsd = 0;
    static struct option opts[] = {
        {"pulses", required_argument, 0, 'p'}, {"sweep", required_argument, 0, 'S'},
        {"concept", required_argument, 0, 'c'}, {"debris", required_argument, 0, 'd'},
        {"seed", required_argument, 0, 's'}, {"prt", required_argument, 0, 't'},
        {"lambda", required_argument, 0, 'l'}, {"pulsewidth", required_argument, 0, 'w'},
        {"beamwidth", required_argument, 0, 'b'}, {"density", required_argument, 0, 'D'},
        {"mp-dsd", required_argument, 0, 'g'}, {"les", required_argument, 0, 'L'},
        {"warm-up", required_argument, 0, 'W'}, {"output", no_argument, 0, 'o'},
        {"out-dir", required_argument, 0, 'O'}, {"quiet", no_argument, 0, 'q'},
        {"verbose", no_argument, 0, 'v'}, {"no-run", no_argument, 0, 'N'}, {"help", no_argument, 0, 'h'},
        {"population", required_argument, 0, 1001}, {"gate", required_argument, 0, 1002},
        {"moments", required_argument, 0, 1003}, {"phase32", no_argument, 0, 1004},
        {"hilo", no_argument, 0, 1005}, {"range-cap", required_argument, 0, 1006}, {0, 0, 0, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "p:S:c:d:s:t:l:w:b:D:g:L:W:oO:qvNh", opts, NULL)) != -1) {
        switch (c) {
            case 'p': pulses = atoi(optarg); break;
            case 'S': snprintf(sweep, sizeof(sweep), "%s", optarg); break;
            case 'c':
                concept = 0;
                if (strchr(optarg, 'B') || strchr(optarg, 'b')) concept |= RSSimulationConceptBoundedParticleVelocity;
                if (strchr(optarg, 'D') || strchr(optarg, 'd')) concept |= RSSimulationConceptDraggedBackground;
                if (strchr(optarg, 'U') || strchr(optarg, 'u')) concept |= RSSimulationConceptUniformDSDScaledRCS;
                if (strchr(optarg, 'T') || strchr(optarg, 't')) concept |= RSSimulationConceptTransparentBackground;
                break;
            case 'd':
                if (ngroups >= RS_MAX_DEBRIS_TYPES - 1 || sscanf(optarg, "%d,%d", &dtype[ngroups], &dcount[ngroups]) != 2) {
                    fprintf(stderr, "Each debris group should be specified as -dTYPE,COUNT.\n");
                    return EXIT_FAILURE;
                }
                ngroups++;
                break;
            case 's': seed = atoi(optarg); break;
            case 't': prt = (float)atof(optarg); break;
            case 'l': lambda = (float)atof(optarg); break;
            case 'w': pw = (float)atof(optarg); break;
            case 'b': bw = (float)atof(optarg); break;
            case 'D': density = (float)atof(optarg); break;
            case 'g': {
                const char *p = optarg;
                while (p && *p && ndsd < 32) {
                    dsd[ndsd++] = 1.0e-3f * (float)atof(p);
                    p = strchr(p, ':');
                    if (p) p++;
                }
                break;
            }
            case 'L': snprintf(les, sizeof(les), "%s", optarg); break;
            case 'W': warm = atoi(optarg); break;
            case 'o': output = 1; break;
            case 'O': snprintf(outdir, sizeof(outdir), "%s", optarg); break;
            case 'q': quiet = 1; break;
            case 'v': verb++; break;
            case 'N': no_run = 1; break;
            case 1001: population = atol(optarg); break;
            case 1002: gate = (float)atof(optarg); break;
            case 1003: moments = atoi(optarg); break;
            case 1004: phase32 = 1; break;
            case 1005: hilo = 1; break;
            case 1006: range_cap = (float)atof(optarg); break;
            case 'h': usage(argv[0]); return EXIT_SUCCESS;
            default: usage(argv[0]); return EXIT_FAILURE;
        }
    }
    /* defaults of simradar.c:746-757 */
    if (concept == 0)
        concept = RSSimulationConceptDraggedBackground | RSSimulationConceptBoundedParticleVelocity |
                  RSSimulationConceptUniformDSDScaledRCS;
    if (!sweep[0]) snprintf(sweep, sizeof(sweep), "P:3.0,-12:12:0.01");
    if (pulses < 0) pulses = 1000;
    if (warm < 0) warm = pulses > 1000 ? 2000 : 0;
    POSPattern *scan = POS_init_with_string(sweep);
    if (!scan) {
        fprintf(stderr, "Invalid scan pattern '%s'.\n", sweep);
        return EXIT_FAILURE;
    }
    if (no_run) {
        for (int k = 0; k < pulses; k++) {
            POS_get_next_angles(scan);
            printf("%6d  E%6.2f  A%7.2f\n", k, scan->el, scan->az);
        }
        POS_free(scan);
        return EXIT_SUCCESS;
    }
    RSHandle *S = RS_init_verbose((char)verb);
    if (!S) {
        POS_free(scan);
        return EXIT_FAILURE;
    }
    RS_set_concept(S, concept);
    RS_set_scan_pattern(S, scan);
    if (bw > 0.0f) RS_set_antenna_params(S, bw, 44.5f);
    if (density > 0.0f) RS_set_density(S, density);
    if (pw > 0.0f) RS_set_tx_params(S, pw, 50.0e3f);
    if (lambda > 0.0f) RS_set_lambda(S, lambda);
    if (seed >= 0) RS_set_random_seed(S, (uint32_t)seed);
    if (ndsd > 0) RS_set_dsd_to_mp_with_sizes(S, dsd, ndsd);
    else RS_set_dsd_to_mp(S);
    if (les[0]) RS_set_vel_data_to_config(S, les);
    RS_set_sampling_spacing(S, gate, 1.0f, 1.0f);
    if (range_cap > 0.0f) RS_set_range_cap(S, range_cap);
    for (int k = 0; k < ngroups; k++) RS_add_debris(S, (OBJConfig)dtype[k], (size_t)dcount[k]);
    RS_set_scan_box(S, RS_suggest_scan_domain(S));
    if (prt > 0.0f) RS_set_prt(S, prt);
    if (population > 0) RS_set_population(S, (size_t)population);
    if (phase32) RS_set_phase_precision(S, 1);
    RS_populate(S);
    if (S->num_scats == 0) {
        RS_free(S);
        POS_free(scan);
        return EXIT_FAILURE;
    }
    if (hilo && rsd_set_position_precision(RS_engine(S), 1) != RSD_OK) {
        fprintf(stderr, "hi/lo positions: %s\n", rsd_last_error(RS_engine(S)));
        RS_free(S);
        POS_free(scan);
        return EXIT_FAILURE;
    }
    if (moments > 1) RS_set_moment_pulses(S, moments);
    if (!quiet) printf("Emulating %d pulse%s with %zu scatter bodies\n", pulses, pulses > 1 ? "s" : "", S->num_scats);
    if (warm > 0) {   /* simradar.c:975-995: longer PRT so that the bodies move */
        const float prt_user = S->params.prt;
        RS_set_prt(S, 1.0f / 60.0f);
        for (int k = 0; k < warm; k++) RS_advance_time(S);
        RS_set_prt(S, prt_user);
    }
    const uint32_t R = S->params.range_count;
    IQPulseHeader *ph = NULL;
    float *cache = NULL;
    if (output) {
        ph = (IQPulseHeader *)calloc((size_t)pulses, sizeof(IQPulseHeader));
        cache = (float *)calloc((size_t)pulses * R * 4, sizeof(float));
        if (!ph || !cache) {
            fprintf(stderr, "Out of memory for the pulse cache.\n");
            RS_free(S);
            POS_free(scan);
            return EXIT_FAILURE;
        }
    }
    const double t0 = now_s();
    for (int k = 0; k < pulses; k++) {
        RS_make_pulse(S);
        if (output) {
            RS_download_pulse_only(S);
            ph[k].time = (float)S->sim_tic;
            ph[k].az_deg = scan->az;
            ph[k].el_deg = scan->el;
            memcpy(cache + (size_t)k * R * 4, S->pulse, (size_t)R * 4 * sizeof(float));
        }
        RS_advance_time(S);
        RS_advance_beam(S);
    }
    RS_download_pulse_only(S);   /* waits for the last pulse */
    const double dt = now_s() - t0;
    if (!quiet) printf("%d pulses in %.3f s: %.1f pulses/s, %.3e scatterer-updates/s\n", pulses, dt, pulses / dt,
                       pulses / dt * (double)S->num_scats);
    int rc = EXIT_SUCCESS;
    if (output) {
        IQFileHeader fh;
        memset(&fh, 0, sizeof(fh));
        fh.params = S->params;
        for (int k = 0; k < RS_MAX_DEBRIS_TYPES; k++) fh.counts[k] = (uint32_t)S->counts[k];
        snprintf(fh.scan_mode, sizeof(fh.scan_mode), "%c", scan->mode);
        fh.scan_start = scan->sweeps[0].azStart;
        fh.scan_end = scan->sweeps[0].azEnd;
        fh.scan_delta = scan->sweeps[0].azDelta;
        fh.simulation_seed = S->random_seed;
        char name[1024], stamp[64];
        time_t tt = time(NULL);
        strftime(stamp, sizeof(stamp), "%Y%m%d-%H%M%S", localtime(&tt));
        snprintf(name, sizeof(name), "%s/sim-%s-%s%04.1f.iq", outdir, stamp, POS_is_ppi(scan) ? "E" : (POS_is_rhi(scan) ? "A" : "S"),
                 POS_is_ppi(scan) ? scan->sweeps[0].elStart : (POS_is_rhi(scan) ? scan->sweeps[0].azStart : (float)pulses));
        FILE *f = fopen(name, "wb");
        if (!f) {
            fprintf(stderr, "Error creating file %s for IQ data.\n", name);
            rc = EXIT_FAILURE;
        } else {
            fwrite(&fh, sizeof(fh), 1, f);
            for (int k = 0; k < pulses; k++) {
                fwrite(&ph[k], sizeof(IQPulseHeader), 1, f);
                fwrite(cache + (size_t)k * R * 4, sizeof(float) * 4, R, f);
            }
            if (!quiet) printf("Output file : %s (%ld B, seed %u)\n", name, ftell(f), fh.simulation_seed);
            fclose(f);
        }
    }
    if (moments > 1) {
        float *m = (float *)malloc((size_t)R * 4 * sizeof(float));
        if (m && RS_get_moments(S, moments, m) == (int)R) {
            printf("  gate   range (m)     Z (dB)    V (m/s)   ZDR (dB)   rho_hv\n");
            for (uint32_t g = 0; g < R; g++)
                printf("  %4u  %10.2f  %9.3f  %9.3f  %9.3f  %7.4f\n", g, S->params.range_start + g * S->params.range_delta,
                       m[4 * g], m[4 * g + 1], m[4 * g + 2], m[4 * g + 3]);
        }
        free(m);
    }
    free(ph);
    free(cache);
    RS_free(S);
    POS_free(scan);
    return rc;
}
