/*
 * rs_providers.c -- in-memory table providers, scan patterns and logging for
 * the RS_* host runtime (plain C).
 *
 * Interfaces as les.h / adm.h / rcs.h / obj.h / pos.h / log.h of the reference;
 * the tables themselves are synthetic (see rs_providers.h for the fields).
 */
#include "rs_providers.h"

#include <pthread.h>
#include <stdarg.h>
#include <strings.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

/* ------------------------------------------------------------------ logging */

char *now(void) {
    static char buf[64];
    struct timeval tv;
    gettimeofday(&tv, NULL);
    struct tm tm;
    localtime_r(&tv.tv_sec, &tm);
    snprintf(buf, sizeof(buf), "%02d:%02d:%02d.%03d", tm.tm_hour, tm.tm_min, tm.tm_sec, (int)(tv.tv_usec / 1000));
    return buf;
}

char *commaint(long long num) {
    static char bufs[8][64];
    static int k = 0;
    char *out = bufs[k];
    k = (k + 1) % 8;
    char tmp[64];
    snprintf(tmp, sizeof(tmp), "%lld", num < 0 ? -num : num);
    int len = (int)strlen(tmp), j = 0;
    if (num < 0) out[j++] = '-';
    for (int i = 0; i < len; i++) {
        out[j++] = tmp[i];
        int left = len - 1 - i;
        if (left > 0 && left % 3 == 0) out[j++] = ',';
    }
    out[j] = '\0';
    return out;
}

char *commafloat(float num) {
    static char bufs[8][64];
    static int k = 0;
    char *out = bufs[k];
    k = (k + 1) % 8;
    long long ip = (long long)num;
    float frac = fabsf(num - (float)ip);
    snprintf(out, 64, "%s.%02d", commaint(ip), (int)(frac * 100.0f + 0.5f) % 100);
    return out;
}

void rsprint(const char *format, ...) {
    char msg[4096];
    va_list args;
    va_start(args, format);
    vsnprintf(msg, sizeof(msg), format, args);
    va_end(args);
    size_t n = strlen(msg);
    while (n > 0 && msg[n - 1] == '\n') msg[--n] = '\0';
    FILE *out = (strncmp(msg, "ERROR", 5) == 0 || strncmp(msg, "WARNING", 7) == 0) ? stderr : stdout;
    fprintf(out, "%s : RS : %s\n", now(), msg);
}

/* ------------------------------------------------------- seeded generator */

uint64_t rs_splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double rs_uniform(uint64_t *s) { return ((double)(rs_splitmix64(s) >> 11) + 0.5) * 0x1.0p-53; }

double rs_gauss(uint64_t *s) {
    double u1 = rs_uniform(s), u2 = rs_uniform(s);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* ----------------------------------------------------------------- LES */

typedef struct {
    uint32_t nx, ny, nz, frames;
    float ax, rx, az, rz, core, vmax, sigma, tp, tr;
    float cx, cy;           /* vortex centre offset from the grid centre (m) */
    uint64_t seed;
    LESTable **tables;
    /* The frame after the one last handed out is generated ahead on a thread
     * of its own, as the reference reads the next LES file ahead (les.c:268,
     * les.c:666-692). */
    pthread_t worker;
    int worker_on;
    uint32_t ahead;
    /* Frames read from the reference's LES files instead of generated. */
    int file_mode;
    char dir[2200];         /* <tables>/les/<config> */
    float v0, p0, t0, ay, ry;
    int stretched, p_for_cn2;
} LESMem;

/* ---------------------------------------------------------- table files */

/* First directory of dirs[0..n) holding the file probe (relative to it). */
static int find_table_dir(char dirs[][1024], int n, const char *probe, char *out) {
    struct stat st;
    char f[2200];
    for (int i = 0; i < n; i++) {
        if (!dirs[i][0]) continue;
        snprintf(f, sizeof(f), "%s/%s", dirs[i], probe);
        if (stat(dirs[i], &st) == 0 && S_ISDIR(st.st_mode) && stat(f, &st) == 0 && S_ISREG(st.st_mode)) {
            snprintf(out, 1024, "%s", dirs[i]);
            return 1;
        }
    }
    return 0;
}

/* Search list of adm.c:56-79 / rcs.c:92-114 (sub "les" for ADM, "rcs" for
 * RCS) and les.c:74-91 (les = 1): the given path (LES: path/tables) or
 * <cwd>/Contents/Resources/<sub>, then folders under $HOME. */
static int table_search_list(const char *path, const char *sub, int les, char dirs[8][1024]) {
    int k = 0;
    if (path) {
        snprintf(dirs[k++], 1024, les ? "%s/tables" : "%s", path);
    } else {
        char cwd[960];
        if (getcwd(cwd, sizeof(cwd))) snprintf(dirs[k++], 1024, "%s/Contents/Resources/%s", cwd, sub);
    }
    const char *home = getenv("HOME");
    if (home) {
        snprintf(dirs[k++], 1024, "%s/Documents/tables", home);
        snprintf(dirs[k++], 1024, "%s/Downloads/tables", home);
        if (les) {
            snprintf(dirs[k++], 1024, "%s/Desktop/les", home);
            snprintf(dirs[k++], 1024, "%s/Documents/les", home);
            snprintf(dirs[k++], 1024, "%s/Downloads/les", home);
        } else {
            snprintf(dirs[k++], 1024, "%s/tables", home);
        }
    }
    return k;
}

#define LES_FILE_NBLOCK 10      /* volumes per LES data file (les.c:12) */

LESHandle LES_init_synthetic(uint32_t nx, uint32_t ny, uint32_t nz, float a_xy, float r_xy, float a_z, float r_z,
                             float core_m, float vmax, float sigma, uint64_t seed, uint32_t frames) {
    if (nx < 2 || ny < 2 || nz < 2 || frames < 1) return NULL;
    LESMem *h = (LESMem *)calloc(1, sizeof(LESMem));
    if (!h) return NULL;
    h->nx = nx; h->ny = ny; h->nz = nz; h->frames = frames;
    h->ax = a_xy; h->rx = r_xy; h->az = a_z; h->rz = r_z;
    h->core = core_m; h->vmax = vmax; h->sigma = sigma;
    h->tp = 2.0f;
    h->tr = 50.0f;
    h->seed = seed;
    h->tables = (LESTable **)calloc(frames, sizeof(LESTable *));
    if (!h->tables) {
        free(h);
        return NULL;
    }
    return (LESHandle)h;
}

/* LES from the reference's files (les.c:71-291, 318-422, 474-517): the grid
 * record <dir>/les/<config>/fort.10_2 (uint32 rev, nx, ny, nz, then the x, y and
 * z planes, each after 8 bytes) and the data files LES_mean_1_6_fnum<k>.dat
 * (uint32 version, then per volume a float time stamp and the u, v, w, p, t
 * planes, each record followed by 8 bytes), scaled by v0 / p0 / t0 and given
 * the grid spacings of the configuration as les.c:154-199 does. */
static LESHandle les_init_files(const char *config, const char *dir) {
    char f[2400];
    snprintf(f, sizeof(f), "%s/les/%s/fort.10_2", dir, config);
    FILE *g = fopen(f, "rb");
    if (!g) return NULL;
    uint32_t hdr[4];
    if (fread(hdr, sizeof(uint32_t), 4, g) != 4 || hdr[1] < 2 || hdr[2] < 2 || hdr[3] < 2) {
        fclose(g);
        return NULL;
    }
    LESMem *h = (LESMem *)calloc(1, sizeof(LESMem));
    if (!h) {
        fclose(g);
        return NULL;
    }
    h->file_mode = 1;
    snprintf(h->dir, sizeof(h->dir), "%s/les/%s", dir, config);
    h->nx = hdr[1]; h->ny = hdr[2]; h->nz = hdr[3];
    const size_t nn = (size_t)h->nx * h->ny * h->nz;
    h->tr = 50.0f;
    h->v0 = 100.0f;
    h->p0 = h->v0 * h->v0;
    h->t0 = h->v0 * h->v0;
    if (!strcasecmp(config, LESConfigSuctionVortices) || !strcasecmp(config, LESConfigSuctionVorticesLarge)) {
        h->ax = h->ay = h->az = 2.0f;
        h->rx = h->ry = 1.0212f;
        h->rz = 1.05f;
        h->tp = 2.0f;
        h->stretched = 1;
    } else if (!strcasecmp(config, LESConfigTwoCell)) {
        h->v0 = 225.0f;
        h->p0 = h->t0 = h->v0 * h->v0;
        h->ax = h->ay = h->az = 1.0f;
        h->rx = h->ry = h->rz = 1.0f;
        h->tp = 5.0f;
    } else if (!strcasecmp(config, LESConfigFlat)) {
        /* spacings from the grid record: x[1] - x[0], y[nx] - y[0], z[nx ny] - z[0] (km) */
        float *xyz = (float *)malloc(3 * nn * sizeof(float));
        int ok = xyz != NULL;
        for (int c = 0; ok && c < 3; c++)
            ok = fseek(g, 2 * sizeof(float), SEEK_CUR) == 0 && fread(xyz + c * nn, sizeof(float), nn, g) == nn;
        if (!ok) {
            free(xyz);
            fclose(g);
            free(h);
            return NULL;
        }
        h->p0 = 1.0f;
        h->ax = h->ay = h->az = 0.0f;
        h->rx = 1.0e3f * (xyz[1] - xyz[0]);
        h->ry = 1.0e3f * (xyz[nn + h->nx] - xyz[nn]);
        h->rz = 1.0e3f * (xyz[2 * nn + (size_t)h->nx * h->ny] - xyz[2 * nn]);
        h->tp = 60.0f;
        h->p_for_cn2 = 1;
        free(xyz);
    } else {
        h->ax = h->ay = h->az = 1.0f;
        h->rx = h->ry = h->rz = 1.0f;
        h->tp = 5.0f;
    }
    fclose(g);
    /* volumes per file from the first file's size; files numbered from 1 */
    const size_t vol = 4 + 8 + 5 * (nn * sizeof(float) + 8);
    size_t nvol = 0, nfiles = 0;
    for (;;) {
        snprintf(f, sizeof(f), "%s/LES_mean_1_6_fnum%zu.dat", h->dir, nfiles + 1);
        struct stat st;
        if (stat(f, &st) != 0 || !S_ISREG(st.st_mode)) break;
        if (nfiles == 0) nvol = st.st_size > 4 ? ((size_t)st.st_size - 4) / vol : 0;
        nfiles++;
    }
    if (nfiles == 0 || nvol == 0) {
        rsprint("ERROR: LES data files missing or short in %s.", h->dir);
        free(h);
        return NULL;
    }
    if (nvol != LES_FILE_NBLOCK) rsprint("WARNING: each LES data file holds %zu volumes, expected %d.", nvol, LES_FILE_NBLOCK);
    h->seed = nvol;                    /* volumes per file */
    h->frames = (uint32_t)(nfiles * nvol);
    h->tables = (LESTable **)calloc(h->frames, sizeof(LESTable *));
    if (!h->tables) {
        free(h);
        return NULL;
    }
    return (LESHandle)h;
}

LESHandle LES_init_with_config_path(const LESConfig config, const char *path) {
    if (config && config[0]) {
        char dirs[8][1024], dir[1024], probe[1100];
        const int n = table_search_list(path, "les", 1, dirs);
        snprintf(probe, sizeof(probe), "les/%s/fort.10_2", config);
        if (find_table_dir(dirs, n, probe, dir)) {
            LESHandle h = les_init_files(config, dir);
            if (h) return h;
        }
    }
    if (config && (!strcmp(config, LESConfigSuctionVortices) || !strcmp(config, LESConfigSuctionVorticesLarge))) {
        return LES_init_synthetic(176, 176, 80, 2.0f, 1.0212f, 2.0f, 1.05f, 50.0f, 40.0f, 2.0f, 20250101ull, 10);
    }
    /* "suctvort:cx,cy,seed": the same grid with the vortex centre moved by
     * (cx, cy) metres and its own noise seed (independent domains of a batch) */
    if (config && !strncmp(config, "suctvort:", 9)) {
        float cx = 0.0f, cy = 0.0f;
        unsigned long long seed = 20250101ull;
        if (sscanf(config + 9, "%f,%f,%llu", &cx, &cy, &seed) >= 2) {
            LESMem *h = (LESMem *)LES_init_synthetic(176, 176, 80, 2.0f, 1.0212f, 2.0f, 1.05f, 50.0f, 40.0f, 2.0f,
                                                     (uint64_t)seed, 10);
            if (h) {
                h->cx = cx;
                h->cy = cy;
            }
            return (LESHandle)h;
        }
    }
    rsprint("ERROR: LES configuration '%s' needs external data; only '%s' is generated in memory.",
            config ? config : "(null)", LESConfigSuctionVortices);
    return NULL;
}

LESHandle LES_init(void) { return LES_init_with_config_path(LESConfigSuctionVortices, NULL); }

static void les_free_table(LESTable *t) {
    if (!t) return;
    free(t->uvwt);
    free(t->cpxx);
    free(t);
}

static void les_join(LESMem *h) {
    if (h->worker_on) {
        pthread_join(h->worker, NULL);
        h->worker_on = 0;
    }
}

void LES_free(LESHandle i) {
    LESMem *h = (LESMem *)i;
    if (!h) return;
    les_join(h);
    for (uint32_t f = 0; f < h->frames; f++) les_free_table(h->tables[f]);
    free(h->tables);
    free(h);
}

/* Position of texel index i along a stretched axis: the stretched map sends
 * |x| = a (r^m - 1) / (r - 1) to coordinate m + offset and texel i is read at
 * coordinate i + 0.5 (rs.cl:130-151 with the offsets of rs.c:3058-3061). */
static double texel_position(int i, double offset, double a, double r) {
    double m = (double)i + 0.5 - offset;
    double s = m < 0.0 ? -1.0 : 1.0;
    return s * a * (pow(r, fabs(m)) - 1.0) / (r - 1.0);
}

/* One slab of z slices of a frame.  splitmix64 advances its state by a fixed
 * increment per draw and every texel takes six draws, so a slab starts from the
 * state the sequential walk would have there: the frame is the same for any
 * number of slabs. */
typedef struct {
    const LESMem *h;
    LESTable *t;
    const double *xs, *ys;      /* texel positions along x and y */
    uint64_t s0;                /* generator state before texel 0 */
    uint32_t k0, k1;            /* z slices [k0, k1) */
} LESSlab;

static void *les_slab_run(void *arg) {
    const LESSlab *p = (const LESSlab *)arg;
    const LESMem *h = p->h;
    for (uint32_t k = p->k0; k < p->k1; k++) {
        size_t idx = (size_t)k * h->nx * h->ny;
        uint64_t s = p->s0 + (uint64_t)(6 * idx) * 0x9E3779B97F4A7C15ull;
        for (uint32_t j = 0; j < h->ny; j++) {
            const double y = p->ys[j];
            for (uint32_t i = 0; i < h->nx; i++, idx++) {
                const double x = p->xs[i];
                double xc = x - h->cx, yc = y - h->cy;
                double rho = sqrt(xc * xc + yc * yc);
                double vt = rho <= h->core ? h->vmax * rho / h->core : h->vmax * h->core / rho;
                double u = rho > 0.0 ? -vt * yc / rho : 0.0;
                double v = rho > 0.0 ? vt * xc / rho : 0.0;
                p->t->uvwt[idx][0] = (float)(u + h->sigma * rs_gauss(&s));
                p->t->uvwt[idx][1] = (float)(v + h->sigma * rs_gauss(&s));
                p->t->uvwt[idx][2] = (float)(h->sigma * rs_gauss(&s));
                p->t->uvwt[idx][3] = 0.0f;
            }
        }
    }
    return NULL;
}

#define LES_SLABS 8

static LESTable *les_make_frame(LESMem *h, uint32_t f) {
    LESTable *t = (LESTable *)calloc(1, sizeof(LESTable));
    if (!t) return NULL;
    t->nx = h->nx; t->ny = h->ny; t->nz = h->nz;
    t->nn = h->nx * h->ny * h->nz;
    t->nt = 1;
    t->nc = h->frames;
    t->is_stretched = true;
    t->tr = h->tr;
    t->tp = h->tp;
    t->ax = h->ax; t->ay = h->ax; t->az = h->az;
    t->rx = h->rx; t->ry = h->rx; t->rz = h->rz;
    t->uvwt = (LESFloat4 *)malloc((size_t)t->nn * sizeof(LESFloat4));
    double *xs = (double *)malloc((size_t)(h->nx + h->ny) * sizeof(double));
    if (!t->uvwt || !xs) {
        free(t->uvwt);
        free(xs);
        free(t);
        return NULL;
    }
    double *ys = xs + h->nx;
    const double offx = 0.5 * (double)(h->nx - 1), offy = 0.5 * (double)(h->ny - 1);
    for (uint32_t i = 0; i < h->nx; i++) xs[i] = texel_position((int)i, offx, h->ax, h->rx);
    for (uint32_t j = 0; j < h->ny; j++) ys[j] = texel_position((int)j, offy, h->ax, h->rx);
    const uint64_t s0 = h->seed * 0x9E3779B97F4A7C15ull + (uint64_t)(f + 1) * 0xD1B54A32D192ED03ull;
    const uint32_t nslab = h->nz < LES_SLABS ? h->nz : LES_SLABS;
    LESSlab slab[LES_SLABS];
    pthread_t tid[LES_SLABS];
    int started[LES_SLABS] = {0};
    for (uint32_t b = 0; b < nslab; b++) {
        slab[b].h = h;
        slab[b].t = t;
        slab[b].xs = xs;
        slab[b].ys = ys;
        slab[b].s0 = s0;
        slab[b].k0 = (uint32_t)((uint64_t)h->nz * b / nslab);
        slab[b].k1 = (uint32_t)((uint64_t)h->nz * (b + 1) / nslab);
        if (b > 0) started[b] = pthread_create(&tid[b], NULL, les_slab_run, &slab[b]) == 0;
    }
    les_slab_run(&slab[0]);
    for (uint32_t b = 1; b < nslab; b++) {
        if (started[b]) pthread_join(tid[b], NULL);
        else les_slab_run(&slab[b]);
    }
    free(xs);
    return t;
}

/* Volume f of a file-backed LES (les.c:318-384). */
static LESTable *les_read_frame(LESMem *h, uint32_t f) {
    const size_t nn = (size_t)h->nx * h->ny * h->nz, nvol = (size_t)h->seed;
    char name[2300];
    snprintf(name, sizeof(name), "%s/LES_mean_1_6_fnum%zu.dat", h->dir, f / nvol + 1);
    FILE *fid = fopen(name, "rb");
    if (!fid) return NULL;
    const long off = (long)(4 + (f % nvol) * (4 + 8 + 5 * (nn * sizeof(float) + 8)));
    float *buf = (float *)malloc(5 * nn * sizeof(float));
    LESTable *t = (LESTable *)calloc(1, sizeof(LESTable));
    int ok = buf && t && fseek(fid, off + 4 + 8, SEEK_SET) == 0;
    for (int c = 0; ok && c < 5; c++) {
        ok = fread(buf + c * nn, sizeof(float), nn, fid) == nn && fseek(fid, 8, SEEK_CUR) == 0;
    }
    fclose(fid);
    if (ok) {
        t->uvwt = (LESFloat4 *)malloc(nn * sizeof(LESFloat4));
        t->cpxx = (LESFloat4 *)malloc(nn * sizeof(LESFloat4));
        ok = t->uvwt && t->cpxx;
    }
    if (!ok) {
        free(buf);
        if (t) {
            free(t->uvwt);
            free(t->cpxx);
            free(t);
        }
        return NULL;
    }
    t->nx = h->nx; t->ny = h->ny; t->nz = h->nz;
    t->nn = (uint32_t)nn;
    t->nt = (uint32_t)nvol;
    t->nc = h->frames;
    t->is_stretched = h->stretched != 0;
    t->tr = h->tr;
    t->tp = h->tp;
    t->ax = h->ax; t->ay = h->ay; t->az = h->az;
    t->rx = h->rx; t->ry = h->ry; t->rz = h->rz;
    for (size_t k = 0; k < nn; k++) {
        const float u = buf[k] * h->v0, v = buf[nn + k] * h->v0, w = buf[2 * nn + k] * h->v0;
        const float pp = buf[3 * nn + k] * h->p0, tt = buf[4 * nn + k] * h->t0;
        t->uvwt[k][0] = u; t->uvwt[k][1] = v; t->uvwt[k][2] = w; t->uvwt[k][3] = tt;
        t->cpxx[k][0] = h->p_for_cn2 ? pp : 0.0f;
        t->cpxx[k][1] = h->p_for_cn2 ? 0.0f : pp;
        t->cpxx[k][2] = 0.0f;
        t->cpxx[k][3] = 0.0f;
    }
    free(buf);
    return t;
}

static LESTable *les_frame(LESMem *h, uint32_t f) { return h->file_mode ? les_read_frame(h, f) : les_make_frame(h, f); }

/* The (cn2, p, 0, 0) volume of a frame (les.c:389-405; read by fp_atts,
 * rs.cl:590-648).  File frames carry it; a synthetic frame gets it on the
 * first request: log10 Cn^2 = -13 - z / 1 km + 0.25 sin(2 pi x / 180 m)
 * cos(2 pi y / 240 m) at the texel positions of the stretched grid, p = 0. */
const LESFloat4 *LES_get_cpxx(const LESHandle L, LESTable *t) {
    LESMem *h = (LESMem *)L;
    if (!h || !t) return NULL;
    if (t->cpxx) return t->cpxx;
    LESFloat4 *c = (LESFloat4 *)malloc((size_t)t->nn * sizeof(LESFloat4));
    if (!c) return NULL;
    const double offx = 0.5 * (double)(t->nx - 1), offy = 0.5 * (double)(t->ny - 1);
    size_t idx = 0;
    for (uint32_t k = 0; k < t->nz; k++) {
        const double z = texel_position((int)k, 0.0, t->az, t->rz);
        for (uint32_t j = 0; j < t->ny; j++) {
            const double y = texel_position((int)j, offy, t->ay, t->ry);
            for (uint32_t i = 0; i < t->nx; i++, idx++) {
                const double x = texel_position((int)i, offx, t->ax, t->rx);
                c[idx][0] = (float)(-13.0 - 1.0e-3 * z + 0.25 * sin(2.0 * M_PI * x / 180.0) * cos(2.0 * M_PI * y / 240.0));
                c[idx][1] = c[idx][2] = c[idx][3] = 0.0f;
            }
        }
    }
    t->cpxx = c;
    return c;
}

static void *les_ahead_run(void *arg) {
    LESMem *h = (LESMem *)arg;
    h->tables[h->ahead] = les_frame(h, h->ahead);
    return NULL;
}

LESTable *LES_get_frame(const LESHandle i, const int n) {
    LESMem *h = (LESMem *)i;
    if (!h || n < 0 || (uint32_t)n >= h->frames) return NULL;
    les_join(h);
    if (!h->tables[n]) h->tables[n] = les_frame(h, (uint32_t)n);
    const uint32_t next = ((uint32_t)n + 1) % h->frames;
    if (h->file_mode) {
        /* file-backed: keep the frame handed out and the next one only (the
         * caller has copied the previous frame when it asks for a new one) */
        for (uint32_t f = 0; f < h->frames; f++) {
            if (f != (uint32_t)n && f != next && h->tables[f]) {
                les_free_table(h->tables[f]);
                h->tables[f] = NULL;
            }
        }
    }
    if (!h->tables[next]) {
        h->ahead = next;
        h->worker_on = pthread_create(&h->worker, NULL, les_ahead_run, h) == 0;
    }
    return h->tables[n];
}

float LES_get_table_period(const LESHandle i) { return i ? ((LESMem *)i)->tp : 0.0f; }
size_t LES_get_table_count(const LESHandle i) { return i ? ((LESMem *)i)->frames : 0; }

/* ----------------------------------------------------------------- ADM */

void ADM_compute_properties(ADMBase *phys) {
    /* adm.c:279-304: mass, Tachikawa number and inverse inertia lengths,
     * non-dimensionalised with v0 = 100 m/s and g = 9.8 m/s^2 */
    const float g = 9.8f, v0 = 100.0f, rho_air = 1.21f;
    phys->mass = phys->x * phys->y * phys->z * phys->rho;
    float ix = (phys->y * phys->y + phys->z * phys->z) * phys->mass / 12.0f;
    float iy = (phys->x * phys->x + phys->z * phys->z) * phys->mass / 12.0f;
    float iz = (phys->x * phys->x + phys->y * phys->y) * phys->mass / 12.0f;
    phys->Ta = rho_air * (phys->y * phys->z * v0 * v0) / (2.0f * phys->mass * g);
    phys->inv_inln_x = (phys->mass * phys->x) / (ix * g * g);
    phys->inv_inln_y = (phys->mass * phys->y) / (iy * g * g);
    phys->inv_inln_z = (phys->mass * phys->z) / (iz * g * g);
    phys->Ta *= g / (v0 * v0);
}

void ADM_dimension_set(ADMTable *T, const float x, const float y, const float z, const float r) {
    T->phys.x = x;
    T->phys.y = y;
    T->phys.z = z;
    T->phys.rho = r;
    ADM_compute_properties(&T->phys);
}

/* Drag along the body-frame flow direction u = (cos b, sin b cos a, sin b sin a)
 * with a shape factor C(b) = 1 + 0.3 cos 2b, small torques, plus N(0, noise^2). */
ADMTable *ADM_table_synthetic(const char *name, uint32_t nb, uint32_t na, float noise, uint64_t seed) {
    ADMTable *T = (ADMTable *)calloc(1, sizeof(ADMTable));
    if (!T) return NULL;
    T->nb = nb;
    T->na = na;
    T->nn = nb * na;
    snprintf(T->name, sizeof(T->name), "%s", name ? name : "synthetic");
    float **planes[8] = {&T->data.b, &T->data.a, &T->data.cdx, &T->data.cdy, &T->data.cdz, &T->data.cmx, &T->data.cmy, &T->data.cmz};
    for (int p = 0; p < 8; p++) {
        *planes[p] = (float *)malloc((size_t)T->nn * sizeof(float));
        if (!*planes[p]) {
            ADM_table_free(T);
            return NULL;
        }
    }
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x243F6A8885A308D3ull;
    for (uint32_t j = 0; j < na; j++) {
        double a = M_PI * (double)j / (double)(na - 1);
        for (uint32_t i = 0; i < nb; i++) {
            size_t k = (size_t)j * nb + i;
            double b = -M_PI + 2.0 * M_PI * (double)i / (double)(nb - 1);
            double C = 1.0 + 0.3 * cos(2.0 * b);
            T->data.b[k] = (float)b;
            T->data.a[k] = (float)a;
            T->data.cdx[k] = (float)(C * cos(b) + noise * rs_gauss(&s));
            T->data.cdy[k] = (float)(C * sin(b) * cos(a) + noise * rs_gauss(&s));
            T->data.cdz[k] = (float)(C * sin(b) * sin(a) + noise * rs_gauss(&s));
            T->data.cmx[k] = (float)(0.02 * sin(2.0 * b) * sin(a) + noise * rs_gauss(&s));
            T->data.cmy[k] = (float)(-0.05 * sin(2.0 * b) * cos(a) + noise * rs_gauss(&s));
            T->data.cmz[k] = (float)(0.03 * sin(b) * sin(2.0 * a) + noise * rs_gauss(&s));
        }
    }
    return T;
}

void ADM_table_free(ADMTable *T) {
    if (!T) return;
    free(T->data.b); free(T->data.a); free(T->data.cdx); free(T->data.cdy); free(T->data.cdz);
    free(T->data.cmx); free(T->data.cmy); free(T->data.cmz);
    free(T);
}

/* ----------------------------------------------------------------- RCS */

/* hh/vv/hv amplitudes over alpha in [-pi, pi] (x) and beta in [0, pi] (y):
 * hh = s (0.6 + 0.4 cos^2 b) e^{j 0.5 sin a}, vv = s (0.5 + 0.3 sin^2 b cos^2 a)
 * e^{-j 0.3 cos b}, hv = 0.1 s sin 2b sin a, plus N(0, (noise s)^2) per value. */
RCSTable *RCS_table_synthetic(const char *name, uint32_t na, uint32_t nb, float scale, float noise, uint64_t seed) {
    RCSTable *T = (RCSTable *)calloc(1, sizeof(RCSTable));
    if (!T) return NULL;
    T->na = na;
    T->nb = nb;
    T->nn = na * nb;
    T->lambda = 0.1f;
    snprintf(T->name, sizeof(T->name), "%s", name ? name : "synthetic");
    float **planes[8] = {&T->data.a, &T->data.b, &T->data.hh_real, &T->data.vv_real, &T->data.hv_real,
                         &T->data.hh_imag, &T->data.vv_imag, &T->data.hv_imag};
    for (int p = 0; p < 8; p++) {
        *planes[p] = (float *)malloc((size_t)T->nn * sizeof(float));
        if (!*planes[p]) {
            RCS_table_free(T);
            return NULL;
        }
    }
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x13198A2E03707344ull;
    const double sc = scale, ns = noise * scale;
    for (uint32_t j = 0; j < nb; j++) {
        double b = M_PI * (double)j / (double)(nb - 1);
        for (uint32_t i = 0; i < na; i++) {
            size_t k = (size_t)j * na + i;
            double a = -M_PI + 2.0 * M_PI * (double)i / (double)(na - 1);
            double hh = sc * (0.6 + 0.4 * cos(b) * cos(b)), ph = 0.5 * sin(a);
            double vv = sc * (0.5 + 0.3 * sin(b) * sin(b) * cos(a) * cos(a)), pv = -0.3 * cos(b);
            double hv = 0.1 * sc * sin(2.0 * b) * sin(a);
            T->data.a[k] = (float)a;
            T->data.b[k] = (float)b;
            T->data.hh_real[k] = (float)(hh * cos(ph) + ns * rs_gauss(&s));
            T->data.hh_imag[k] = (float)(hh * sin(ph) + ns * rs_gauss(&s));
            T->data.vv_real[k] = (float)(vv * cos(pv) + ns * rs_gauss(&s));
            T->data.vv_imag[k] = (float)(vv * sin(pv) + ns * rs_gauss(&s));
            T->data.hv_real[k] = (float)(hv + ns * rs_gauss(&s));
            T->data.hv_imag[k] = (float)(0.3 * hv + ns * rs_gauss(&s));
        }
    }
    return T;
}

void RCS_table_free(RCSTable *T) {
    if (!T) return;
    free(T->data.a); free(T->data.b); free(T->data.hh_real); free(T->data.vv_real); free(T->data.hv_real);
    free(T->data.hh_imag); free(T->data.vv_imag); free(T->data.hv_imag);
    free(T);
}

/* ------------------------------------------------------ ADM / RCS files */

/* <path> of adm.c:152-225 / rcs.c:181-246: two uint16 grid sizes, then six
 * float planes of nn values.  ADM: nb (beta, x) then na, planes cdx cdy cdz
 * cmx cmy cmz, axes beta in [-180, 180] and alpha in [0, 180] degrees, body
 * 0.002 x 0.04 x 0.04 m at 1120 kg m^-3 until ADM_dimension_set.  RCS: na
 * (alpha, x) then nb, planes hh vv hv real then imaginary, lambda 0.1 m. */
static int read_planes(const char *path, uint16_t dims[2], float **planes, int np) {
    FILE *fid = fopen(path, "rb");
    if (!fid) return 0;
    int ok = fread(dims, sizeof(uint16_t), 2, fid) == 2 && dims[0] > 1 && dims[1] > 1;
    const size_t nn = ok ? (size_t)dims[0] * dims[1] : 0;
    for (int p = 0; ok && p < np; p++) {
        planes[p] = (float *)malloc(nn * sizeof(float));
        ok = planes[p] && fread(planes[p], sizeof(float), nn, fid) == nn;
    }
    fclose(fid);
    return ok;
}

ADMTable *ADM_table_read(const char *path) {
    ADMTable *T = (ADMTable *)calloc(1, sizeof(ADMTable));
    if (!T) return NULL;
    uint16_t d[2];
    float *pl[6] = {0};
    if (!read_planes(path, d, pl, 6)) {
        for (int p = 0; p < 6; p++) free(pl[p]);
        free(T);
        rsprint("ERROR: unable to read ADM table %s.", path);
        return NULL;
    }
    T->nb = d[0];
    T->na = d[1];
    T->nn = T->nb * T->na;
    T->data.cdx = pl[0]; T->data.cdy = pl[1]; T->data.cdz = pl[2];
    T->data.cmx = pl[3]; T->data.cmy = pl[4]; T->data.cmz = pl[5];
    T->data.b = (float *)malloc(T->nb * sizeof(float));
    T->data.a = (float *)malloc(T->na * sizeof(float));
    if (!T->data.b || !T->data.a) {
        ADM_table_free(T);
        return NULL;
    }
    for (uint32_t i = 0; i < T->nb; i++) T->data.b[i] = (float)i / (float)(T->nb - 1) * 360.0f - 180.0f;
    for (uint32_t i = 0; i < T->na; i++) T->data.a[i] = (float)i / (float)(T->na - 1) * 180.0f;
    const char *base = strrchr(path, '/');
    snprintf(T->name, sizeof(T->name), "%s", base ? base + 1 : path);
    snprintf(T->path, sizeof(T->path), "%s", path);
    ADM_dimension_set(T, 0.002f, 0.040f, 0.040f, 1120.0f);
    return T;
}

RCSTable *RCS_table_read(const char *path) {
    RCSTable *T = (RCSTable *)calloc(1, sizeof(RCSTable));
    if (!T) return NULL;
    uint16_t d[2];
    float *pl[6] = {0};
    if (!read_planes(path, d, pl, 6)) {
        for (int p = 0; p < 6; p++) free(pl[p]);
        free(T);
        rsprint("ERROR: unable to read RCS table %s.", path);
        return NULL;
    }
    T->na = d[0];
    T->nb = d[1];
    T->nn = T->na * T->nb;
    T->lambda = 0.1f;
    T->data.hh_real = pl[0]; T->data.vv_real = pl[1]; T->data.hv_real = pl[2];
    T->data.hh_imag = pl[3]; T->data.vv_imag = pl[4]; T->data.hv_imag = pl[5];
    T->data.a = (float *)malloc(T->na * sizeof(float));
    T->data.b = (float *)malloc(T->nb * sizeof(float));
    if (!T->data.a || !T->data.b) {
        RCS_table_free(T);
        return NULL;
    }
    for (uint32_t i = 0; i < T->na; i++) T->data.a[i] = (float)i / (float)(T->na - 1) * 360.0f - 180.0f;
    for (uint32_t i = 0; i < T->nb; i++) T->data.b[i] = (float)i / (float)(T->nb - 1) * 180.0f;
    const char *base = strrchr(path, '/');
    snprintf(T->name, sizeof(T->name), "%s", base ? base + 1 : path);
    snprintf(T->path, sizeof(T->path), "%s", path);
    return T;
}

/* Handles of adm.h / rcs.h: a table folder found on the search list, tables
 * read from it on request and freed with the handle. */
#define TABLES_MAX 64
typedef struct {
    char dir[1024];
    int count;
    void *tables[TABLES_MAX];
} TableMem;

static TableMem *table_mem(const char *path, const char *sub, const char *probe) {
    char dirs[8][1024];
    const int n = table_search_list(path, sub, 0, dirs);
    TableMem *h = (TableMem *)calloc(1, sizeof(TableMem));
    if (!h) return NULL;
    if (!find_table_dir(dirs, n, probe, h->dir)) {
        free(h);
        rsprint("ERROR: Unable to find the %s data folder.", sub);
        return NULL;
    }
    return h;
}

ADMHandle ADM_init_with_path(const char *path) {
    return (ADMHandle)table_mem(path, "les", "adm/" ADMConfigSquarePlate ".adm");
}
ADMHandle ADM_init(void) { return ADM_init_with_path(NULL); }

ADMTable *ADM_get_table(const ADMHandle in, const ADMConfig config) {
    TableMem *h = (TableMem *)in;
    if (!h || !config || h->count >= TABLES_MAX) return NULL;
    char f[1200];
    snprintf(f, sizeof(f), "%s/adm/%s.adm", h->dir, config);
    ADMTable *T = ADM_table_read(f);
    if (T) h->tables[h->count++] = T;
    return T;
}

void ADM_free(ADMHandle in) {
    TableMem *h = (TableMem *)in;
    if (!h) return;
    for (int i = 0; i < h->count; i++) ADM_table_free((ADMTable *)h->tables[i]);
    free(h);
}

RCSHandle RCS_init_with_path(const char *path) {
    return (RCSHandle)table_mem(path, "rcs", "rcs/" RCSConfigLeaf ".rcs");
}
RCSHandle RCS_init(void) { return RCS_init_with_path(NULL); }

RCSTable *RCS_get_table(const RCSHandle in, const RCSConfig config) {
    TableMem *h = (TableMem *)in;
    if (!h || !config || h->count >= TABLES_MAX) return NULL;
    char f[1200];
    snprintf(f, sizeof(f), "%s/rcs/%s.rcs", h->dir, config);
    RCSTable *T = RCS_table_read(f);
    if (T) h->tables[h->count++] = T;
    return T;
}

void RCS_free(RCSHandle in) {
    TableMem *h = (TableMem *)in;
    if (!h) return;
    for (int i = 0; i < h->count; i++) RCS_table_free((RCSTable *)h->tables[i]);
    free(h);
}

/* ----------------------------------------------------------------- OBJ */

#define OBJ_MAX 64

typedef struct {
    int count;
    OBJTable obj_table[OBJ_MAX];
    char adm_dir[1024];     /* folder holding adm/<config>.adm, "" if none found */
    char rcs_dir[1024];     /* folder holding rcs/<config>.rcs */
} OBJMem;

/* The ADM and RCS folders are looked up as adm.c:56-111 and rcs.c:92-145 do
 * (probing adm/square_plate.adm and rcs/leaf.rcs); a type whose files are
 * found is read from them, any other uses the synthetic tables. */
OBJHandle OBJ_init_with_path(const char *path) {
    OBJMem *O = (OBJMem *)calloc(1, sizeof(OBJMem));
    if (!O) return NULL;
    char dirs[8][1024];
    int n = table_search_list(path, "les", 0, dirs);
    if (!find_table_dir(dirs, n, "adm/" ADMConfigSquarePlate ".adm", O->adm_dir)) O->adm_dir[0] = '\0';
    n = table_search_list(path, "rcs", 0, dirs);
    if (!find_table_dir(dirs, n, "rcs/" RCSConfigLeaf ".rcs", O->rcs_dir)) O->rcs_dir[0] = '\0';
    return (OBJHandle)O;
}

OBJHandle OBJ_init(void) { return OBJ_init_with_path(NULL); }

void OBJ_free(OBJHandle in) {
    OBJMem *O = (OBJMem *)in;
    if (!O) return;
    for (int i = 0; i < O->count; i++) {
        ADM_table_free(O->obj_table[i].adm_table);
        RCS_table_free(O->obj_table[i].rcs_table);
    }
    free(O);
}

/* Dimensions (m) and densities (kg m^-3) as obj.c:102-148. */
OBJTable *OBJ_get_table(const OBJHandle in, OBJConfig type) {
    OBJMem *O = (OBJMem *)in;
    if (!O || O->count >= OBJ_MAX) return NULL;
    OBJTable *t = &O->obj_table[O->count];
    const char *adm = ADMConfigSquarePlate, *rcs = RCSConfigPlate;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f, rho = 0.0f, scale = 0.05f;
    switch (type) {
        case OBJConfigWoodboard2x4: rcs = RCSConfigWoodBoard; dx = 0.0508f; dy = 0.3048f; dz = 0.1016f; rho = 500.0f; scale = 0.05f; break;
        case OBJConfigWoodboard4x8: rcs = RCSConfigWoodBoard; dx = 0.1016f; dy = 0.3048f; dz = 0.2032f; rho = 500.0f; scale = 0.08f; break;
        case OBJConfigLeaf: adm = ADMConfigModelPlate; rcs = RCSConfigLeaf; dx = 0.002f; dy = 0.04f; dz = 0.04f; rho = 1120.0f; scale = 0.01f; break;
        case OBJConfigLeafBig: adm = ADMConfigModelPlate; rcs = RCSConfigLeaf; dx = 0.001f; dy = 0.08f; dz = 0.06f; rho = 350.0f; scale = 0.02f; break;
        case OBJConfigMetalSheet: rcs = RCSConfigPlate; dx = 0.001f; dy = 1.0f; dz = 1.0f; rho = 7850.0f; scale = 0.1f; break;
        case OBJConfigBrick: rcs = RCSConfigBrick; dx = 0.065f; dy = 0.215f; dz = 0.1125f; rho = 2200.0f; scale = 0.05f; break;
        default: return NULL;
    }
    char f[1200];
    if (O->adm_dir[0]) {
        snprintf(f, sizeof(f), "%s/adm/%s.adm", O->adm_dir, adm);
        t->adm_table = ADM_table_read(f);
    } else {
        t->adm_table = ADM_table_synthetic(adm, 361, 181, 0.01f, 1000ull + (uint64_t)type);
    }
    if (O->rcs_dir[0]) {
        snprintf(f, sizeof(f), "%s/rcs/%s.rcs", O->rcs_dir, rcs);
        t->rcs_table = RCS_table_read(f);
    } else {
        t->rcs_table = RCS_table_synthetic(rcs, 361, 181, scale, 0.02f, 2000ull + (uint64_t)type);
    }
    if (!t->adm_table || !t->rcs_table) {
        ADM_table_free(t->adm_table);
        RCS_table_free(t->rcs_table);
        t->adm_table = NULL;
        t->rcs_table = NULL;
        return NULL;
    }
    ADM_dimension_set(t->adm_table, dx, dy, dz, rho);
    O->count++;
    return t;
}

OBJTable *OBJ_table_at(const OBJHandle in, int t) {
    OBJMem *O = (OBJMem *)in;
    if (!O || t < 0 || t >= O->count) return NULL;
    return &O->obj_table[t];
}

/* ----------------------------------------------------------------- POS */

POSPattern *POS_init(void) { return POS_init_with_string("P:5,-12:12:0.05/10,-12:12:0.05/15,-12:12:0.05"); }

void POS_free(POSPattern *p) { free(p); }

POSPattern *POS_init_with_string(const char *scan_string) {
    POSPattern *p = (POSPattern *)calloc(1, sizeof(POSPattern));
    if (!p) return NULL;
    POS_parse_from_string(p, scan_string);
    return p;
}

/* pos.c:34-56: emit the current position, then step (with repetitions). */
int POS_get_next_angles(POSPattern *scan) {
    int k = (int)scan->index;
    scan->az = scan->positions[k].az;
    scan->el = scan->positions[k].el;
    scan->positions[k].index++;
    if (scan->positions[k].index == scan->positions[k].count) {
        scan->positions[k].index = 0;
        scan->index++;
        if (scan->index == scan->count) scan->index = 0;
    }
    scan->tic++;
    return 0;
}

/* pos.c:58-184: "P:el,az0:az1:daz/...", "R:az,el0:el1:del/...", "D:az,el,count/...". */
int POS_parse_from_string(POSPattern *scan, const char *string) {
    if (!string || strlen(string) < 2 || string[1] != ':') {
        rsprint("ERROR: Expected : after the scan mode character.");
        return 1;
    }
    char buf[1024];
    snprintf(buf, sizeof(buf), "%s", string + 2);
    scan->mode = string[0];
    int i = 0, j = 0, k = 0;
    float f1, f2, f3, f4;
    char *save = NULL;
    char *token = strtok_r(buf, "/", &save);
    const char m = (char)(scan->mode | 0x20);
    if (m == 'p' || m == 'r') {
        while (token && j < POS_MAX_PATTERN_COUNT && i < POS_MAX_SWEEP_COUNT) {
            k = sscanf(token, "%f,%f:%f:%f", &f1, &f2, &f3, &f4);
            POSSweep *sw = &scan->sweeps[i];
            if (m == 'p') {
                sw->elStart = f1; sw->elEnd = f1; sw->elDelta = 0.0f;
                sw->azStart = f2; sw->azEnd = f3; sw->azDelta = f4;
            } else {
                sw->azStart = f1; sw->azEnd = f1; sw->azDelta = 0.0f;
                sw->elStart = f2; sw->elEnd = f3; sw->elDelta = f4;
            }
            do {
                if (m == 'p') {
                    scan->positions[j].el = f1;
                    scan->positions[j].az = f2;
                } else {
                    scan->positions[j].az = f1;
                    scan->positions[j].el = f2;
                }
                scan->positions[j].index = 0;
                scan->positions[j].count = 1;
                f2 += f4;
                j++;
            } while (fabs((m == 'p' ? scan->positions[j - 1].az : scan->positions[j - 1].el) - f3) > 0.1f * f4 &&
                     j < POS_MAX_PATTERN_COUNT);
            token = strtok_r(NULL, "/", &save);
            i++;
        }
        scan->sweepCount = (uint32_t)i;
        scan->sweepIndex = 0;
    } else if (m == 'd') {
        while (token && j < POS_MAX_PATTERN_COUNT) {
            k = sscanf(token, "%f,%f,%f", &f1, &f2, &f3);
            scan->positions[j].az = f1;
            scan->positions[j].el = f2;
            scan->positions[j].index = 0;
            scan->positions[j].count = (uint32_t)f3;
            token = strtok_r(NULL, "/", &save);
            j++;
        }
    }
    scan->count = (uint32_t)j;
    scan->index = 0;
    if (k > 0) {
        scan->el = scan->positions[0].el;
        scan->az = scan->positions[0].az;
    }
    return 0;
}

bool POS_is_ppi(const POSPattern *s) { return s->mode == 'p' || s->mode == 'P'; }
bool POS_is_rhi(const POSPattern *s) { return s->mode == 'r' || s->mode == 'R'; }
bool POS_is_dbs(const POSPattern *s) { return s->mode == 'd' || s->mode == 'D'; }
bool POS_is_empty(const POSPattern *s) { return !s->count; }
