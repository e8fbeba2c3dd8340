/*
 * les_solver.c - the solver program: the reference's main() flow
 * (les_combustion_solver.c:4016-4636) on the gfx950 engine's C ABI.
 *
 * What stays as in the reference, so bayesian_optimizer.py and analysis.py
 * read the results unchanged:
 *   - output files in the output directory (OUTPUT_DIR, :93):
 *     probe_data_flame_region_loc{n}.csv (header :3393, rows :3453, one row
 *       per probe line cell every step, PROBE_DATA_INTERVAL 1 :283),
 *     simulation_monitor.csv every 10 steps (:3874-3924),
 *     LES_point_data_phi%.2f_i%d_j%d_k%d.csv flushed at checkpoints (:3824-3850),
 *     Premixed_C3H8_Air_%04d.vtk every OUTPUT_INTERVAL steps (:3473-3668),
 *     checkpoint.bin + restart_info.txt + stats_backup.bin every
 *       CHECKPOINT_INTERVAL steps (:3671-3698, :3739-3757);
 *   - the restart rules: "-restart file.vtk" (:4144-4148) or auto-resume
 *     from restart_info.txt ("BINARY\n<step>" -> checkpoint.bin, or a VTK
 *     path + step) (:4150-4181), and the total step count rule
 *     total_steps = (int)(SIMULATION_TIME / DT_0), extended by 1000 steps
 *     when the resume step is past it (:4366-4376);
 *   - the per-step console line (:4598-4599) and Ctrl+C handling (:21-28).
 *
 * What changes:
 *   - run parameters are runtime values: --preset plus --set NAME=VALUE
 *     (les_params field names or the reference's #define names such as
 *     FLAME_ACCEL_STRENGTH) and --params FILE, which also reads the
 *     "#define NAME VALUE" lines the optimiser prints (print_c_defines,
 *     bayesian_optimizer.py:703-716), so a trial needs no recompilation;
 *   - checkpoint_aux.bin beside checkpoint.bin holds the state the reference
 *     leaves out (SEM/RNG state, DT history, step): a resumed run continues
 *     exactly; the final checkpoint records the real step (the reference
 *     writes step 0 there, :4623); a checkpoint.bin without a matching
 *     checkpoint_aux.bin (one the reference wrote, or a stale aux blob)
 *     resumes as the reference's own resume does: main()'s attribute map
 *     (:4123-4136), fresh SEM eddies (:4190), no DT history;
 *   - duct-mode runs (mode B) write the same files; their checkpoint.bin
 *     holds rho*u, rho*v, rho*w, phi, T and restart_info.txt starts "DUCT".
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <signal.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../include/les_gpu.h"

#define CHECKPOINT_INTERVAL 50      /* :18 */
#define OUTPUT_INTERVAL 500         /* :86 */
#define MONITOR_INTERVAL 10         /* :3877 */

static volatile sig_atomic_t stop_requested = 0;
static void handle_sigint(int sig)
{
    (void)sig;
    stop_requested = 1;
}

/* ---------------------------------------------------------------- options */
typedef struct {
    const char* name;      /* les_params field or the reference's #define */
    size_t off;
    char type;             /* 'd' double, 'i' int, 'u' uint64 */
} param_name;

#define PD(n, f) {n, offsetof(les_params, f), 'd'}
#define PI(n, f) {n, offsetof(les_params, f), 'i'}
static const param_name PARAMS[] = {
    PI("nx", nx), PI("ny", ny), PI("nz", nz), PD("lx", lx), PD("ly", ly), PD("lz", lz),
    PI("mode", mode), PI("integrator", integrator), PI("advection", advection), PI("precision", precision),
    PD("inlet_radius", inlet_radius), PD("INLET_RADIUS", inlet_radius),
    PD("inlet_velocity", inlet_velocity), PD("INLET_VELOCITY", inlet_velocity),
    PD("inlet_temperature", inlet_temperature), PD("INLET_TEMPERATURE", inlet_temperature),
    PD("ambient_temperature", ambient_temperature), PD("AMBIENT_TEMPERATURE", ambient_temperature),
    PD("turbulence_intensity", turbulence_intensity), PD("TURBULENCE_INTENSITY", turbulence_intensity),
    PI("num_vortices", num_vortices), PI("NUM_VORTICES", num_vortices),
    PD("turbulent_length_scale", turbulent_length_scale), PD("TURBULENT_LENGTH_SCALE", turbulent_length_scale),
    PD("inlet_noise_amp", inlet_noise_amp),
    PD("s_l", s_l), PD("S_L", s_l), PD("burnt_temperature", burnt_temperature),
    PD("BURNT_GAS_TEMPERATURE", burnt_temperature), PI("initial_flame_front_i", initial_flame_front_i),
    PD("equivalence_ratio", equivalence_ratio),
    PD("pr_t", pr_t), PD("PR_T", pr_t), PD("sc_t", sc_t), PD("SC_T", sc_t), PD("fractal_d3", fractal_d3),
    PD("filter_weight_center", filter_weight_center), PD("nu_unburnt_approx", nu_unburnt_approx),
    PD("NU_UNBURNT_AIR_APPROX", nu_unburnt_approx), PD("rho_unburnt_approx", rho_unburnt_approx),
    PD("rho_burnt_approx", rho_burnt_approx),
    PD("courant", courant), PD("COURANT_NUMBER", courant), PD("courant_diffusion", courant_diffusion),
    PD("flame_accel_strength", flame_accel_strength), PD("FLAME_ACCEL_STRENGTH", flame_accel_strength),
    PD("flame_accel_width", flame_accel_width), PD("FLAME_ACCEL_WIDTH", flame_accel_width),
    PD("flame_accel_center", flame_accel_center), PD("FLAME_ACCEL_CENTER", flame_accel_center),
    PD("cs_unburned", cs_unburned), PD("CS_UNBURNED", cs_unburned),
    PD("cs_burnt", cs_burnt), PD("CS_BURNT", cs_burnt),
    PD("cs_flame_boost", cs_flame_boost), PD("CS_FLAME_BOOST", cs_flame_boost),
    PD("tip_protection_radius", tip_protection_radius), PD("max_flame_force", max_flame_force),
    PD("flame_ramp_distance", flame_ramp_distance), PD("weight_x", weight_x),
    PD("relaxation_coeff", relaxation_coeff), PI("scheme_transition_step", scheme_transition_step),
    PD("rho_unburnt", rho_unburnt), PD("prandtl", prandtl), PD("flame_x0", flame_x0),
    PI("reinit_interval", reinit_interval), PI("reinit_iters", reinit_iters), PI("proj_iters", proj_iters),
    PI("stats_start_step", stats_start_step), PI("stats_per_cell", stats_per_cell),
    {"seed", offsetof(les_params, seed), 'u'},
    PD("fixed_dt", fixed_dt), PD("FIXED_DT", fixed_dt),
};

static int set_param(les_params* p, const char* name, const char* value)
{
    for (size_t a = 0; a < sizeof(PARAMS) / sizeof(PARAMS[0]); ++a) {
        if (strcmp(PARAMS[a].name, name) != 0) continue;
        char* end = NULL;
        errno = 0;
        if (PARAMS[a].type == 'd') {
            double v = strtod(value, &end);
            if (end == value || errno) return -1;
            *(double*)((char*)p + PARAMS[a].off) = v;
        } else if (PARAMS[a].type == 'i') {
            long v = strtol(value, &end, 10);
            if (end == value || errno) return -1;
            *(int*)((char*)p + PARAMS[a].off) = (int)v;
        } else {
            unsigned long long v = strtoull(value, &end, 10);
            if (end == value || errno) return -1;
            *(uint64_t*)((char*)p + PARAMS[a].off) = (uint64_t)v;
        }
        return 0;
    }
    return -2;
}

/* "NAME=VALUE", "NAME VALUE" or "#define NAME VALUE" lines; '#' comments */
static int read_params_file(les_params* p, const char* path)
{
    FILE* fp = fopen(path, "r");
    if (!fp) { fprintf(stderr, "Error: cannot open parameter file %s\n", path); return -1; }
    char line[512];
    int lineno = 0;
    while (fgets(line, sizeof(line), fp)) {
        ++lineno;
        char* s = line;
        while (*s == ' ' || *s == '\t') ++s;
        if (strncmp(s, "#define", 7) == 0) s += 7;
        else if (*s == '#' || *s == '\n' || *s == 0) continue;
        char name[128], value[128];
        for (char* q = s; *q; ++q) if (*q == '=') *q = ' ';
        if (sscanf(s, "%127s %127s", name, value) != 2) continue;
        int e = set_param(p, name, value);
        if (e == -2) fprintf(stderr, "Warning: %s:%d: unknown parameter %s (ignored)\n", path, lineno, name);
        else if (e) { fprintf(stderr, "Error: %s:%d: bad value for %s\n", path, lineno, name); fclose(fp); return -1; }
    }
    fclose(fp);
    return 0;
}

typedef struct {
    const char* out;          /* output directory, with a trailing '/' */
    const char* restart_vtk;
    int steps;                /* <0: SIMULATION_TIME rule */
    double sim_time;
    int checkpoint_interval, output_interval, monitor_interval;
    int synthetic_mixture;
    const char* chem;
    const char* tran;
    int quiet;
} options;

static void usage(void)
{
    fprintf(stderr,
            "usage: les_solver [-restart FILE.vtk] [--out DIR] [--preset reference|duct64|duct128|ensemble|large]\n"
            "                  [--set NAME=VALUE]... [--params FILE] [--steps N] [--sim-time S]\n"
            "                  [--checkpoint-interval N] [--output-interval N] [--chem chem.bin] [--tran tran.bin]\n"
            "                  [--synthetic-mixture] [--quiet]\n");
}

/* ---------------------------------------------------------------- files */
static void path_in(char* buf, size_t n, const options* o, const char* name)
{
    snprintf(buf, n, "%s%s", o->out, name);
}

static int write_all(const char* path, const void* a, size_t na, const void* b, size_t nb)
{
    char tmp[1100];
    snprintf(tmp, sizeof(tmp), "%s.tmp", path);
    FILE* fp = fopen(tmp, "wb");
    if (!fp) return -1;
    int ok = fwrite(a, 1, na, fp) == na && (nb == 0 || fwrite(b, 1, nb, fp) == nb);
    ok = (fclose(fp) == 0) && ok;
    if (!ok || rename(tmp, path) != 0) { remove(tmp); return -1; }
    return 0;
}

static double* read_file(const char* path, size_t* bytes)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) return NULL;
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    double* b = (n > 0) ? (double*)malloc((size_t)n) : NULL;
    if (b && fread(b, 1, (size_t)n, fp) != (size_t)n) { free(b); b = NULL; }
    fclose(fp);
    if (b) *bytes = (size_t)n;
    return b;
}

/* ---------------------------------------------------------------- run state */
typedef struct {
    les_params p;
    les_gpu* g;
    size_t n;                 /* cells */
    int nf;                   /* checkpoint fields: 11 (mode R) or 5 (duct) */
    double* f;                /* nf x n host fields */
    int npts;
    int pts[3 * LES_NPOINT_MAX];
    double* pbuf;             /* point samples pending a flush */
    int pbuf_steps, pbuf_cap;
    les_mixture mix;
} run;

static int duct(const run* r) { return r->p.mode == LES_MODE_DUCT; }

static int download_fields(run* r)
{
    double* ptr[LES_NFIELDS];
    for (int a = 0; a < r->nf; ++a) ptr[a] = r->f + (size_t)a * r->n;
    return les_gpu_download_state(r->g, ptr, r->nf);
}

static int upload_fields(run* r)
{
    const double* ptr[LES_NFIELDS];
    for (int a = 0; a < r->nf; ++a) ptr[a] = r->f + (size_t)a * r->n;
    return les_gpu_upload_state(r->g, ptr, r->nf);
}

/* save_quick_checkpoint (:3671-3698) + our aux sidecar */
static int save_checkpoint(run* r, const options* o, int t)
{
    char path[1024];
    if (download_fields(r) != LES_OK) return -1;
    path_in(path, sizeof(path), o, "checkpoint.bin");
    if (write_all(path, &t, sizeof(int), r->f, (size_t)r->nf * r->n * 8)) return -1;
    size_t ab = 0;
    les_gpu_export_aux(r->g, NULL, &ab);
    size_t extra = duct(r) ? 0 : 2 * r->n * 8;          /* mode R: nu, alpha (next DT) */
    char* aux = (char*)malloc(ab + extra);
    if (!aux) return -1;
    size_t cap = ab;
    if (les_gpu_export_aux(r->g, aux, &cap) != LES_OK) { free(aux); return -1; }
    if (!duct(r)) {
        les_gpu_download_field(r->g, "nu", (double*)(aux + ab));
        les_gpu_download_field(r->g, "alpha", (double*)(aux + ab + r->n * 8));
    }
    path_in(path, sizeof(path), o, "checkpoint_aux.bin");
    int e = write_all(path, aux, ab + extra, NULL, 0);
    free(aux);
    if (e) return -1;
    path_in(path, sizeof(path), o, "restart_info.txt");
    FILE* fp = fopen(path, "w");
    if (!fp) return -1;
    fprintf(fp, "%s\n%d\n", duct(r) ? "DUCT" : "BINARY", t);
    fclose(fp);
    return 0;
}

/* save_statistics_backup (:3739-3757): count, then u, v, w, uu, vv, ww, uv, uw, vw, G sums */
static int save_stats(run* r, const options* o)
{
    static const int order[10] = {0, 1, 2, 4, 5, 6, 7, 8, 9, 3};
    int count = les_gpu_stats_count(r->g);
    double* buf = (double*)malloc(10 * r->n * 8);
    if (!buf) return -1;
    for (int a = 0; a < 10; ++a) {
        char nm[8];
        snprintf(nm, sizeof(nm), "stat%d", order[a]);
        if (les_gpu_download_field(r->g, nm, buf + (size_t)a * r->n) != LES_OK)
            memset(buf + (size_t)a * r->n, 0, r->n * 8);    /* no per-cell accumulators in this run */
    }
    char path[1024];
    path_in(path, sizeof(path), o, "stats_backup.bin");
    int e = write_all(path, &count, sizeof(int), buf, 10 * r->n * 8);
    free(buf);
    return e;
}

static void load_stats(run* r, const options* o)
{
    static const int order[10] = {0, 1, 2, 4, 5, 6, 7, 8, 9, 3};
    char path[1024];
    path_in(path, sizeof(path), o, "stats_backup.bin");
    size_t bytes = 0;
    char* b = (char*)read_file(path, &bytes);
    if (!b || bytes != sizeof(int) + 10 * r->n * 8) {
        printf("No stats backup found. Resetting averages.\n");
        free(b);
        return;
    }
    int count;
    memcpy(&count, b, sizeof(int));
    const double* s = (const double*)(b + sizeof(int));
    for (int a = 0; a < 10; ++a) {
        char nm[8];
        snprintf(nm, sizeof(nm), "stat%d", order[a]);
        les_gpu_upload_field(r->g, nm, s + (size_t)a * r->n);
    }
    les_gpu_set_stats_count(r->g, count);
    printf("Statistics restored. Avg steps: %d\n", count);
    free(b);
}

/* load_quick_checkpoint (:3701-3737) + aux; returns the step or -1 */
static int load_checkpoint(run* r, const options* o)
{
    char path[1024];
    path_in(path, sizeof(path), o, "checkpoint.bin");
    size_t bytes = 0;
    char* b = (char*)read_file(path, &bytes);
    if (!b || bytes != sizeof(int) + (size_t)r->nf * r->n * 8) { free(b); return -1; }
    int step;
    memcpy(&step, b, sizeof(int));
    memcpy(r->f, b + sizeof(int), (size_t)r->nf * r->n * 8);
    free(b);
    if (upload_fields(r) != LES_OK) return -1;
    path_in(path, sizeof(path), o, "checkpoint_aux.bin");
    char* aux = (char*)read_file(path, &bytes);
    size_t ab = 0;
    les_gpu_export_aux(r->g, NULL, &ab);
    size_t extra = duct(r) ? 0 : 2 * r->n * 8;
    int aux_step = -1;
    if (aux && bytes == ab + extra && les_aux_step(aux, ab, &aux_step) == LES_OK && aux_step == step &&
        les_gpu_import_aux(r->g, aux, ab) == LES_OK) {
        if (!duct(r)) {
            les_gpu_upload_field(r->g, "nu", (const double*)(aux + ab));
            les_gpu_upload_field(r->g, "alpha", (const double*)(aux + ab + r->n * 8));
        }
    } else {
        /* a checkpoint without (or with a stale) aux blob, e.g. one the
         * reference wrote: resume as the reference does (:4184-4190) --
         * main()'s attribute map, fresh SEM eddies, no DT history */
        if (aux && aux_step >= 0 && aux_step != step)
            printf("checkpoint_aux.bin is from step %d, checkpoint.bin from step %d: ignoring the aux state.\n",
                   aux_step, step);
        printf("No matching checkpoint_aux.bin: SEM state and DT history start fresh.\n");
        if (!duct(r)) les_gpu_use_resume_attributes(r->g);
        les_gpu_init_sem(r->g);
    }
    free(aux);
    return step;
}

/* output_vtk (:3473-3668): ASCII structured grid, k-j-i order */
static void vtk_scalar(FILE* fp, const char* name, const double* a, const run* r, const char* fmt)
{
    fprintf(fp, "SCALARS %s float\nLOOKUP_TABLE default\n", name);
    const int nx = r->p.nx, ny = r->p.ny, nz = r->p.nz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double v = a ? a[((size_t)i * ny + j) * nz + k] : 0.0;
                fprintf(fp, fmt, isfinite(v) ? v : 0.0);
            }
}

static int output_vtk(run* r, const options* o, int t)
{
    char name[128], path[1024];
    snprintf(name, sizeof(name), "Premixed_C3H8_Air_%04d.vtk", t);
    path_in(path, sizeof(path), o, name);
    FILE* fp = fopen(path, "w");
    if (!fp) { fprintf(stderr, "Error: Could not open file %s\n", path); return -1; }
    const int nx = r->p.nx, ny = r->p.ny, nz = r->p.nz;
    const size_t n = r->n;
    const double dx = r->p.lx / nx, dy = r->p.ly / ny, dz = r->p.lz / nz;
    double* u = (double*)malloc(n * 8);
    double* v = (double*)malloc(n * 8);
    double* w = (double*)malloc(n * 8);
    double* T = (double*)malloc(n * 8);
    double* G = (double*)malloc(n * 8);
    double* rho = (double*)malloc(n * 8);
    double* tmp = (double*)malloc(n * 8);
    double* st[10];
    for (int a = 0; a < 10; ++a) st[a] = (double*)malloc(n * 8);
    if (download_fields(r) != LES_OK) { fclose(fp); return -1; }
    if (duct(r)) {
        les_gpu_download_field(r->g, "T", T);
        les_gpu_download_field(r->g, "phi", tmp);
        for (size_t q = 0; q < n; ++q) {
            rho[q] = r->p.rho_unburnt * r->p.inlet_temperature / T[q];
            u[q] = r->f[q] / rho[q]; v[q] = r->f[n + q] / rho[q]; w[q] = r->f[2 * n + q] / rho[q];
            G[q] = 0.5 * (1.0 + tanh(tmp[q] / (2.0 * dx)));
        }
    } else {
        memcpy(u, r->f, n * 8); memcpy(v, r->f + n, n * 8); memcpy(w, r->f + 2 * n, n * 8);
        memcpy(T, r->f + 3 * n, n * 8); memcpy(G, r->f + 4 * n, n * 8); memcpy(rho, r->f + 5 * n, n * 8);
    }
    fprintf(fp, "# vtk DataFile Version 3.0\nVTK output\nASCII\nDATASET STRUCTURED_GRID\n");
    fprintf(fp, "DIMENSIONS %d %d %d\n", nx, ny, nz);
    fprintf(fp, "POINTS %d float\n", nx * ny * nz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) fprintf(fp, "%.6f %.6f %.6f\n", i * dx, j * dy, k * dz);
    fprintf(fp, "\nPOINT_DATA %d\n", nx * ny * nz);
    fprintf(fp, "VECTORS velocity float\n");
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t q = ((size_t)i * ny + j) * nz + k;
                fprintf(fp, "%.6f %.6f %.6f\n", isfinite(u[q]) ? u[q] : 0.0, isfinite(v[q]) ? v[q] : 0.0,
                        isfinite(w[q]) ? w[q] : 0.0);
            }
    vtk_scalar(fp, "temperature", T, r, "%.6f\n");
    vtk_scalar(fp, "G_value", G, r, "%.6f\n");
    static const char* spn[5] = {"Y_C3H8", "Y_O2", "Y_N2", "Y_CO2", "Y_H2O"};
    for (int s = 0; s < 5; ++s) vtk_scalar(fp, spn[s], duct(r) ? NULL : r->f + (size_t)(6 + s) * n, r, "%.6f\n");
    vtk_scalar(fp, "density", rho, r, "%.6f\n");
    vtk_scalar(fp, "omega", NULL, r, "%.6e\n");           /* the reference never sets omega */
    const char* fields[5][2] = {{"mu_mix", "mu_mix"}, {"nu_mix", "nu"}, {"ST_turbulent", "ST"}, {"nu_t", "nu_t"},
                                {"grad_G_magnitude", "gradG"}};
    for (int a = 0; a < 5; ++a) {
        int ok = les_gpu_download_field(r->g, fields[a][1], tmp) == LES_OK;
        vtk_scalar(fp, fields[a][0], ok ? tmp : NULL, r, "%.6e\n");
    }
    /* calculate_resolved_reynolds_stresses (:3195-3219) from the FP64 accumulators */
    int cnt = les_gpu_stats_count(r->g);
    int have = 1;
    for (int a = 0; a < 10 && have; ++a) {
        char nm[8];
        snprintf(nm, sizeof(nm), "stat%d", a);
        have = les_gpu_download_field(r->g, nm, st[a]) == LES_OK;
    }
    const int pr[6][3] = {{4, 0, 0}, {5, 1, 1}, {6, 2, 2}, {7, 0, 1}, {8, 0, 2}, {9, 1, 2}};
    const char* rn[6] = {"re_stress_uu", "re_stress_vv", "re_stress_ww", "re_stress_uv", "re_stress_uw",
                         "re_stress_vw"};
    for (int a = 0; a < 6; ++a) {
        for (size_t q = 0; q < n; ++q)
            tmp[q] = (have && cnt > 0) ? st[pr[a][0]][q] / cnt - (st[pr[a][1]][q] / cnt) * (st[pr[a][2]][q] / cnt)
                                       : 0.0;
        vtk_scalar(fp, rn[a], tmp, r, "%.6e\n");
    }
    for (size_t q = 0; q < n; ++q) tmp[q] = (have && cnt > 0) ? st[3][q] / cnt : 0.0;
    vtk_scalar(fp, "G_mean", tmp, r, "%.6f\n");
    fclose(fp);
    free(u); free(v); free(w); free(T); free(G); free(rho); free(tmp);
    for (int a = 0; a < 10; ++a) free(st[a]);
    return 0;
}

/* load_vtk_for_restart (:3229-3368): velocity, temperature, G_value and the
 * species blocks (k-j-i order); density from the EOS inside the engine's
 * upload (mode R); mode B is not restartable from VTK (phi is not in it) */
static int load_vtk(run* r, const char* filename)
{
    if (duct(r)) { fprintf(stderr, "Error: duct runs resume from checkpoint.bin, not from VTK\n"); return 0; }
    FILE* fp = fopen(filename, "r");
    if (!fp) { fprintf(stderr, "Error: Cannot open restart file: %s\n", filename); return 0; }
    const int nx = r->p.nx, ny = r->p.ny, nz = r->p.nz;
    const size_t n = r->n;
    char line[256];
    int dx_, dy_, dz_, cnt = 0, ok = 0;
    while (fgets(line, sizeof(line), fp)) {
        if (strncmp(line, "DIMENSIONS", 10) == 0 && sscanf(line, "DIMENSIONS %d %d %d", &dx_, &dy_, &dz_) == 3 &&
            (dx_ != nx || dy_ != ny || dz_ != nz)) {
            fprintf(stderr, "FATAL ERROR: VTK dimensions (%d, %d, %d) do not match the run (%d, %d, %d).\n", dx_,
                    dy_, dz_, nx, ny, nz);
            fclose(fp);
            return 0;
        }
        if (strncmp(line, "POINT_DATA", 10) == 0) {
            if (sscanf(line, "POINT_DATA %d", &cnt) != 1 || cnt != nx * ny * nz) { fclose(fp); return 0; }
            ok = 1;
            break;
        }
    }
    if (!ok) { fclose(fp); return 0; }
    static const char* sc[7] = {"SCALARS temperature", "SCALARS G_value", "SCALARS Y_C3H8", "SCALARS Y_O2",
                                "SCALARS Y_N2", "SCALARS Y_CO2", "SCALARS Y_H2O"};
    static const int dst[7] = {3, 4, 6, 7, 8, 9, 10};
    while (fgets(line, sizeof(line), fp)) {
        if (strncmp(line, "VECTORS velocity", 16) == 0) {
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        size_t q = ((size_t)i * ny + j) * nz + k;
                        if (fscanf(fp, "%lf %lf %lf", &r->f[q], &r->f[n + q], &r->f[2 * n + q]) != 3) {
                            fclose(fp);
                            return 0;
                        }
                    }
            continue;
        }
        for (int a = 0; a < 7; ++a) {
            if (strncmp(line, sc[a], strlen(sc[a])) != 0) continue;
            if (!fgets(line, sizeof(line), fp)) { fclose(fp); return 0; }
            double* f = r->f + (size_t)dst[a] * n;
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i)
                        if (fscanf(fp, "%lf", &f[((size_t)i * ny + j) * nz + k]) != 1) { fclose(fp); return 0; }
        }
    }
    fclose(fp);
    /* update_rho_from_EOS (:1883-1910): rho = p0 / (R_univ sum(Y/W) T) where
     * sum(Y/W) > 1e-9 and T > 1 K; the reference applies it to attributes 2
     * and 3 and leaves the other cells' (uninitialised) rho; here every cell */
    const double* T = r->f + 3 * n;
    for (size_t q = 0; q < n; ++q) {
        double s = 0.0;
        for (int k = 0; k < 5; ++k) s += r->f[(size_t)(6 + k) * n + q] / (r->mix.sp[k].mw / 1000.0);
        if (s > 1e-9 && T[q] > 1.0) r->f[5 * n + q] = 101325.0 / ((8.314462 / (1.0 / s)) * T[q]);
        else r->f[5 * n + q] = 0.0;
    }
    return upload_fields(r) == LES_OK;
}

/* flush_point_data (:3824-3850) */
static void flush_points(run* r, const options* o)
{
    for (int p = 0; p < r->npts && r->pbuf_steps > 0; ++p) {
        char name[256], path[1024];
        snprintf(name, sizeof(name), "LES_point_data_phi%.2f_i%d_j%d_k%d.csv", r->p.equivalence_ratio,
                 r->pts[3 * p], r->pts[3 * p + 1], r->pts[3 * p + 2]);
        path_in(path, sizeof(path), o, name);
        FILE* fp = fopen(path, "a");
        if (!fp) continue;
        fseek(fp, 0, SEEK_END);
        if (ftell(fp) == 0) fprintf(fp, "u,v,w,G\n");
        for (int s = 0; s < r->pbuf_steps; ++s) {
            const double* q = r->pbuf + ((size_t)s * r->npts + p) * 4;
            fprintf(fp, "%.6f,%.6f,%.6f,%.6f\n", q[0], q[1], q[2], q[3]);
        }
        fclose(fp);
    }
    r->pbuf_steps = 0;
}

static void collect_points(run* r)
{
    int ns = 0, np = 0;
    int room = r->pbuf_cap - r->pbuf_steps;
    if (room <= 0 || r->npts == 0) return;
    les_gpu_point_fetch(r->g, r->pbuf + (size_t)r->pbuf_steps * r->npts * 4, room, &ns, &np);
    r->pbuf_steps += ns;
}

/* initialize_probe_file (:3373-3400) / record_flame_region_data rows (:3453) */
static void init_probe_files(const run* r, const options* o)
{
    for (int n = 0; n < r->p.num_probe_x; ++n) {
        char name[128], path[1024];
        snprintf(name, sizeof(name), "probe_data_flame_region_loc%d.csv", n);
        path_in(path, sizeof(path), o, name);
        remove(path);
        FILE* fp = fopen(path, "w");
        if (!fp) { fprintf(stderr, "Warning: Could not create probe file: %s\n", path); continue; }
        fprintf(fp, "timestep,x_location,y_flame,y_physical,y_relative,u,v,w,G,nu_t,S_mag\n");
        fclose(fp);
    }
}

static void write_probe_rows(run* r, const options* o)
{
    static les_probe_row rows[1 << 14];
    FILE* fps[LES_NPROBE_MAX] = {0};
    int n;
    do {
        n = 0;
        if (les_gpu_probe_fetch(r->g, rows, (int)(sizeof(rows) / sizeof(rows[0])), &n) != LES_OK) break;
        for (int a = 0; a < n; ++a) {
            const les_probe_row* x = &rows[a];
            if (x->station < 0 || x->station >= LES_NPROBE_MAX) continue;
            if (!fps[x->station]) {
                char name[128], path[1024];
                snprintf(name, sizeof(name), "probe_data_flame_region_loc%d.csv", x->station);
                path_in(path, sizeof(path), o, name);
                fps[x->station] = fopen(path, "a");
                if (!fps[x->station]) continue;
            }
            fprintf(fps[x->station], "%d,%.3f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6e,%.6e\n", x->timestep,
                    x->x_location, x->y_flame, x->y_physical, x->y_relative, x->u, x->v, x->w, x->G, x->nu_t,
                    x->S_mag);
        }
    } while (n == (int)(sizeof(rows) / sizeof(rows[0])));
    for (int s = 0; s < LES_NPROBE_MAX; ++s) if (fps[s]) fclose(fps[s]);
}

/* write_monitor_log (:3874-3924) */
static void write_monitor(const run* r, const options* o, int step, const les_step_report* rep)
{
    char path[1024];
    path_in(path, sizeof(path), o, "simulation_monitor.csv");
    FILE* fp = fopen(path, "a");
    if (!fp) { fprintf(stderr, "Warning: Could not open monitor log at %s\n", path); return; }
    fseek(fp, 0, SEEK_END);
    if (ftell(fp) == 0) fprintf(fp, "step,dt,total_mass,min_T,max_T,min_G,max_G,min_Y_C3H8\n");
    fprintf(fp, "%d,%.6e,%.6e,%.2f,%.2f,%.4f,%.4f,%.2e\n", step, rep->dt, rep->total_mass, rep->min_T, rep->max_T,
            rep->min_G, rep->max_G, rep->min_Y_fuel);
    fclose(fp);
    if (rep->min_T < 200.0 || rep->max_T > 3000.0 || rep->min_Y_fuel < -1e-5)
        printf(" [WARNING] Unstable values detected at step %d! T_range=[%.1f, %.1f], Min_Fuel=%.2e\n", step,
               rep->min_T, rep->max_T, rep->min_Y_fuel);
    (void)r;
}

/* initial DT (:4366-4370): C * DX / max |u| (calculate_max_velocity
 * :1472-1485: attributes 1 and 3 in mode R; planes 0..nx-2 in duct mode) */
static double initial_dt(run* r)
{
    if (download_fields(r) != LES_OK) return -1.0;
    const size_t n = r->n;
    double* attr = NULL;
    if (!duct(r)) {
        attr = (double*)malloc(n * 8);
        if (!attr || les_gpu_download_field(r->g, "attribute", attr) != LES_OK) { free(attr); return -1.0; }
    }
    double mv = 0.0;
    const double dx = r->p.lx / r->p.nx;
    for (size_t q = 0; q < n; ++q) {
        double u, v, w;
        if (duct(r)) {
            double rho = r->p.rho_unburnt * r->p.inlet_temperature / r->f[4 * n + q];
            u = r->f[q] / rho; v = r->f[n + q] / rho; w = r->f[2 * n + q] / rho;
        } else {
            u = r->f[q]; v = r->f[n + q]; w = r->f[2 * n + q];
        }
        if (attr && attr[q] != 1.0 && attr[q] != 3.0) continue;
        if (!attr && q / ((size_t)r->p.ny * r->p.nz) == (size_t)r->p.nx - 1) continue;
        double s = sqrt(u * u + v * v + w * w);
        if (s > mv) mv = s;
    }
    free(attr);
    return r->p.courant * dx / (mv + 1e-10);
}

int main(int argc, char* argv[])
{
    options o;
    memset(&o, 0, sizeof(o));
    o.out = "./";
    o.steps = -1;
    o.sim_time = 20.0;                     /* SIMULATION_TIME (:81) */
    o.checkpoint_interval = CHECKPOINT_INTERVAL;
    o.output_interval = OUTPUT_INTERVAL;
    o.monitor_interval = MONITOR_INTERVAL;
    o.chem = "chem.bin";
    o.tran = "tran.bin";
    les_params p;
    les_params_default(&p, LES_PRESET_REFERENCE);
    static const char* presets[5] = {"reference", "duct64", "duct128", "ensemble", "large"};
    char outbuf[1024];
    /* the preset comes first so --set/--params override it wherever they appear */
    for (int a = 1; a + 1 < argc; ++a)
        if (strcmp(argv[a], "--preset") == 0) {
            int found = -1;
            for (int k = 0; k < 5; ++k) if (strcmp(argv[a + 1], presets[k]) == 0) found = k;
            if (found < 0) { usage(); return 2; }
            les_params_default(&p, found);
        }
    for (int a = 1; a < argc; ++a) {
        const char* s = argv[a];
        const char* nxt = a + 1 < argc ? argv[a + 1] : NULL;
        if (strcmp(s, "-restart") == 0 && nxt) { o.restart_vtk = nxt; ++a; }
        else if (strcmp(s, "--preset") == 0 && nxt) ++a;
        else if (strcmp(s, "--out") == 0 && nxt) {
            size_t L = strlen(nxt);
            snprintf(outbuf, sizeof(outbuf), "%s%s", nxt, (L && nxt[L - 1] == '/') ? "" : "/");
            o.out = outbuf;
            mkdir(nxt, 0755);
            ++a;
        } else if (strcmp(s, "--set") == 0 && nxt) {
            char name[128];
            const char* eq = strchr(nxt, '=');
            if (!eq || (size_t)(eq - nxt) >= sizeof(name)) { usage(); return 2; }
            memcpy(name, nxt, (size_t)(eq - nxt));
            name[eq - nxt] = 0;
            int e = set_param(&p, name, eq + 1);
            if (e) { fprintf(stderr, "Error: %s parameter %s\n", e == -2 ? "unknown" : "bad value for", name); return 2; }
            ++a;
        } else if (strcmp(s, "--params") == 0 && nxt) {
            if (read_params_file(&p, nxt)) return 2;
            ++a;
        } else if (strcmp(s, "--steps") == 0 && nxt) { o.steps = atoi(nxt); ++a; }
        else if (strcmp(s, "--sim-time") == 0 && nxt) { o.sim_time = atof(nxt); ++a; }
        else if (strcmp(s, "--checkpoint-interval") == 0 && nxt) { o.checkpoint_interval = atoi(nxt); ++a; }
        else if (strcmp(s, "--output-interval") == 0 && nxt) { o.output_interval = atoi(nxt); ++a; }
        else if (strcmp(s, "--chem") == 0 && nxt) { o.chem = nxt; ++a; }
        else if (strcmp(s, "--tran") == 0 && nxt) { o.tran = nxt; ++a; }
        else if (strcmp(s, "--synthetic-mixture") == 0) o.synthetic_mixture = 1;
        else if (strcmp(s, "--quiet") == 0) o.quiet = 1;
        else { usage(); return 2; }
    }
    if (les_params_validate(&p) != LES_OK) { fprintf(stderr, "Error: invalid run parameters\n"); return 2; }
    if (o.checkpoint_interval < 1) o.checkpoint_interval = CHECKPOINT_INTERVAL;
    if (o.output_interval < 1) o.output_interval = OUTPUT_INTERVAL;

    run r;
    memset(&r, 0, sizeof(r));
    r.p = p;
    r.n = (size_t)p.nx * p.ny * p.nz;
    r.nf = duct(&r) ? 5 : LES_NFIELDS;
    r.f = (double*)malloc((size_t)r.nf * r.n * 8);
    printf("Allocating memory for grid size %dx%dx%d (Approx %.2f GB)...\n", p.nx, p.ny, p.nz,
           (double)sizeof(double) * r.n * r.nf / 1024 / 1024 / 1024);
    if (!r.f) { fprintf(stderr, "Memory allocation failed for the host fields\n"); return 1; }
    signal(SIGINT, handle_sigint);
    clock_t start_time = clock();

    /* property data (:4093-4098): chem.bin / tran.bin, as the reference */
    les_mixture* mixp = &r.mix;
    if (o.synthetic_mixture || duct(&r)) les_mixture_synthetic(mixp);
    else if (les_mixture_load(mixp, o.chem, o.tran) != LES_OK) {
        fprintf(stderr, "Error: cannot read %s / %s (use --synthetic-mixture for the built-in set)\n", o.chem, o.tran);
        return 1;
    }
    if (les_gpu_create(&r.p, mixp, &r.g) != LES_OK) return 1;
    les_gpu_get_points(r.g, r.pts, LES_NPOINT_MAX, &r.npts);
    printf("Measurement points setup: %d points.\n", r.npts);
    r.pbuf_cap = 4096;
    r.pbuf = (double*)malloc((size_t)r.pbuf_cap * (r.npts ? r.npts : 1) * 4 * 8);

    /* restart detection (:4139-4181) */
    int start_step = 0, restart_mode = 0;
    char restart_file[1024] = "";
    if (o.restart_vtk) {
        restart_mode = 1;
        snprintf(restart_file, sizeof(restart_file), "%s", o.restart_vtk);
        const char* us = strrchr(restart_file, '_');
        if (us) start_step = atoi(us + 1);
    } else {
        char path[1024], line[1024];
        path_in(path, sizeof(path), &o, "restart_info.txt");
        FILE* fp = fopen(path, "r");
        if (fp) {
            int saved = 0;
            if (fgets(line, sizeof(line), fp)) {
                line[strcspn(line, "\n")] = 0;
                if (strncmp(line, "BINARY", 6) == 0 || strncmp(line, "DUCT", 4) == 0) {
                    if (fgets(line, sizeof(line), fp) && sscanf(line, "%d", &saved) == 1) {
                        start_step = saved;
                        restart_mode = 2;
                        printf("\n!!! AUTO-RESUME DETECTED (From Checkpoint Step %d) !!!\n", saved);
                    }
                } else {
                    snprintf(restart_file, sizeof(restart_file), "%s", line);
                    if (fgets(line, sizeof(line), fp) && sscanf(line, "%d", &saved) == 1) {
                        start_step = saved;
                        restart_mode = 1;
                        printf("\n!!! AUTO-RESUME DETECTED (From VTK Step %d) !!!\n", saved);
                    }
                }
            }
            fclose(fp);
        }
    }
    if (restart_mode == 2) {
        int s = load_checkpoint(&r, &o);
        if (s < 0) {
            fprintf(stderr, "Binary resume failed. Starting from scratch.\n");
            restart_mode = 0;
            start_step = 0;
        } else {
            start_step = s;
            load_stats(&r, &o);
            les_gpu_set_step(r.g, start_step);
        }
    } else if (restart_mode == 1) {
        printf("Resuming from VTK: %s\n", restart_file);
        if (!load_vtk(&r, restart_file)) {
            fprintf(stderr, "VTK resume failed. Starting from scratch.\n");
            restart_mode = 0;
            start_step = 0;
        } else {
            load_stats(&r, &o);
            les_gpu_set_step(r.g, start_step);
            if (!duct(&r)) les_gpu_use_resume_attributes(r.g);   /* setting() is skipped (:4269-4280) */
            les_gpu_init_sem(r.g);                        /* no SEM state in a VTK: fresh eddies */
        }
    }
    if (restart_mode == 0) {
        printf("Starting NEW simulation...\n");
        if (les_gpu_init_state(r.g) != LES_OK) { fprintf(stderr, "%s\n", les_gpu_last_error(r.g)); return 1; }
        init_probe_files(&r, &o);
        output_vtk(&r, &o, 0);
    }

    /* total steps (:4366-4376) */
    int total_steps;
    if (o.steps >= 0) total_steps = start_step + o.steps;
    else {
        double dt0 = initial_dt(&r);
        total_steps = dt0 > 0 ? (int)(o.sim_time / dt0) : 0;
        if (start_step >= total_steps) {
            printf("Simulation already reached target time. Extending by 1000 steps...\n");
            total_steps = start_step + 1000;
        }
    }
    printf("\n--- Simulation Loop Start ---\nFrom Step %d to %d\n", start_step, total_steps);

    les_step_report rep;
    memset(&rep, 0, sizeof(rep));
    int t = start_step;
    int status = 0;
    while (t < total_steps && !stop_requested) {
        /* run up to the next step that writes something: the monitor every
         * monitor_interval steps, checkpoints, VTK */
        int last = t;
        while (last + 1 < total_steps && (last % o.monitor_interval) != 0 &&
               !(last > 0 && (last % o.checkpoint_interval == 0 || last % o.output_interval == 0)))
            ++last;
        int n = last - t + 1;
        int e = les_gpu_step(r.g, n, &rep);
        if (e != LES_OK) {
            fprintf(stderr, "Error at step %d: %s\n", last, les_gpu_last_error(r.g));
            status = 1;
            break;
        }
        write_probe_rows(&r, &o);
        collect_points(&r);
        if (r.pbuf_steps >= r.pbuf_cap) flush_points(&r, &o);
        if (last > 0 && last % o.checkpoint_interval == 0) {
            printf("Checkpointing at step %d... ", last);
            save_checkpoint(&r, &o, last + 1);
            save_stats(&r, &o);
            flush_points(&r, &o);
            printf("Done.\n");
        }
        if (last % o.monitor_interval == 0) write_monitor(&r, &o, last, &rep);
        if (last > 0 && last % o.output_interval == 0) output_vtk(&r, &o, last);
        if (!o.quiet)
            printf("Step: %d/%d (%.1f%%), DT: %.2e, MaxVel: %.2f, AvgSteps: %d\n", last, total_steps,
                   (last * 100.0 / total_steps), rep.dt, rep.max_velocity, les_gpu_stats_count(r.g));
        t = last + 1;
    }

    printf("\nFinalizing output...\n");
    if (!stop_requested && status == 0) output_vtk(&r, &o, total_steps);
    save_checkpoint(&r, &o, t);
    save_stats(&r, &o);
    flush_points(&r, &o);
    double elapsed = (double)(clock() - start_time) / CLOCKS_PER_SEC;
    printf("Simulation finished. Total time: %.2f seconds\n", elapsed);
    les_gpu_destroy(r.g);
    free(r.f);
    free(r.pbuf);
    return status;
}
