/*
 * ref_harness.c - driver for the reference solver compiled from its own
 * source (les_combustion_solver.c, left where it lies under the reference
 * tree).  TEST INFRASTRUCTURE ONLY: it produces golden vectors that pin the
 * CPU oracle, and CPU baseline timings.
 *
 * The reference translation unit is compiled with three renames
 * (see oracle/Makefile):
 *   -Dmain=les_reference_main   the harness owns main() and calls it
 *   -Dprintf=les_ref_printf     every printf of the solver goes through the
 *                               hook below; the per-step progress line
 *                               ("Step: ...", :4601) marks the end of a step
 *   -Dsrand=les_ref_srand       srand(time(NULL)) (:4085) becomes a fixed seed
 * Nothing else of the reference is changed.  At the end of step t the hook
 * dumps the solver's global arrays (they are plain extern symbols of the
 * reference TU) and, after the requested number of steps, raises the
 * reference's own stop flag (stop_requested, :21) so that its loop breaks at
 * the next step and runs its final save (:4605-4625).
 *
 * Environment:
 *   LES_REF_SEED        seed handed to srand (default 0)
 *   LES_REF_STEPS       number of steps to run (default 3)
 *   LES_REF_DUMP        comma-separated step indices to dump (default all)
 *   LES_REF_OUT         output directory for dumps (default ".")
 *   LES_REF_RHS_IN      directory with u,v,w,T,G,rho,Y0..Y4 .bin inputs: after
 *                       the last step, calculate_RHS is evaluated on them
 *   LES_REF_RHS_T       t_loop passed to that calculate_RHS call (default 600)
 *   LES_REF_VERBOSE     1 = forward the solver's printf output
 *   LES_REF_TIME        1 = print wall time per step (JSON line at the end)
 */
#define _GNU_SOURCE
#include <signal.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "../include/les_gpu.h"

#define REF_NX 60
#define REF_NY 20
#define REF_NZ 40
#define REF_NSP 5
typedef double (*ref_grid)[REF_NY][REF_NZ];
typedef int (*ref_igrid)[REF_NY][REF_NZ];

/* globals of the reference TU (les_combustion_solver.c:458-612) */
extern ref_grid u, v, w, T, G, rho, P;
extern ref_grid Yk[REF_NSP];
extern ref_grid u_filtered, v_filtered, w_filtered, T_filtered, G_filtered, rho_filtered, P_filtered;
extern ref_grid Yk_filtered[REF_NSP];
extern ref_grid nu, mu_mix, alpha, nu_t, S_mag, SijSij, div_u, ST_turbulent, Cp_mix, lambda_mix;
extern ref_grid Dk[REF_NSP];
extern ref_grid grad_G_magnitude_for_debug;
extern ref_grid k1_rhou, k1_rhov, k1_rhow, k1_T, k1_G;
extern ref_grid k1_Yk[REF_NSP];
extern ref_grid rho_temp;
extern ref_igrid attribute;
extern double u_inlet_fluct[REF_NY][REF_NZ], v_inlet_fluct[REF_NY][REF_NZ], w_inlet_fluct[REF_NY][REF_NZ];
typedef struct { double x, y, z, radius, su, sv, sw; } ref_vortex;
extern ref_vortex vortices[2000];
extern volatile sig_atomic_t stop_requested;

void calculate_RHS(int t_loop, const ref_grid u_in, const ref_grid v_in, const ref_grid w_in,
                   const ref_grid T_in, const ref_grid G_in, const ref_grid* Yk_in, const ref_grid rho_in,
                   ref_grid drhou_dt, ref_grid drhov_dt, ref_grid drhow_dt, ref_grid dT_dt, ref_grid dG_dt,
                   ref_grid* dYk_dt);

int les_reference_main(int argc, char* argv[]);
void record_flame_region_data(int t_loop, const ref_grid u, const ref_grid v, const ref_grid w, const ref_grid G,
                              const ref_grid nu_t, const ref_grid S_mag);

static unsigned g_seed = 0;
static int g_steps = 3;
static int g_done = 0;
static int g_verbose = 0;
static int g_time = 0;
static char g_out[1024] = ".";
static char g_dump[4096] = "";
static double g_t0 = 0.0, g_tlast = 0.0;
static double g_step_times[100000];

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void les_ref_srand(unsigned int seed_from_solver)
{
    (void)seed_from_solver;
    srand(g_seed);
}

static void dump(const char* name, int step, const void* data, size_t bytes)
{
    char path[1400];
    if (step >= 0) snprintf(path, sizeof(path), "%s/%s_s%d.bin", g_out, name, step);
    else snprintf(path, sizeof(path), "%s/%s.bin", g_out, name);
    FILE* fp = fopen(path, "wb");
    if (!fp) { fprintf(stderr, "harness: cannot write %s\n", path); exit(2); }
    fwrite(data, 1, bytes, fp);
    fclose(fp);
}

static void load(const char* dir, const char* name, void* data, size_t bytes)
{
    char path[1400];
    snprintf(path, sizeof(path), "%s/%s.bin", dir, name);
    FILE* fp = fopen(path, "rb");
    if (!fp || fread(data, 1, bytes, fp) != bytes) { fprintf(stderr, "harness: cannot read %s\n", path); exit(2); }
    fclose(fp);
}

static int want_dump(int step)
{
    if (!g_dump[0]) return 1;
    char buf[4096];
    snprintf(buf, sizeof(buf), "%s", g_dump);
    for (char* t = strtok(buf, ","); t; t = strtok(NULL, ","))
        if (atoi(t) == step) return 1;
    return 0;
}

static const size_t NB = sizeof(double) * REF_NX * REF_NY * REF_NZ;

static void dump_state(int step)
{
    dump("u", step, u, NB); dump("v", step, v, NB); dump("w", step, w, NB);
    dump("T", step, T, NB); dump("G", step, G, NB); dump("rho", step, rho, NB);
    char nm[16];
    for (int s = 0; s < REF_NSP; ++s) { snprintf(nm, sizeof nm, "Y%d", s); dump(nm, step, Yk[s], NB); }
    dump("nu_t", step, nu_t, NB); dump("S_mag", step, S_mag, NB);
    dump("nu", step, nu, NB); dump("alpha", step, alpha, NB);
    dump("rho_temp", step, rho_temp, NB);
    dump("k1_rhou", step, k1_rhou, NB); dump("k1_rhov", step, k1_rhov, NB); dump("k1_rhow", step, k1_rhow, NB);
    dump("k1_T", step, k1_T, NB); dump("k1_G", step, k1_G, NB);
    for (int s = 0; s < REF_NSP; ++s) { snprintf(nm, sizeof nm, "k1_Y%d", s); dump(nm, step, k1_Yk[s], NB); }
    dump("u_inlet", step, u_inlet_fluct, sizeof(u_inlet_fluct));
    dump("v_inlet", step, v_inlet_fluct, sizeof(v_inlet_fluct));
    dump("w_inlet", step, w_inlet_fluct, sizeof(w_inlet_fluct));
    dump("vortices", step, vortices, sizeof(vortices));
}

static void rhs_on_inputs(const char* dir, int t_rhs)
{
    load(dir, "u", u, NB); load(dir, "v", v, NB); load(dir, "w", w, NB);
    load(dir, "T", T, NB); load(dir, "G", G, NB); load(dir, "rho", rho, NB);
    char nm[16];
    for (int s = 0; s < REF_NSP; ++s) { snprintf(nm, sizeof nm, "Y%d", s); load(dir, nm, Yk[s], NB); }
    calculate_RHS(t_rhs, u, v, w, T, G, Yk, rho, k1_rhou, k1_rhov, k1_rhow, k1_T, k1_G, k1_Yk);
    dump("rhs_rhou", -1, k1_rhou, NB); dump("rhs_rhov", -1, k1_rhov, NB); dump("rhs_rhow", -1, k1_rhow, NB);
    dump("rhs_T", -1, k1_T, NB); dump("rhs_G", -1, k1_G, NB);
    for (int s = 0; s < REF_NSP; ++s) { snprintf(nm, sizeof nm, "rhs_Y%d", s); dump(nm, -1, k1_Yk[s], NB); }
    dump("rhs_u_f", -1, u_filtered, NB); dump("rhs_v_f", -1, v_filtered, NB); dump("rhs_w_f", -1, w_filtered, NB);
    dump("rhs_T_f", -1, T_filtered, NB); dump("rhs_G_f", -1, G_filtered, NB); dump("rhs_rho_f", -1, rho_filtered, NB);
    dump("rhs_nu", -1, nu, NB); dump("rhs_mu_mix", -1, mu_mix, NB); dump("rhs_alpha", -1, alpha, NB);
    dump("rhs_nu_t", -1, nu_t, NB); dump("rhs_S_mag", -1, S_mag, NB); dump("rhs_ST", -1, ST_turbulent, NB);
    dump("rhs_cp_mix", -1, Cp_mix, NB); dump("rhs_lambda_mix", -1, lambda_mix, NB);
    dump("rhs_gradG", -1, grad_G_magnitude_for_debug, NB);
    for (int s = 0; s < REF_NSP; ++s) { snprintf(nm, sizeof nm, "rhs_D%d", s); dump(nm, -1, Dk[s], NB); }
    int attr_d[REF_NX * REF_NY * REF_NZ];
    memcpy(attr_d, attribute, sizeof(attr_d));
    dump("attribute", -1, attr_d, sizeof(attr_d));
    /* flame-region probe rows of the loaded state (appended to the solver's
     * probe CSV files with timestep 999999) */
    record_flame_region_data(999999, u, v, w, G, nu_t, S_mag);
}

int les_ref_printf(const char* fmt, ...)
{
    int is_step = (strncmp(fmt, "Step: ", 6) == 0);
    if (g_verbose) {
        va_list ap;
        va_start(ap, fmt);
        vprintf(fmt, ap);
        va_end(ap);
    }
    if (is_step && !g_done) {
        int step;
        va_list ap;
        va_start(ap, fmt);
        step = va_arg(ap, int);         /* t_loop is the first argument (:4601-4602) */
        va_end(ap);
        double now = wall();
        if (step < 100000) g_step_times[step] = now - g_tlast;
        g_tlast = now;
        if (!g_time && want_dump(step)) dump_state(step);
        if (step + 1 >= g_steps) {
            g_done = 1;
            stop_requested = 1;
            const char* rin = getenv("LES_REF_RHS_IN");
            if (rin && rin[0]) {
                const char* tr = getenv("LES_REF_RHS_T");
                rhs_on_inputs(rin, tr ? atoi(tr) : 600);
            }
        }
    }
    return 0;
}

int main(int argc, char** argv)
{
    const char* e;
    if ((e = getenv("LES_REF_SEED"))) g_seed = (unsigned)strtoul(e, NULL, 10);
    if ((e = getenv("LES_REF_STEPS"))) g_steps = atoi(e);
    if ((e = getenv("LES_REF_DUMP"))) snprintf(g_dump, sizeof g_dump, "%s", e);
    if ((e = getenv("LES_REF_OUT"))) snprintf(g_out, sizeof g_out, "%s", e);
    if ((e = getenv("LES_REF_VERBOSE"))) g_verbose = atoi(e);
    if ((e = getenv("LES_REF_TIME"))) g_time = atoi(e);
    if (g_steps < 1) g_steps = 1;
    /* the solver reads chem.bin / tran.bin from its working directory (:4097-4098) */
    if (access("chem.bin", R_OK) != 0 || access("tran.bin", R_OK) != 0) {
        les_mixture m;
        les_mixture_synthetic(&m);
        if (les_mixture_save(&m, "chem.bin", "tran.bin") != LES_OK) {
            fprintf(stderr, "harness: cannot write property files\n");
            return 2;
        }
    }
    g_t0 = g_tlast = wall();
    char* args[1] = {argv[0]};
    (void)argc;
    int rc = les_reference_main(1, args);
    double tot = wall() - g_t0;
    if (g_time) {
        /* steady-state per-step time: skip step 0 (first-touch, VTK write at init) */
        double s = 0.0;
        int n = 0;
        for (int t = 1; t < g_steps && t < 100000; ++t) { s += g_step_times[t]; n++; }
        fprintf(stdout, "{\"steps\": %d, \"total_s\": %.6f, \"s_per_step\": %.6f}\n",
                g_steps, tot, n ? s / n : tot / g_steps);
    }
    return rc;
}
