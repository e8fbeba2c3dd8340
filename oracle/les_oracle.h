/*
 * les_oracle.h - CPU oracle for the LES time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C (FP64, OpenMP) restatement of
 * the reference's algorithm used as the checker for the HIP engine.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product path never calls it.
 *
 * Runtime-sized grid; same [i][j][k] layout as the reference's Grid3D.
 */
#ifndef LES_ORACLE_H
#define LES_ORACLE_H

#include "../include/les_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_ctx ora_ctx;

ora_ctx* ora_create(const les_params* p, const les_mixture* mix);
void     ora_destroy(ora_ctx* c);

int  ora_init_state(ora_ctx* c);
int  ora_set_state(ora_ctx* c, const double* const* fields, int n_fields);
int  ora_get_state(const ora_ctx* c, double* const* fields, int n_fields);
int  ora_get_field(const ora_ctx* c, const char* name, double* out);
int  ora_set_field(ora_ctx* c, const char* name, const double* in);
int  ora_step(ora_ctx* c, int n_steps, les_step_report* rep);
int  ora_rhs(ora_ctx* c, int t_loop, double* const* out, int n_out);
int  ora_probe_fetch(ora_ctx* c, les_probe_row* rows, int max_rows, int* n_rows);
int  ora_get_step(const ora_ctx* c);
/* Flame-region probe rows of the current state (record_flame_region_data). */
int  ora_probe_now(ora_ctx* c, int t_loop);
void ora_set_step(ora_ctx* c, int step);
/* SEM eddies: n x 7 doubles (x, y, z, radius, su, sv, sw). */
int  ora_get_vortices(const ora_ctx* c, double* out, int max_n);
int  ora_set_vortices(ora_ctx* c, const double* in, int n);
/* Inlet fluctuation planes (ny*nz each) of the last step. */
int  ora_get_inlet(const ora_ctx* c, double* uf, double* vf, double* wf);
/* Max nu / alpha over fluid cells carried into the next DT. */
void ora_get_dt_state(const ora_ctx* c, double* max_nu, double* max_alpha);
void ora_set_dt_state(ora_ctx* c, double max_nu, double max_alpha);
/* Individual operators for per-operator parity. */
int  ora_filter_and_props(ora_ctx* c, const double* const* in);
int  ora_sem_plane(ora_ctx* c, int t_loop, double dt);
int  ora_boundary(ora_ctx* c);
/* glibc rand() stream helper: n draws from srand(seed) (for the RNG pin). */
int  ora_rand_stream(unsigned seed, int n, int* out);
/* Duct-mode operators and hooks (les_oracle_duct.c). */
double ora_weno5_face(double v0, double v1, double v2, double v3, double v4);
double ora_minmod(double a, double b);
double ora_muscl_face(double v1, double v2, double v3);
double ora_duct_table(const double* tab, double t0, double dt, int n, double T);
double ora_duct_progress(double phi, double dx);
double ora_duct_uniform(uint64_t seed, uint64_t a, uint64_t b, uint64_t slot);
int  ora_duct_sem(ora_ctx* c, double dt);
int  ora_duct_reinit(ora_ctx* c, int iters);
/* projection relaxation of rho*u with `iters` sweeps; FP64 RMS residuals */
int  ora_duct_project(ora_ctx* c, int iters, double* res0, double* res1);
int  ora_duct_set_eddies(ora_ctx* c, const double* e, int n);
int  ora_duct_set_dt_state(ora_ctx* c, double max_nu, double max_alpha);
int  ora_duct_profile_count(const ora_ctx* c);
/* Statistics: 10 FP64 accumulators (u,v,w,G,uu,vv,ww,uv,uw,vw) and count. */
int  ora_get_stats(const ora_ctx* c, double* const* acc, int* count);

#ifdef __cplusplus
}
#endif
#endif
