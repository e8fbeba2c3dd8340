/*
 * les_oracle_internal.h - private state of the CPU oracle (test infrastructure).
 */
#ifndef LES_ORACLE_INTERNAL_H
#define LES_ORACLE_INTERNAL_H

#include <stdlib.h>
#include "les_oracle.h"

#define NSP LES_NSPECIES
#define NRHS (5 + LES_NSPECIES)      /* drhou, drhov, drhow, dT, dG, dY0..dY4 */

/* physical constants of the reference (les_combustion_solver.c:136-141) */
#define R_UNIV   8.314462
#define PATM     101325.0
#define P_CONST  101325.0
#define G_GRAV   (-9.81)
#define SMALL    1e-10

#define IX(c, i, j, k) ((((size_t)(i)) * (size_t)(c)->ny + (size_t)(j)) * (size_t)(c)->nz + (size_t)(k))

typedef struct { double x, y, z, radius, su, sv, sw; } ora_vortex;

struct ora_duct;     /* duct-mode state (les_oracle_duct.c) */

struct ora_ctx {
    les_params p;
    les_mixture m;
    int nx, ny, nz;
    size_t n;
    double dx, dy, dz;
    int* attr;
    /* state */
    double *u, *v, *w, *T, *G, *rho, *Y[NSP], *rhou, *rhov, *rhow;
    /* Phi_n */
    double *rhou0, *rhov0, *rhow0, *T0, *G0, *Y0[NSP];
    /* intermediate RK state */
    double *ut, *vt, *wt, *Tt, *Gt, *Yt[NSP], *rhot;
    /* stage derivatives */
    double* k[4][NRHS];
    /* filtered fields and properties */
    double *uf, *vf, *wf, *Tf, *Gf, *rhof, *Yf[NSP];
    double *nu, *mu_mix, *Dk[NSP], *cp_mix, *lam_mix, *alpha, *nu_t, *S_mag, *SijSij, *div_u, *ST, *gradG;
    double last_rho_avg;
    /* SEM */
    ora_vortex* vort;
    int nvort;
    double *ufl, *vfl, *wfl;
    /* composition */
    double inletY[NSP], ambY[NSP], st_o2, st_co2, st_h2o;
    /* glibc rand() stream */
    struct random_data rd;
    char rstate[128];
    int step;
    /* probe rows */
    les_probe_row* probe;
    int n_probe, probe_cap;
    /* statistics */
    double* stat[10];
    int stat_count;
    /* duct mode */
    struct ora_duct* duct;
};

int    ora_rand(ora_ctx* c);
void   ora_composition(ora_ctx* c);
double ora_property_from_fit(double temp, const double co[4]);
double ora_eos(const ora_ctx* c, const double Yc[NSP], double T, int* ok);
void   ora_props_cell(const ora_ctx* c, size_t q);
double ora_dt_ref(ora_ctx* c, double* max_vel_out);

double ora_conv_nonconservative(const ora_ctx* c, int i, int j, int k, const double* f,
                                const double* u, const double* v, const double* w);
double ora_conv_conservative(const ora_ctx* c, int i, int j, int k, const double* f,
                             const double* u, const double* v, const double* w, const double* rho);
double ora_conv_muscl(const ora_ctx* c, int i, int j, int k, const double* f,
                      const double* u, const double* v, const double* w);
double ora_conv_teno(const ora_ctx* c, int i, int j, int k, const double* f,
                     const double* u, const double* v, const double* w);
double ora_teno5(double v0, double v1, double v2, double v3, double v4);
double ora_superbee(double r);
double ora_weno5_gradient(const double p[6], double dx);

/* duct mode */
int  ora_duct_create(ora_ctx* c);
void ora_duct_destroy(ora_ctx* c);
int  ora_duct_init_state(ora_ctx* c);
int  ora_duct_step(ora_ctx* c, les_step_report* rep);
int  ora_duct_rhs(ora_ctx* c, double* const* out, int n_out);
int  ora_duct_set_state(ora_ctx* c, const double* const* f, int nf);
int  ora_duct_get_state(const ora_ctx* c, double* const* f, int nf);
int  ora_duct_get_field(const ora_ctx* c, const char* name, double* out);

#endif
