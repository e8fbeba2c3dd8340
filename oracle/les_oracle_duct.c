/*
 * les_oracle_duct.c - CPU oracle of the duct-mode scheme (FP64, C + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the engine's duct mode.  The duct
 * mode has no counterpart in the reference (its configs ask for a periodic
 * duct, SSP-RK3, WENO5/MUSCL-minmod, cp/mu tables, Gaussian SEM eddies and a
 * level-set G), so this file is the specification; every operator is pinned by
 * closed forms in tests/test_oracle_duct.py, tests/test_oracle_duct_closed_forms.py
 * and tests/test_oracle_projection.py, and oracle/mutation_check.py shows that
 * each of 39 planted mistakes fails at least one of them.  The RHS keeps the reference's
 * structure (calculate_RHS, les_combustion_solver.c:2494-2963): box filter,
 * Smagorinsky nu_t with Cs(G) (:1569-1633), SSR turbulent flame speed
 * (:1913-1965), conservative momentum with buoyancy and the flame-acceleration
 * force (:2552-2691), G equation with S_T|grad G|, SGS and artificial diffusion
 * (:2706-2818), T with nudging to T*(G) (:2912-2955).
 *
 * Duct state: rho*u, rho*v, rho*w, phi (signed distance, m; phi<0 unburnt),
 * T.  rho = rho_u T_u / T.  The reference's G in [0,1] is the progress
 * c(phi) = (1 + tanh(phi / (2 dx))) / 2.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "les_oracle_internal.h"

#define DNF 5   /* rho*u, rho*v, rho*w, phi, T */

typedef struct ora_duct {
    double* q[DNF];       /* state */
    double* q0[DNF];      /* q^n */
    double* qs[DNF];      /* RK stage state */
    double* L[DNF];       /* RHS */
    double *uf, *vf, *wf, *Tf, *pf;          /* filtered primitives */
    double *nu_t, *S_mag, *ST, *nu, *alpha;  /* SGS / properties */
    double* cp_tab;
    double* mu_tab;
    double* eddy;         /* num_vortices x 7: x, y, z, eu, ev, ew, generation */
    double* inlet;        /* 3 planes ny*nz */
    double max_nu, max_alpha;
    double rho_avg;
    double* prof;         /* nx x 10 plane accumulators (u,v,w,c,uu,vv,ww,uv,uw,vw) */
    int prof_count;
} ora_duct;

/* ---- counter-based RNG shared with the device (splitmix64 finaliser) ---- */
static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double ora_duct_uniform(uint64_t seed, uint64_t a, uint64_t b, uint64_t slot)
{
    uint64_t h = mix64(seed ^ mix64(a * 0x9E3779B97F4A7C15ull ^ mix64(b * 0xC2B2AE3D27D4EB4Full + slot)));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);   /* [0, 1) */
}

/* ---- model functions ---- */
static inline double duct_rho(const ora_ctx* c, double T) { return c->p.rho_unburnt * c->p.inlet_temperature / T; }
double ora_duct_progress(double phi, double dx) { return 0.5 * (1.0 + tanh(phi / (2.0 * dx))); }

double ora_duct_table(const double* tab, double t0, double dt, int n, double T)
{
    double s = (T - t0) / dt;
    if (s < 0.0) s = 0.0;
    if (s > (double)(n - 1)) s = (double)(n - 1);
    int i0 = (int)s;
    if (i0 > n - 2) i0 = n - 2;
    double f = s - (double)i0;
    return tab[i0] + f * (tab[i0 + 1] - tab[i0]);
}

/* periodic index helpers */
static inline int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
#define QX(c, i, j, k) IX(c, i, wrap(j, (c)->ny), wrap(k, (c)->nz))

int ora_duct_create(ora_ctx* c)
{
    if (c->p.integrator != LES_INT_SSPRK3) return LES_ERR_PARAM;
    if (c->p.advection != LES_ADV_WENO5 && c->p.advection != LES_ADV_MUSCL_MINMOD) return LES_ERR_PARAM;
    ora_duct* d = (ora_duct*)calloc(1, sizeof(ora_duct));
    if (!d) return LES_ERR_PARAM;
    c->duct = d;
    size_t n = c->n;
    for (int f = 0; f < DNF; ++f) {
        d->q[f] = calloc(n, 8); d->q0[f] = calloc(n, 8); d->qs[f] = calloc(n, 8); d->L[f] = calloc(n, 8);
    }
    d->uf = calloc(n, 8); d->vf = calloc(n, 8); d->wf = calloc(n, 8); d->Tf = calloc(n, 8); d->pf = calloc(n, 8);
    d->nu_t = calloc(n, 8); d->S_mag = calloc(n, 8); d->ST = calloc(n, 8); d->nu = calloc(n, 8); d->alpha = calloc(n, 8);
    int tn = c->p.table_n;
    d->cp_tab = calloc(tn, 8);
    d->mu_tab = calloc(tn, 8);
    for (int t = 0; t < tn; ++t) {
        double T = c->p.table_t0 + t * c->p.table_dt;
        d->cp_tab[t] = 1000.0 + 0.2 * (T - 300.0);
        d->mu_tab[t] = 1.8e-5 * pow(T / 300.0, 0.7);
    }
    d->eddy = calloc(7 * (size_t)(c->p.num_vortices > 0 ? c->p.num_vortices : 1), 8);
    d->inlet = calloc(3 * (size_t)c->ny * c->nz, 8);
    d->prof = calloc(10 * (size_t)c->nx, 8);
    return LES_OK;
}

void ora_duct_destroy(ora_ctx* c)
{
    ora_duct* d = c->duct;
    if (!d) return;
    for (int f = 0; f < DNF; ++f) { free(d->q[f]); free(d->q0[f]); free(d->qs[f]); free(d->L[f]); }
    free(d->uf); free(d->vf); free(d->wf); free(d->Tf); free(d->pf);
    free(d->nu_t); free(d->S_mag); free(d->ST); free(d->nu); free(d->alpha);
    free(d->cp_tab); free(d->mu_tab); free(d->eddy); free(d->inlet); free(d->prof);
    free(d);
    c->duct = NULL;
}

/* ---- SEM with Gaussian eddies (structure of :861-1075) ---- */
static void eddy_draw(const ora_ctx* c, double* e, int idx, uint64_t gen, int initial)
{
    const uint64_t s = c->p.seed;
    const double L = c->p.turbulent_length_scale;
    if (initial) e[0] = (2.0 * ora_duct_uniform(s, idx, gen, 0) - 1.0) * L;
    e[1] = ora_duct_uniform(s, idx, gen, 1) * c->p.ly;
    e[2] = ora_duct_uniform(s, idx, gen, 2) * c->p.lz;
    e[3] = 2.0 * ora_duct_uniform(s, idx, gen, 3) - 1.0;
    e[4] = 2.0 * ora_duct_uniform(s, idx, gen, 4) - 1.0;
    e[5] = 2.0 * ora_duct_uniform(s, idx, gen, 5) - 1.0;
    e[6] = (double)gen;
}

static void duct_init_eddies(ora_ctx* c)
{
    for (int e = 0; e < c->p.num_vortices; ++e) eddy_draw(c, c->duct->eddy + 7 * (size_t)e, e, 0, 1);
}

/* Inflow plane: eddy advection/recycling, Gaussian sum, mean removal and rms
 * rescale to U*I (same normalisation as :1005-1074). */
int ora_duct_sem(ora_ctx* c, double dt)
{
    ora_duct* d = c->duct;
    const double L = c->p.turbulent_length_scale, U = c->p.inlet_velocity;
    for (int e = 0; e < c->p.num_vortices; ++e) {
        double* E = d->eddy + 7 * (size_t)e;
        E[0] += U * dt;
        if (E[0] > L) {
            E[0] -= 2.0 * L;
            eddy_draw(c, E, e, (uint64_t)E[6] + 1, 0);
        }
    }
    const int ny = c->ny, nz = c->nz;
    const size_t P = (size_t)ny * nz;
    const double ly = c->p.ly, lz = c->p.lz, inv_l2 = 1.0 / (L * L);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            double y = j * c->dy, z = k * c->dz, su = 0.0, sv = 0.0, sw = 0.0;
            for (int e = 0; e < c->p.num_vortices; ++e) {
                const double* E = d->eddy + 7 * (size_t)e;
                double ddx = 0.0 - E[0];
                double ddy = y - E[1];
                ddy -= ly * rint(ddy / ly);
                double ddz = z - E[2];
                ddz -= lz * rint(ddz / lz);
                double f = exp(-M_PI * (ddx * ddx + ddy * ddy + ddz * ddz) * inv_l2);
                su += E[3] * f; sv += E[4] * f; sw += E[5] * f;
            }
            size_t p = (size_t)j * nz + k;
            d->inlet[p] = su; d->inlet[P + p] = sv; d->inlet[2 * P + p] = sw;
        }
    double tgt = U * c->p.turbulence_intensity;
    for (int a = 0; a < 3; ++a) {
        double* f = d->inlet + a * P;
        double s = 0.0;
        for (size_t p = 0; p < P; ++p) s += f[p];
        double mean = s / (double)P;
        double q = 0.0;
        for (size_t p = 0; p < P; ++p) { f[p] -= mean; q += f[p] * f[p]; }
        double rms = sqrt(q / (double)P);
        double sc = (rms > SMALL) ? tgt / rms : 0.0;
        for (size_t p = 0; p < P; ++p) f[p] *= sc;
    }
    return LES_OK;
}

/* ---- initial state: plug flow, planar flame at x0 ---- */
int ora_duct_init_state(ora_ctx* c)
{
    ora_duct* d = c->duct;
    const double U = c->p.inlet_velocity, Tu = c->p.inlet_temperature, Tb = c->p.burnt_temperature;
    for (int i = 0; i < c->nx; ++i)
        for (int j = 0; j < c->ny; ++j)
            for (int k = 0; k < c->nz; ++k) {
                size_t q = IX(c, i, j, k);
                double phi = i * c->dx - c->p.flame_x0;
                double cc = ora_duct_progress(phi, c->dx);
                double T = Tu + cc * (Tb - Tu);
                double rho = duct_rho(c, T);
                d->q[0][q] = rho * U; d->q[1][q] = 0.0; d->q[2][q] = 0.0;
                d->q[3][q] = phi; d->q[4][q] = T;
                c->attr[q] = (i == 0) ? 1 : (i == c->nx - 1 ? 2 : 3);
            }
    memset(d->nu_t, 0, c->n * 8);
    memset(d->S_mag, 0, c->n * 8);
    d->max_nu = 0.0; d->max_alpha = 0.0;
    duct_init_eddies(c);
    memset(d->inlet, 0, 3 * (size_t)c->ny * c->nz * 8);
    memset(d->prof, 0, 10 * (size_t)c->nx * 8);
    d->prof_count = 0;
    c->step = 0;
    c->n_probe = 0;
    c->stat_count = 0;
    for (int a = 0; a < 10; ++a) memset(c->stat[a], 0, c->n * 8);
    return LES_OK;
}

/* ---- reconstruction ---- */
/* Jiang-Shu WENO5 value at the face between v2 and v3 from the left-biased
 * stencil v0..v4: linear weights (0.1, 0.6, 0.3), eps = 1e-6, nonlinear weights
 * alpha_k = d_k / (eps + IS_k)^2.  Multiplying every alpha_k by the product of
 * the three (eps + IS)^2 leaves the normalised weights unchanged and needs one
 * division per face instead of four; the three candidate values share the
 * common factor 1/6. */
double ora_weno5_face(double v0, double v1, double v2, double v3, double v4)
{
    double a0 = v0 - 2.0 * v1 + v2, b0 = v0 - 4.0 * v1 + 3.0 * v2;
    double a1 = v1 - 2.0 * v2 + v3, b1 = v1 - v3;
    double a2 = v2 - 2.0 * v3 + v4, b2 = 3.0 * v2 - 4.0 * v3 + v4;
    double e0 = 1.0e-6 + ((13.0 / 12.0) * (a0 * a0) + 0.25 * (b0 * b0));
    double e1 = 1.0e-6 + ((13.0 / 12.0) * (a1 * a1) + 0.25 * (b1 * b1));
    double e2 = 1.0e-6 + ((13.0 / 12.0) * (a2 * a2) + 0.25 * (b2 * b2));
    double s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
    double w0 = 0.1 * (s1 * s2), w1 = 0.6 * (s0 * s2), w2 = 0.3 * (s0 * s1);
    double p0 = 2.0 * v0 - 7.0 * v1 + 11.0 * v2;
    double p1 = -v1 + 5.0 * v2 + 2.0 * v3;
    double p2 = 2.0 * v2 + 5.0 * v3 - v4;
    return (w0 * p0 + w1 * p1 + w2 * p2) / (6.0 * (w0 + w1 + w2));
}

double ora_minmod(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}

/* MUSCL-minmod value at the face between v2 and v3 from the left-biased stencil. */
double ora_muscl_face(double v1, double v2, double v3)
{
    return v2 + 0.5 * ora_minmod(v2 - v1, v3 - v2);
}

static inline double face_left(int scheme, double v0, double v1, double v2, double v3, double v4)
{
    return scheme == LES_ADV_WENO5 ? ora_weno5_face(v0, v1, v2, v3, v4) : ora_muscl_face(v1, v2, v3);
}

/* Gather the 7 values f[o], o = -3..3, along direction a through (i, j, k). */
static void line7(const ora_ctx* c, const double* f, int i, int j, int k, int a, double* v)
{
    for (int o = -3; o <= 3; ++o) {
        int ii = i, jj = j, kk = k;
        if (a == 0) ii += o; else if (a == 1) jj += o; else kk += o;
        if (ii < 0) ii = 0;
        if (ii > c->nx - 1) ii = c->nx - 1;
        v[o + 3] = f[QX(c, ii, jj, kk)];
    }
}

/* Upwind face values phi_{+1/2} and phi_{-1/2} of the cell for velocity sign
 * s >= 0 (left-biased) or s < 0 (mirrored stencil). */
static void faces(int scheme, const double* v, int positive, double* fp, double* fm)
{
    if (positive) {
        *fp = face_left(scheme, v[1], v[2], v[3], v[4], v[5]);
        *fm = face_left(scheme, v[0], v[1], v[2], v[3], v[4]);
    } else {
        *fp = face_left(scheme, v[6], v[5], v[4], v[3], v[2]);
        *fm = face_left(scheme, v[5], v[4], v[3], v[2], v[1]);
    }
}

/* ---- one RHS evaluation on state Q ---- */
static void duct_rhs_eval(ora_ctx* c, double* const* Q, double* const* out)
{
    ora_duct* d = c->duct;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double DX = c->dx, DY = c->dy, DZ = c->dz;
    const les_params* P = &c->p;
    const int scheme = P->advection;
    const double Tu = P->inlet_temperature, Tb = P->burnt_temperature;

    /* (1) box filter of the primitives u, v, w, T, phi (:1307-1470), periodic in y/z */
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t q = IX(c, i, j, k);
                double pu[7], pv[7], pw[7], pT[7], pp[7];
                const int di[7] = {-1, 0, 0, 0, 0, 0, 1}, dj[7] = {0, -1, 0, 0, 0, 1, 0}, dk[7] = {0, 0, -1, 0, 1, 0, 0};
                int npt = (i == 0 || i == nx - 1) ? 1 : 7;
                const double irut = 1.0 / (P->rho_unburnt * P->inlet_temperature);
                for (int t = 0; t < npt; ++t) {
                    size_t r = (npt == 1) ? q : QX(c, i + di[t], j + dj[t], k + dk[t]);
                    int slot = (npt == 1) ? 3 : t;
                    double T = Q[4][r], ri = T * irut;      /* 1/rho */
                    pu[slot] = Q[0][r] * ri; pv[slot] = Q[1][r] * ri; pw[slot] = Q[2][r] * ri;
                    pT[slot] = T; pp[slot] = Q[3][r];
                }
                if (npt == 1) {
                    d->uf[q] = pu[3]; d->vf[q] = pv[3]; d->wf[q] = pw[3]; d->Tf[q] = pT[3]; d->pf[q] = pp[3];
                } else {
                    const double wc = P->filter_weight_center, iden = 1.0 / (6.0 + wc);
                    d->uf[q] = (pu[0] + pu[1] + pu[2] + wc * pu[3] + pu[4] + pu[5] + pu[6]) * iden;
                    d->vf[q] = (pv[0] + pv[1] + pv[2] + wc * pv[3] + pv[4] + pv[5] + pv[6]) * iden;
                    d->wf[q] = (pw[0] + pw[1] + pw[2] + wc * pw[3] + pw[4] + pw[5] + pw[6]) * iden;
                    d->Tf[q] = (pT[0] + pT[1] + pT[2] + wc * pT[3] + pT[4] + pT[5] + pT[6]) * iden;
                    d->pf[q] = (pp[0] + pp[1] + pp[2] + wc * pp[3] + pp[4] + pp[5] + pp[6]) * iden;
                }
            }

    /* (2) properties, nu_t (:1569-1633), S_T (:1913-1965) */
    const double cu = P->cs_unburned, cb = P->cs_burnt, cf = P->cs_flame_boost;
    const double SL = P->s_l, nuu = P->nu_unburnt_approx, D3 = P->fractal_d3;
    const double tau = Tb / Tu - 1.0;
    double mnu = 0.0, mal = 0.0;
    #pragma omp parallel for schedule(static) reduction(max : mnu, mal)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t q = IX(c, i, j, k);
                double Tf = d->Tf[q], rf = duct_rho(c, Tf);
                double mu = ora_duct_table(d->mu_tab, P->table_t0, P->table_dt, P->table_n, Tf);
                double cp = ora_duct_table(d->cp_tab, P->table_t0, P->table_dt, P->table_n, Tf);
                double nu = mu / rf;
                double lam = mu * cp / P->prandtl;
                double al = lam / (rf * cp);
                d->nu[q] = nu; d->alpha[q] = al;
                if (i == 0 || i == nx - 1) { d->nu_t[q] = 0.0; d->S_mag[q] = 0.0; d->ST[q] = SL; continue; }
                if (nu > mnu) mnu = nu;
                if (al > mal) mal = al;
                const double hx = 0.5 / DX, hy = 0.5 / DY, hz = 0.5 / DZ;
                double du_dx = (d->uf[QX(c, i + 1, j, k)] - d->uf[QX(c, i - 1, j, k)]) * hx;
                double dv_dy = (d->vf[QX(c, i, j + 1, k)] - d->vf[QX(c, i, j - 1, k)]) * hy;
                double dw_dz = (d->wf[QX(c, i, j, k + 1)] - d->wf[QX(c, i, j, k - 1)]) * hz;
                double du_dy = (d->uf[QX(c, i, j + 1, k)] - d->uf[QX(c, i, j - 1, k)]) * hy;
                double du_dz = (d->uf[QX(c, i, j, k + 1)] - d->uf[QX(c, i, j, k - 1)]) * hz;
                double dv_dx = (d->vf[QX(c, i + 1, j, k)] - d->vf[QX(c, i - 1, j, k)]) * hx;
                double dv_dz = (d->vf[QX(c, i, j, k + 1)] - d->vf[QX(c, i, j, k - 1)]) * hz;
                double dw_dx = (d->wf[QX(c, i + 1, j, k)] - d->wf[QX(c, i - 1, j, k)]) * hx;
                double dw_dy = (d->wf[QX(c, i, j + 1, k)] - d->wf[QX(c, i, j - 1, k)]) * hy;
                double S12 = 0.5 * (du_dy + dv_dx), S13 = 0.5 * (du_dz + dw_dx), S23 = 0.5 * (dv_dz + dw_dy);
                double ss = 2.0 * (S12 * S12 + S13 * S13 + S23 * S23) + du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz;
                double smag = sqrt(2.0 * ss);
                double div = du_dx + dv_dy + dw_dz;
                double g = ora_duct_progress(d->pf[q], DX);
                double cs = cu + (cb - cu) * g;
                cs += cf * 4.0 * g * (1.0 - g);
                d->S_mag[q] = smag;
                d->nu_t[q] = (cs * DX) * (cs * DX) * smag;
                double num = ss - (div * div) * (1.0 / 3.0);
                double X = (pow(DX, 4.0) / (nu * nu)) * num;
                double d_eta;
                if (X < SMALL) d_eta = 1.0;
                else {
                    double px = pow(X, 0.75);
                    double alg = 0.2187 * px;
                    d_eta = 0.2340 * px * (alg > 1.0 ? log(alg) : 0.0);
                }
                double eta = DX / (d_eta + SMALL);
                double dF = nuu / (SL + SMALL);
                double asf = 8.0 * exp(10.0 * dF / (8.0 * eta + SMALL));
                double wr = pow(asf * eta / DX, 2.0 - D3);
                double te = (tau * SL > SMALL) ? DX * fmax(0.0, div) / (tau * SL) : 0.0;
                d->ST[q] = fmax(SL, SL * fmax(0.0, wr + te));
            }
    /* the diffusive DT limit of the next step uses the maxima over all stages */
    if (mnu > d->max_nu) d->max_nu = mnu;
    if (mal > d->max_alpha) d->max_alpha = mal;
    double srho = 0.0;
    for (size_t q = 0; q < c->n; ++q) srho += duct_rho(c, d->Tf[q]);
    d->rho_avg = srho / (double)c->n;

    /* (3) RHS on interior planes 1..nx-2 */
    const double Fm = P->max_flame_force, wdt = P->flame_accel_width, ctr = P->flame_accel_center;
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t q = IX(c, i, j, k);
                const int xband = (i < 3 || i > nx - 4);
                const double dd[3] = {DX, DY, DZ};
                const double* vel[3] = {d->uf, d->vf, d->wf};
                double rf0 = duct_rho(c, d->Tf[q]);
                /* neighbours along each axis (central values) */
                size_t nb[3][2];
                nb[0][0] = QX(c, i - 1, j, k); nb[0][1] = QX(c, i + 1, j, k);
                nb[1][0] = QX(c, i, j - 1, k); nb[1][1] = QX(c, i, j + 1, k);
                nb[2][0] = QX(c, i, j, k - 1); nb[2][1] = QX(c, i, j, k + 1);
                /* momentum: conservative flux with upwind face values */
                double conv[3] = {0, 0, 0};
                for (int a = 0; a < 3; ++a) {
                    const double* U = vel[a];
                    double up = 0.5 * (U[nb[a][1]] + U[q]);
                    double um = 0.5 * (U[q] + U[nb[a][0]]);
                    double rp = 0.5 * (duct_rho(c, d->Tf[nb[a][1]]) + rf0);
                    double rm = 0.5 * (rf0 + duct_rho(c, d->Tf[nb[a][0]]));
                    double mp = rp * up, mm = rm * um;
                    for (int m = 0; m < 3; ++m) {
                        double v[7];
                        line7(c, vel[m], i, j, k, a, v);
                        double fp, fm;
                        if (a == 0 && xband) {
                            fp = (up >= 0.0) ? v[3] : v[4];
                            fm = (um >= 0.0) ? v[2] : v[3];
                        } else {
                            double t1, t2;
                            faces(scheme, v, up >= 0.0, &fp, &t1);
                            faces(scheme, v, um >= 0.0, &t2, &fm);
                        }
                        double term = (mp * fp - mm * fm) * (1.0 / dd[a]);
                        conv[m] = (a == 0) ? term : conv[m] + term;
                    }
                }
                double mu = ora_duct_table(d->mu_tab, P->table_t0, P->table_dt, P->table_n, d->Tf[q]);
                double mu_eff = mu + rf0 * d->nu_t[q];
                double diff[3];
                for (int m = 0; m < 3; ++m) {
                    const double* U = vel[m];
                    double lx = (U[nb[0][1]] - 2.0 * U[q] + U[nb[0][0]]) * (1.0 / (DX * DX));
                    double ly = (U[nb[1][1]] - 2.0 * U[q] + U[nb[1][0]]) * (1.0 / (DY * DY));
                    double lz = (U[nb[2][1]] - 2.0 * U[q] + U[nb[2][0]]) * (1.0 / (DZ * DZ));
                    diff[m] = mu_eff * (lx + ly + lz);
                }
                double grav = (rf0 - d->rho_avg) * G_GRAV;
                double g0 = ora_duct_progress(d->pf[q], DX);
                double fx = 0.0, fy = 0.0, fz = 0.0;
                if (g0 > 0.01 && g0 < 0.99) {
                    double gx = (ora_duct_progress(d->pf[nb[0][1]], DX) - ora_duct_progress(d->pf[nb[0][0]], DX)) * (0.5 / DX);
                    double gy = (ora_duct_progress(d->pf[nb[1][1]], DX) - ora_duct_progress(d->pf[nb[1][0]], DX)) * (0.5 / DY);
                    double gz = (ora_duct_progress(d->pf[nb[2][1]], DX) - ora_duct_progress(d->pf[nb[2][0]], DX)) * (0.5 / DZ);
                    double df = g0 - ctr;
                    double win = exp(-(df * df) / (2.0 * wdt * wdt));
                    double common = P->flame_accel_strength * rf0 * win;
                    fx = common * gx * P->weight_x;
                    fy = common * gy;
                    fz = common * gz;
                }
                fx = fmax(-Fm, fmin(Fm, fx));
                fy = fmax(-Fm, fmin(Fm, fy));
                fz = fmax(-Fm, fmin(Fm, fz));
                out[0][q] = -conv[0] + diff[0] + grav + fx;
                out[1][q] = -conv[1] + diff[1] + fy;
                out[2][q] = -conv[2] + diff[2] + fz;

                /* scalars phi (level-set G equation) and T: non-conservative upwind */
                double nu0 = d->nu_t[q];
                double ntp[3], ntm[3];
                for (int a = 0; a < 3; ++a) { ntp[a] = d->nu_t[nb[a][1]]; ntm[a] = d->nu_t[nb[a][0]]; }
                const double* S2[2] = {d->pf, d->Tf};
                double cv[2], sg[2], lap[2], grad[3];
                for (int s = 0; s < 2; ++s) {
                    const double* F = S2[s];
                    const double isc2 = 0.5 / ((s == 0) ? P->sc_t : P->pr_t);
                    double acc = 0.0, sacc = 0.0, lacc = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        const double id = 1.0 / dd[a], id2 = 1.0 / (dd[a] * dd[a]), hid = 0.5 / dd[a];
                        double v[7];
                        line7(c, F, i, j, k, a, v);
                        double ua = vel[a][q];
                        double fp, fm;
                        if (a == 0 && xband) {
                            double t = (ua >= 0.0) ? ua * (v[3] - v[2]) * id : ua * (v[4] - v[3]) * id;
                            acc = (a == 0) ? t : acc + t;
                        } else {
                            faces(scheme, v, ua >= 0.0, &fp, &fm);
                            double t = ua * (fp - fm) * id;
                            acc = (a == 0) ? t : acc + t;
                        }
                        double gp = (ntp[a] + nu0) * isc2, gm = (nu0 + ntm[a]) * isc2;
                        double t2 = (gp * (v[4] - v[3]) - gm * (v[3] - v[2])) * id2;
                        sacc = (a == 0) ? t2 : sacc + t2;
                        double t3 = (v[4] - 2.0 * v[3] + v[2]) * id2;
                        lacc = (a == 0) ? t3 : lacc + t3;
                        if (s == 0) grad[a] = (v[4] - v[2]) * hid;
                    }
                    cv[s] = acc; sg[s] = sacc; lap[s] = lacc;
                }
                double gmag = sqrt(grad[0] * grad[0] + grad[1] * grad[1] + grad[2] * grad[2]);
                double art = 0.0;
                if (g0 > 0.1 && g0 < 0.99) {
                    double df = g0 - ctr;
                    double win = exp(-(df * df) / (2.0 * wdt * wdt));
                    art = 5.0 * (nu0 + 1.0e-5) * win * lap[0];
                }
                out[3][q] = -cv[0] + d->ST[q] * gmag + sg[0] + art;
                double Ttg = Tu + g0 * (Tb - Tu);
                out[4][q] = -cv[1] + d->alpha[q] * lap[1] + sg[1] + P->relaxation_coeff * (Ttg - d->Tf[q]);
            }
}

/* ---- SSP-RK3 stage update on interior planes; boundary planes carry over ---- */
static void duct_stage_update(ora_ctx* c, int stage, double DT, double* const* Qin, double* const* Qout)
{
    ora_duct* d = c->duct;
    const int nx = c->nx;
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        int i = (int)(q / ((size_t)c->ny * c->nz));
        for (int f = 0; f < DNF; ++f) {
            double v;
            if (i == 0 || i == nx - 1) v = Qin[f][q];
            else {
                double y = Qin[f][q] + DT * d->L[f][q];
                if (stage == 0) v = y;
                else if (stage == 1) v = 0.75 * d->q0[f][q] + 0.25 * y;
                else v = (1.0 / 3.0) * d->q0[f][q] + (2.0 / 3.0) * y;
                if (f == 4) v = fmax(200.0, fmin(3000.0, v));
            }
            Qout[f][q] = v;
        }
    }
}

/* ---- inflow / outflow planes ---- */
static void duct_boundary(ora_ctx* c)
{
    ora_duct* d = c->duct;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t P = (size_t)ny * nz;
    const double U = c->p.inlet_velocity, Tu = c->p.inlet_temperature;
    const double rho_u = duct_rho(c, Tu);
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            size_t p = (size_t)j * nz + k;
            size_t q0 = IX(c, 0, j, k), q1 = IX(c, 1, j, k);
            d->q[0][q0] = rho_u * (U + d->inlet[p]);
            d->q[1][q0] = rho_u * d->inlet[P + p];
            d->q[2][q0] = rho_u * d->inlet[2 * P + p];
            d->q[3][q0] = d->q[3][q1] - c->dx;
            d->q[4][q0] = Tu;
            size_t qn = IX(c, nx - 1, j, k), qm = IX(c, nx - 2, j, k);
            for (int f = 0; f < DNF; ++f) d->q[f][qn] = d->q[f][qm];
        }
}

/* ---- pressure/projection relaxation of the mass flux m = rho*u (this
 * engine's own specification: the reference has no pressure solve).
 * b = div m (central differences, periodic y/z) on interior planes;
 * `iters` Richardson sweeps phi <- phi + (L phi - b) / d from phi = 0 on the
 * compact 7-point Laplacian L, d = 2/dx^2 + 2/dy^2 + 2/dz^2, with the ghost
 * rules phi(plane 0) = phi(plane 1) (Neumann, inflow) and phi(plane nx-1) = 0
 * (Dirichlet, outflow); then m <- m - grad phi (central, same ghosts) on
 * interior planes.  FP64 RMS of the residual b - L phi over the interior
 * before the sweeps (res0 = RMS of b) and after them (res1). ---- */
static inline double proj_phi(const ora_ctx* c, const double* phi, int i, int j, int k)
{
    if (i <= 0) i = 1;
    if (i >= c->nx - 1) return 0.0;
    return phi[QX(c, i, j, k)];
}

static inline double proj_lap(const ora_ctx* c, const double* phi, int i, int j, int k)
{
    const double ix2 = 1.0 / (c->dx * c->dx), iy2 = 1.0 / (c->dy * c->dy), iz2 = 1.0 / (c->dz * c->dz);
    const double p0 = phi[IX(c, i, j, k)];
    double lx = (proj_phi(c, phi, i + 1, j, k) - 2.0 * p0 + proj_phi(c, phi, i - 1, j, k)) * ix2;
    double ly = (proj_phi(c, phi, i, j + 1, k) - 2.0 * p0 + proj_phi(c, phi, i, j - 1, k)) * iy2;
    double lz = (proj_phi(c, phi, i, j, k + 1) - 2.0 * p0 + proj_phi(c, phi, i, j, k - 1)) * iz2;
    return lx + ly + lz;
}

int ora_duct_project(ora_ctx* c, int iters, double* res0, double* res1)
{
    ora_duct* d = c->duct;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    double* b = d->L[0];
    double* ph[2] = {d->L[1], d->L[2]};
    const double hx = 0.5 / c->dx, hy = 0.5 / c->dy, hz = 0.5 / c->dz;
    const double inv_d = 1.0 / (2.0 / (c->dx * c->dx) + 2.0 / (c->dy * c->dy) + 2.0 / (c->dz * c->dz));
    const double n_int = (double)(nx - 2) * ny * nz;
    double s0 = 0.0;
    memset(ph[0], 0, c->n * 8);
    memset(ph[1], 0, c->n * 8);
    memset(b, 0, c->n * 8);
    #pragma omp parallel for schedule(static) reduction(+ : s0)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double v = (d->q[0][QX(c, i + 1, j, k)] - d->q[0][QX(c, i - 1, j, k)]) * hx +
                           (d->q[1][QX(c, i, j + 1, k)] - d->q[1][QX(c, i, j - 1, k)]) * hy +
                           (d->q[2][QX(c, i, j, k + 1)] - d->q[2][QX(c, i, j, k - 1)]) * hz;
                b[IX(c, i, j, k)] = v;
                s0 += v * v;
            }
    for (int it = 0; it < iters; ++it) {
        const double* po = ph[it & 1];
        double* pn = ph[(it + 1) & 1];
        #pragma omp parallel for schedule(static)
        for (int i = 1; i < nx - 1; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    size_t q = IX(c, i, j, k);
                    pn[q] = po[q] + (proj_lap(c, po, i, j, k) - b[q]) * inv_d;
                }
    }
    const double* phi = ph[iters & 1];
    double s1 = 0.0;
    #pragma omp parallel for schedule(static) reduction(+ : s1)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t q = IX(c, i, j, k);
                double r = b[q] - proj_lap(c, phi, i, j, k);
                s1 += r * r;
                d->q[0][q] -= (proj_phi(c, phi, i + 1, j, k) - proj_phi(c, phi, i - 1, j, k)) * hx;
                d->q[1][q] -= (proj_phi(c, phi, i, j + 1, k) - proj_phi(c, phi, i, j - 1, k)) * hy;
                d->q[2][q] -= (proj_phi(c, phi, i, j, k + 1) - proj_phi(c, phi, i, j, k - 1)) * hz;
            }
    if (res0) *res0 = sqrt(s0 / n_int);
    if (res1) *res1 = sqrt(s1 / n_int);
    return LES_OK;
}

/* ---- level-set reinitialisation: Sussman-Smereka-Osher with first-order
 * Godunov upwinding, pseudo time step 0.3*dx, interior planes only ---- */
int ora_duct_reinit(ora_ctx* c, int iters)
{
    ora_duct* d = c->duct;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    double* phi = d->q[3];
    double* phi0 = d->qs[3];
    double* tmp = d->L[3];
    const double dx = c->dx, dy = c->dy, dz = c->dz, dtau = 0.3 * dx;
    memcpy(phi0, phi, c->n * 8);
    for (int it = 0; it < iters; ++it) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    size_t q = IX(c, i, j, k);
                    if (i == 0 || i == nx - 1) { tmp[q] = phi[q]; continue; }
                    double p0 = phi0[q];
                    double s = p0 / sqrt(p0 * p0 + dx * dx);
                    double ax = (phi[q] - phi[QX(c, i - 1, j, k)]) / dx, bx = (phi[QX(c, i + 1, j, k)] - phi[q]) / dx;
                    double ay = (phi[q] - phi[QX(c, i, j - 1, k)]) / dy, by = (phi[QX(c, i, j + 1, k)] - phi[q]) / dy;
                    double az = (phi[q] - phi[QX(c, i, j, k - 1)]) / dz, bz = (phi[QX(c, i, j, k + 1)] - phi[q]) / dz;
                    double gx, gy, gz;
                    if (p0 > 0.0) {
                        double a1 = fmax(ax, 0.0), b1 = fmin(bx, 0.0);
                        double a2 = fmax(ay, 0.0), b2 = fmin(by, 0.0);
                        double a3 = fmax(az, 0.0), b3 = fmin(bz, 0.0);
                        gx = fmax(a1 * a1, b1 * b1); gy = fmax(a2 * a2, b2 * b2); gz = fmax(a3 * a3, b3 * b3);
                    } else {
                        double a1 = fmin(ax, 0.0), b1 = fmax(bx, 0.0);
                        double a2 = fmin(ay, 0.0), b2 = fmax(by, 0.0);
                        double a3 = fmin(az, 0.0), b3 = fmax(bz, 0.0);
                        gx = fmax(a1 * a1, b1 * b1); gy = fmax(a2 * a2, b2 * b2); gz = fmax(a3 * a3, b3 * b3);
                    }
                    double gm = sqrt(gx + gy + gz);
                    tmp[q] = phi[q] - dtau * s * (gm - 1.0);
                }
        memcpy(phi, tmp, c->n * 8);
    }
    return LES_OK;
}

/* ---- per-step statistics: per-cell FP64 accumulators (:3167-3192) and
 * y-z plane profiles of the resolved stresses at every x ---- */
static void duct_stats(ora_ctx* c)
{
    ora_duct* d = c->duct;
    const int nx = c->nx;
    const size_t plane = (size_t)c->ny * c->nz;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i) {
        double acc[10] = {0};
        for (size_t p = 0; p < plane; ++p) {
            size_t q = (size_t)i * plane + p;
            double rho = duct_rho(c, d->q[4][q]);
            double u = d->q[0][q] / rho, v = d->q[1][q] / rho, w = d->q[2][q] / rho;
            double g = ora_duct_progress(d->q[3][q], c->dx);
            if (c->p.stats_per_cell) {
            c->stat[0][q] += u; c->stat[1][q] += v; c->stat[2][q] += w; c->stat[3][q] += g;
            c->stat[4][q] += u * u; c->stat[5][q] += v * v; c->stat[6][q] += w * w;
            c->stat[7][q] += u * v; c->stat[8][q] += u * w; c->stat[9][q] += v * w;
            }
            acc[0] += u; acc[1] += v; acc[2] += w; acc[3] += g;
            acc[4] += u * u; acc[5] += v * v; acc[6] += w * w; acc[7] += u * v; acc[8] += u * w; acc[9] += v * w;
        }
        for (int a = 0; a < 10; ++a) d->prof[10 * (size_t)i + a] += acc[a];
    }
    c->stat_count++;
    d->prof_count++;
}

static double duct_dt(ora_ctx* c, double* maxvel_out)
{
    ora_duct* d = c->duct;
    double mv = 0.0;
    for (size_t q = 0; q < c->n; ++q) {
        int i = (int)(q / ((size_t)c->ny * c->nz));
        if (i == c->nx - 1) continue;
        double rho = duct_rho(c, d->q[4][q]);
        double u = d->q[0][q] / rho, v = d->q[1][q] / rho, w = d->q[2][q] / rho;
        double vel = sqrt(u * u + v * v + w * w);
        if (vel > mv) mv = vel;
    }
    double maxvel = mv > SMALL ? mv : c->p.inlet_velocity;
    double dtc = c->p.courant * c->dx / (maxvel + SMALL);
    double dtn = c->p.courant_diffusion * (c->dx * c->dx) / (d->max_nu + SMALL);
    double dta = c->p.courant_diffusion * (c->dx * c->dx) / (d->max_alpha + SMALL);
    *maxvel_out = maxvel;
    if (c->p.fixed_dt > 0) return c->p.fixed_dt;      /* prescribed DT_sim */
    return fmin(dtc, fmin(dtn, dta));
}

/* ---- flame-region probe rows (record_flame_region_data :3403-3471
 * structure) on the duct state: G = c(phi), u = rho*u / rho, nu_t and |S| of
 * the last stage ---- */
static void duct_probes(ora_ctx* c, int t_loop)
{
    ora_duct* d = c->duct;
    const int kp = c->nz / 2;
    for (int n = 0; n < c->p.num_probe_x; ++n) {
        double xp = c->p.probe_x[n];
        int ip = (int)(xp / c->dx);
        if (ip <= 0 || ip >= c->nx - 1) continue;
        double yf = -1.0;
        for (int j = 0; j < c->ny - 1; ++j) {
            double g1 = ora_duct_progress(d->q[3][IX(c, ip, j, kp)], c->dx);
            double g2 = ora_duct_progress(d->q[3][IX(c, ip, j + 1, kp)], c->dx);
            if ((g1 - 0.5) * (g2 - 0.5) <= 0.0) {
                double y1 = j * c->dy, y2 = (j + 1) * c->dy;
                if (fabs(g2 - g1) < 1e-9) yf = y1;
                else yf = y1 + (0.5 - g1) / (g2 - g1) * (y2 - y1);
                break;
            }
        }
        if (!(yf > 0.0)) continue;
        for (int j = 0; j < c->ny; ++j) {
            double yp = j * c->dy, yr = yp - yf;
            if (fabs(yr) > 1.0) continue;
            if (c->n_probe >= c->probe_cap) return;
            size_t q = IX(c, ip, j, kp);
            double rho = duct_rho(c, d->q[4][q]);
            les_probe_row* r = &c->probe[c->n_probe++];
            r->timestep = t_loop; r->station = n;
            r->x_location = xp; r->y_flame = yf; r->y_physical = yp; r->y_relative = yr;
            r->u = d->q[0][q] / rho; r->v = d->q[1][q] / rho; r->w = d->q[2][q] / rho;
            r->G = ora_duct_progress(d->q[3][q], c->dx);
            r->nu_t = d->nu_t[q]; r->S_mag = d->S_mag[q];
        }
    }
}

int ora_duct_step(ora_ctx* c, les_step_report* rep)
{
    ora_duct* d = c->duct;
    double maxvel;
    double DT = duct_dt(c, &maxvel);
    d->max_nu = 0.0;
    d->max_alpha = 0.0;
    ora_duct_sem(c, DT);
    for (int f = 0; f < DNF; ++f) memcpy(d->q0[f], d->q[f], c->n * 8);
    duct_rhs_eval(c, d->q, d->L);
    duct_stage_update(c, 0, DT, d->q, d->qs);
    duct_rhs_eval(c, d->qs, d->L);
    duct_stage_update(c, 1, DT, d->qs, d->qs);
    duct_rhs_eval(c, d->qs, d->L);
    duct_stage_update(c, 2, DT, d->qs, d->q);
    double pr0 = 0.0, pr1 = 0.0;
    if (c->p.proj_iters > 0) ora_duct_project(c, c->p.proj_iters, &pr0, &pr1);
    duct_boundary(c);
    if (c->p.reinit_interval > 0 && (c->step + 1) % c->p.reinit_interval == 0)
        ora_duct_reinit(c, c->p.reinit_iters);
    if (c->p.stats_start_step >= 0 && c->step >= c->p.stats_start_step) duct_stats(c);
    duct_probes(c, c->step);
    if (rep) {
        rep->step = c->step; rep->dt = DT; rep->max_velocity = maxvel;
        double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9;
        int bad = 0;
        const double dV = c->dx * c->dy * c->dz;
        for (size_t q = 0; q < c->n; ++q) {
            double T = d->q[4][q];
            tm += duct_rho(c, T) * dV;
            mnT = fmin(mnT, T); mxT = fmax(mxT, T);
            double g = ora_duct_progress(d->q[3][q], c->dx);
            mnG = fmin(mnG, g); mxG = fmax(mxG, g);
            for (int f = 0; f < DNF; ++f) if (!isfinite(d->q[f][q])) bad = 1;
        }
        rep->total_mass = tm; rep->min_T = mnT; rep->max_T = mxT; rep->min_G = mnG; rep->max_G = mxG;
        rep->min_Y_fuel = 0.0; rep->nonfinite = bad; rep->proj_residual = pr1; rep->proj_residual0 = pr0;
    }
    c->step++;
    return LES_OK;
}

int ora_duct_rhs(ora_ctx* c, double* const* out, int n_out)
{
    if (n_out < DNF) return LES_ERR_PARAM;
    duct_rhs_eval(c, c->duct->q, out);
    return LES_OK;
}

/* duct state exchange: fields rho*u, rho*v, rho*w, phi, T */
int ora_duct_set_state(ora_ctx* c, const double* const* f, int nf)
{
    if (nf < DNF) return LES_ERR_PARAM;
    for (int a = 0; a < DNF; ++a) memcpy(c->duct->q[a], f[a], c->n * 8);
    /* a new state starts without a diffusive DT history, as after init */
    c->duct->max_nu = 0.0;
    c->duct->max_alpha = 0.0;
    return LES_OK;
}

int ora_duct_get_state(const ora_ctx* c, double* const* f, int nf)
{
    for (int a = 0; a < nf && a < DNF; ++a) memcpy(f[a], c->duct->q[a], c->n * 8);
    return LES_OK;
}

int ora_duct_get_field(const ora_ctx* c, const char* name, double* out)
{
    ora_duct* d = c->duct;
    struct { const char* n; const double* p; } tab[] = {
        {"nu_t", d->nu_t}, {"S_mag", d->S_mag}, {"ST", d->ST}, {"nu", d->nu}, {"alpha", d->alpha},
        {"u_f", d->uf}, {"v_f", d->vf}, {"w_f", d->wf}, {"T_f", d->Tf}, {"phi_f", d->pf},
        {"rhou", d->q[0]}, {"rhov", d->q[1]}, {"rhow", d->q[2]}, {"phi", d->q[3]}, {"T", d->q[4]},
    };
    for (size_t a = 0; a < sizeof(tab) / sizeof(tab[0]); ++a)
        if (strcmp(tab[a].n, name) == 0) { memcpy(out, tab[a].p, c->n * 8); return LES_OK; }
    if (strcmp(name, "inlet") == 0) { memcpy(out, d->inlet, 3 * (size_t)c->ny * c->nz * 8); return LES_OK; }
    if (strcmp(name, "eddies") == 0) { memcpy(out, d->eddy, 7 * (size_t)c->p.num_vortices * 8); return LES_OK; }
    if (strcmp(name, "profiles") == 0) { memcpy(out, d->prof, 10 * (size_t)c->nx * 8); return LES_OK; }
    if (strcmp(name, "dt_state") == 0) { out[0] = d->max_nu; out[1] = d->max_alpha; out[2] = d->rho_avg; return LES_OK; }
    if (strcmp(name, "u") == 0 || strcmp(name, "v") == 0 || strcmp(name, "w") == 0 || strcmp(name, "G") == 0 ||
        strcmp(name, "rho") == 0) {
        for (size_t q = 0; q < c->n; ++q) {
            double rho = duct_rho(c, d->q[4][q]);
            if (name[0] == 'u') out[q] = d->q[0][q] / rho;
            else if (name[0] == 'v') out[q] = d->q[1][q] / rho;
            else if (name[0] == 'w') out[q] = d->q[2][q] / rho;
            else if (name[0] == 'G') out[q] = ora_duct_progress(d->q[3][q], c->dx);
            else out[q] = rho;
        }
        return LES_OK;
    }
    return LES_ERR_PARAM;
}

/* test hooks */
int ora_duct_set_eddies(ora_ctx* c, const double* e, int n)
{
    if (!c->duct || n != c->p.num_vortices) return LES_ERR_PARAM;
    memcpy(c->duct->eddy, e, 7 * (size_t)n * 8);
    return LES_OK;
}

int ora_duct_set_dt_state(ora_ctx* c, double max_nu, double max_alpha)
{
    if (!c->duct) return LES_ERR_PARAM;
    c->duct->max_nu = max_nu;
    c->duct->max_alpha = max_alpha;
    return LES_OK;
}

int ora_duct_profile_count(const ora_ctx* c) { return c->duct ? c->duct->prof_count : 0; }
