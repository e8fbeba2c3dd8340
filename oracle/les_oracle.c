/*
 * les_oracle.c - CPU oracle of the LES time-step hot path (FP64, C + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Reference-scheme mode (LES_MODE_REFERENCE) restates the reference step of
 * les_combustion_solver.c with its operation order, its branch/clip/fallback
 * bands and its quirks (uninitialised rho_temp read as 0.0 at attributes 1/4/5,
 * nu/alpha read as 0.0 for the first DT, 3 discarded rand() draws before the
 * SEM eddies, stage-4 nu_t/S_mag in the probe rows).  Each function cites the
 * reference lines it follows.  It is pinned against the reference compiled
 * from its own sources (oracle/ref_harness.c) and against closed forms
 * (tests/test_oracle_*.py).
 *
 * Duct mode (LES_MODE_DUCT) is the engine's own scheme (no reference
 * counterpart); it lives in les_oracle_duct.c and is pinned by closed forms.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off (no fast-math: the reference
 * relies on IEEE behaviour, e.g. exp(-inf) at :2807).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "les_oracle.h"
#include "les_oracle_internal.h"

/* ------------------------------------------------------------------------- */
/* allocation                                                                 */
/* ------------------------------------------------------------------------- */
static double* zalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

ora_ctx* ora_create(const les_params* p, const les_mixture* mix)
{
    if (!p || !mix || les_params_validate(p) != LES_OK) return NULL;
    ora_ctx* c = (ora_ctx*)calloc(1, sizeof(ora_ctx));
    if (!c) return NULL;
    c->p = *p;
    c->m = *mix;
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz;
    c->n = (size_t)p->nx * p->ny * p->nz;
    c->dx = p->lx / (double)p->nx;      /* :104-106 */
    c->dy = p->ly / (double)p->ny;
    c->dz = p->lz / (double)p->nz;
    size_t n = c->n;
    c->attr = (int*)calloc(n, sizeof(int));
    double** all[] = {&c->u, &c->v, &c->w, &c->T, &c->G, &c->rho, &c->rhou, &c->rhov, &c->rhow,
                      &c->rhou0, &c->rhov0, &c->rhow0, &c->T0, &c->G0,
                      &c->ut, &c->vt, &c->wt, &c->Tt, &c->Gt, &c->rhot,
                      &c->uf, &c->vf, &c->wf, &c->Tf, &c->Gf, &c->rhof,
                      &c->nu, &c->mu_mix, &c->cp_mix, &c->lam_mix, &c->alpha, &c->nu_t,
                      &c->S_mag, &c->SijSij, &c->div_u, &c->ST, &c->gradG};
    for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); ++a) *all[a] = zalloc(n);
    for (int s = 0; s < NSP; ++s) {
        c->Y[s] = zalloc(n); c->Y0[s] = zalloc(n); c->Yt[s] = zalloc(n);
        c->Yf[s] = zalloc(n); c->Dk[s] = zalloc(n);
    }
    for (int st = 0; st < 4; ++st)
        for (int f = 0; f < NRHS; ++f) c->k[st][f] = zalloc(n);
    for (int a = 0; a < 10; ++a) c->stat[a] = zalloc(n);
    size_t plane = (size_t)p->ny * p->nz;
    c->ufl = zalloc(plane); c->vfl = zalloc(plane); c->wfl = zalloc(plane);
    c->nvort = p->num_vortices;
    c->vort = (ora_vortex*)calloc(c->nvort > 0 ? c->nvort : 1, sizeof(ora_vortex));
    c->probe_cap = 1 << 16;
    c->probe = (les_probe_row*)calloc(c->probe_cap, sizeof(les_probe_row));
    memset(&c->rd, 0, sizeof(c->rd));
    initstate_r((unsigned)c->p.seed, c->rstate, sizeof(c->rstate), &c->rd);
    ora_composition(c);
    if (c->p.mode == LES_MODE_DUCT && ora_duct_create(c) != LES_OK) { ora_destroy(c); return NULL; }
    return c;
}

void ora_destroy(ora_ctx* c)
{
    if (!c) return;
    double* all[] = {c->u, c->v, c->w, c->T, c->G, c->rho, c->rhou, c->rhov, c->rhow,
                     c->rhou0, c->rhov0, c->rhow0, c->T0, c->G0,
                     c->ut, c->vt, c->wt, c->Tt, c->Gt, c->rhot,
                     c->uf, c->vf, c->wf, c->Tf, c->Gf, c->rhof,
                     c->nu, c->mu_mix, c->cp_mix, c->lam_mix, c->alpha, c->nu_t,
                     c->S_mag, c->SijSij, c->div_u, c->ST, c->gradG, c->ufl, c->vfl, c->wfl};
    for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); ++a) free(all[a]);
    for (int s = 0; s < NSP; ++s) { free(c->Y[s]); free(c->Y0[s]); free(c->Yt[s]); free(c->Yf[s]); free(c->Dk[s]); }
    for (int st = 0; st < 4; ++st) for (int f = 0; f < NRHS; ++f) free(c->k[st][f]);
    for (int a = 0; a < 10; ++a) free(c->stat[a]);
    free(c->attr); free(c->vort); free(c->probe);
    ora_duct_destroy(c);
    free(c);
}

/* glibc rand() drawn from a private state: random_r over a TYPE_3 (128-byte)
 * state reproduces rand() after srand(seed) (glibc stdlib/random_r.c). */
int ora_rand(ora_ctx* c)
{
    int32_t r = 0;
    random_r(&c->rd, &r);
    return (int)r;
}

int ora_rand_stream(unsigned seed, int n, int* out)
{
    struct random_data rd;
    char st[128];
    memset(&rd, 0, sizeof(rd));
    initstate_r(seed, st, sizeof(st), &rd);
    for (int a = 0; a < n; ++a) { int32_t r; random_r(&rd, &r); out[a] = (int)r; }
    return LES_OK;
}

/* ------------------------------------------------------------------------- */
/* property evaluators (:385-447)                                            */
/* ------------------------------------------------------------------------- */
double ora_property_from_fit(double temp, const double co[4])     /* :385-391 */
{
    if (temp <= 0) return 0.0;
    if (co[0] == 0.0 && co[1] == 0.0 && co[2] == 0.0 && co[3] == 0.0) return 0.0;
    double lt = log(temp);
    double lp = co[0] + co[1] * lt + co[2] * pow(lt, 2) + co[3] * pow(lt, 3);
    return exp(lp);
}

static double mu_species(const ora_ctx* c, int s, double t)         /* :394-398 */
{
    return ora_property_from_fit(t, c->m.sp[s].eta_fit) * 0.1;
}

static double cp_species(const ora_ctx* c, int s, double t)         /* :401-428 */
{
    const les_species* sp = &c->m.sp[s];
    double tt = t;
    if (tt <= 0.0) tt = 1.0;
    const double* a = (tt > sp->t_mid) ? sp->nasa_high : sp->nasa_low;
    double cpr = a[0] + a[1] * tt + a[2] * pow(tt, 2) + a[3] * pow(tt, 3) + a[4] * pow(tt, 4);
    double cp_mole = cpr * R_UNIV;
    double mw_kg = sp->mw / 1000.0;
    if (mw_kg < 1e-9) return 1000.0;
    return cp_mole / mw_kg;
}

static double lambda_species(const ora_ctx* c, int s, double t)     /* :431-437 */
{
    return ora_property_from_fit(t, c->m.sp[s].lambda_fit) * 1e-5;
}

static double dkj(const ora_ctx* c, int a, int b, double t)         /* :440-447 */
{
    return ora_property_from_fit(t, c->m.diff_fit[a][b]) * 1e-4;
}

/* Inlet composition for the equivalence ratio and stoichiometric ratios (:4101-4120). */
void ora_composition(ora_ctx* c)
{
    const les_species* sp = c->m.sp;
    double phi = c->p.equivalence_ratio;
    double n_f = 1.0, n_o2 = 5.0 / phi, n_n2 = n_o2 * 3.76;
    double mt = n_f * sp[0].mw + n_o2 * sp[1].mw + n_n2 * sp[2].mw;
    c->inletY[0] = (n_f * sp[0].mw) / mt;
    c->inletY[1] = (n_o2 * sp[1].mw) / mt;
    c->inletY[2] = (n_n2 * sp[2].mw) / mt;
    c->inletY[3] = 0.0; c->inletY[4] = 0.0;
    c->ambY[0] = 0.0; c->ambY[1] = 0.232; c->ambY[2] = 0.768; c->ambY[3] = 0.0; c->ambY[4] = 0.0;
    c->st_o2 = (5.0 * sp[1].mw) / sp[0].mw;
    c->st_co2 = (3.0 * sp[3].mw) / sp[0].mw;
    c->st_h2o = (4.0 * sp[4].mw) / sp[0].mw;
}

/* ------------------------------------------------------------------------- */
/* initial state: attribute map + initial fields (setting :1135-1285)         */
/* ------------------------------------------------------------------------- */
static void attribute_map(ora_ctx* c)                               /* :1152-1181 */
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double R = c->p.inlet_radius, zc = c->p.lz / 2.0, r_rim = 0.005;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int a;
                if (i == 0) {
                    double y = j * c->dy, z = k * c->dz;
                    double d2 = pow(y - 0.0, 2) + pow(z - zc, 2);
                    if (d2 <= pow(R, 2)) a = 1;
                    else if (d2 <= pow(r_rim, 2)) a = 6;
                    else a = 4;
                } else if (i == nx - 1) a = 2;
                else if (j == 0) a = 5;
                else if (j == ny - 1 || k == 0 || k == nz - 1) a = 4;
                else a = 3;
                c->attr[IX(c, i, j, k)] = a;
            }
}

double ora_eos(const ora_ctx* c, const double Yc[NSP], double T, int* ok)
{
    double s = 0.0;
    for (int q = 0; q < NSP; ++q) s += Yc[q] / (c->m.sp[q].mw / 1000.0);
    if (s > 1e-9 && T > 1.0) {
        double wtm = 1.0 / s;
        double rmix = R_UNIV / wtm;
        *ok = 1;
        return P_CONST / (rmix * T);
    }
    *ok = 0;
    return 0.0;
}

static void init_vortices_ref(ora_ctx* c)                           /* :861-897 */
{
    /* three draws consumed by the reference's debug printout (:864-866) */
    (void)ora_rand(c); (void)ora_rand(c); (void)ora_rand(c);
    double urms = c->p.inlet_velocity * c->p.turbulence_intensity;
    double L = c->p.turbulent_length_scale;
    double box_x = L, box_y = c->p.ly, box_z = c->p.lz;
    for (int v = 0; v < c->nvort; ++v) {
        double r[7];
        for (int q = 0; q < 7; ++q) r[q] = ora_rand(c) / (double)RAND_MAX;
        ora_vortex* e = &c->vort[v];
        e->x = (r[0] - 0.5) * box_x;
        e->y = r[1] * box_y;
        e->z = r[2] * box_z;
        double rmin = L * 0.5, rmax = L * 1.5;
        e->radius = rmin + (rmax - rmin) * r[3];
        e->su = (r[4] * 2.0 - 1.0) * urms;
        e->sv = (r[5] * 2.0 - 1.0) * urms;
        e->sw = (r[6] * 2.0 - 1.0) * urms;
    }
}

int ora_init_state(ora_ctx* c)
{
    if (c->p.mode == LES_MODE_DUCT) return ora_duct_init_state(c);
    const size_t n = c->n;
    attribute_map(c);
    for (size_t q = 0; q < n; ++q) {                                 /* :1184-1261 */
        int a = c->attr[q];
        int i = (int)(q / ((size_t)c->ny * c->nz));
        double Yc[NSP];
        double u = 0.0, T = 0.0, G = 0.0;
        switch (a) {
        case 1:
            u = c->p.inlet_velocity; T = c->p.inlet_temperature; G = 0.0;
            for (int s = 0; s < NSP; ++s) Yc[s] = c->inletY[s];
            break;
        case 4: case 2:
            T = c->p.ambient_temperature; G = 0.0;
            for (int s = 0; s < NSP; ++s) Yc[s] = c->ambY[s];
            break;
        case 5: case 3:
            if (i <= c->p.initial_flame_front_i) {
                double cons = c->inletY[0];
                G = 1.0; T = c->p.burnt_temperature;
                Yc[0] = c->inletY[0] - cons;
                Yc[1] = c->inletY[1] - cons * c->st_o2;
                Yc[2] = c->inletY[2];
                Yc[3] = c->inletY[3] + cons * c->st_co2;
                Yc[4] = c->inletY[4] + cons * c->st_h2o;
            } else {
                G = 0.0; T = c->p.inlet_temperature;
                for (int s = 0; s < NSP; ++s) Yc[s] = c->inletY[s];
            }
            break;
        default:
            for (int s = 0; s < NSP; ++s) Yc[s] = 0.0;
        }
        c->u[q] = u; c->v[q] = 0.0; c->w[q] = 0.0;
        c->T[q] = T; c->G[q] = G;
        for (int s = 0; s < NSP; ++s) c->Y[s][q] = Yc[s];
        /* setting()'s EOS uses SMALL_NUMBER (1e-10) as the threshold (:1248) */
        double sy = 0.0;
        for (int s = 0; s < NSP; ++s) sy += Yc[s] / (c->m.sp[s].mw / 1000.0);
        if (sy > SMALL && T > 1.0) {
            double wmix = 1.0 / sy;
            double rmix = R_UNIV / wmix;
            c->rho[q] = P_CONST / (rmix * T);
        } else {
            c->rho[q] = 1.2;
        }
        c->rhou[q] = c->rho[q] * c->u[q];
        c->rhov[q] = c->rho[q] * c->v[q];
        c->rhow[q] = c->rho[q] * c->w[q];
    }
    /* initialize() zeroes nu_t, S_mag etc.; nu/alpha/rho_temp are left as
     * fresh (zero) allocations (quirks: first DT, rho_temp at attr 1/4/5). */
    memset(c->nu, 0, n * sizeof(double));
    memset(c->alpha, 0, n * sizeof(double));
    memset(c->nu_t, 0, n * sizeof(double));
    memset(c->S_mag, 0, n * sizeof(double));
    memset(c->rhot, 0, n * sizeof(double));
    init_vortices_ref(c);
    c->step = 0;
    c->n_probe = 0;
    c->stat_count = 0;
    for (int a = 0; a < 10; ++a) memset(c->stat[a], 0, n * sizeof(double));
    return LES_OK;
}

/* ------------------------------------------------------------------------- */
/* box filter (:1307-1470) + mixture properties (:1489-1565, :1684-1881)     */
/* ------------------------------------------------------------------------- */
static inline double box7(const ora_ctx* c, const double* f, int i, int j, int k, double wc)
{
    /* neighbour order of :1329-1331: (i-1), (j-1), (k-1), centre, (k+1), (j+1), (i+1) */
    double den = 6.0 + wc;
    return (f[IX(c, i - 1, j, k)] + f[IX(c, i, j - 1, k)] + f[IX(c, i, j, k - 1)] +
            wc * f[IX(c, i, j, k)] +
            f[IX(c, i, j, k + 1)] + f[IX(c, i, j + 1, k)] + f[IX(c, i + 1, j, k)]) / den;
}

static void filter_ref(ora_ctx* c, const double* u, const double* v, const double* w,
                       const double* T, const double* G, double* const* Y, const double* rho)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double wc = c->p.filter_weight_center;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t q = IX(c, i, j, k);
                int interior = (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1);
                if (interior && c->attr[q] == 3) {
                    c->uf[q] = box7(c, u, i, j, k, wc);
                    c->vf[q] = box7(c, v, i, j, k, wc);
                    c->wf[q] = box7(c, w, i, j, k, wc);
                    c->Tf[q] = box7(c, T, i, j, k, wc);
                    c->rhof[q] = box7(c, rho, i, j, k, wc);
                    for (int s = 0; s < NSP; ++s) c->Yf[s][q] = box7(c, Y[s], i, j, k, wc);
                    c->Gf[q] = box7(c, G, i, j, k, wc);
                } else {
                    /* copies at non-fluid and boundary cells (:1385-1469); P is
                     * identically zero in the reference and is not carried */
                    c->uf[q] = u[q]; c->vf[q] = v[q]; c->wf[q] = w[q];
                    c->Tf[q] = T[q]; c->rhof[q] = rho[q]; c->Gf[q] = G[q];
                    for (int s = 0; s < NSP; ++s) c->Yf[s][q] = Y[s][q];
                }
            }
}

static int mole_fractions(const ora_ctx* c, size_t q, double X[NSP])
{
    double s = 0.0;
    for (int a = 0; a < NSP; ++a) s += c->Yf[a][q] / (c->m.sp[a].mw / 1000.0);
    if (s < 1e-9) return 0;
    for (int a = 0; a < NSP; ++a) X[a] = (c->Yf[a][q] / (c->m.sp[a].mw / 1000.0)) / s;
    return 1;
}

static double wilke_mix(const ora_ctx* c, const double X[NSP], const double mu_k[NSP],
                        const double val_k[NSP])
{
    /* Wilke / Wassiljew mixing (:1530-1546, :1842-1863) */
    double acc = 0.0;
    for (int a = 0; a < NSP; ++a) {
        if (X[a] < 1e-12) continue;
        double mu_i = mu_k[a], M_i = c->m.sp[a].mw, den = 0.0;
        for (int b = 0; b < NSP; ++b) {
            if (X[b] < 1e-12) continue;
            double mu_j = mu_k[b], M_j = c->m.sp[b].mw;
            double phi = pow(1.0 + pow(mu_i / mu_j, 0.5) * pow(M_j / M_i, 0.25), 2) /
                         (sqrt(8.0) * sqrt(1.0 + M_i / M_j));
            den += X[b] * phi;
        }
        if (den > 1e-9) acc += (X[a] * val_k[a]) / den;
    }
    return acc;
}

void ora_props_cell(const ora_ctx* c, size_t q)
{
    /* update_nu_mix (:1497-1562) */
    double* nu = c->nu; double* mu_mix = c->mu_mix;
    double Tl = c->Tf[q];
    double X[NSP];
    if (c->attr[q] != 3) {
        nu[q] = 1.5e-5; mu_mix[q] = 1.8e-5;
        for (int s = 0; s < NSP; ++s) c->Dk[s][q] = 0.0;             /* :1751 */
        c->cp_mix[q] = 1000.0;                                       /* :1790 */
        c->lam_mix[q] = 0.0; c->alpha[q] = 0.0;                      /* :1875-1876 */
        return;
    }
    int have_x = (Tl > 0) ? mole_fractions(c, q, X) : 0;
    if (Tl <= 0 || !have_x) {
        nu[q] = 1.5e-5; mu_mix[q] = 1.8e-5;
    } else {
        double mu_k[NSP];
        for (int s = 0; s < NSP; ++s) mu_k[s] = mu_species(c, s, Tl);
        double mu = wilke_mix(c, X, mu_k, mu_k);
        mu_mix[q] = fmax(1e-7, mu);
        if (c->rhof[q] > 1e-3) nu[q] = mu_mix[q] / c->rhof[q];
        else nu[q] = 1.5e-5;
        nu[q] = fmax(1e-7, nu[q]);
    }
    /* update_Dk (:1691-1749) */
    if (Tl <= 0 || !have_x) {
        for (int s = 0; s < NSP; ++s) c->Dk[s][q] = 1e-5;
    } else {
        for (int a = 0; a < NSP; ++a) {
            double sum = 0.0;
            for (int b = 0; b < NSP; ++b) {
                if (a == b) continue;
                if (X[b] > 1e-12) {
                    double d = dkj(c, a, b, Tl);
                    if (d > 1e-12) sum += X[b] / d;
                }
            }
            double D;
            if (fabs(1.0 - X[a]) < 1e-9) {
                int other = (a != 2) ? 2 : 1;
                D = dkj(c, a, other, Tl);
            } else if (sum > 1e-12) {
                D = (1.0 - X[a]) / sum;
            } else {
                D = 1e-5;
            }
            /* pressure correction with P_filtered = 0: PATM/(p0 + 0) == 1.0 */
            double ptot = P_CONST + 0.0;
            if (ptot > 1.0) D *= (PATM / ptot);
            c->Dk[a][q] = fmax(1e-9, fmin(1e-2, D));
        }
    }
    /* update_Cp_mix (:1768-1792) */
    if (Tl <= 0) {
        c->cp_mix[q] = 1000.0;
    } else {
        double cp = 0.0;
        for (int s = 0; s < NSP; ++s) cp += c->Yf[s][q] * cp_species(c, s, Tl);
        c->cp_mix[q] = fmax(100.0, cp);
    }
    /* update_lambda_mix_alpha (:1807-1877) */
    if (Tl <= 0 || !have_x) {
        c->lam_mix[q] = 0.025; c->alpha[q] = 1.8e-5;
    } else {
        double lam_k[NSP], mu_k[NSP];
        for (int s = 0; s < NSP; ++s) { lam_k[s] = lambda_species(c, s, Tl); mu_k[s] = mu_species(c, s, Tl); }
        double lam = wilke_mix(c, X, mu_k, lam_k);
        c->lam_mix[q] = fmax(1e-3, lam);
        if (c->rhof[q] > 1e-6 && c->cp_mix[q] > 1e-6) c->alpha[q] = c->lam_mix[q] / (c->rhof[q] * c->cp_mix[q]);
        else c->alpha[q] = 1.8e-5;
        c->alpha[q] = fmax(1e-7, c->alpha[q]);
    }
}

/* ------------------------------------------------------------------------- */
/* SGS viscosity (:1569-1633) and SSR turbulent flame speed (:1913-1965)       */
/* ------------------------------------------------------------------------- */
static void nu_t_and_st_ref(ora_ctx* c)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double DX = c->dx, DY = c->dy, DZ = c->dz;
    const double cu = c->p.cs_unburned, cb = c->p.cs_burnt, cf = c->p.cs_flame_boost;
    const double* uf = c->uf; const double* vf = c->vf; const double* wf = c->wf;
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                size_t q = IX(c, i, j, k);
                if (c->attr[q] != 3) {
                    c->SijSij[q] = 0.0; c->div_u[q] = 0.0; c->S_mag[q] = 0.0; c->nu_t[q] = 0.0;
                    continue;
                }
                double du_dx = (uf[IX(c, i + 1, j, k)] - uf[IX(c, i - 1, j, k)]) / (2.0 * DX);
                double dv_dy = (vf[IX(c, i, j + 1, k)] - vf[IX(c, i, j - 1, k)]) / (2.0 * DY);
                double dw_dz = (wf[IX(c, i, j, k + 1)] - wf[IX(c, i, j, k - 1)]) / (2.0 * DZ);
                double du_dy = (uf[IX(c, i, j + 1, k)] - uf[IX(c, i, j - 1, k)]) / (2.0 * DY);
                double du_dz = (uf[IX(c, i, j, k + 1)] - uf[IX(c, i, j, k - 1)]) / (2.0 * DZ);
                double dv_dx = (vf[IX(c, i + 1, j, k)] - vf[IX(c, i - 1, j, k)]) / (2.0 * DX);
                double dv_dz = (vf[IX(c, i, j, k + 1)] - vf[IX(c, i, j, k - 1)]) / (2.0 * DZ);
                double dw_dx = (wf[IX(c, i + 1, j, k)] - wf[IX(c, i - 1, j, k)]) / (2.0 * DX);
                double dw_dy = (wf[IX(c, i, j + 1, k)] - wf[IX(c, i, j - 1, k)]) / (2.0 * DY);
                double S12 = 0.5 * (du_dy + dv_dx);
                double S13 = 0.5 * (du_dz + dw_dx);
                double S23 = 0.5 * (dv_dz + dw_dy);
                double ss = 2.0 * (S12 * S12 + S13 * S13 + S23 * S23) +
                            du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz;
                c->SijSij[q] = ss;
                c->S_mag[q] = sqrt(2.0 * ss);
                c->div_u[q] = du_dx + dv_dy + dw_dz;
                double g = c->Gf[q];
                if (g < 0.0) g = 0.0;
                if (g > 1.0) g = 1.0;
                double cs = cu + (cb - cu) * g;
                cs += cf * 4.0 * g * (1.0 - g);
                c->nu_t[q] = (cs * DX) * (cs * DX) * c->S_mag[q];
            }
    /* SSR-SGS turbulent flame speed (:1921-1964) with the RHS's arguments (:2700-2704) */
    const double SL = c->p.s_l, nuu = c->p.nu_unburnt_approx, D3 = c->p.fractal_d3;
    const double tau = (c->p.rho_unburnt_approx / c->p.rho_burnt_approx) - 1.0;
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                size_t q = IX(c, i, j, k);
                if (c->attr[q] != 3) { c->ST[q] = SL; continue; }
                double nul = c->nu[q];
                double num = c->SijSij[q] - (c->div_u[q] * c->div_u[q]) * (1.0 / 3.0);
                double X = (pow(DX, 4.0) / (nul * nul)) * num;
                double d_eta;
                if (X < SMALL) d_eta = 1.0;
                else {
                    double px = pow(X, 0.75);
                    double al = 0.2187 * px;
                    d_eta = 0.2340 * px * (al > 1.0 ? log(al) : 0.0);
                }
                double eta = DX / (d_eta + SMALL);
                double dF = nuu / (SL + SMALL);
                double Dc = 8.0 * eta;
                double asf = 8.0 * exp(10.0 * dF / (Dc + SMALL));
                double base = asf * eta / DX;
                double wr = pow(base, (2.0 - D3));
                double te = 0.0;
                if (tau * SL > SMALL) te = DX * fmax(0.0, c->div_u[q]) / (tau * SL);
                double ratio = wr + te;
                double st = SL * fmax(0.0, ratio);
                c->ST[q] = fmax(SL, st);
            }
}

/* ------------------------------------------------------------------------- */
/* convection schemes (:637-727, :735-827, :2017-2159, :2190-2298)            */
/* ------------------------------------------------------------------------- */
/* stride of the three directions */
static inline size_t sx(const ora_ctx* c) { return (size_t)c->ny * c->nz; }
static inline size_t sy(const ora_ctx* c) { return (size_t)c->nz; }

/* 1st-order upwind fallback shared by :642-651 and :2043-2052 */
static double upwind1(const ora_ctx* c, int i, int j, int k, const double* f,
                      const double* u, const double* v, const double* w)
{
    size_t q = IX(c, i, j, k), X = sx(c), Y = sy(c);
    double r;
    if (u[q] >= 0.0) r = u[q] * (f[q] - f[q - X]) / c->dx;
    else r = u[q] * (f[q + X] - f[q]) / c->dx;
    if (v[q] >= 0.0) r += v[q] * (f[q] - f[q - Y]) / c->dy;
    else r += v[q] * (f[q + Y] - f[q]) / c->dy;
    if (w[q] >= 0.0) r += w[q] * (f[q] - f[q - 1]) / c->dz;
    else r += w[q] * (f[q + 1] - f[q]) / c->dz;
    return r;
}

/* QUICK-coefficient face value with the reference's coefficients
 * (UW2 = 0, UW1 = DW1 = 0.5: a central average, :164-166). */
static inline double quick(double uw2, double uw1, double dw1)
{
    return (0.0 * uw2 + 0.5 * uw1 + 0.5 * dw1);
}

double ora_conv_nonconservative(const ora_ctx* c, int i, int j, int k, const double* f,
                                const double* u, const double* v, const double* w)
{   /* :637-678 */
    if (i < 2 || i >= c->nx - 2 || j < 2 || j >= c->ny - 2 || k < 2 || k >= c->nz - 2)
        return upwind1(c, i, j, k, f, u, v, w);
    size_t q = IX(c, i, j, k);
    const size_t S[3] = {sx(c), sy(c), 1};
    const double* vel[3] = {u, v, w};
    const double d[3] = {c->dx, c->dy, c->dz};
    double out[3];
    for (int a = 0; a < 3; ++a) {
        size_t s = S[a];
        double ph, mh;
        if (vel[a][q] >= 0.0) {
            ph = quick(f[q - s], f[q], f[q + s]);
            mh = quick(f[q - 2 * s], f[q - s], f[q]);
        } else {
            ph = quick(f[q + 2 * s], f[q + s], f[q]);
            mh = quick(f[q + s], f[q], f[q - s]);
        }
        out[a] = vel[a][q] * (ph - mh) / d[a];
    }
    return out[0] + out[1] + out[2];
}

double ora_conv_conservative(const ora_ctx* c, int i, int j, int k, const double* f,
                             const double* u, const double* v, const double* w,
                             const double* rho)
{   /* :683-727 */
    size_t q = IX(c, i, j, k);
    const size_t S[3] = {sx(c), sy(c), 1};
    const double* vel[3] = {u, v, w};
    const double d[3] = {c->dx, c->dy, c->dz};
    if (i < 2 || i >= c->nx - 2 || j < 2 || j >= c->ny - 2 || k < 2 || k >= c->nz - 2) {
        double acc = 0.0;
        for (int a = 0; a < 3; ++a) {
            size_t s = S[a];
            const double* U = vel[a];
            double mp = 0.5 * (rho[q + s] + rho[q]) * 0.5 * (U[q + s] + U[q]);
            double mm = 0.5 * (rho[q] + rho[q - s]) * 0.5 * (U[q] + U[q - s]);
            double fp = (mp > 0) ? f[q] : f[q + s];
            double fm = (mm > 0) ? f[q - s] : f[q];
            double term = (mp * fp - mm * fm) / d[a];
            acc = (a == 0) ? term : acc + term;
        }
        return acc;
    }
    double acc = 0.0;
    for (int a = 0; a < 3; ++a) {
        size_t s = S[a];
        const double* U = vel[a];
        double up = 0.5 * (U[q + s] + U[q]);
        double um = 0.5 * (U[q] + U[q - s]);
        double rp = 0.5 * (rho[q + s] + rho[q]);
        double rm = 0.5 * (rho[q] + rho[q - s]);
        double mp = rp * up, mm = rm * um;
        double fp, fm;
        if (up >= 0.0) fp = quick(f[q - s], f[q], f[q + s]);
        else fp = quick(f[q + 2 * s], f[q + s], f[q]);
        if (um >= 0.0) fm = quick(f[q - 2 * s], f[q - s], f[q]);
        else fm = quick(f[q + s], f[q], f[q - s]);
        double term = (mp * fp - mm * fm) / d[a];
        acc = (a == 0) ? term : acc + term;
    }
    return acc;
}

double ora_superbee(double r)                                       /* :2017-2028 */
{
    if (r <= 0.0) return 0.0;
    double t1 = fmin(2.0 * r, 1.0);
    double t2 = fmin(r, 2.0);
    return fmax(t1, t2);
}

static inline double slope_ratio(double up, double down)
{
    return (fabs(down) < SMALL) ? 1e9 : up / down;                  /* :2062 */
}

double ora_conv_muscl(const ora_ctx* c, int i, int j, int k, const double* f,
                      const double* u, const double* v, const double* w)
{   /* :2037-2159, superbee limiter */
    if (i < 2 || i >= c->nx - 2 || j < 2 || j >= c->ny - 2 || k < 2 || k >= c->nz - 2)
        return upwind1(c, i, j, k, f, u, v, w);
    size_t q = IX(c, i, j, k);
    const size_t S[3] = {sx(c), sy(c), 1};
    const double* vel[3] = {u, v, w};
    const double d[3] = {c->dx, c->dy, c->dz};
    double out[3];
    for (int a = 0; a < 3; ++a) {
        size_t s = S[a];
        double fm2 = f[q - 2 * s], fm1 = f[q - s], f0 = f[q], fp1 = f[q + s], fp2 = f[q + 2 * s];
        if (vel[a][q] >= 0.0) {
            double dd_m = f0 - fm1;
            double lim_m = ora_superbee(slope_ratio(fm1 - fm2, dd_m));
            double mh = fm1 + 0.5 * lim_m * dd_m;
            double dd_p = fp1 - f0;
            double lim_p = ora_superbee(slope_ratio(f0 - fm1, dd_p));
            double ph = f0 + 0.5 * lim_p * dd_p;
            out[a] = vel[a][q] * (ph - mh) / d[a];
        } else {
            double dd_m = f0 - fm1;
            double lim_m = ora_superbee(slope_ratio(fp1 - f0, dd_m));
            double mh = f0 - 0.5 * lim_m * dd_m;
            double dd_p = fp1 - f0;
            double lim_p = ora_superbee(slope_ratio(fp2 - fp1, dd_p));
            double ph = fp1 - 0.5 * lim_p * dd_p;
            out[a] = vel[a][q] * (ph - mh) / d[a];
        }
    }
    return out[0] + out[1] + out[2];
}

double ora_teno5(double v0, double v1, double v2, double v3, double v4)   /* :2190-2236 */
{
    double is0 = (13.0 / 12.0) * pow(v0 - 2.0 * v1 + v2, 2) + (1.0 / 4.0) * pow(v0 - 4.0 * v1 + 3.0 * v2, 2);
    double is1 = (13.0 / 12.0) * pow(v1 - 2.0 * v2 + v3, 2) + (1.0 / 4.0) * pow(v1 - v3, 2);
    double is2 = (13.0 / 12.0) * pow(v2 - 2.0 * v3 + v4, 2) + (1.0 / 4.0) * pow(3.0 * v2 - 4.0 * v3 + v4, 2);
    int d0 = is0 < 1.0e-6, d1 = is1 < 1.0e-6, d2 = is2 < 1.0e-6;
    double p0 = (2.0 * v0 - 7.0 * v1 + 11.0 * v2) / 6.0;
    double p1 = (-v1 + 5.0 * v2 + 2.0 * v3) / 6.0;
    double p2 = (2.0 * v2 + 5.0 * v3 - v4) / 6.0;
    if (d0 + d1 + d2 == 0) {
        if (is0 <= is1 && is0 <= is2) return p0;
        else if (is1 <= is0 && is1 <= is2) return p1;
        else return p2;
    }
    double w0 = d0 * 0.1, w1 = d1 * 0.6, w2 = d2 * 0.3;
    double ws = w0 + w1 + w2;
    if (ws < 1e-40) return p1;
    return (w0 * p0 + w1 * p1 + w2 * p2) / ws;
}

double ora_conv_teno(const ora_ctx* c, int i, int j, int k, const double* f,
                     const double* u, const double* v, const double* w)
{   /* :2245-2298 */
    if (i < 3 || i >= c->nx - 3 || j < 3 || j >= c->ny - 3 || k < 3 || k >= c->nz - 3)
        return ora_conv_nonconservative(c, i, j, k, f, u, v, w);
    size_t q = IX(c, i, j, k);
    const size_t S[3] = {sx(c), sy(c), 1};
    const double* vel[3] = {u, v, w};
    const double d[3] = {c->dx, c->dy, c->dz};
    double out[3];
    for (int a = 0; a < 3; ++a) {
        size_t s = S[a];
#define F(o) f[(size_t)((ptrdiff_t)q + (ptrdiff_t)(o) * (ptrdiff_t)s)]
        double ph, mh;
        if (vel[a][q] >= 0.0) {
            ph = ora_teno5(F(-2), F(-1), F(0), F(1), F(2));
            mh = ora_teno5(F(-3), F(-2), F(-1), F(0), F(1));
        } else {
            mh = ora_teno5(F(2), F(1), F(0), F(-1), F(-2));
            ph = ora_teno5(F(3), F(2), F(1), F(0), F(-1));
        }
#undef F
        out[a] = vel[a][q] * (ph - mh) / d[a];
    }
    return out[0] + out[1] + out[2];
}

double ora_weno5_gradient(const double p[6], double dx)             /* :749-827 */
{
    const double eps = 1.0e-6;
    double is0, is1, is2;
    is0 = (13.0 / 12.0) * pow(p[1] - 2.0 * p[2] + p[3], 2) + (1.0 / 4.0) * pow(p[1] - 4.0 * p[2] + 3.0 * p[3], 2);
    is1 = (13.0 / 12.0) * pow(p[2] - 2.0 * p[3] + p[4], 2) + (1.0 / 4.0) * pow(p[2] - p[4], 2);
    is2 = (13.0 / 12.0) * pow(p[3] - 2.0 * p[4] + p[5], 2) + (1.0 / 4.0) * pow(3.0 * p[3] - 4.0 * p[4] + p[5], 2);
    double a0 = (3.0 / 10.0) / pow(eps + is0, 2);
    double a1 = (6.0 / 10.0) / pow(eps + is1, 2);
    double a2 = (1.0 / 10.0) / pow(eps + is2, 2);
    double sa = a0 + a1 + a2;
    double w0 = a0 / sa, w1 = a1 / sa, w2 = a2 / sa;
    double s0 = (2.0 * p[1] - 7.0 * p[2] + 11.0 * p[3]) / 6.0;
    double s1 = (-p[2] + 5.0 * p[3] + 2.0 * p[4]) / 6.0;
    double s2 = (2.0 * p[3] + 5.0 * p[4] - p[5]) / 6.0;
    double ph = w0 * s0 + w1 * s1 + w2 * s2;
    is0 = (13.0 / 12.0) * pow(p[0] - 2.0 * p[1] + p[2], 2) + (1.0 / 4.0) * pow(p[0] - 4.0 * p[1] + 3.0 * p[2], 2);
    is1 = (13.0 / 12.0) * pow(p[1] - 2.0 * p[2] + p[3], 2) + (1.0 / 4.0) * pow(p[1] - p[3], 2);
    is2 = (13.0 / 12.0) * pow(p[2] - 2.0 * p[3] + p[4], 2) + (1.0 / 4.0) * pow(3.0 * p[2] - 4.0 * p[3] + p[4], 2);
    a0 = (1.0 / 10.0) / pow(eps + is0, 2);
    a1 = (6.0 / 10.0) / pow(eps + is1, 2);
    a2 = (3.0 / 10.0) / pow(eps + is2, 2);
    sa = a0 + a1 + a2;
    w0 = a0 / sa; w1 = a1 / sa; w2 = a2 / sa;
    s0 = (-p[0] + 5.0 * p[1] + 2.0 * p[2]) / 6.0;
    s1 = (2.0 * p[1] + 5.0 * p[2] - p[3]) / 6.0;
    s2 = (11.0 * p[2] - 7.0 * p[3] + 2.0 * p[4]) / 6.0;
    double mh = w0 * s0 + w1 * s1 + w2 * s2;
    return (ph - mh) / dx;
}

/* ------------------------------------------------------------------------- */
/* RHS (:2494-2963)                                                           */
/* ------------------------------------------------------------------------- */
static inline double flux_div(double gp, double gm, double fp, double f0, double fm, double d)
{
    /* (gamma_+ (f+ - f0)/d - gamma_- (f0 - f-)/d)/d  (:2757-2759) */
    return (gp * (fp - f0) / d - gm * (f0 - fm) / d) / d;
}

static void rhs_ref(ora_ctx* c, int t_loop, double* const* out)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double DX = c->dx, DY = c->dy, DZ = c->dz;
    const size_t X = sx(c), Y = sy(c);
    const les_params* P = &c->p;

    /* mixture properties on the filtered fields */
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) ora_props_cell(c, q);
    nu_t_and_st_ref(c);

    /* domain-mean density over attributes 1, 2, 3, 5 (:2536-2550), serial order */
    double tm = 0.0, tv = 0.0;
    const double dV = DX * DY * DZ;
    for (size_t q = 0; q < c->n; ++q) {
        int a = c->attr[q];
        if (a == 1 || a == 2 || a == 3 || a == 5) { tm += c->rhof[q] * dV; tv += dV; }
    }
    double rho_avg = (tv > 0) ? (tm / tv) : 1.2;
    c->last_rho_avg = rho_avg;

    const double* uf = c->uf; const double* vf = c->vf; const double* wf = c->wf;
    const double* rf = c->rhof; const double* Gf = c->Gf; const double* Tf = c->Tf;
    const double* nut = c->nu_t;
    const double zc = P->lz / 2.0;
    const double wdt = P->flame_accel_width, ctr = P->flame_accel_center;
    const double Fmax = P->max_flame_force;

    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                size_t q = IX(c, i, j, k);
                if (c->attr[q] != 3) {
                    for (int f = 0; f < NRHS; ++f) out[f][q] = 0.0;
                    c->gradG[q] = 0.0;
                    continue;
                }
                /* (a) momentum (:2556-2691) */
                double cu = ora_conv_conservative(c, i, j, k, uf, uf, vf, wf, rf);
                double cv = ora_conv_conservative(c, i, j, k, vf, uf, vf, wf, rf);
                double cw = ora_conv_conservative(c, i, j, k, wf, uf, vf, wf, rf);
                double mu_eff = c->mu_mix[q] + rf[q] * nut[q];
                const double* vel[3] = {uf, vf, wf};
                double diff[3];
                for (int a = 0; a < 3; ++a) {
                    const double* U = vel[a];
                    double l_x = (U[q + X] - 2.0 * U[q] + U[q - X]) / (DX * DX);
                    double l_y = (U[q + Y] - 2.0 * U[q] + U[q - Y]) / (DY * DY);
                    double l_z = (U[q + 1] - 2.0 * U[q] + U[q - 1]) / (DZ * DZ);
                    diff[a] = mu_eff * (l_x + l_y + l_z);
                }
                double grav = (rf[q] - rho_avg) * G_GRAV;
                double fx = 0.0, fy = 0.0, fz = 0.0;
                double Gl = Gf[q];
                if (Gl > 0.01 && Gl < 0.99) {
                    double gx = (Gf[q + X] - Gf[q - X]) / (2.0 * DX);
                    double gy = (Gf[q + Y] - Gf[q - Y]) / (2.0 * DY);
                    double gz = (Gf[q + 1] - Gf[q - 1]) / (2.0 * DZ);
                    if (gx < 0.0) gx = 0.0;
                    if (gy < 0.0) gy = 0.0;
                    double zrel = (k * DZ) - zc;
                    if (zrel > 0.0) { if (gz < 0.0) gz = 0.0; }
                    else { if (gz > 0.0) gz = 0.0; }
                    double df = Gl - ctr;
                    double win = exp(-(df * df) / (2.0 * wdt * wdt));
                    double tip = 1.0;
                    double common = P->flame_accel_strength * rf[q] * win * tip;
                    fx = common * gx * P->weight_x;
                    fy = common * gy * 1.0;
                    fz = common * gz * 1.0;
                    double zp = k * DZ, ramp = 1.0;
                    if (zp < P->flame_ramp_distance) ramp = zp / P->flame_ramp_distance;
                    fx *= ramp; fy *= ramp; fz *= ramp;
                }
                if (fx > Fmax) fx = Fmax;
                if (fx < -Fmax) fx = -Fmax;
                if (fy > Fmax) fy = Fmax;
                if (fy < -Fmax) fy = -Fmax;
                if (fz > Fmax) fz = Fmax;
                if (fz < -Fmax) fz = -Fmax;
                out[0][q] = -cu + diff[0] + grav + fx;
                out[1][q] = -cv + diff[1] + fy;
                out[2][q] = -cw + diff[2] + fz;

                /* (b) G (:2706-2818) */
                double conv = ora_conv_teno(c, i, j, k, Gf, uf, vf, wf);
                double g_ip = (nut[q + X] + nut[q]) / (2.0 * P->sc_t);
                double g_im = (nut[q] + nut[q - X]) / (2.0 * P->sc_t);
                double g_jp = (nut[q + Y] + nut[q]) / (2.0 * P->sc_t);
                double g_jm = (nut[q] + nut[q - Y]) / (2.0 * P->sc_t);
                double g_kp = (nut[q + 1] + nut[q]) / (2.0 * P->sc_t);
                double g_km = (nut[q] + nut[q - 1]) / (2.0 * P->sc_t);
                double sgs = flux_div(g_ip, g_im, Gf[q + X], Gf[q], Gf[q - X], DX) +
                             flux_div(g_jp, g_jm, Gf[q + Y], Gf[q], Gf[q - Y], DY) +
                             flux_div(g_kp, g_km, Gf[q + 1], Gf[q], Gf[q - 1], DZ);
                double gx, gy, gz;
                int weno = (t_loop >= P->scheme_transition_step) &&
                           (i >= 3 && i < nx - 3 && j >= 3 && j < ny - 3 && k >= 3 && k < nz - 3);
                if (!weno) {
                    gx = (Gf[q + X] - Gf[q - X]) / (2.0 * DX);
                    gy = (Gf[q + Y] - Gf[q - Y]) / (2.0 * DY);
                    gz = (Gf[q + 1] - Gf[q - 1]) / (2.0 * DZ);
                } else {
                    double pp[6];
                    for (int o = 0; o < 6; ++o) pp[o] = Gf[q + (o - 2) * (ptrdiff_t)X];
                    gx = ora_weno5_gradient(pp, DX);
                    for (int o = 0; o < 6; ++o) pp[o] = Gf[q + (o - 2) * (ptrdiff_t)Y];
                    gy = ora_weno5_gradient(pp, DY);
                    for (int o = 0; o < 6; ++o) pp[o] = Gf[q + (o - 2)];
                    gz = ora_weno5_gradient(pp, DZ);
                }
                double gmag = sqrt(gx * gx + gy * gy + gz * gz);
                c->gradG[q] = gmag;
                double prop = c->ST[q] * gmag;
                double art = 0.0;
                if (Gl > 0.1 && Gl < 0.99) {
                    double df = Gl - ctr;
                    double win = exp(-(df * df) / (2.0 * wdt * wdt));
                    double lx = (Gf[q + X] - 2.0 * Gl + Gf[q - X]) / (DX * DX);
                    double ly = (Gf[q + Y] - 2.0 * Gl + Gf[q - Y]) / (DY * DY);
                    double lz = (Gf[q + 1] - 2.0 * Gl + Gf[q - 1]) / (DZ * DZ);
                    double rp = fabs(j * DY);
                    double tr = P->tip_protection_radius;
                    double tip = 1.0 - exp(-(rp * rp) / (tr * tr));
                    double coef = 5.0 * (nut[q] + 1.0e-5) * win * tip;
                    art = coef * (lx + ly + lz);
                }
                out[4][q] = -conv + prop + sgs + art;

                /* (c) species (:2837-2896) */
                double gv = Gl;
                if (gv < 0.0) gv = 0.0;
                if (gv > 1.0) gv = 1.0;
                double d_ip = (nut[q + X] + nut[q]) / (2.0 * P->sc_t), d_im = (nut[q] + nut[q - X]) / (2.0 * P->sc_t);
                double d_jp = (nut[q + Y] + nut[q]) / (2.0 * P->sc_t), d_jm = (nut[q] + nut[q - Y]) / (2.0 * P->sc_t);
                double d_kp = (nut[q + 1] + nut[q]) / (2.0 * P->sc_t), d_km = (nut[q] + nut[q - 1]) / (2.0 * P->sc_t);
                for (int s = 0; s < NSP; ++s) {
                    const double* Yf = c->Yf[s];
                    double cy = ora_conv_muscl(c, i, j, k, Yf, uf, vf, wf);
                    double lx = (Yf[q + X] - 2.0 * Yf[q] + Yf[q - X]) / (DX * DX);
                    double ly = (Yf[q + Y] - 2.0 * Yf[q] + Yf[q - Y]) / (DY * DY);
                    double lz = (Yf[q + 1] - 2.0 * Yf[q] + Yf[q - 1]) / (DZ * DZ);
                    double mol = c->Dk[s][q] * (lx + ly + lz);
                    double sg = flux_div(d_ip, d_im, Yf[q + X], Yf[q], Yf[q - X], DX) +
                                flux_div(d_jp, d_jm, Yf[q + Y], Yf[q], Yf[q - Y], DY) +
                                flux_div(d_kp, d_km, Yf[q + 1], Yf[q], Yf[q - 1], DZ);
                    double tgt;
                    if (s == 0) tgt = c->inletY[0] * (1.0 - gv);
                    else if (s == 1) { double cons = c->inletY[0] * c->st_o2; tgt = fmax(0.0, c->inletY[1] - cons * gv); }
                    else if (s == 3) { double pr = c->inletY[0] * c->st_co2; tgt = c->inletY[3] + pr * gv; }
                    else if (s == 4) { double pr = c->inletY[0] * c->st_h2o; tgt = c->inletY[4] + pr * gv; }
                    else tgt = c->inletY[2];
                    double nud = P->relaxation_coeff * (tgt - Yf[q]);
                    out[5 + s][q] = -cy + mol + sg + nud;
                }

                /* (d) temperature (:2912-2955) */
                double ct = ora_conv_muscl(c, i, j, k, Tf, uf, vf, wf);
                double tx = (Tf[q + X] - 2.0 * Tf[q] + Tf[q - X]) / (DX * DX);
                double ty = (Tf[q + Y] - 2.0 * Tf[q] + Tf[q - Y]) / (DY * DY);
                double tz = (Tf[q + 1] - 2.0 * Tf[q] + Tf[q - 1]) / (DZ * DZ);
                double md = 0.0;
                if (rf[q] > 1e-6 && c->cp_mix[q] > 1e-6)
                    md = c->lam_mix[q] / (rf[q] * c->cp_mix[q]) * (tx + ty + tz);
                double k_ip = (nut[q + X] + nut[q]) / (2.0 * P->pr_t), k_im = (nut[q] + nut[q - X]) / (2.0 * P->pr_t);
                double k_jp = (nut[q + Y] + nut[q]) / (2.0 * P->pr_t), k_jm = (nut[q] + nut[q - Y]) / (2.0 * P->pr_t);
                double k_kp = (nut[q + 1] + nut[q]) / (2.0 * P->pr_t), k_km = (nut[q] + nut[q - 1]) / (2.0 * P->pr_t);
                double sd = flux_div(k_ip, k_im, Tf[q + X], Tf[q], Tf[q - X], DX) +
                            flux_div(k_jp, k_jm, Tf[q + Y], Tf[q], Tf[q - Y], DY) +
                            flux_div(k_kp, k_km, Tf[q + 1], Tf[q], Tf[q - 1], DZ);
                double Ttgt = P->inlet_temperature + gv * (P->burnt_temperature - P->inlet_temperature);
                double nT = P->relaxation_coeff * (Ttgt - Tf[q]);
                out[3][q] = -ct + md + sd + nT;
            }
}

/* Full RHS on a given state (filter + properties + RHS). */
static void rhs_full_ref(ora_ctx* c, int t_loop, const double* u, const double* v, const double* w,
                         const double* T, const double* G, double* const* Y, const double* rho,
                         double* const* out)
{
    filter_ref(c, u, v, w, T, G, Y, rho);
    rhs_ref(c, t_loop, out);
}

/* ------------------------------------------------------------------------- */
/* SEM inflow (:900-1075)                                                     */
/* ------------------------------------------------------------------------- */
static void update_vortices_ref(ora_ctx* c, double dt)              /* :900-939 */
{
    double urms = c->p.inlet_velocity * c->p.turbulence_intensity;
    double L = c->p.turbulent_length_scale;
    double box_x = L, box_y = c->p.ly, box_z = c->p.lz;
    for (int v = 0; v < c->nvort; ++v) {
        ora_vortex* e = &c->vort[v];
        e->x += c->p.inlet_velocity * dt;
        if (e->x > box_x / 2.0) {
            e->x -= box_x;
            double ry = (double)ora_rand(c) / (double)RAND_MAX;
            double rz = (double)ora_rand(c) / (double)RAND_MAX;
            e->y = ry * box_y;
            e->z = rz * box_z;
            double ru = (double)ora_rand(c) / (double)RAND_MAX;
            double rv = (double)ora_rand(c) / (double)RAND_MAX;
            double rw = (double)ora_rand(c) / (double)RAND_MAX;
            e->su = (ru * 2.0 - 1.0) * urms;
            e->sv = (rv * 2.0 - 1.0) * urms;
            e->sw = (rw * 2.0 - 1.0) * urms;
            double rr = (double)ora_rand(c) / (double)RAND_MAX;
            double rmin = L * 0.5, rmax = L * 1.5;
            e->radius = rmin + (rmax - rmin) * rr;
        }
    }
}

int ora_sem_plane(ora_ctx* c, int t_loop, double dt)                /* :943-1075 */
{
    (void)t_loop;
    const int ny = c->ny, nz = c->nz;
    update_vortices_ref(c, dt);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            size_t p = (size_t)j * nz + k;
            double su = 0.0, sv = 0.0, sw = 0.0;
            if (c->attr[IX(c, 0, j, k)] == 1) {
                double yp = j * c->dy, zp = k * c->dz;
                for (int v = 0; v < c->nvort; ++v) {
                    const ora_vortex* e = &c->vort[v];
                    double ddx = 0.0 - e->x, ddy = yp - e->y, ddz = zp - e->z;
                    double dist = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                    double inf = 0.0;
                    if (dist < e->radius) inf = 1.0 - (dist / e->radius);
                    su += e->su * inf;
                    sv += e->sv * inf;
                    sw += e->sw * inf;
                }
            }
            c->ufl[p] = su; c->vfl[p] = sv; c->wfl[p] = sw;
        }
    /* mean removal and rms rescale in serial (j, k) order (:1006-1074) */
    double tu = 0, tv = 0, tw = 0;
    int nin = 0;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
            if (c->attr[IX(c, 0, j, k)] == 1) {
                size_t p = (size_t)j * nz + k;
                tu += c->ufl[p]; tv += c->vfl[p]; tw += c->wfl[p]; nin++;
            }
    if (nin > 0) {
        double au = tu / nin, av = tv / nin, aw = tw / nin;
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                if (c->attr[IX(c, 0, j, k)] == 1) {
                    size_t p = (size_t)j * nz + k;
                    c->ufl[p] -= au; c->vfl[p] -= av; c->wfl[p] -= aw;
                }
    }
    double qu = 0, qv = 0, qw = 0;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
            if (c->attr[IX(c, 0, j, k)] == 1) {
                size_t p = (size_t)j * nz + k;
                qu += c->ufl[p] * c->ufl[p]; qv += c->vfl[p] * c->vfl[p]; qw += c->wfl[p] * c->wfl[p];
            }
    if (nin > 0) {
        double ru = sqrt(qu / nin), rv = sqrt(qv / nin), rw = sqrt(qw / nin);
        double tgt = c->p.inlet_velocity * c->p.turbulence_intensity;
        double s_u = (ru > SMALL) ? tgt / ru : 0.0;
        double s_v = (rv > SMALL) ? tgt / rv : 0.0;
        double s_w = (rw > SMALL) ? tgt / rw : 0.0;
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                if (c->attr[IX(c, 0, j, k)] == 1) {
                    size_t p = (size_t)j * nz + k;
                    c->ufl[p] *= s_u; c->vfl[p] *= s_v; c->wfl[p] *= s_w;
                }
    }
    return LES_OK;
}

/* ------------------------------------------------------------------------- */
/* boundary conditions (:2967-3161)                                           */
/* ------------------------------------------------------------------------- */
static void copy_cell(ora_ctx* c, size_t dst, size_t src, int with_velocity)
{
    if (with_velocity) { c->u[dst] = c->u[src]; c->v[dst] = c->v[src]; c->w[dst] = c->w[src]; }
    c->T[dst] = c->T[src]; c->G[dst] = c->G[src];
    for (int s = 0; s < NSP; ++s) c->Y[s][dst] = c->Y[s][src];
}

static void ambient_cell(ora_ctx* c, size_t q)
{
    c->u[q] = 0.0; c->v[q] = 0.0; c->w[q] = 0.0;
    c->T[q] = c->p.ambient_temperature; c->G[q] = 0.0;
    for (int s = 0; s < NSP; ++s) c->Y[s][q] = c->ambY[s];
}

int ora_boundary(ora_ctx* c)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t X = sx(c), Y = sy(c);
    const double U = c->p.inlet_velocity;
    /* 1. inflow / outflow faces (:2978-3043); serial (j, k) order keeps the
     * rand() stream of the inlet noise deterministic */
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            size_t q0 = IX(c, 0, j, k);
            int a = c->attr[q0];
            if (a == 1) {
                double amp = c->p.inlet_noise_amp * U;
                double r1 = ((double)ora_rand(c) / RAND_MAX) * 2.0 - 1.0;
                double r2 = ((double)ora_rand(c) / RAND_MAX) * 2.0 - 1.0;
                double r3 = ((double)ora_rand(c) / RAND_MAX) * 2.0 - 1.0;
                size_t p = (size_t)j * nz + k;
                c->u[q0] = fmax(0.0, U + c->ufl[p] + amp * r1);
                c->v[q0] = c->vfl[p] + amp * r2;
                c->w[q0] = c->wfl[p] + amp * r3;
                c->T[q0] = c->p.inlet_temperature;
                c->G[q0] = 0.0;
                for (int s = 0; s < NSP; ++s) c->Y[s][q0] = c->inletY[s];
            } else if (a == 4) {
                if (c->u[q0 + X] < 0) ambient_cell(c, q0);
                else copy_cell(c, q0, q0 + X, 1);
            } else if (a == 6) {
                c->u[q0] = 0.0; c->v[q0] = 0.0; c->w[q0] = 0.0;
                copy_cell(c, q0, q0 + X, 0);
            }
            size_t qn = IX(c, nx - 1, j, k);
            if (c->attr[qn] == 2) copy_cell(c, qn, qn - X, 1);
        }
    /* 2. side faces (:3049-3123): y faces for all k, then z faces for all j */
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i) {
        for (int k = 0; k < nz; ++k) {
            size_t q = IX(c, i, 0, k);
            if (c->attr[q] == 5) {
                c->T[q] = c->T[q + Y]; c->G[q] = c->G[q + Y];
                for (int s = 0; s < NSP; ++s) c->Y[s][q] = c->Y[s][q + Y];
                c->u[q] = c->u[q + Y]; c->v[q] = 0.0; c->w[q] = c->w[q + Y];
            }
            q = IX(c, i, ny - 1, k);
            if (c->attr[q] == 4) {
                if (c->v[q - Y] < 0) ambient_cell(c, q);
                else copy_cell(c, q, q - Y, 1);
            }
        }
        for (int j = 0; j < ny; ++j) {
            size_t q = IX(c, i, j, 0);
            if (c->attr[q] == 4) {
                if (c->w[q + 1] > 0) ambient_cell(c, q);
                else copy_cell(c, q, q + 1, 1);
            }
            q = IX(c, i, j, nz - 1);
            if (c->attr[q] == 4) {
                if (c->w[q - 1] < 0) ambient_cell(c, q);
                else copy_cell(c, q, q - 1, 1);
            }
        }
    }
    /* 3. EOS, momentum and mass-fraction normalisation on every cell (:3130-3160) */
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        double s = 0;
        for (int a = 0; a < NSP; ++a) s += c->Y[a][q] / (c->m.sp[a].mw / 1000.0);
        if (s > 1e-9 && c->T[q] > 1.0) c->rho[q] = P_CONST / (R_UNIV / (1.0 / s) * c->T[q]);
        else c->rho[q] = 1.2;
        c->rhou[q] = c->rho[q] * c->u[q];
        c->rhov[q] = c->rho[q] * c->v[q];
        c->rhow[q] = c->rho[q] * c->w[q];
        int a = c->attr[q];
        if (a == 3 || a == 2 || a == 5) {
            double sy_ = 0.0;
            for (int b = 0; b < NSP; ++b) sy_ += c->Y[b][q];
            if (sy_ > 1e-6) for (int b = 0; b < NSP; ++b) c->Y[b][q] /= sy_;
        }
    }
    return LES_OK;
}

/* ------------------------------------------------------------------------- */
/* RK4 step (:4390-4603)                                                      */
/* ------------------------------------------------------------------------- */
static void eos_fields(ora_ctx* c, const double* T, double* const* Y, double* rho_out)
{   /* update_rho_from_EOS (:1883-1910); fallback reads the global rho */
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        int a = c->attr[q];
        if (a != 3 && a != 2) continue;
        double s = 0.0;
        for (int b = 0; b < NSP; ++b) s += Y[b][q] / (c->m.sp[b].mw / 1000.0);
        if (s > 1e-9 && T[q] > 1.0) {
            double wtm = 1.0 / s;
            double rmix = R_UNIV / wtm;
            rho_out[q] = P_CONST / (rmix * T[q]);
        } else {
            rho_out[q] = c->rho[q];
        }
    }
}

static void stage_combine(ora_ctx* c, double coef_dt, int stage)
{   /* intermediate states (:4432-4455, :4461-4483, :4489-4511) */
    double* const* kk = c->k[stage];
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        if (c->attr[q] == 3) {
            c->Tt[q] = c->T0[q] + coef_dt * kk[3][q];
            c->Gt[q] = c->G0[q] + coef_dt * kk[4][q];
            for (int s = 0; s < NSP; ++s) c->Yt[s][q] = c->Y0[s][q] + coef_dt * kk[5 + s][q];
            c->Tt[q] = fmax(200.0, fmin(3000.0, c->Tt[q]));
            c->Gt[q] = fmax(0.0, fmin(1.0, c->Gt[q]));
            double sum = 0;
            for (int s = 0; s < NSP; ++s) { c->Yt[s][q] = fmax(0.0, fmin(1.0, c->Yt[s][q])); sum += c->Yt[s][q]; }
            if (sum > SMALL) for (int s = 0; s < NSP; ++s) c->Yt[s][q] /= sum;
        } else {
            c->Tt[q] = c->T0[q]; c->Gt[q] = c->G0[q];
            for (int s = 0; s < NSP; ++s) c->Yt[s][q] = c->Y0[s][q];
        }
    }
    eos_fields(c, c->Tt, c->Yt, c->rhot);
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        if (c->attr[q] == 3) {
            double ru = c->rhou0[q] + coef_dt * kk[0][q];
            double rv = c->rhov0[q] + coef_dt * kk[1][q];
            double rw = c->rhow0[q] + coef_dt * kk[2][q];
            if (c->rhot[q] > SMALL) { c->ut[q] = ru / c->rhot[q]; c->vt[q] = rv / c->rhot[q]; c->wt[q] = rw / c->rhot[q]; }
            else { c->ut[q] = 0; c->vt[q] = 0; c->wt[q] = 0; }
        } else {
            c->ut[q] = c->u[q]; c->vt[q] = c->v[q]; c->wt[q] = c->w[q];
        }
    }
}

static void final_combine(ora_ctx* c, double DT)
{   /* (:4517-4544) */
    double* const* k1 = c->k[0]; double* const* k2 = c->k[1];
    double* const* k3 = c->k[2]; double* const* k4 = c->k[3];
    const double h = DT / 6.0;
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        if (c->attr[q] != 3) continue;
        c->T[q] = c->T0[q] + h * (k1[3][q] + 2.0 * k2[3][q] + 2.0 * k3[3][q] + k4[3][q]);
        c->G[q] = c->G0[q] + h * (k1[4][q] + 2.0 * k2[4][q] + 2.0 * k3[4][q] + k4[4][q]);
        for (int s = 0; s < NSP; ++s)
            c->Y[s][q] = c->Y0[s][q] + h * (k1[5 + s][q] + 2.0 * k2[5 + s][q] + 2.0 * k3[5 + s][q] + k4[5 + s][q]);
        c->T[q] = fmax(200.0, fmin(3000.0, c->T[q]));
        c->G[q] = fmax(0.0, fmin(1.0, c->G[q]));
        double sum = 0;
        for (int s = 0; s < NSP; ++s) { c->Y[s][q] = fmax(0.0, fmin(1.0, c->Y[s][q])); sum += c->Y[s][q]; }
        if (sum > SMALL) for (int s = 0; s < NSP; ++s) c->Y[s][q] /= sum;
    }
    eos_fields(c, c->T, c->Y, c->rho);
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        if (c->attr[q] != 3) continue;
        c->rhou[q] = c->rhou0[q] + h * (k1[0][q] + 2.0 * k2[0][q] + 2.0 * k3[0][q] + k4[0][q]);
        c->rhov[q] = c->rhov0[q] + h * (k1[1][q] + 2.0 * k2[1][q] + 2.0 * k3[1][q] + k4[1][q]);
        c->rhow[q] = c->rhow0[q] + h * (k1[2][q] + 2.0 * k2[2][q] + 2.0 * k3[2][q] + k4[2][q]);
        if (c->rho[q] > SMALL) {
            c->u[q] = c->rhou[q] / c->rho[q]; c->v[q] = c->rhov[q] / c->rho[q]; c->w[q] = c->rhow[q] / c->rho[q];
        } else { c->u[q] = 0; c->v[q] = 0; c->w[q] = 0; }
    }
}

double ora_dt_ref(ora_ctx* c, double* max_vel_out)                  /* :4399-4411 */
{
    double mv = 0.0;
    for (size_t q = 0; q < c->n; ++q) {
        int a = c->attr[q];
        if (a == 1 || a == 3) {
            double vel = sqrt(c->u[q] * c->u[q] + c->v[q] * c->v[q] + c->w[q] * c->w[q]);
            if (vel > mv) mv = vel;
        }
    }
    double maxvel = mv > SMALL ? mv : c->p.inlet_velocity;         /* :1484 */
    double dtc = c->p.courant * c->dx / (maxvel + SMALL);
    double mn = 0.0, ma = 0.0;
    for (size_t q = 0; q < c->n; ++q)
        if (c->attr[q] == 3) {
            if (c->nu[q] > mn) mn = c->nu[q];
            if (c->alpha[q] > ma) ma = c->alpha[q];
        }
    double dtn = c->p.courant_diffusion * (c->dx * c->dx) / (mn + SMALL);
    double dta = c->p.courant_diffusion * (c->dx * c->dx) / (ma + SMALL);
    if (max_vel_out) *max_vel_out = maxvel;
    return fmin(dtc, fmin(dtn, dta));
}

static void probes_ref(ora_ctx* c, int t_loop)                      /* :3403-3471 */
{
    const int kp = c->nz / 2;
    for (int n = 0; n < c->p.num_probe_x; ++n) {
        double xp = c->p.probe_x[n];
        int ip = (int)(xp / c->dx);
        if (ip <= 0 || ip >= c->nx - 1) continue;
        double yf = -1.0;
        for (int j = 0; j < c->ny - 1; ++j) {
            double g1 = c->G[IX(c, ip, j, kp)], g2 = c->G[IX(c, ip, j + 1, kp)];
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
            if (fabs(yr) > 1.0) continue;                           /* SAMPLING_RANGE 1 m */
            if (c->n_probe >= c->probe_cap) return;
            size_t q = IX(c, ip, j, kp);
            les_probe_row* r = &c->probe[c->n_probe++];
            r->timestep = t_loop; r->station = n;
            r->x_location = xp; r->y_flame = yf; r->y_physical = yp; r->y_relative = yr;
            r->u = c->u[q]; r->v = c->v[q]; r->w = c->w[q]; r->G = c->G[q];
            r->nu_t = c->nu_t[q]; r->S_mag = c->S_mag[q];
        }
    }
}

static void stats_ref(ora_ctx* c)                                   /* :3167-3192 */
{
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) {
        double u = c->u[q], v = c->v[q], w = c->w[q];
        c->stat[0][q] += u; c->stat[1][q] += v; c->stat[2][q] += w; c->stat[3][q] += c->G[q];
        c->stat[4][q] += u * u; c->stat[5][q] += v * v; c->stat[6][q] += w * w;
        c->stat[7][q] += u * v; c->stat[8][q] += u * w; c->stat[9][q] += v * w;
    }
    c->stat_count++;
}

static void monitor(const ora_ctx* c, les_step_report* rep)        /* :3880-3897 */
{
    const double dV = c->dx * c->dy * c->dz;
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, mnY = 1e9;
    int bad = 0;
    for (size_t q = 0; q < c->n; ++q) {
        tm += c->rho[q] * dV;
        if (c->T[q] < mnT) mnT = c->T[q];
        if (c->T[q] > mxT) mxT = c->T[q];
        if (c->G[q] < mnG) mnG = c->G[q];
        if (c->G[q] > mxG) mxG = c->G[q];
        if (c->Y[0][q] < mnY) mnY = c->Y[0][q];
        if (!isfinite(c->u[q]) || !isfinite(c->v[q]) || !isfinite(c->w[q]) || !isfinite(c->T[q]) || !isfinite(c->G[q])) bad = 1;
    }
    rep->total_mass = tm; rep->min_T = mnT; rep->max_T = mxT;
    rep->min_G = mnG; rep->max_G = mxG; rep->min_Y_fuel = mnY; rep->nonfinite = bad;
}

static int step_ref(ora_ctx* c, les_step_report* rep)
{
    const int t = c->step;
    double maxvel = 0.0;
    double DT = ora_dt_ref(c, &maxvel);
    ora_sem_plane(c, t, DT);
    /* save Phi_n (:4421-4426) */
    memcpy(c->rhou0, c->rhou, c->n * sizeof(double));
    memcpy(c->rhov0, c->rhov, c->n * sizeof(double));
    memcpy(c->rhow0, c->rhow, c->n * sizeof(double));
    memcpy(c->T0, c->T, c->n * sizeof(double));
    memcpy(c->G0, c->G, c->n * sizeof(double));
    for (int s = 0; s < NSP; ++s) memcpy(c->Y0[s], c->Y[s], c->n * sizeof(double));

    rhs_full_ref(c, t, c->u, c->v, c->w, c->T, c->G, c->Y, c->rho, c->k[0]);
    stage_combine(c, 0.5 * DT, 0);
    rhs_full_ref(c, t, c->ut, c->vt, c->wt, c->Tt, c->Gt, c->Yt, c->rhot, c->k[1]);
    stage_combine(c, 0.5 * DT, 1);
    rhs_full_ref(c, t, c->ut, c->vt, c->wt, c->Tt, c->Gt, c->Yt, c->rhot, c->k[2]);
    stage_combine(c, 1.0 * DT, 2);
    rhs_full_ref(c, t, c->ut, c->vt, c->wt, c->Tt, c->Gt, c->Yt, c->rhot, c->k[3]);
    final_combine(c, DT);
    ora_boundary(c);
    probes_ref(c, t);
    if (c->p.stats_start_step >= 0 && t >= c->p.stats_start_step) stats_ref(c);
    if (rep) {
        rep->step = t; rep->dt = DT; rep->max_velocity = maxvel;
        monitor(c, rep);
        rep->proj_residual = 0.0;
        rep->proj_residual0 = 0.0;
    }
    c->step++;
    return LES_OK;
}

int ora_step(ora_ctx* c, int n_steps, les_step_report* rep)
{
    for (int s = 0; s < n_steps; ++s) {
        int rc = (c->p.mode == LES_MODE_DUCT) ? ora_duct_step(c, rep) : step_ref(c, rep);
        if (rc != LES_OK) return rc;
    }
    return LES_OK;
}

int ora_rhs(ora_ctx* c, int t_loop, double* const* out, int n_out)
{
    if (n_out < NRHS) return LES_ERR_PARAM;
    if (c->p.mode == LES_MODE_DUCT) return ora_duct_rhs(c, out, n_out);
    rhs_full_ref(c, t_loop, c->u, c->v, c->w, c->T, c->G, c->Y, c->rho, out);
    return LES_OK;
}

int ora_filter_and_props(ora_ctx* c, const double* const* in)
{
    /* in: u, v, w, T, G, rho, Y0..Y4 */
    double* Yin[NSP];
    for (int s = 0; s < NSP; ++s) Yin[s] = (double*)in[6 + s];
    filter_ref(c, in[0], in[1], in[2], in[3], in[4], Yin, in[5]);
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < c->n; ++q) ora_props_cell(c, q);
    nu_t_and_st_ref(c);
    return LES_OK;
}

/* ------------------------------------------------------------------------- */
/* accessors                                                                  */
/* ------------------------------------------------------------------------- */
int ora_set_state(ora_ctx* c, const double* const* f, int nf)
{
    if (c->p.mode == LES_MODE_DUCT) return ora_duct_set_state(c, f, nf);
    if (nf < LES_NFIELDS) return LES_ERR_PARAM;
    size_t b = c->n * sizeof(double);
    memcpy(c->u, f[0], b); memcpy(c->v, f[1], b); memcpy(c->w, f[2], b);
    memcpy(c->T, f[3], b); memcpy(c->G, f[4], b); memcpy(c->rho, f[5], b);
    for (int s = 0; s < NSP; ++s) memcpy(c->Y[s], f[6 + s], b);
    for (size_t q = 0; q < c->n; ++q) {
        c->rhou[q] = c->rho[q] * c->u[q];
        c->rhov[q] = c->rho[q] * c->v[q];
        c->rhow[q] = c->rho[q] * c->w[q];
    }
    if (c->attr[0] == 0) attribute_map(c);
    return LES_OK;
}

int ora_get_state(const ora_ctx* c, double* const* f, int nf)
{
    if (c->p.mode == LES_MODE_DUCT) return ora_duct_get_state(c, f, nf);
    const double* src[LES_NFIELDS] = {c->u, c->v, c->w, c->T, c->G, c->rho,
                                      c->Y[0], c->Y[1], c->Y[2], c->Y[3], c->Y[4]};
    for (int a = 0; a < nf && a < LES_NFIELDS; ++a) memcpy(f[a], src[a], c->n * sizeof(double));
    return LES_OK;
}

static double* field_by_name(const ora_ctx* c, const char* name)
{
    struct { const char* n; double* p; } tab[] = {
        {"u", c->u}, {"v", c->v}, {"w", c->w}, {"T", c->T}, {"G", c->G}, {"rho", c->rho},
        {"rhou", c->rhou}, {"rhov", c->rhov}, {"rhow", c->rhow},
        {"nu_t", c->nu_t}, {"S_mag", c->S_mag}, {"nu", c->nu}, {"alpha", c->alpha},
        {"mu_mix", c->mu_mix}, {"cp_mix", c->cp_mix}, {"lambda_mix", c->lam_mix},
        {"ST", c->ST}, {"SijSij", c->SijSij}, {"div_u", c->div_u}, {"gradG", c->gradG},
        {"u_f", c->uf}, {"v_f", c->vf}, {"w_f", c->wf}, {"T_f", c->Tf}, {"G_f", c->Gf},
        {"rho_f", c->rhof}, {"rho_temp", c->rhot},
        {"Y0", c->Y[0]}, {"Y1", c->Y[1]}, {"Y2", c->Y[2]}, {"Y3", c->Y[3]}, {"Y4", c->Y[4]},
        {"Y0_f", c->Yf[0]}, {"Y1_f", c->Yf[1]}, {"Y2_f", c->Yf[2]}, {"Y3_f", c->Yf[3]}, {"Y4_f", c->Yf[4]},
        {"D0", c->Dk[0]}, {"D1", c->Dk[1]}, {"D2", c->Dk[2]}, {"D3", c->Dk[3]}, {"D4", c->Dk[4]},
    };
    for (size_t a = 0; a < sizeof(tab) / sizeof(tab[0]); ++a)
        if (strcmp(tab[a].n, name) == 0) return tab[a].p;
    return NULL;
}

int ora_get_field(const ora_ctx* c, const char* name, double* out)
{
    if (strcmp(name, "attribute") == 0) {
        for (size_t q = 0; q < c->n; ++q) out[q] = (double)c->attr[q];
        return LES_OK;
    }
    if (c->p.mode == LES_MODE_DUCT) {
        int rc = ora_duct_get_field(c, name, out);
        if (rc != LES_ERR_PARAM) return rc;
    }
    double* f = field_by_name(c, name);
    if (!f) return LES_ERR_PARAM;
    memcpy(out, f, c->n * sizeof(double));
    return LES_OK;
}

int ora_set_field(ora_ctx* c, const char* name, const double* in)
{
    if (strcmp(name, "attribute") == 0) {
        for (size_t q = 0; q < c->n; ++q) c->attr[q] = (int)in[q];
        return LES_OK;
    }
    double* f = field_by_name(c, name);
    if (!f) return LES_ERR_PARAM;
    memcpy(f, in, c->n * sizeof(double));
    return LES_OK;
}

int ora_probe_fetch(ora_ctx* c, les_probe_row* rows, int max_rows, int* n_rows)
{
    int n = c->n_probe < max_rows ? c->n_probe : max_rows;
    memcpy(rows, c->probe, (size_t)n * sizeof(les_probe_row));
    if (n < c->n_probe)
        memmove(c->probe, c->probe + n, (size_t)(c->n_probe - n) * sizeof(les_probe_row));
    c->n_probe -= n;
    *n_rows = n;
    return LES_OK;
}

int ora_probe_now(ora_ctx* c, int t_loop)
{
    if (c->p.mode != LES_MODE_REFERENCE) return LES_ERR_PARAM;
    probes_ref(c, t_loop);
    return LES_OK;
}

int ora_get_step(const ora_ctx* c) { return c->step; }
void ora_set_step(ora_ctx* c, int step) { c->step = step; }

int ora_get_vortices(const ora_ctx* c, double* out, int max_n)
{
    int n = c->nvort < max_n ? c->nvort : max_n;
    for (int v = 0; v < n; ++v) {
        const ora_vortex* e = &c->vort[v];
        double* o = out + 7 * v;
        o[0] = e->x; o[1] = e->y; o[2] = e->z; o[3] = e->radius; o[4] = e->su; o[5] = e->sv; o[6] = e->sw;
    }
    return n;
}

int ora_set_vortices(ora_ctx* c, const double* in, int n)
{
    if (n != c->nvort) return LES_ERR_PARAM;
    for (int v = 0; v < n; ++v) {
        ora_vortex* e = &c->vort[v];
        const double* o = in + 7 * v;
        e->x = o[0]; e->y = o[1]; e->z = o[2]; e->radius = o[3]; e->su = o[4]; e->sv = o[5]; e->sw = o[6];
    }
    return LES_OK;
}

int ora_get_inlet(const ora_ctx* c, double* uf, double* vf, double* wf)
{
    size_t b = (size_t)c->ny * c->nz * sizeof(double);
    memcpy(uf, c->ufl, b); memcpy(vf, c->vfl, b); memcpy(wf, c->wfl, b);
    return LES_OK;
}

void ora_get_dt_state(const ora_ctx* c, double* max_nu, double* max_alpha)
{
    double mn = 0.0, ma = 0.0;
    for (size_t q = 0; q < c->n; ++q)
        if (c->attr[q] == 3) {
            if (c->nu[q] > mn) mn = c->nu[q];
            if (c->alpha[q] > ma) ma = c->alpha[q];
        }
    *max_nu = mn; *max_alpha = ma;
}

void ora_set_dt_state(ora_ctx* c, double max_nu, double max_alpha)
{
    /* place the maxima on the fluid cells so that ora_dt_ref reads them */
    for (size_t q = 0; q < c->n; ++q) {
        c->nu[q] = (c->attr[q] == 3) ? max_nu : 0.0;
        c->alpha[q] = (c->attr[q] == 3) ? max_alpha : 0.0;
    }
}

int ora_get_stats(const ora_ctx* c, double* const* acc, int* count)
{
    for (int a = 0; a < 10; ++a) if (acc[a]) memcpy(acc[a], c->stat[a], c->n * sizeof(double));
    *count = c->stat_count;
    return LES_OK;
}
