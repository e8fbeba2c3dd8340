// les_ref_kernels.hpp - pointer bundles shared by the reference-scheme kernels
// and the engine that launches them (all passed to kernels by value).
#pragma once
#include "les_device.hpp"

// Inputs / outputs of one RHS evaluation (calculate_RHS, :2494-2963).
struct RefPtrs {
    const int* attr;
    const les_mixture* mix;
    // RHS input state (the step state or the RK intermediate state)
    const double *u, *v, *w, *T, *G, *rho;
    const double* Y[LES_NSP];
    // filtered fields
    double *uf, *vf, *wf, *Tf, *Gf, *rhof;
    double* Yf[LES_NSP];
    // properties and SGS quantities
    double *nu, *mu_mix, *cp, *lam, *alpha, *nu_t, *S_mag, *ST, *gradG;
    double* Dk[LES_NSP];
    // RHS outputs: drhou, drhov, drhow, dT, dG, dY0..dY4
    double* out[LES_NRHS];
    const double* rho_avg;
};

// State, Phi_n, intermediate state and stage derivatives of the RK update.
struct RefStagePtrs {
    const int* attr;
    double *u, *v, *w, *T, *G, *rho, *rhou, *rhov, *rhow;
    double* Y[LES_NSP];
    double *rhou0, *rhov0, *rhow0, *T0, *G0;
    double* Y0[LES_NSP];
    double *ut, *vt, *wt, *Tt, *Gt, *rhot;
    double* Yt[LES_NSP];
    double* k[4][LES_NRHS];
};

// Output groups of a split RHS or property launch (small grids): n slabs of
// nx blocks, slab b computing the outputs in mask[b].  RHS bits: the ten
// outputs drhou, drhov, drhow, dT, dG, dY0..dY4.  Property bits: bit s the
// diffusivity of species s, then the filtered fields with cp, then the
// transport properties nu, mu, lambda, alpha.
#define REF_GROUPS_MAX 12
#define REF_RHS_ALL ((1u << LES_NRHS) - 1u)
#define REF_PROPS_FIELDS (1u << LES_NSP)
#define REF_PROPS_TRANSPORT (1u << (LES_NSP + 1))
#define REF_PROPS_ALL ((1u << (LES_NSP + 2)) - 1u)
struct RefGroups {
    int n;
    unsigned mask[REF_GROUPS_MAX];
};

// The DT rule (:4399-4411) from the three exact maxima; the host (for the
// eddies) and the device (for the RK stages) evaluate this same code, so both
// hold the same DT.
__host__ __device__ inline double ref_dt_rule(double mv, double mn, double ma, double inlet_velocity,
                                              double courant, double courant_diffusion, double dx)
{
    double maxvel = mv > LES_SMALL ? mv : inlet_velocity;
    double dtc = courant * dx / (maxvel + LES_SMALL);
    double dtn = courant_diffusion * (dx * dx) / (mn + LES_SMALL);
    double dta = courant_diffusion * (dx * dx) / (ma + LES_SMALL);
    return fmin(dtc, fmin(dtn, dta));
}

struct ProbeArgs {
    int n_station;
    int i_probe[LES_NPROBE_MAX];
};

struct StatPtrs {
    double* acc[10];   // sums of u, v, w, G, uu, vv, ww, uv, uw, vw
};
