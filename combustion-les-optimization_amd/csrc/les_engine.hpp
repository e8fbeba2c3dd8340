// les_engine.hpp - context of the gfx950 LES engine (host side, internal).
#pragma once
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include <new>
#include <stdio.h>
#include <type_traits>
#include <vector>

#include "les_ref_kernels.hpp"

struct DuctState;   // duct-mode device state (les_duct_kernels.hip)

// Device buffers of the reference scheme.
struct R {
    int* attr = nullptr;
    double *u, *v, *w, *T, *G, *rho, *rhou, *rhov, *rhow;
    double* Y[LES_NSP];
    double *rhou0, *rhov0, *rhow0, *T0, *G0;
    double* Y0[LES_NSP];
    double *ut, *vt, *wt, *Tt, *Gt, *rhot;
    double* Yt[LES_NSP];
    double* k[4][LES_NRHS];
    double *uf, *vf, *wf, *Tf, *Gf, *rhof;
    double* Yf[LES_NSP];
    double *nu, *mu_mix, *cp, *lam, *alpha, *nu_t, *S_mag, *ST, *gradG;
    double* Dk[LES_NSP];
    double *inlet, *noise, *vort, *partial, *rho_avg;
    int* inlet_idx;          // plane indices of the inlet cells (attribute 1), plane order
    unsigned long long* maxbits;
    double* dt;              // the step's DT (k_ref_dt_final)
    double* probe_rows;
    int* probe_flags;
    double* stat[10];
    int* pt_idx;             // point-series cells
    double* pt_ring;         // [cap][point][u, v, w, G]
};

struct les_gpu {
    les_params p;
    les_mixture mix;
    GridDims g;
    hipStream_t stream = nullptr;
    les_mixture* d_mix = nullptr;
    std::vector<void*> allocs;
    char err[512] = {0};
    int step = 0;
    // composition (:4101-4120)
    double inletY[LES_NSP], ambY[LES_NSP], st_o2, st_co2, st_h2o;
    // host-side glibc rand() stream (reference scheme)
    struct random_data rd;
    char rstate[128];
    std::vector<double> vort;      // num_vortices x 7
    std::vector<double> noise_h;   // 3 per inlet-plane cell
    std::vector<int> attr_h;
    int n_inlet = 0;               // inlet cells of the reference scheme's plane
    int attr_map = 0;              // 0: setting()'s map (:1152-1181); 1: main()'s, kept by a binary resume (:4123-4136)
    std::vector<les_probe_row> probe;
    int stat_count = 0;
    // timing (HIP events on the engine's stream)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev_dt = nullptr;    // a reference-scheme step's DT maxima are on the host
    int timing = 0;
    double rhs_ms_sum = 0.0;
    int rhs_launches = 0;
    double step_ms = 0.0;
    hipGraphExec_t graph_exec = nullptr;
    double* mon_partial = nullptr;
    // point series of the reference scheme (point sampling :4557-4569)
    int npts_r = 0;
    int pt_cap_r = 4096;
    int pt_base_r = 0;
    // pinned host staging of the reference scheme's step (hipHostMalloc): the
    // DT maxima read back, the eddies and inlet noise sent, and one step's
    // probe rows, which are turned into rows at the next synchronisation
    unsigned long long* h_maxbits = nullptr;
    double* h_vort = nullptr;
    void* h_vort_dev = nullptr;    // h_vort's device address (read by k_pull_doubles)
    double* h_noise = nullptr;
    int* h_probe_flags = nullptr;
    double* h_probe_rows = nullptr;
    int probe_pending = -1;        // step whose rows are in h_probe_rows, not yet turned into rows
    std::vector<void*> host_allocs;
    R r;
    DuctState* d = nullptr;
};

int monitor(les_gpu* c, const double* rho, const double* T, const double* G, const double* Y0, const double* u,
            les_step_report* rep);

// Runs the body of a C ABI entry point: a C++ exception from host code (a
// failed host allocation) becomes an error code and message instead of
// crossing the C boundary.
template <typename F>
static inline int abi_guard(const les_gpu* c, F&& body)
{
    const char* what = nullptr;
    int code = LES_OK;
    try {
        return body();
    } catch (const std::bad_alloc&) {
        what = "host memory allocation failed";
        code = LES_ERR_HIP;
    } catch (...) {
        what = "unexpected host error";
        code = LES_ERR_STATE;
    }
    if (c) snprintf(const_cast<les_gpu*>(c)->err, sizeof(c->err), "%s", what);
    return code;
}

// duct mode (les_duct_kernels.hip, les_duct_host.hip)
int  duct_create(les_gpu* c);
void duct_destroy(les_gpu* c);
int  duct_init_state(les_gpu* c);
int  duct_upload_state(les_gpu* c, const double* const* f, int nf);
int  duct_download_state(les_gpu* c, double* const* f, int nf);
int  duct_download_field(les_gpu* c, const char* name, double* out);
int  duct_step(les_gpu* c, int n_steps, les_step_report* rep);
int  duct_rhs(les_gpu* c, double* const* out, int n_out);
int  duct_rhs_noalloc(les_gpu* c, double* const* out, int n_out);
void duct_params_changed(les_gpu* c);
int  duct_point_fetch(les_gpu* c, double* out, int max_steps, int* n_steps);
int  duct_set_points(les_gpu* c, const int* cells, int n);
int  duct_upload_field(les_gpu* c, const char* name, const double* data);
int  duct_export_aux(les_gpu* c, void* buf, size_t* bytes);
int  duct_import_aux(les_gpu* c, const void* buf, size_t bytes);
int  duct_set_step(les_gpu* c, int step);
int  duct_init_sem(les_gpu* c);
int  duct_launch_config(const les_gpu* c, int* tile_k, int* xcd_map);
int  duct_npoints(const les_gpu* c);
