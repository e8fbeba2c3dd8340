// les_engine.hip - C ABI (include/les_gpu.h) of the gfx950 LES engine: device
// state, per-step orchestration on one HIP stream, host-side pieces that the
// reference keeps serial on purpose (glibc rand() stream of the SEM eddies and
// inlet noise, :879-933, :2990-2992) and the probe/statistics read-back.
//
// One context = one run.  All device buffers live for the context's lifetime;
// a step allocates nothing.
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "les_engine.hpp"
#include "les_ref_kernels.hip"
#include "les_duct_kernels.hip"
#include "les_duct_host.hip"

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static int set_err(les_gpu* c, int code, const char* fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPC(call)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_err(c, LES_ERR_HIP, "%s failed: %s (%s:%d)", #call,                 \
                           hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

#define LAUNCH_CHECK()                                                                     \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return set_err(c, LES_ERR_HIP, "kernel launch failed: %s (%s:%d)",             \
                           hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

static double* dalloc(les_gpu* c, size_t n, bool zero = true)
{
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(double) + 256) != hipSuccess) return nullptr;
    c->allocs.push_back(p);
    if (zero) (void)hipMemset(p, 0, n * sizeof(double));
    return (double*)p;
}

// ---------------------------------------------------------------------------
// host-side restatements of the one-shot set-up (no kernel: runs once)
// ---------------------------------------------------------------------------
// Inlet composition for the equivalence ratio (:4101-4120).
static void composition(les_gpu* c)
{
    const les_species* sp = c->mix.sp;
    double phi = c->p.equivalence_ratio;
    double nf = 1.0, no2 = 5.0 / phi, nn2 = no2 * 3.76;
    double mt = nf * sp[0].mw + no2 * sp[1].mw + nn2 * sp[2].mw;
    c->inletY[0] = (nf * sp[0].mw) / mt;
    c->inletY[1] = (no2 * sp[1].mw) / mt;
    c->inletY[2] = (nn2 * sp[2].mw) / mt;
    c->inletY[3] = 0.0; c->inletY[4] = 0.0;
    double amb[5] = {0.0, 0.232, 0.768, 0.0, 0.0};
    for (int s = 0; s < 5; ++s) c->ambY[s] = amb[s];
    c->st_o2 = (5.0 * sp[1].mw) / sp[0].mw;
    c->st_co2 = (3.0 * sp[3].mw) / sp[0].mw;
    c->st_h2o = (4.0 * sp[4].mw) / sp[0].mw;
}

// Attribute map of a new run (setting, :1152-1181), or with kind 1 the map
// main() sets before its restart detection (:4123-4136: no symmetry plane at
// j = 0, no rim), which the reference's binary resume keeps because it skips
// setting() (:4184-4190).
static void attribute_map(les_gpu* c, std::vector<int>& a, int kind = 0)
{
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const double R = c->p.inlet_radius, zc = c->p.lz / 2.0, rrim = 0.005;
    a.resize((size_t)nx * ny * nz);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int v;
                if (i == 0) {
                    double y = j * c->g.dy, z = k * c->g.dz;
                    double d2 = pow(y - 0.0, 2) + pow(z - zc, 2);
                    if (d2 <= pow(R, 2)) v = 1;
                    else if (d2 <= pow(rrim, 2)) v = 6;
                    else v = 4;
                    if (kind == 1 && v == 6) v = 4;
                } else if (i == nx - 1) v = 2;
                else if (j == 0) v = (kind == 1) ? 4 : 5;
                else if (j == ny - 1 || k == 0 || k == nz - 1) v = 4;
                else v = 3;
                a[((size_t)i * ny + j) * nz + k] = v;
            }
}

// Plane indices of the inlet cells (attribute 1 on the inflow plane) in
// plane order: the cells the inflow normalisation sums over.
static int upload_inlet_list(les_gpu* c)
{
    const size_t plane = (size_t)c->p.ny * c->p.nz;
    std::vector<int> idx;
    for (size_t pl = 0; pl < plane; ++pl)
        if (c->attr_h[pl] == 1) idx.push_back((int)pl);
    c->n_inlet = (int)idx.size();
    if (!idx.empty() &&
        hipMemcpy(c->r.inlet_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        snprintf(c->err, sizeof(c->err), "inlet list upload failed");
        return LES_ERR_HIP;
    }
    return LES_OK;
}

static int rng_draw(les_gpu* c)
{
    int32_t r = 0;
    random_r(&c->rd, &r);
    return (int)r;
}

// SEM eddies of a new run (:861-897): three draws go to the reference's debug
// printout first.
static void init_vortices(les_gpu* c)
{
    (void)rng_draw(c); (void)rng_draw(c); (void)rng_draw(c);
    double urms = c->p.inlet_velocity * c->p.turbulence_intensity;
    double L = c->p.turbulent_length_scale;
    for (int v = 0; v < c->p.num_vortices; ++v) {
        double r[7];
        for (int q = 0; q < 7; ++q) r[q] = rng_draw(c) / (double)RAND_MAX;
        double* e = &c->vort[7 * (size_t)v];
        e[0] = (r[0] - 0.5) * L;
        e[1] = r[1] * c->p.ly;
        e[2] = r[2] * c->p.lz;
        double rmin = L * 0.5, rmax = L * 1.5;
        e[3] = rmin + (rmax - rmin) * r[3];
        e[4] = (r[4] * 2.0 - 1.0) * urms;
        e[5] = (r[5] * 2.0 - 1.0) * urms;
        e[6] = (r[6] * 2.0 - 1.0) * urms;
    }
}

// Eddy advection and recycling (:900-939), then the inlet white noise of the
// boundary update (:2990-2992) in the reference's serial (j, k) order.
static void host_sem_and_noise(les_gpu* c, double dt)
{
    double urms = c->p.inlet_velocity * c->p.turbulence_intensity;
    double L = c->p.turbulent_length_scale;
    for (int v = 0; v < c->p.num_vortices; ++v) {
        double* e = &c->vort[7 * (size_t)v];
        e[0] += c->p.inlet_velocity * dt;
        if (e[0] > L / 2.0) {
            e[0] -= L;
            double ry = (double)rng_draw(c) / (double)RAND_MAX;
            double rz = (double)rng_draw(c) / (double)RAND_MAX;
            e[1] = ry * c->p.ly;
            e[2] = rz * c->p.lz;
            double ru = (double)rng_draw(c) / (double)RAND_MAX;
            double rv = (double)rng_draw(c) / (double)RAND_MAX;
            double rw = (double)rng_draw(c) / (double)RAND_MAX;
            e[4] = (ru * 2.0 - 1.0) * urms;
            e[5] = (rv * 2.0 - 1.0) * urms;
            e[6] = (rw * 2.0 - 1.0) * urms;
            double rr = (double)rng_draw(c) / (double)RAND_MAX;
            double rmin = L * 0.5, rmax = L * 1.5;
            e[3] = rmin + (rmax - rmin) * rr;
        }
    }
    const int ny = c->p.ny, nz = c->p.nz;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            size_t pl = (size_t)j * nz + k;
            if (c->attr_h[pl] == 1) {
                for (int t = 0; t < 3; ++t)
                    c->noise_h[3 * pl + t] = ((double)rng_draw(c) / RAND_MAX) * 2.0 - 1.0;
            }
        }
}

static RefConsts make_ref_consts(const les_gpu* c)
{
    RefConsts rc;
    memset(&rc, 0, sizeof(rc));
    const les_params& p = c->p;
    for (int s = 0; s < LES_NSP; ++s) { rc.inletY[s] = c->inletY[s]; rc.ambY[s] = c->ambY[s]; rc.mw[s] = c->mix.sp[s].mw; }
    rc.st_o2 = c->st_o2; rc.st_co2 = c->st_co2; rc.st_h2o = c->st_h2o;
    rc.inlet_velocity = p.inlet_velocity;
    rc.inlet_temperature = p.inlet_temperature;
    rc.ambient_temperature = p.ambient_temperature;
    rc.burnt_temperature = p.burnt_temperature;
    rc.noise_amp = p.inlet_noise_amp * p.inlet_velocity;
    rc.filter_wc = p.filter_weight_center;
    rc.cs_u = p.cs_unburned; rc.cs_b = p.cs_burnt; rc.cs_f = p.cs_flame_boost;
    rc.s_l = p.s_l; rc.nu_unburnt = p.nu_unburnt_approx;
    rc.tau = (p.rho_unburnt_approx / p.rho_burnt_approx) - 1.0;
    rc.d3 = p.fractal_d3;
    rc.pr_t = p.pr_t; rc.sc_t = p.sc_t;
    rc.accel_strength = p.flame_accel_strength; rc.accel_width = p.flame_accel_width;
    rc.accel_center = p.flame_accel_center; rc.weight_x = p.weight_x;
    rc.tip_radius = p.tip_protection_radius; rc.max_force = p.max_flame_force;
    rc.ramp_distance = p.flame_ramp_distance; rc.lz_half = p.lz / 2.0;
    rc.relaxation = p.relaxation_coeff;
    rc.transition_step = p.scheme_transition_step;
    return rc;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static int create_ref(les_gpu* c)
{
    const size_t n = (size_t)c->g.n;
    R& r = c->r;
    double** st[] = {&r.u, &r.v, &r.w, &r.T, &r.G, &r.rho, &r.rhou, &r.rhov, &r.rhow,
                     &r.rhou0, &r.rhov0, &r.rhow0, &r.T0, &r.G0,
                     &r.ut, &r.vt, &r.wt, &r.Tt, &r.Gt, &r.rhot,
                     &r.uf, &r.vf, &r.wf, &r.Tf, &r.Gf, &r.rhof,
                     &r.nu, &r.mu_mix, &r.cp, &r.lam, &r.alpha, &r.nu_t, &r.S_mag, &r.ST, &r.gradG};
    for (auto pp : st) if (!(*pp = dalloc(c, n))) return set_err(c, LES_ERR_HIP, "device allocation failed");
    for (int s = 0; s < LES_NSP; ++s) {
        if (!(r.Y[s] = dalloc(c, n)) || !(r.Y0[s] = dalloc(c, n)) || !(r.Yt[s] = dalloc(c, n)) ||
            !(r.Yf[s] = dalloc(c, n)) || !(r.Dk[s] = dalloc(c, n)))
            return set_err(c, LES_ERR_HIP, "device allocation failed");
    }
    for (int st4 = 0; st4 < 4; ++st4)
        for (int f = 0; f < LES_NRHS; ++f)
            if (!(r.k[st4][f] = dalloc(c, n))) return set_err(c, LES_ERR_HIP, "device allocation failed");
    size_t plane = (size_t)c->p.ny * c->p.nz;
    r.inlet = dalloc(c, 3 * plane);
    // eddies and inlet noise in one block, sent to the device in one copy a step
    const size_t nvort = 7 * (size_t)(c->p.num_vortices > 0 ? c->p.num_vortices : 1);
    r.vort = dalloc(c, nvort + 3 * plane);
    r.noise = r.vort ? r.vort + nvort : nullptr;
    r.partial = dalloc(c, 2 * 1024);
    c->mon_partial = dalloc(c, 7 * 512);
    r.rho_avg = dalloc(c, 1);
    r.maxbits = (unsigned long long*)dalloc(c, 4);
    r.dt = dalloc(c, 1);
    r.probe_rows = dalloc(c, (size_t)LES_NPROBE_MAX * c->p.ny * 8);
    void* fl = nullptr;
    HIPC(hipMalloc(&fl, 64 * sizeof(int)));
    c->allocs.push_back(fl);
    r.probe_flags = (int*)fl;
    if (!r.inlet || !r.noise || !r.vort || !r.partial || !r.rho_avg || !r.maxbits || !r.dt || !r.probe_rows ||
        !c->mon_partial)
        return set_err(c, LES_ERR_HIP, "device allocation failed");
    void* at = nullptr;
    HIPC(hipMalloc(&at, n * sizeof(int)));
    c->allocs.push_back(at);
    r.attr = (int*)at;
    void* il = nullptr;
    HIPC(hipMalloc(&il, plane * sizeof(int)));
    c->allocs.push_back(il);
    r.inlet_idx = (int*)il;
    for (int a = 0; a < 10; ++a) if (!(r.stat[a] = dalloc(c, n))) return set_err(c, LES_ERR_HIP, "device allocation failed");
    {
        void* pi = nullptr;
        HIPC(hipMalloc(&pi, LES_NPOINT_MAX * sizeof(int)));
        c->allocs.push_back(pi);
        r.pt_idx = (int*)pi;
        r.pt_ring = dalloc(c, (size_t)c->pt_cap_r * LES_NPOINT_MAX * 4);
        if (!r.pt_ring) return set_err(c, LES_ERR_HIP, "device allocation failed");
        // setup_measurement_points (:3780-3822): the centre axis and the shear
        // layer (j = R_inlet/DY) at x = 4 R_inlet and 8 R_inlet, k = NZ/2
        int cells[4], np = 0;
        const int js = (int)(c->p.inlet_radius / c->g.dy);
        const int il[2] = {(int)(4.0 * c->p.inlet_radius / c->g.dx), (int)(8.0 * c->p.inlet_radius / c->g.dx)};
        for (int a = 0; a < 2; ++a) {
            if (il[a] >= c->p.nx) continue;
            cells[np++] = (int)lidx(c->g, il[a], 0, c->p.nz / 2);
            cells[np++] = (int)lidx(c->g, il[a], js, c->p.nz / 2);
        }
        HIPC(hipMemcpy(r.pt_idx, cells, np * sizeof(int), hipMemcpyHostToDevice));
        c->npts_r = np;
    }
    {
        void* hp[4] = {nullptr, nullptr, nullptr, nullptr};
        const size_t hb[4] = {4 * sizeof(unsigned long long), (nvort + 3 * plane) * sizeof(double),
                              64 * sizeof(int), (size_t)LES_NPROBE_MAX * c->p.ny * 8 * sizeof(double)};
        for (int a = 0; a < 4; ++a) {
            HIPC(hipHostMalloc(&hp[a], hb[a], hipHostMallocDefault));
            c->host_allocs.push_back(hp[a]);
        }
        c->h_maxbits = (unsigned long long*)hp[0];
        c->h_vort = (double*)hp[1];
        c->h_noise = c->h_vort + nvort;
        HIPC(hipHostGetDevicePointer(&c->h_vort_dev, c->h_vort, 0));
        c->h_probe_flags = (int*)hp[2];
        c->h_probe_rows = (double*)hp[3];
    }
    c->vort.assign(7 * (size_t)(c->p.num_vortices > 0 ? c->p.num_vortices : 1), 0.0);
    c->noise_h.assign(3 * plane, 0.0);
    attribute_map(c, c->attr_h);
    HIPC(hipMemcpy(r.attr, c->attr_h.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return upload_inlet_list(c);
}

extern "C" int les_gpu_create(const les_params* p, const les_mixture* mix, les_gpu** out)
{
    return abi_guard(nullptr, [&]() -> int {
        if (!p || !mix || !out) return LES_ERR_PARAM;
        *out = nullptr;
        int v = les_params_validate(p);
        if (v != LES_OK) return v;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LES_ERR_HIP;
        les_gpu* c = new (std::nothrow) les_gpu();
        if (!c) return LES_ERR_HIP;
        c->p = *p;
        c->mix = *mix;
        c->g.nx = p->nx; c->g.ny = p->ny; c->g.nz = p->nz;
        c->g.n = (long long)p->nx * p->ny * p->nz;
        c->g.sx = (long long)p->ny * p->nz;
        c->g.sy = p->nz;
        c->g.dx = p->lx / (double)p->nx;
        c->g.dy = p->ly / (double)p->ny;
        c->g.dz = p->lz / (double)p->nz;
        memset(&c->rd, 0, sizeof(c->rd));
        initstate_r((unsigned)p->seed, c->rstate, sizeof(c->rstate), &c->rd);
        composition(c);
        int rc = LES_OK;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = LES_ERR_HIP;
        if (rc == LES_OK) {
            void* m = nullptr;
            if (hipMalloc(&m, sizeof(les_mixture)) != hipSuccess) rc = LES_ERR_HIP;
            else {
                c->allocs.push_back(m);
                c->d_mix = (les_mixture*)m;
                if (hipMemcpy(m, mix, sizeof(les_mixture), hipMemcpyHostToDevice) != hipSuccess) rc = LES_ERR_HIP;
            }
        }
        if (rc == LES_OK) {
            if (hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess ||
                hipEventCreate(&c->ev[2]) != hipSuccess || hipEventCreate(&c->ev[3]) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_dt, hipEventDisableTiming) != hipSuccess)
                rc = LES_ERR_HIP;
        }
        if (rc == LES_OK) rc = (p->mode == LES_MODE_REFERENCE) ? create_ref(c) : duct_create(c);
        if (rc != LES_OK) {
            fprintf(stderr, "les_gpu_create: %s\n", c->err);
            les_gpu_destroy(c);
            return rc;
        }
        *out = c;
        return LES_OK;
    });
}

extern "C" void les_gpu_destroy(les_gpu* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    duct_destroy(c);
    for (void* p : c->allocs) (void)hipFree(p);
    for (void* p : c->host_allocs) (void)hipHostFree(p);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->ev_dt) (void)hipEventDestroy(c->ev_dt);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* les_gpu_last_error(const les_gpu* c) { return c ? c->err : "null context"; }
extern "C" int les_gpu_get_step(const les_gpu* c) { return c ? c->step : -1; }

// ---------------------------------------------------------------------------
// reference-scheme step
// ---------------------------------------------------------------------------
static RefPtrs ref_ptrs(les_gpu* c, bool temp_state, int stage)
{
    R& r = c->r;
    RefPtrs p;
    memset(&p, 0, sizeof(p));
    p.attr = r.attr;
    p.mix = c->d_mix;
    p.u = temp_state ? r.ut : r.u; p.v = temp_state ? r.vt : r.v; p.w = temp_state ? r.wt : r.w;
    p.T = temp_state ? r.Tt : r.T; p.G = temp_state ? r.Gt : r.G; p.rho = temp_state ? r.rhot : r.rho;
    for (int s = 0; s < LES_NSP; ++s) { p.Y[s] = temp_state ? r.Yt[s] : r.Y[s]; p.Yf[s] = r.Yf[s]; p.Dk[s] = r.Dk[s]; }
    p.uf = r.uf; p.vf = r.vf; p.wf = r.wf; p.Tf = r.Tf; p.Gf = r.Gf; p.rhof = r.rhof;
    p.nu = r.nu; p.mu_mix = r.mu_mix; p.cp = r.cp; p.lam = r.lam; p.alpha = r.alpha;
    p.nu_t = r.nu_t; p.S_mag = r.S_mag; p.ST = r.ST; p.gradG = r.gradG;
    for (int f = 0; f < LES_NRHS; ++f) p.out[f] = r.k[stage][f];
    p.rho_avg = r.rho_avg;
    return p;
}

static RefStagePtrs stage_ptrs(les_gpu* c)
{
    R& r = c->r;
    RefStagePtrs s;
    s.attr = r.attr;
    s.u = r.u; s.v = r.v; s.w = r.w; s.T = r.T; s.G = r.G; s.rho = r.rho;
    s.rhou = r.rhou; s.rhov = r.rhov; s.rhow = r.rhow;
    s.rhou0 = r.rhou0; s.rhov0 = r.rhov0; s.rhow0 = r.rhow0; s.T0 = r.T0; s.G0 = r.G0;
    s.ut = r.ut; s.vt = r.vt; s.wt = r.wt; s.Tt = r.Tt; s.Gt = r.Gt; s.rhot = r.rhot;
    for (int b = 0; b < LES_NSP; ++b) { s.Y[b] = r.Y[b]; s.Y0[b] = r.Y0[b]; s.Yt[b] = r.Yt[b]; }
    for (int a = 0; a < 4; ++a) for (int f = 0; f < LES_NRHS; ++f) s.k[a][f] = r.k[a][f];
    return s;
}

// One RHS evaluation on the step state (temp_state=false) or the RK state.
static int ref_rhs(les_gpu* c, bool temp_state, int stage, int t_loop)
{
    const GridDims& g = c->g;
    RefConsts rc = make_ref_consts(c);
    RefPtrs p = ref_ptrs(c, temp_state, stage);
    Launch3 L = launch3(g.nx, g.ny, g.nz);
    // a grid of fewer than two line waves per SIMD (1024 SIMDs) runs the
    // property and RHS launches split into output groups: more waves, each
    // with a shorter dependent instruction stream (same values either way);
    // the heaviest group first, as blocks are dispatched in slab order
    const long long line_waves = (long long)L.grid.x * g.ny * g.nx;
    const bool split = line_waves < 2048 && (long long)g.nx * REF_GROUPS_MAX <= 65535;
    RefGroups gp{}, gr{};
    gp.n = gr.n = 1;
    if (split) {
        // transport; fields and cp; each diffusivity
        const unsigned pm[] = {REF_PROPS_TRANSPORT, REF_PROPS_FIELDS, 1u, 2u, 4u, 8u, 16u};
        // momentum; G; T and Y0; Y1 and Y2; Y3 and Y4
        const unsigned rm[] = {7u, 16u, 8u | 32u, 64u | 128u, 256u | 512u};
        gp.n = (int)(sizeof pm / sizeof pm[0]);
        gr.n = (int)(sizeof rm / sizeof rm[0]);
        for (int b = 0; b < gp.n; ++b) gp.mask[b] = pm[b];
        for (int b = 0; b < gr.n; ++b) gr.mask[b] = rm[b];
    }
    dim3 gridp = L.grid, gridr = L.grid;
    gridp.z = (unsigned)(g.nx * gp.n);
    gridr.z = (unsigned)(g.nx * gr.n);
    hipLaunchKernelGGL(k_ref_filter_props, gridp, L.block, 0, c->stream, g, rc, p, gp);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_ref_nut_st, L.grid, L.block, 0, c->stream, g, rc, p);
    LAUNCH_CHECK();
    int nb = (int)((g.n + 255) / 256);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_ref_rho_partial, dim3(nb), dim3(256), 0, c->stream, g, c->r.attr, c->r.rhof, c->r.partial);
    hipLaunchKernelGGL(k_ref_rho_final, dim3(1), dim3(256), 0, c->stream, nb, c->r.partial, c->r.rho_avg);
    LAUNCH_CHECK();
    if (c->timing) HIPC(hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_ref_rhs, gridr, L.block, 0, c->stream, g, rc, p, t_loop, gr);
    LAUNCH_CHECK();
    if (c->timing) {
        HIPC(hipEventRecord(c->ev[1], c->stream));
        HIPC(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->rhs_ms_sum += ms;
        c->rhs_launches++;
    }
    return LES_OK;
}

__global__ void k_ref_points(const int* idx, int npts, const double* u, const double* v, const double* w,
                             const double* G, double* out)
{
    const int p = threadIdx.x;
    if (p >= npts) return;
    const int q = idx[p];
    out[4 * p] = u[q]; out[4 * p + 1] = v[q]; out[4 * p + 2] = w[q]; out[4 * p + 3] = G[q];
}

// The probe rows of step c->probe_pending, copied into pinned memory by that
// step and complete once the stream has been synchronised since, become rows
// (record_flame_region_data :3403-3471, SAMPLING_RANGE :280).
static void drain_probes(les_gpu* c)
{
    if (c->probe_pending < 0) return;
    const int t = c->probe_pending;
    const GridDims& g = c->g;
    for (int n = 0; n < c->p.num_probe_x; ++n) {
        if (!c->h_probe_flags[n]) continue;
        for (int j = 0; j < g.ny; ++j) {
            const double* q = &c->h_probe_rows[((size_t)n * g.ny + j) * 8];
            double yp = j * g.dy, yr = yp - q[0];
            if (fabs(yr) > 1.0) continue;
            les_probe_row pr;
            pr.timestep = t; pr.station = n; pr.x_location = c->p.probe_x[n];
            pr.y_flame = q[0]; pr.y_physical = yp; pr.y_relative = yr;
            pr.u = q[1]; pr.v = q[2]; pr.w = q[3]; pr.G = q[4]; pr.nu_t = q[5]; pr.S_mag = q[6];
            c->probe.push_back(pr);
        }
    }
    c->probe_pending = -1;
}

// n doubles from device-visible pinned host memory into device memory
__global__ void __launch_bounds__(256) k_pull_doubles(const double* src, double* dst, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

static int ref_step(les_gpu* c, les_step_report* rep, bool last)
{
    const GridDims& g = c->g;
    R& r = c->r;
    RefConsts rc = make_ref_consts(c);
    int t = c->step;
    // DT (:4399-4411): exact device maxima, DT computed on the device for the
    // RK stages and, from the same maxima read back, on the host for the
    // eddies; the RK stages are queued before the host waits, so the host's
    // eddy update overlaps them
    HIPC(hipMemsetAsync(r.maxbits, 0, 4 * sizeof(unsigned long long), c->stream));
    int nb = (int)((g.n + 255) / 256);
    if (nb > 256) nb = 256;                 // one block per CU: 3 atomics each
    hipLaunchKernelGGL(k_ref_dt_max, dim3(nb), dim3(256), 0, c->stream, g, r.attr, r.u, r.v, r.w, r.nu, r.alpha,
                       r.maxbits);
    hipLaunchKernelGGL(k_ref_dt_final, dim3(1), dim3(64), 0, c->stream, r.maxbits, c->p.inlet_velocity, c->p.courant,
                       c->p.courant_diffusion, g.dx, r.dt);
    LAUNCH_CHECK();
    HIPC(hipMemcpyAsync(c->h_maxbits, r.maxbits, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipEventRecord(c->ev_dt, c->stream));

    // RK4 (:4420-4544)
    RefStagePtrs sp = stage_ptrs(c);
    int nbc = (int)((g.n + 255) / 256);
    hipLaunchKernelGGL(k_ref_save, dim3(nbc), dim3(256), 0, c->stream, g, sp);
    for (int s = 0; s < 4; ++s) {
        int e = ref_rhs(c, s > 0, s, t);
        if (e != LES_OK) return e;
        if (s < 3) hipLaunchKernelGGL(k_ref_stage, dim3(nbc), dim3(256), 0, c->stream, g, rc, sp, s, r.dt);
        else hipLaunchKernelGGL(k_ref_final, dim3(nbc), dim3(256), 0, c->stream, g, rc, sp, r.dt);
        LAUNCH_CHECK();
    }

    // the step's one host wait: DT's maxima (and everything queued before
    // them, the previous step's probe rows and copies included)
    HIPC(hipEventSynchronize(c->ev_dt));
    drain_probes(c);
    double mv, mn, ma;
    memcpy(&mv, &c->h_maxbits[0], 8); memcpy(&mn, &c->h_maxbits[1], 8); memcpy(&ma, &c->h_maxbits[2], 8);
    const double maxvel = mv > LES_SMALL ? mv : c->p.inlet_velocity;
    const double DT = ref_dt_rule(mv, mn, ma, c->p.inlet_velocity, c->p.courant, c->p.courant_diffusion, g.dx);

    // SEM eddies + inlet noise on the host (glibc rand() stream), plane on the
    // device: the plane is read only by the inflow boundary condition, so it
    // is built after the RK stages are queued
    host_sem_and_noise(c, DT);
    size_t plane = (size_t)c->p.ny * c->p.nz;
    // staged in pinned memory, which the previous step's copy has finished
    // reading (it was queued before the event waited on above)
    memcpy(c->h_vort, c->vort.data(), c->vort.size() * sizeof(double));
    memcpy(c->h_noise, c->noise_h.data(), 3 * plane * sizeof(double));
    {   // pulled by a kernel from the pinned block: a DMA copy here would wait
        // for the queue hand-over between the copy engine and the kernels
        const long long nn = (long long)(c->vort.size() + 3 * plane);
        hipLaunchKernelGGL(k_pull_doubles, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)c->h_vort_dev, r.vort, nn);
    }
    hipLaunchKernelGGL(k_ref_sem_plane, dim3((unsigned)((plane + 3) / 4)), dim3(256), 0, c->stream, g, r.attr, r.vort,
                       c->p.num_vortices, r.inlet);
    hipLaunchKernelGGL(k_ref_sem_normalise, dim3(1), dim3(SEM_NORM_THREADS), 0, c->stream, g, r.inlet_idx, c->n_inlet, r.inlet,
                       c->p.inlet_velocity * c->p.turbulence_intensity);
    LAUNCH_CHECK();

    // boundary conditions (:2967-3161)
    hipLaunchKernelGGL(k_ref_bc_x, dim3((g.nz + 63) / 64, g.ny), dim3(64), 0, c->stream, g, rc, sp, r.inlet, r.noise);
    hipLaunchKernelGGL(k_ref_bc_y, dim3((g.nz + 63) / 64, g.nx - 2), dim3(64), 0, c->stream, g, rc, sp);
    hipLaunchKernelGGL(k_ref_bc_z, dim3((g.ny + 63) / 64, g.nx - 2), dim3(64), 0, c->stream, g, rc, sp);
    hipLaunchKernelGGL(k_ref_bc_eos, dim3(nbc), dim3(256), 0, c->stream, g, rc, sp);
    LAUNCH_CHECK();

    // point series (:4557-4569) into the device ring
    if (c->npts_r > 0) {
        hipLaunchKernelGGL(k_ref_points, dim3(1), dim3(64), 0, c->stream, r.pt_idx, c->npts_r, r.u, r.v, r.w, r.G,
                           r.pt_ring + (size_t)(t % c->pt_cap_r) * LES_NPOINT_MAX * 4);
        LAUNCH_CHECK();
    }
    // probes (:4572-4574): rows copied into pinned memory, turned into rows
    // at the next synchronisation (the next step's DT wait or the end of
    // les_gpu_step)
    if (c->p.num_probe_x > 0) {
        ProbeArgs pa;
        memset(&pa, 0, sizeof(pa));
        pa.n_station = c->p.num_probe_x;
        for (int n = 0; n < pa.n_station; ++n) pa.i_probe[n] = (int)(c->p.probe_x[n] / g.dx);
        hipLaunchKernelGGL(k_ref_probe, dim3(1), dim3(64), 0, c->stream, g, pa, r.u, r.v, r.w, r.G, r.nu_t, r.S_mag,
                           r.probe_rows, r.probe_flags);
        LAUNCH_CHECK();
        HIPC(hipMemcpyAsync(c->h_probe_flags, r.probe_flags, pa.n_station * sizeof(int), hipMemcpyDeviceToHost,
                            c->stream));
        HIPC(hipMemcpyAsync(c->h_probe_rows, r.probe_rows, (size_t)pa.n_station * g.ny * 8 * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
        c->probe_pending = t;
    }
    if (c->p.stats_start_step >= 0 && t >= c->p.stats_start_step) {
        StatPtrs st;
        for (int a = 0; a < 10; ++a) st.acc[a] = r.stat[a];
        hipLaunchKernelGGL(k_stats_accum, dim3(nbc), dim3(256), 0, c->stream, g.n, r.u, r.v, r.w, r.G, st);
        LAUNCH_CHECK();
        c->stat_count++;
    }
    if (rep) { rep->step = t; rep->dt = DT; rep->max_velocity = maxvel; }
    if (last && rep) {
        int e = monitor(c, r.rho, r.T, r.G, r.Y[0], r.u, rep);
        if (e != LES_OK) return e;
    }
    c->step++;
    return LES_OK;
}

// Monitor reductions (:3880-3897) and the non-finite flag.
int monitor(les_gpu* c, const double* rho, const double* T, const double* G, const double* Y0, const double* u,
            les_step_report* rep)
{
    const GridDims& g = c->g;
    int nb = (int)((g.n + 255) / 256);
    if (nb > 512) nb = 512;
    double* part = c->mon_partial;
    hipLaunchKernelGGL(k_monitor, dim3(nb), dim3(256), 0, c->stream, g.n, g.dx * g.dy * g.dz, rho, T, G, Y0, u, part);
    LAUNCH_CHECK();
    std::vector<double> h((size_t)nb * 7);
    HIPC(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, mnY = 1e9, bad = 0;
    for (int b = 0; b < nb; ++b) {
        const double* q = &h[7 * (size_t)b];
        tm += q[0]; mnT = fmin(mnT, q[1]); mxT = fmax(mxT, q[2]);
        mnG = fmin(mnG, q[3]); mxG = fmax(mxG, q[4]); mnY = fmin(mnY, q[5]); bad = fmax(bad, q[6]);
    }
    rep->total_mass = tm; rep->min_T = mnT; rep->max_T = mxT; rep->min_G = mnG; rep->max_G = mxG;
    rep->min_Y_fuel = mnY; rep->nonfinite = bad > 0.0;
    return LES_OK;
}

// ---------------------------------------------------------------------------
// ABI entry points
// ---------------------------------------------------------------------------
// Rebuilds the host and device attribute maps of a reference-scheme run.
static int set_attribute_map(les_gpu* c, int kind)
{
    attribute_map(c, c->attr_h, kind);
    HIPC(hipMemcpy(c->r.attr, c->attr_h.data(), c->attr_h.size() * sizeof(int), hipMemcpyHostToDevice));
    c->attr_map = kind;
    return upload_inlet_list(c);
}

extern "C" int les_gpu_init_state(les_gpu* c)
{
    return abi_guard(c, [&]() -> int {
        if (!c) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_init_state(c);
        // a new run takes setting()'s attribute map (:1152-1181)
        if (c->attr_map != 0) {
            int e = set_attribute_map(c, 0);
            if (e != LES_OK) return e;
        }
        // initial fields of a new run (setting :1184-1261) on the host, then upload
        const size_t n = (size_t)c->g.n;
        const int ny = c->p.ny, nz = c->p.nz;
        std::vector<double> f(n * LES_NFIELDS);
        double* F[LES_NFIELDS];
        for (int a = 0; a < LES_NFIELDS; ++a) F[a] = &f[a * n];
        for (size_t q = 0; q < n; ++q) {
            int a = c->attr_h[q];
            int i = (int)(q / ((size_t)ny * nz));
            double Yc[5] = {0, 0, 0, 0, 0}, u = 0.0, T = 0.0, G = 0.0;
            if (a == 1) {
                u = c->p.inlet_velocity; T = c->p.inlet_temperature;
                for (int s = 0; s < 5; ++s) Yc[s] = c->inletY[s];
            } else if (a == 4 || a == 2) {
                T = c->p.ambient_temperature;
                for (int s = 0; s < 5; ++s) Yc[s] = c->ambY[s];
            } else if (a == 5 || a == 3) {
                if (i <= c->p.initial_flame_front_i) {
                    double cons = c->inletY[0];
                    G = 1.0; T = c->p.burnt_temperature;
                    Yc[0] = c->inletY[0] - cons;
                    Yc[1] = c->inletY[1] - cons * c->st_o2;
                    Yc[2] = c->inletY[2];
                    Yc[3] = c->inletY[3] + cons * c->st_co2;
                    Yc[4] = c->inletY[4] + cons * c->st_h2o;
                } else {
                    T = c->p.inlet_temperature;
                    for (int s = 0; s < 5; ++s) Yc[s] = c->inletY[s];
                }
            }
            F[0][q] = u; F[1][q] = 0.0; F[2][q] = 0.0; F[3][q] = T; F[4][q] = G;
            double sy = 0.0;
            for (int s = 0; s < 5; ++s) sy += Yc[s] / (c->mix.sp[s].mw / 1000.0);
            if (sy > LES_SMALL && T > 1.0) {
                double wmix = 1.0 / sy;
                double rmix = LES_R_UNIV / wmix;
                F[5][q] = LES_P_CONST / (rmix * T);
            } else F[5][q] = 1.2;
            for (int s = 0; s < 5; ++s) F[6 + s][q] = Yc[s];
        }
        int e = les_gpu_upload_state(c, (const double* const*)F, LES_NFIELDS);
        if (e != LES_OK) return e;
        R& r = c->r;
        size_t b = n * sizeof(double);
        HIPC(hipMemsetAsync(r.nu, 0, b, c->stream));
        HIPC(hipMemsetAsync(r.alpha, 0, b, c->stream));
        HIPC(hipMemsetAsync(r.nu_t, 0, b, c->stream));
        HIPC(hipMemsetAsync(r.S_mag, 0, b, c->stream));
        HIPC(hipMemsetAsync(r.rhot, 0, b, c->stream));
        for (int a = 0; a < 10; ++a) HIPC(hipMemsetAsync(r.stat[a], 0, b, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        init_vortices(c);
        c->step = 0;
        c->stat_count = 0;
        c->probe.clear();
        c->probe_pending = -1;
        return LES_OK;
    });
}

__global__ void k_momentum_from_prim(long long n, const double* rho, const double* u, const double* v,
                                     const double* w, double* ru, double* rv, double* rw)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    ru[q] = rho[q] * u[q]; rv[q] = rho[q] * v[q]; rw[q] = rho[q] * w[q];
}

extern "C" int les_gpu_upload_state(les_gpu* c, const double* const* f, int nf)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !f) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_upload_state(c, f, nf);
        if (nf < LES_NFIELDS) return set_err(c, LES_ERR_PARAM, "need %d fields", LES_NFIELDS);
        R& r = c->r;
        size_t b = (size_t)c->g.n * sizeof(double);
        double* dst[LES_NFIELDS] = {r.u, r.v, r.w, r.T, r.G, r.rho, r.Y[0], r.Y[1], r.Y[2], r.Y[3], r.Y[4]};
        for (int a = 0; a < LES_NFIELDS; ++a) HIPC(hipMemcpyAsync(dst[a], f[a], b, hipMemcpyHostToDevice, c->stream));
        int nb = (int)((c->g.n + 255) / 256);
        hipLaunchKernelGGL(k_momentum_from_prim, dim3(nb), dim3(256), 0, c->stream, c->g.n, r.rho, r.u, r.v, r.w, r.rhou,
                           r.rhov, r.rhow);
        LAUNCH_CHECK();
        HIPC(hipStreamSynchronize(c->stream));
        return LES_OK;
    });
}

extern "C" int les_gpu_download_state(les_gpu* c, double* const* f, int nf)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !f) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_download_state(c, f, nf);
        R& r = c->r;
        size_t b = (size_t)c->g.n * sizeof(double);
        const double* src[LES_NFIELDS] = {r.u, r.v, r.w, r.T, r.G, r.rho, r.Y[0], r.Y[1], r.Y[2], r.Y[3], r.Y[4]};
        for (int a = 0; a < nf && a < LES_NFIELDS; ++a)
            HIPC(hipMemcpyAsync(f[a], src[a], b, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        return LES_OK;
    });
}

extern "C" int les_gpu_download_field(les_gpu* c, const char* name, double* out)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !name || !out) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_download_field(c, name, out);
        R& r = c->r;
        size_t n = (size_t)c->g.n;
        if (strcmp(name, "attribute") == 0) {
            for (size_t q = 0; q < n; ++q) out[q] = (double)c->attr_h[q];
            return LES_OK;
        }
        struct { const char* nm; const double* p; } tab[] = {
            {"nu_t", r.nu_t}, {"S_mag", r.S_mag}, {"nu", r.nu}, {"alpha", r.alpha}, {"mu_mix", r.mu_mix},
            {"cp_mix", r.cp}, {"lambda_mix", r.lam}, {"ST", r.ST}, {"gradG", r.gradG},
            {"u_f", r.uf}, {"v_f", r.vf}, {"w_f", r.wf}, {"T_f", r.Tf}, {"G_f", r.Gf}, {"rho_f", r.rhof},
            {"rho_temp", r.rhot}, {"rhou", r.rhou}, {"rhov", r.rhov}, {"rhow", r.rhow},
            {"D0", r.Dk[0]}, {"D1", r.Dk[1]}, {"D2", r.Dk[2]}, {"D3", r.Dk[3]}, {"D4", r.Dk[4]},
            {"inlet", r.inlet}};
        for (auto& t : tab) {
            if (strcmp(t.nm, name) == 0) {
                size_t cnt = strcmp(name, "inlet") == 0 ? 3 * (size_t)c->p.ny * c->p.nz : n;
                HIPC(hipMemcpyAsync(out, t.p, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                HIPC(hipStreamSynchronize(c->stream));
                return LES_OK;
            }
        }
        if (strncmp(name, "stat", 4) == 0) {
            int a = atoi(name + 4);
            if (a < 0 || a > 9) return LES_ERR_PARAM;
            HIPC(hipMemcpyAsync(out, r.stat[a], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPC(hipStreamSynchronize(c->stream));
            return LES_OK;
        }
        return set_err(c, LES_ERR_PARAM, "unknown field %s", name);
    });
}

extern "C" int les_gpu_step(les_gpu* c, int n_steps, les_step_report* rep)
{
    return abi_guard(c, [&]() -> int {
        if (!c || n_steps < 0) return LES_ERR_PARAM;
        c->rhs_ms_sum = 0.0;
        c->rhs_launches = 0;
        if (c->p.mode == LES_MODE_DUCT) return duct_step(c, n_steps, rep);
        HIPC(hipEventRecord(c->ev[2], c->stream));
        les_step_report tmp;
        for (int s = 0; s < n_steps; ++s) {
            int e = ref_step(c, rep ? rep : &tmp, s == n_steps - 1);
            if (e != LES_OK) return e;
        }
        HIPC(hipEventRecord(c->ev[3], c->stream));
        HIPC(hipEventSynchronize(c->ev[3]));
        drain_probes(c);
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        c->step_ms = n_steps > 0 ? ms / n_steps : 0.0;
        if (rep && rep->nonfinite) return set_err(c, LES_ERR_NAN, "non-finite value in the state at step %d", rep->step);
        return LES_OK;
    });
}

extern "C" int les_gpu_rhs(les_gpu* c, int t_loop, double* const* out, int n_out)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !out || n_out < LES_NRHS) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_rhs(c, out, n_out);
        int e = ref_rhs(c, false, 0, t_loop);
        if (e != LES_OK) return e;
        size_t b = (size_t)c->g.n * sizeof(double);
        for (int f = 0; f < LES_NRHS; ++f) HIPC(hipMemcpyAsync(out[f], c->r.k[0][f], b, hipMemcpyDeviceToHost, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        return LES_OK;
    });
}

extern "C" int les_gpu_rhs_noalloc(les_gpu* c, int t_loop, double* const* out, int n_out)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !out || n_out < LES_NRHS) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_rhs_noalloc(c, out, n_out);
        // the reference scheme evaluates into the context's own stage buffers
        // and allocates nothing per call: the same path as les_gpu_rhs
        return les_gpu_rhs(c, t_loop, out, n_out);
    });
}

extern "C" int les_gpu_probe_fetch(les_gpu* c, les_probe_row* rows, int max_rows, int* n_rows)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !rows || !n_rows || max_rows < 0) return LES_ERR_PARAM;
        int n = (int)c->probe.size() < max_rows ? (int)c->probe.size() : max_rows;
        if (n > 0) memcpy(rows, c->probe.data(), (size_t)n * sizeof(les_probe_row));
        c->probe.erase(c->probe.begin(), c->probe.begin() + n);
        *n_rows = n;
        return LES_OK;
    });
}

extern "C" int les_gpu_set_params(les_gpu* c, const les_params* p)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !p) return LES_ERR_PARAM;
        if (p->nx != c->p.nx || p->ny != c->p.ny || p->nz != c->p.nz || p->mode != c->p.mode ||
            p->num_vortices != c->p.num_vortices || p->precision != c->p.precision)
            return set_err(c, LES_ERR_PARAM, "grid, mode, precision and eddy count are fixed per context");
        // quantities the context derives once at creation: cell sizes, the
        // RNG seed, the attribute map and inlet composition (mode R), the
        // property tables (duct)
        if (p->lx != c->p.lx || p->ly != c->p.ly || p->lz != c->p.lz || p->seed != c->p.seed)
            return set_err(c, LES_ERR_PARAM, "domain size and seed are fixed per context");
        if (p->mode == LES_MODE_REFERENCE &&
            (p->inlet_radius != c->p.inlet_radius || p->equivalence_ratio != c->p.equivalence_ratio))
            return set_err(c, LES_ERR_PARAM, "inlet radius and equivalence ratio are fixed per context");
        if (p->mode == LES_MODE_DUCT &&
            (p->table_t0 != c->p.table_t0 || p->table_dt != c->p.table_dt || p->table_n != c->p.table_n))
            return set_err(c, LES_ERR_PARAM, "the cp/mu tables are fixed per context");
        if (p->integrator != c->p.integrator || p->advection != c->p.advection) {
            // a duct run may switch between its two reconstructions; nothing else
            // changes the scheme of a live context
            const bool duct_ok = p->mode == LES_MODE_DUCT && p->integrator == LES_INT_SSPRK3 &&
                                 (p->advection == LES_ADV_WENO5 || p->advection == LES_ADV_MUSCL_MINMOD);
            if (!duct_ok) return set_err(c, LES_ERR_PARAM, "integrator %d / advection %d not available to this context",
                                         p->integrator, p->advection);
        }
        if (les_params_validate(p) != LES_OK) return LES_ERR_PARAM;
        c->p = *p;
        if (c->p.mode == LES_MODE_DUCT) duct_params_changed(c);
        return LES_OK;
    });
}

extern "C" int les_gpu_kernel_timing(const les_gpu* c, double* rhs_ms, int* launches, double* step_ms)
{
    return abi_guard(c, [&]() -> int {
        if (!c) return LES_ERR_PARAM;
        if (rhs_ms) *rhs_ms = c->rhs_launches ? c->rhs_ms_sum / c->rhs_launches : 0.0;
        if (launches) *launches = c->rhs_launches;
        if (step_ms) *step_ms = c->step_ms;
        return LES_OK;
    });
}

extern "C" int les_gpu_launch_config(const les_gpu* c, int* tile_k, int* xcd_map)
{
    if (!c || !c->d || c->p.mode != LES_MODE_DUCT) return LES_ERR_PARAM;
    return duct_launch_config(c, tile_k, xcd_map);
}

extern "C" int les_gpu_set_timing(les_gpu* c, int on)
{
    return abi_guard(c, [&]() -> int {
        if (!c) return LES_ERR_PARAM;
        c->timing = on ? 1 : 0;
        return LES_OK;
    });
}

extern "C" int les_gpu_stats_count(const les_gpu* c) { return c ? c->stat_count : -1; }

extern "C" int les_gpu_point_fetch(les_gpu* c, double* out, int max_steps, int* n_steps, int* n_points)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !out || !n_steps || max_steps < 0) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) {
            if (n_points) *n_points = duct_npoints(c);
            return duct_point_fetch(c, out, max_steps, n_steps);
        }
        // reference scheme: the newest pt_cap_r steps since the last fetch
        const int total = c->step - c->pt_base_r;
        const int avail = total < c->pt_cap_r ? total : c->pt_cap_r;
        const int n = avail < max_steps ? avail : max_steps;
        const int first = c->step - avail;
        const int np = c->npts_r;
        for (int a = 0; a < n; ++a) {
            const double* src = c->r.pt_ring + (size_t)((first + a) % c->pt_cap_r) * LES_NPOINT_MAX * 4;
            HIPC(hipMemcpyAsync(out + (size_t)a * np * 4, src, (size_t)np * 4 * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
        }
        HIPC(hipStreamSynchronize(c->stream));
        c->pt_base_r = c->step;
        *n_steps = n;
        if (n_points) *n_points = np;
        return LES_OK;
    });
}

extern "C" int les_gpu_set_points(les_gpu* c, const int* ijk, int n)
{
    return abi_guard(c, [&]() -> int {
        if (!c || n < 0 || n > LES_NPOINT_MAX || (n > 0 && !ijk)) return LES_ERR_PARAM;
        int cells[LES_NPOINT_MAX];
        for (int a = 0; a < n; ++a) {
            const int i = ijk[3 * a], j = ijk[3 * a + 1], k = ijk[3 * a + 2];
            if (i < 0 || i >= c->p.nx || j < 0 || j >= c->p.ny || k < 0 || k >= c->p.nz)
                return set_err(c, LES_ERR_PARAM, "point %d (%d, %d, %d) is outside the grid", a, i, j, k);
            cells[a] = (int)lidx(c->g, i, j, k);
        }
        if (c->p.mode == LES_MODE_DUCT) return duct_set_points(c, cells, n);
        if (n > 0) HIPC(hipMemcpy(c->r.pt_idx, cells, n * sizeof(int), hipMemcpyHostToDevice));
        c->npts_r = n;
        c->pt_base_r = c->step;
        return LES_OK;
    });
}

extern "C" int les_gpu_get_points(les_gpu* c, int* ijk, int max_n, int* n)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !n || (max_n > 0 && !ijk)) return LES_ERR_PARAM;
        std::vector<int> cells(LES_NPOINT_MAX);
        int np;
        if (c->p.mode == LES_MODE_DUCT) {
            np = duct_npoints(c);
            memcpy(cells.data(), c->d->desc[0].point_idx, np * sizeof(int));
        } else {
            np = c->npts_r;
            HIPC(hipMemcpy(cells.data(), c->r.pt_idx, np * sizeof(int), hipMemcpyDeviceToHost));
        }
        for (int a = 0; a < np && a < max_n; ++a) {
            const int q = cells[a];
            ijk[3 * a] = q / (c->p.ny * c->p.nz);
            ijk[3 * a + 1] = (q / c->p.nz) % c->p.ny;
            ijk[3 * a + 2] = q % c->p.nz;
        }
        *n = np;
        return LES_OK;
    });
}

// ---------------------------------------------------------------------------
// resume support: fields outside the checkpoint layout, and the engine state
// that is not a field (versioned byte blob)
// ---------------------------------------------------------------------------
extern "C" int les_gpu_upload_field(les_gpu* c, const char* name, const double* data)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !name || !data) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_upload_field(c, name, data);
        const size_t b = (size_t)c->g.n * sizeof(double);
        double* dst = nullptr;
        if (strcmp(name, "nu") == 0) dst = c->r.nu;
        else if (strcmp(name, "alpha") == 0) dst = c->r.alpha;
        else if (strncmp(name, "stat", 4) == 0) {
            int a = atoi(name + 4);
            if (a < 0 || a > 9) return LES_ERR_PARAM;
            dst = c->r.stat[a];
        }
        if (!dst) return set_err(c, LES_ERR_PARAM, "field %s cannot be uploaded", name);
        HIPC(hipMemcpyAsync(dst, data, b, hipMemcpyHostToDevice, c->stream));
        HIPC(hipStreamSynchronize(c->stream));
        return LES_OK;
    });
}

extern "C" int les_gpu_set_stats_count(les_gpu* c, int count)
{
    return abi_guard(c, [&]() -> int {
        if (!c || count < 0) return LES_ERR_PARAM;
        c->stat_count = count;
        return LES_OK;
    });
}

static const char AUX_MAGIC[8] = {'L', 'E', 'S', 'A', 'U', 'X', '2', 0};

// Every aux blob starts with magic[8] "LESAUX<version>" and int mode, int step.
struct AuxRef {
    char magic[8];
    int mode, step, stat_count, nvort;
    int rand_type, rand_deg, rand_sep, attr_map;
    long long fptr, rptr, state, end;     // offsets into the state bytes
    char rstate[128];
};

extern "C" int les_gpu_export_aux(les_gpu* c, void* buf, size_t* bytes)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !bytes) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_export_aux(c, buf, bytes);
        const size_t nv = 7 * (size_t)c->p.num_vortices;
        const size_t need = sizeof(AuxRef) + nv * sizeof(double);
        if (!buf || *bytes < need) { *bytes = need; return buf ? LES_ERR_PARAM : LES_OK; }
        AuxRef a;
        memset(&a, 0, sizeof(a));
        memcpy(a.magic, AUX_MAGIC, 8);
        a.mode = LES_MODE_REFERENCE; a.step = c->step; a.stat_count = c->stat_count; a.nvort = c->p.num_vortices;
        a.rand_type = c->rd.rand_type; a.rand_deg = c->rd.rand_deg; a.rand_sep = c->rd.rand_sep;
        a.attr_map = c->attr_map;
        const char* base = c->rstate;
        a.fptr = (const char*)c->rd.fptr - base; a.rptr = (const char*)c->rd.rptr - base;
        a.state = (const char*)c->rd.state - base; a.end = (const char*)c->rd.end_ptr - base;
        memcpy(a.rstate, c->rstate, sizeof(a.rstate));
        memcpy(buf, &a, sizeof(a));
        memcpy((char*)buf + sizeof(a), c->vort.data(), nv * sizeof(double));
        *bytes = need;
        return LES_OK;
    });
}

extern "C" int les_gpu_import_aux(les_gpu* c, const void* buf, size_t bytes)
{
    return abi_guard(c, [&]() -> int {
        if (!c || !buf) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_import_aux(c, buf, bytes);
        AuxRef a;
        if (bytes < sizeof(a)) return set_err(c, LES_ERR_PARAM, "aux blob too short");
        memcpy(&a, buf, sizeof(a));
        const size_t nv = 7 * (size_t)c->p.num_vortices;
        if (memcmp(a.magic, AUX_MAGIC, 8) != 0 || a.mode != LES_MODE_REFERENCE || a.nvort != c->p.num_vortices ||
            bytes != sizeof(a) + nv * sizeof(double))
            return set_err(c, LES_ERR_PARAM, "aux blob does not match this context");
        const long long lim = (long long)sizeof(c->rstate);
        if (a.fptr < 0 || a.fptr >= lim || a.rptr < 0 || a.rptr >= lim || a.state < 0 || a.state >= lim || a.end < 0 ||
            a.end > lim)
            return set_err(c, LES_ERR_PARAM, "aux blob holds an invalid random-state layout");
        if (a.attr_map != 0 && a.attr_map != 1) return set_err(c, LES_ERR_PARAM, "aux blob holds an invalid attribute map");
        if (a.attr_map != c->attr_map) {
            int e = set_attribute_map(c, a.attr_map);
            if (e != LES_OK) return e;
        }
        memcpy(c->rstate, a.rstate, sizeof(c->rstate));
        c->rd.rand_type = a.rand_type; c->rd.rand_deg = a.rand_deg; c->rd.rand_sep = a.rand_sep;
        c->rd.fptr = (int32_t*)(c->rstate + a.fptr); c->rd.rptr = (int32_t*)(c->rstate + a.rptr);
        c->rd.state = (int32_t*)(c->rstate + a.state); c->rd.end_ptr = (int32_t*)(c->rstate + a.end);
        memcpy(c->vort.data(), (const char*)buf + sizeof(a), nv * sizeof(double));
        HIPC(hipMemcpy(c->r.vort, c->vort.data(), nv * sizeof(double), hipMemcpyHostToDevice));
        c->step = a.step;
        c->stat_count = a.stat_count;
        c->pt_base_r = a.step;
        c->probe.clear();
        c->probe_pending = -1;
        return LES_OK;
    });
}

// restart from a state without engine aux state (VTK restart, :4269-4280):
// step index and fresh SEM eddies (initialize_vortices :861-877)
extern "C" int les_gpu_set_step(les_gpu* c, int step)
{
    return abi_guard(c, [&]() -> int {
        if (!c || step < 0) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_set_step(c, step);
        c->step = step;
        c->pt_base_r = step;
        return LES_OK;
    });
}

extern "C" int les_gpu_init_sem(les_gpu* c)
{
    return abi_guard(c, [&]() -> int {
        if (!c) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT) return duct_init_sem(c);
        init_vortices(c);
        return LES_OK;
    });
}

// The reference's restart paths (binary :4184-4190, VTK :4269-4280) skip
// setting() and keep main()'s attribute map (:4123-4136).
extern "C" int les_gpu_use_resume_attributes(les_gpu* c)
{
    return abi_guard(c, [&]() -> int {
        if (!c) return LES_ERR_PARAM;
        if (c->p.mode == LES_MODE_DUCT)
            return set_err(c, LES_ERR_PARAM, "the duct scheme has one attribute map (inflow, fluid, outflow)");
        return set_attribute_map(c, 1);
    });
}

extern "C" int les_aux_step(const void* buf, size_t bytes, int* step)
{
    if (!buf || !step || bytes < 16 || memcmp(buf, "LESAUX", 6) != 0) return LES_ERR_PARAM;
    int mode = 0;
    memcpy(&mode, (const char*)buf + 8, sizeof(int));
    if (mode != LES_MODE_REFERENCE && mode != LES_MODE_DUCT) return LES_ERR_PARAM;
    memcpy(step, (const char*)buf + 12, sizeof(int));
    return LES_OK;
}

// ---------------------------------------------------------------------------
// HBM reference point: a plain streaming copy of `bytes` (FP64 elements, one
// read and one write each), timed with HIP events over `reps` launches.  Its
// rocprof FETCH_SIZE / WRITE_SIZE readings calibrate the counters the stencil
// kernels are measured with, and its GB/s is the achievable bandwidth.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_copy_stream(long long n, const double2* __restrict__ a,
                                                     double2* __restrict__ b)
{
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        b[q] = a[q];
}

extern "C" int les_copy_bandwidth(size_t bytes, int reps, double* ms_per_copy, double* gbps)
{
    return abi_guard(nullptr, [&]() -> int {
        if (bytes < 4096 || reps < 1 || !ms_per_copy || !gbps) return LES_ERR_PARAM;
        const long long n = (long long)(bytes / 16);
        void *a = nullptr, *b = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int rc = LES_OK;
        if (hipMalloc(&a, (size_t)n * 16) != hipSuccess || hipMalloc(&b, (size_t)n * 16) != hipSuccess ||
            hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
            hipEventCreate(&e1) != hipSuccess)
            rc = LES_ERR_HIP;
        if (rc == LES_OK && hipMemsetAsync(a, 0, (size_t)n * 16, s) != hipSuccess) rc = LES_ERR_HIP;
        const int grid = 4096;
        if (rc == LES_OK) {
            hipLaunchKernelGGL(k_copy_stream, dim3(grid), dim3(256), 0, s, n, (const double2*)a, (double2*)b);  // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps; ++r)
                hipLaunchKernelGGL(k_copy_stream, dim3(grid), dim3(256), 0, s, n, (const double2*)a, (double2*)b);
            (void)hipEventRecord(e1, s);
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) rc = LES_ERR_HIP;
        }
        if (rc == LES_OK) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            *ms_per_copy = ms / reps;
            *gbps = 2.0 * (double)n * 16 / (*ms_per_copy * 1e-3) / 1e9;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        return rc;
    });
}
