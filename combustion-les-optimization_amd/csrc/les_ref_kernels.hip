// les_ref_kernels.hip - gfx950 kernels of the reference scheme (FP64).
//
// These kernels restate the reference step (les_combustion_solver.c) with its
// operation order so that the FP64 device path reproduces the CPU results to
// reduction order (and device-libm ulps in exp/log/pow).  One thread per cell,
// k fastest across a wavefront; stencils read their neighbours straight from
// HBM/L2 (the reference-scheme RHS is FP64-VALU bound by its ~400 libm calls
// per cell, so LDS staging would not move its roofline).
#include "les_ref_kernels.hpp"

// ---------------------------------------------------------------------------
// box filter stencil (:1307-1470)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double box7(const double* f, long long q, long long sx, long long sy, double wc)
{   // :1329-1331 neighbour order
    return (f[q - sx] + f[q - sy] + f[q - 1] + wc * f[q] + f[q + 1] + f[q + sy] + f[q + sx]) / (6.0 + wc);
}

// ---------------------------------------------------------------------------
// K1: box filter (:1307-1470) + mixture properties (:1489-1881), fused per cell
//
// The reference evaluates every property fit from scratch: log T, (ln T)^2 and
// (ln T)^3 in each of the 30 fits (5 viscosities, 5 conductivities, 20
// binary diffusivities, :385-391), T^2..T^4 in each of the 5 NASA
// polynomials (:401-416), and the 25 Wilke factors phi_ab with their row sums
// twice, once for the viscosity and once for the conductivity mixing
// (:1530-1546, :1842-1863), whose inputs (X, mu_k, M) are the same.  Each of
// these is computed once per cell here with the same operations, so every
// value is bit for bit what the repeated evaluation gives; the molar-mass
// factors of phi_ab, constants of the mixture, once per block.
// ---------------------------------------------------------------------------
__device__ static double prop_fit_lt(double t, double lt, double lt2, double lt3, const double* co)
{   // prop_fit with ln t, (ln t)^2, (ln t)^3 given
    if (t <= 0) return 0.0;
    if (co[0] == 0.0 && co[1] == 0.0 && co[2] == 0.0 && co[3] == 0.0) return 0.0;
    double lp = co[0] + co[1] * lt + co[2] * lt2 + co[3] * lt3;
    return exp(lp);
}

__device__ static double cp_sp_pow(const les_mixture* m, int s, double tt, double t2, double t3, double t4)
{   // cp_sp with tt (t clamped as cp_sp does) and its powers given
    const les_species* sp = &m->sp[s];
    const double* a = (tt > sp->t_mid) ? sp->nasa_high : sp->nasa_low;
    double cpr = a[0] + a[1] * tt + a[2] * t2 + a[3] * t3 + a[4] * t4;
    double cpm = cpr * LES_R_UNIV;
    double mwk = sp->mw / 1000.0;
    if (mwk < 1e-9) return 1000.0;
    return cpm / mwk;
}

// update_Dk (:1695-1749) for species s: mixture-averaged diffusivity from the
// binary fits, with X, T and ln T powers given
__device__ static double dk_species(const les_mixture* m, int s, const double* X, double Tf, double lt,
                                    double lt2, double lt3)
{
    double sum = 0.0;
    for (int b = 0; b < LES_NSP; ++b) {
        if (s == b) continue;
        if (X[b] > 1e-12) {
            double d = prop_fit_lt(Tf, lt, lt2, lt3, m->diff_fit[s][b]) * 1e-4;
            if (d > 1e-12) sum += X[b] / d;
        }
    }
    double D;
    if (fabs(1.0 - X[s]) < 1e-9) D = prop_fit_lt(Tf, lt, lt2, lt3, m->diff_fit[s][(s != 2) ? 2 : 1]) * 1e-4;
    else if (sum > 1e-12) D = (1.0 - X[s]) / sum;
    else D = 1e-5;
    double ptot = LES_P_CONST + 0.0;        // P_filtered is identically zero
    if (ptot > 1.0) D *= (LES_PATM / ptot);
    return fmax(1e-9, fmin(1e-2, D));
}

// gs.n == 1: one thread per cell writes every filtered field and property.
// gs.n > 1: gridDim.z = nx * gs.n; slab b writes what gs.mask[b] names: bit s
// (s < LES_NSP) the diffusivity of species s, REF_PROPS_FIELDS the filtered
// fields and cp, REF_PROPS_TRANSPORT nu, mu, lambda and alpha.  A slab filters
// what it needs again and derives X, ln T and cp with the same operations, so
// every value has the same bits as in the one-slab form.
__global__ void __launch_bounds__(256) k_ref_filter_props(GridDims g, RefConsts rc, RefPtrs p, RefGroups gs)
{
    // per-pair factors of phi_ab that depend on the molar masses only
    __shared__ double s_m4[LES_NSP][LES_NSP], s_den[LES_NSP][LES_NSP];
    const les_mixture* m = p.mix;
    {
        const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
        if (tid < LES_NSP * LES_NSP) {
            const int a = tid / LES_NSP, b = tid % LES_NSP;
            const double M_i = m->sp[a].mw, M_j = m->sp[b].mw;
            s_m4[a][b] = pow(M_j / M_i, 0.25);
            s_den[a][b] = sqrt(8.0) * sqrt(1.0 + M_i / M_j);
        }
        __syncthreads();
    }
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int b = gs.n > 1 ? (int)blockIdx.z / g.nx : 0;
    int i = (int)blockIdx.z - b * g.nx;
    const unsigned want = gs.n > 1 ? gs.mask[b] : REF_PROPS_ALL;
    if (k >= g.nz || j >= g.ny) return;
    long long q = lidx(g, i, j, k);
    int a = p.attr[q];
    bool interior = (i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1 && k > 0 && k < g.nz - 1);
    double Yf[LES_NSP];
    double Tf, rf = 0.0;
    const bool w_fields = want & REF_PROPS_FIELDS, w_trans = want & REF_PROPS_TRANSPORT;
    const unsigned w_dk = want & ((1u << LES_NSP) - 1u);
    if (!w_fields && !w_trans) {           // diffusivities only: Y and T filtered again
        if (a != 3) {
            for (int sd = 0; sd < LES_NSP; ++sd)
                if (want & (1u << sd)) p.Dk[sd][q] = 0.0;
            return;
        }
        const double wc = rc.filter_wc;
        if (interior) {
            Tf = box7(p.T, q, g.sx, g.sy, wc);
            for (int s = 0; s < LES_NSP; ++s) Yf[s] = box7(p.Y[s], q, g.sx, g.sy, wc);
        } else {
            Tf = p.T[q];
            for (int s = 0; s < LES_NSP; ++s) Yf[s] = p.Y[s][q];
        }
        double X[LES_NSP];
        bool have_x = false;
        if (Tf > 0) {
            double sm = 0.0;
            for (int s = 0; s < LES_NSP; ++s) sm += Yf[s] / (m->sp[s].mw / 1000.0);
            if (!(sm < 1e-9)) {
                have_x = true;
                for (int s = 0; s < LES_NSP; ++s) X[s] = (Yf[s] / (m->sp[s].mw / 1000.0)) / sm;
            }
        }
        if (!have_x) {
            for (int sd = 0; sd < LES_NSP; ++sd)
                if (want & (1u << sd)) p.Dk[sd][q] = 1e-5;
            return;
        }
        const double lt = log(Tf), lt2 = pow(lt, 2.0), lt3 = pow(lt, 3.0);
        for (int sd = 0; sd < LES_NSP; ++sd)
            if (want & (1u << sd)) p.Dk[sd][q] = dk_species(m, sd, X, Tf, lt, lt2, lt3);
        return;
    }
    if (interior && a == 3) {
        const double wc = rc.filter_wc;
        if (w_fields) {
            p.uf[q] = box7(p.u, q, g.sx, g.sy, wc);
            p.vf[q] = box7(p.v, q, g.sx, g.sy, wc);
            p.wf[q] = box7(p.w, q, g.sx, g.sy, wc);
        }
        Tf = box7(p.T, q, g.sx, g.sy, wc);
        rf = box7(p.rho, q, g.sx, g.sy, wc);
        for (int s = 0; s < LES_NSP; ++s) Yf[s] = box7(p.Y[s], q, g.sx, g.sy, wc);
        if (w_fields) p.Gf[q] = box7(p.G, q, g.sx, g.sy, wc);
    } else {
        if (w_fields) {
            p.uf[q] = p.u[q]; p.vf[q] = p.v[q]; p.wf[q] = p.w[q]; p.Gf[q] = p.G[q];
        }
        Tf = p.T[q]; rf = p.rho[q];
        for (int s = 0; s < LES_NSP; ++s) Yf[s] = p.Y[s][q];
    }
    if (w_fields) {
        p.Tf[q] = Tf; p.rhof[q] = rf;
        for (int s = 0; s < LES_NSP; ++s) p.Yf[s][q] = Yf[s];
    }

    if (a != 3) {
        if (w_trans) {
            p.nu[q] = 1.5e-5; p.mu_mix[q] = 1.8e-5; p.lam[q] = 0.0; p.alpha[q] = 0.0;
        }
        for (int s = 0; s < LES_NSP; ++s)
            if (w_dk & (1u << s)) p.Dk[s][q] = 0.0;
        if (w_fields) p.cp[q] = 1000.0;
        return;
    }
    double X[LES_NSP];
    bool have_x = false;
    if (Tf > 0) {
        double sm = 0.0;
        for (int s = 0; s < LES_NSP; ++s) sm += Yf[s] / (m->sp[s].mw / 1000.0);
        if (!(sm < 1e-9)) {
            have_x = true;
            for (int s = 0; s < LES_NSP; ++s) X[s] = (Yf[s] / (m->sp[s].mw / 1000.0)) / sm;
        }
    }
    // ln T and its powers, shared by the 30 property fits
    double lt = 0.0, lt2 = 0.0, lt3 = 0.0;
    if (Tf > 0 && (w_trans || w_dk)) {
        lt = log(Tf);
        lt2 = pow(lt, 2.0);
        lt3 = pow(lt, 3.0);
    }
    double mu_k[LES_NSP], lk[LES_NSP];
    if (have_x && w_trans)
        for (int s = 0; s < LES_NSP; ++s) {
            mu_k[s] = prop_fit_lt(Tf, lt, lt2, lt3, m->sp[s].eta_fit) * 0.1;
            lk[s] = prop_fit_lt(Tf, lt, lt2, lt3, m->sp[s].lambda_fit) * 1e-5;
        }
    // Wilke mixing of mu and lambda (:1530-1546, :1842-1863): one set of
    // phi_ab and row sums, both accumulations in the reference's order
    double mu_acc = 0.0, lam_acc = 0.0;
    if (have_x && w_trans) {
        for (int aa = 0; aa < LES_NSP; ++aa) {
            if (X[aa] < 1e-12) continue;
            double mu_i = mu_k[aa], den = 0.0;
            for (int bb = 0; bb < LES_NSP; ++bb) {
                if (X[bb] < 1e-12) continue;
                double mu_j = mu_k[bb];
                double phi = pow(1.0 + pow(mu_i / mu_j, 0.5) * s_m4[aa][bb], 2.0) / s_den[aa][bb];
                den += X[bb] * phi;
            }
            if (den > 1e-9) {
                mu_acc += (X[aa] * mu_k[aa]) / den;
                lam_acc += (X[aa] * lk[aa]) / den;
            }
        }
    }
    // update_nu_mix (:1501-1557)
    if (!w_trans) {
    } else if (!have_x) {
        p.nu[q] = 1.5e-5; p.mu_mix[q] = 1.8e-5;
    } else {
        double mm = fmax(1e-7, mu_acc);
        p.mu_mix[q] = mm;
        double nu = (rf > 1e-3) ? mm / rf : 1.5e-5;
        p.nu[q] = fmax(1e-7, nu);
    }
    // update_Dk (:1695-1749)
    for (int s = 0; s < LES_NSP; ++s)
        if (w_dk & (1u << s)) p.Dk[s][q] = have_x ? dk_species(m, s, X, Tf, lt, lt2, lt3) : 1e-5;
    if (!w_fields && !w_trans) return;
    // update_Cp_mix (:1772-1786)
    double cpm;
    if (Tf <= 0) cpm = 1000.0;
    else {
        const double t2 = pow(Tf, 2.0), t3 = pow(Tf, 3.0), t4 = pow(Tf, 4.0);
        double cp = 0.0;
        for (int s = 0; s < LES_NSP; ++s) cp += Yf[s] * cp_sp_pow(m, s, Tf, t2, t3, t4);
        cpm = fmax(100.0, cp);
    }
    if (w_fields) p.cp[q] = cpm;
    // update_lambda_mix_alpha (:1811-1872)
    if (!w_trans) {
    } else if (!have_x) {
        p.lam[q] = 0.025; p.alpha[q] = 1.8e-5;
    } else {
        double lam = fmax(1e-3, lam_acc);
        p.lam[q] = lam;
        double al = (rf > 1e-6 && cpm > 1e-6) ? lam / (rf * cpm) : 1.8e-5;
        p.alpha[q] = fmax(1e-7, al);
    }
}

// ---------------------------------------------------------------------------
// K2: SGS viscosity (:1569-1633) + SSR flame speed (:1913-1965)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ref_nut_st(GridDims g, RefConsts rc, RefPtrs p)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int i = blockIdx.z;
    if (k < 1 || k >= g.nz - 1 || j < 1 || j >= g.ny - 1 || i < 1 || i >= g.nx - 1) return;
    long long q = lidx(g, i, j, k);
    if (p.attr[q] != 3) {
        p.S_mag[q] = 0.0; p.nu_t[q] = 0.0; p.ST[q] = rc.s_l;
        return;
    }
    const double DX = g.dx, DY = g.dy, DZ = g.dz;
    const long long X = g.sx, Y = g.sy;
    const double *uf = p.uf, *vf = p.vf, *wf = p.wf;
    double du_dx = (uf[q + X] - uf[q - X]) / (2.0 * DX);
    double dv_dy = (vf[q + Y] - vf[q - Y]) / (2.0 * DY);
    double dw_dz = (wf[q + 1] - wf[q - 1]) / (2.0 * DZ);
    double du_dy = (uf[q + Y] - uf[q - Y]) / (2.0 * DY);
    double du_dz = (uf[q + 1] - uf[q - 1]) / (2.0 * DZ);
    double dv_dx = (vf[q + X] - vf[q - X]) / (2.0 * DX);
    double dv_dz = (vf[q + 1] - vf[q - 1]) / (2.0 * DZ);
    double dw_dx = (wf[q + X] - wf[q - X]) / (2.0 * DX);
    double dw_dy = (wf[q + Y] - wf[q - Y]) / (2.0 * DY);
    double S12 = 0.5 * (du_dy + dv_dx);
    double S13 = 0.5 * (du_dz + dw_dx);
    double S23 = 0.5 * (dv_dz + dw_dy);
    double ss = 2.0 * (S12 * S12 + S13 * S13 + S23 * S23) + du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz;
    double smag = sqrt(2.0 * ss);
    double div = du_dx + dv_dy + dw_dz;
    p.S_mag[q] = smag;
    double gg = p.Gf[q];
    if (gg < 0.0) gg = 0.0;
    if (gg > 1.0) gg = 1.0;
    double cs = rc.cs_u + (rc.cs_b - rc.cs_u) * gg;
    cs += rc.cs_f * 4.0 * gg * (1.0 - gg);
    p.nu_t[q] = (cs * DX) * (cs * DX) * smag;
    // SSR-SGS S_T (:1926-1957)
    const double SL = rc.s_l;
    double nul = p.nu[q];
    double num = ss - (div * div) * (1.0 / 3.0);
    double Xe = (pow(DX, 4.0) / (nul * nul)) * num;
    double d_eta;
    if (Xe < LES_SMALL) d_eta = 1.0;
    else {
        double px = pow(Xe, 0.75);
        double al = 0.2187 * px;
        d_eta = 0.2340 * px * (al > 1.0 ? log(al) : 0.0);
    }
    double eta = DX / (d_eta + LES_SMALL);
    double dF = rc.nu_unburnt / (SL + LES_SMALL);
    double Dc = 8.0 * eta;
    double asf = 8.0 * exp(10.0 * dF / (Dc + LES_SMALL));
    double base = asf * eta / DX;
    double wr = pow(base, (2.0 - rc.d3));
    double te = 0.0;
    if (rc.tau * SL > LES_SMALL) te = DX * fmax(0.0, div) / (rc.tau * SL);
    double st = SL * fmax(0.0, wr + te);
    p.ST[q] = fmax(SL, st);
}

// ---------------------------------------------------------------------------
// K3: domain-mean density (:2536-2550); deterministic two-pass tree
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ref_rho_partial(GridDims g, const int* attr, const double* rhof,
                                                         double* partial)
{
    __shared__ double sm[256], sv[256];
    double dV = g.dx * g.dy * g.dz;
    double tm = 0.0, tv = 0.0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < g.n; q += (long long)gridDim.x * 256) {
        int a = attr[q];
        if (a == 1 || a == 2 || a == 3 || a == 5) { tm += rhof[q] * dV; tv += dV; }
    }
    sm[threadIdx.x] = tm; sv[threadIdx.x] = tv;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sm[threadIdx.x] += sm[threadIdx.x + s]; sv[threadIdx.x] += sv[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = sm[0]; partial[2 * blockIdx.x + 1] = sv[0]; }
}

__global__ void __launch_bounds__(256) k_ref_rho_final(int nblocks, const double* partial, double* rho_avg)
{
    __shared__ double sm[256], sv[256];
    double tm = 0.0, tv = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) { tm += partial[2 * b]; tv += partial[2 * b + 1]; }
    sm[threadIdx.x] = tm; sv[threadIdx.x] = tv;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sm[threadIdx.x] += sm[threadIdx.x + s]; sv[threadIdx.x] += sv[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) rho_avg[0] = (sv[0] > 0) ? (sm[0] / sv[0]) : 1.2;
}

// ---------------------------------------------------------------------------
// convection schemes (:637-727, :749-827, :2017-2159, :2190-2298)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double quick3(double uw2, double uw1, double dw1)
{
    return (0.0 * uw2 + 0.5 * uw1 + 0.5 * dw1);                                // :164-166
}

__device__ static double upwind1(const double* f, long long q, const long long* S, const double* vel,
                                 const double* d)
{
    double r;
    if (vel[0] >= 0.0) r = vel[0] * (f[q] - f[q - S[0]]) / d[0];
    else r = vel[0] * (f[q + S[0]] - f[q]) / d[0];
    if (vel[1] >= 0.0) r += vel[1] * (f[q] - f[q - S[1]]) / d[1];
    else r += vel[1] * (f[q + S[1]] - f[q]) / d[1];
    if (vel[2] >= 0.0) r += vel[2] * (f[q] - f[q - S[2]]) / d[2];
    else r += vel[2] * (f[q + S[2]] - f[q]) / d[2];
    return r;
}

__device__ static double conv_nonconservative(const double* f, long long q, bool near, const long long* S,
                                              const double* vel, const double* d)
{   // :637-678
    if (near) return upwind1(f, q, S, vel, d);
    double o[3];
    for (int a = 0; a < 3; ++a) {
        long long s = S[a];
        double ph, mh;
        if (vel[a] >= 0.0) { ph = quick3(f[q - s], f[q], f[q + s]); mh = quick3(f[q - 2 * s], f[q - s], f[q]); }
        else { ph = quick3(f[q + 2 * s], f[q + s], f[q]); mh = quick3(f[q + s], f[q], f[q - s]); }
        o[a] = vel[a] * (ph - mh) / d[a];
    }
    return o[0] + o[1] + o[2];
}

__device__ static double conv_conservative(const double* f, long long q, bool near, const long long* S,
                                           const double* const* U, const double* rho, const double* d)
{   // :683-727
    double acc = 0.0;
    for (int a = 0; a < 3; ++a) {
        long long s = S[a];
        const double* V = U[a];
        double term;
        if (near) {
            double mp = 0.5 * (rho[q + s] + rho[q]) * 0.5 * (V[q + s] + V[q]);
            double mm = 0.5 * (rho[q] + rho[q - s]) * 0.5 * (V[q] + V[q - s]);
            double fp = (mp > 0) ? f[q] : f[q + s];
            double fm = (mm > 0) ? f[q - s] : f[q];
            term = (mp * fp - mm * fm) / d[a];
        } else {
            double up = 0.5 * (V[q + s] + V[q]);
            double um = 0.5 * (V[q] + V[q - s]);
            double rp = 0.5 * (rho[q + s] + rho[q]);
            double rm = 0.5 * (rho[q] + rho[q - s]);
            double mp = rp * up, mm = rm * um;
            double fp = (up >= 0.0) ? quick3(f[q - s], f[q], f[q + s]) : quick3(f[q + 2 * s], f[q + s], f[q]);
            double fm = (um >= 0.0) ? quick3(f[q - 2 * s], f[q - s], f[q]) : quick3(f[q + s], f[q], f[q - s]);
            term = (mp * fp - mm * fm) / d[a];
        }
        acc = (a == 0) ? term : acc + term;
    }
    return acc;
}

__device__ __forceinline__ double superbee(double r)                          // :2017-2028
{
    if (r <= 0.0) return 0.0;
    return fmax(fmin(2.0 * r, 1.0), fmin(r, 2.0));
}

__device__ __forceinline__ double sratio(double up, double down)
{
    return (fabs(down) < LES_SMALL) ? 1e9 : up / down;                         // :2062
}

__device__ static double conv_muscl(const double* f, long long q, bool near, const long long* S,
                                    const double* vel, const double* d)
{   // :2037-2159
    if (near) return upwind1(f, q, S, vel, d);
    double o[3];
    for (int a = 0; a < 3; ++a) {
        long long s = S[a];
        double fm2 = f[q - 2 * s], fm1 = f[q - s], f0 = f[q], fp1 = f[q + s], fp2 = f[q + 2 * s];
        double ddm = f0 - fm1, ddp = fp1 - f0;
        double mh, ph;
        if (vel[a] >= 0.0) {
            mh = fm1 + 0.5 * superbee(sratio(fm1 - fm2, ddm)) * ddm;
            ph = f0 + 0.5 * superbee(sratio(f0 - fm1, ddp)) * ddp;
        } else {
            mh = f0 - 0.5 * superbee(sratio(fp1 - f0, ddm)) * ddm;
            ph = fp1 - 0.5 * superbee(sratio(fp2 - fp1, ddp)) * ddp;
        }
        o[a] = vel[a] * (ph - mh) / d[a];
    }
    return o[0] + o[1] + o[2];
}

__device__ static double teno5(double v0, double v1, double v2, double v3, double v4)   // :2190-2236
{
    double is0 = (13.0 / 12.0) * pow(v0 - 2.0 * v1 + v2, 2.0) + (1.0 / 4.0) * pow(v0 - 4.0 * v1 + 3.0 * v2, 2.0);
    double is1 = (13.0 / 12.0) * pow(v1 - 2.0 * v2 + v3, 2.0) + (1.0 / 4.0) * pow(v1 - v3, 2.0);
    double is2 = (13.0 / 12.0) * pow(v2 - 2.0 * v3 + v4, 2.0) + (1.0 / 4.0) * pow(3.0 * v2 - 4.0 * v3 + v4, 2.0);
    int d0 = is0 < 1.0e-6, d1 = is1 < 1.0e-6, d2 = is2 < 1.0e-6;
    double p0 = (2.0 * v0 - 7.0 * v1 + 11.0 * v2) / 6.0;
    double p1 = (-v1 + 5.0 * v2 + 2.0 * v3) / 6.0;
    double p2 = (2.0 * v2 + 5.0 * v3 - v4) / 6.0;
    if (d0 + d1 + d2 == 0) {
        if (is0 <= is1 && is0 <= is2) return p0;
        else if (is1 <= is0 && is1 <= is2) return p1;
        return p2;
    }
    double w0 = d0 * 0.1, w1 = d1 * 0.6, w2 = d2 * 0.3;
    double ws = w0 + w1 + w2;
    if (ws < 1e-40) return p1;
    return (w0 * p0 + w1 * p1 + w2 * p2) / ws;
}

__device__ static double conv_teno(const double* f, long long q, bool near2, bool near3, const long long* S,
                                   const double* vel, const double* d)
{   // :2245-2298
    if (near3) return conv_nonconservative(f, q, near2, S, vel, d);
    double o[3];
    for (int a = 0; a < 3; ++a) {
        long long s = S[a];
        double ph, mh;
        if (vel[a] >= 0.0) {
            ph = teno5(f[q - 2 * s], f[q - s], f[q], f[q + s], f[q + 2 * s]);
            mh = teno5(f[q - 3 * s], f[q - 2 * s], f[q - s], f[q], f[q + s]);
        } else {
            mh = teno5(f[q + 2 * s], f[q + s], f[q], f[q - s], f[q - 2 * s]);
            ph = teno5(f[q + 3 * s], f[q + 2 * s], f[q + s], f[q], f[q - s]);
        }
        o[a] = vel[a] * (ph - mh) / d[a];
    }
    return o[0] + o[1] + o[2];
}

__device__ static double weno5_grad(const double* f, long long q, long long s, double dx)   // :749-827
{
    double p[6];
    for (int o = 0; o < 6; ++o) p[o] = f[q + (o - 2) * s];
    const double eps = 1.0e-6;
    double is0 = (13.0 / 12.0) * pow(p[1] - 2.0 * p[2] + p[3], 2.0) + (1.0 / 4.0) * pow(p[1] - 4.0 * p[2] + 3.0 * p[3], 2.0);
    double is1 = (13.0 / 12.0) * pow(p[2] - 2.0 * p[3] + p[4], 2.0) + (1.0 / 4.0) * pow(p[2] - p[4], 2.0);
    double is2 = (13.0 / 12.0) * pow(p[3] - 2.0 * p[4] + p[5], 2.0) + (1.0 / 4.0) * pow(3.0 * p[3] - 4.0 * p[4] + p[5], 2.0);
    double a0 = (3.0 / 10.0) / pow(eps + is0, 2.0);
    double a1 = (6.0 / 10.0) / pow(eps + is1, 2.0);
    double a2 = (1.0 / 10.0) / pow(eps + is2, 2.0);
    double sa = a0 + a1 + a2;
    double ph = (a0 / sa) * ((2.0 * p[1] - 7.0 * p[2] + 11.0 * p[3]) / 6.0) +
                (a1 / sa) * ((-p[2] + 5.0 * p[3] + 2.0 * p[4]) / 6.0) +
                (a2 / sa) * ((2.0 * p[3] + 5.0 * p[4] - p[5]) / 6.0);
    is0 = (13.0 / 12.0) * pow(p[0] - 2.0 * p[1] + p[2], 2.0) + (1.0 / 4.0) * pow(p[0] - 4.0 * p[1] + 3.0 * p[2], 2.0);
    is1 = (13.0 / 12.0) * pow(p[1] - 2.0 * p[2] + p[3], 2.0) + (1.0 / 4.0) * pow(p[1] - p[3], 2.0);
    is2 = (13.0 / 12.0) * pow(p[2] - 2.0 * p[3] + p[4], 2.0) + (1.0 / 4.0) * pow(3.0 * p[2] - 4.0 * p[3] + p[4], 2.0);
    a0 = (1.0 / 10.0) / pow(eps + is0, 2.0);
    a1 = (6.0 / 10.0) / pow(eps + is1, 2.0);
    a2 = (3.0 / 10.0) / pow(eps + is2, 2.0);
    sa = a0 + a1 + a2;
    double mh = (a0 / sa) * ((-p[0] + 5.0 * p[1] + 2.0 * p[2]) / 6.0) +
                (a1 / sa) * ((2.0 * p[1] + 5.0 * p[2] - p[3]) / 6.0) +
                (a2 / sa) * ((11.0 * p[2] - 7.0 * p[3] + 2.0 * p[4]) / 6.0);
    return (ph - mh) / dx;
}

__device__ __forceinline__ double fluxdiv(double gp, double gm, double fp, double f0, double fm, double d)
{
    return (gp * (fp - f0) / d - gm * (f0 - fm) / d) / d;                      // :2757-2759
}

// ---------------------------------------------------------------------------
// K4: RHS of momentum, G, species and T (:2552-2962)
// ---------------------------------------------------------------------------
// gs.n == 1: one thread computes all ten outputs of its cell.  gs.n > 1:
// gridDim.z = nx * gs.n and slab b of nx blocks computes the outputs in
// gs.mask[b] (bit f: drhou, drhov, drhow, dT, dG with |grad G|, dY0..dY4), so
// a small grid runs several times the waves, each with a part of the
// instruction stream.  Every output is the same arithmetic either way (no FMA
// contraction in this file), so every grouping gives identical bits.
__global__ void __launch_bounds__(256) k_ref_rhs(GridDims g, RefConsts rc, RefPtrs p, int t_loop, RefGroups gs)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int b = gs.n > 1 ? (int)blockIdx.z / g.nx : 0;
    int i = (int)blockIdx.z - b * g.nx;
    const unsigned want = gs.n > 1 ? gs.mask[b] : REF_RHS_ALL;
    const bool do_T = want & (1u << 3), do_G = want & (1u << 4);
    if (k < 1 || k >= g.nz - 1 || j < 1 || j >= g.ny - 1 || i < 1 || i >= g.nx - 1) return;
    long long q = lidx(g, i, j, k);
    if (p.attr[q] != 3) {
        for (int f = 0; f < LES_NRHS; ++f)
            if (want & (1u << f)) p.out[f][q] = 0.0;
        if (do_G) p.gradG[q] = 0.0;
        return;
    }
    const double DX = g.dx, DY = g.dy, DZ = g.dz;
    const long long X = g.sx, Y = g.sy;
    const long long S[3] = {X, Y, 1};
    const double d[3] = {DX, DY, DZ};
    const double *uf = p.uf, *vf = p.vf, *wf = p.wf, *rf = p.rhof, *Gf = p.Gf, *Tf = p.Tf, *nut = p.nu_t;
    const bool near2 = (i < 2 || i >= g.nx - 2 || j < 2 || j >= g.ny - 2 || k < 2 || k >= g.nz - 2);
    const bool near3 = (i < 3 || i >= g.nx - 3 || j < 3 || j >= g.ny - 3 || k < 3 || k >= g.nz - 3);
    const double vel[3] = {uf[q], vf[q], wf[q]};
    const double* U[3] = {uf, vf, wf};
    const double rho_avg = p.rho_avg[0];

    const double Gl = Gf[q];
    const double wdt = rc.accel_width, ctr = rc.accel_center;
    // (a) momentum (:2556-2691)
    if (want & 7u) {
        double conv[3] = {0.0, 0.0, 0.0}, diff[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a)
            if (want & (1u << a)) conv[a] = conv_conservative(U[a], q, near2, S, U, rf, d);
        double mu_eff = p.mu_mix[q] + rf[q] * nut[q];
        for (int a = 0; a < 3; ++a) {
            if (!(want & (1u << a))) continue;
            const double* V = U[a];
            double lx = (V[q + X] - 2.0 * V[q] + V[q - X]) / (DX * DX);
            double ly = (V[q + Y] - 2.0 * V[q] + V[q - Y]) / (DY * DY);
            double lz = (V[q + 1] - 2.0 * V[q] + V[q - 1]) / (DZ * DZ);
            diff[a] = mu_eff * (lx + ly + lz);
        }
        double grav = (rf[q] - rho_avg) * LES_G_GRAV;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        if (Gl > 0.01 && Gl < 0.99) {
            double gx = (Gf[q + X] - Gf[q - X]) / (2.0 * DX);
            double gy = (Gf[q + Y] - Gf[q - Y]) / (2.0 * DY);
            double gz = (Gf[q + 1] - Gf[q - 1]) / (2.0 * DZ);
            if (gx < 0.0) gx = 0.0;
            if (gy < 0.0) gy = 0.0;
            double zrel = (k * DZ) - rc.lz_half;
            if (zrel > 0.0) { if (gz < 0.0) gz = 0.0; }
            else { if (gz > 0.0) gz = 0.0; }
            double df = Gl - ctr;
            double win = exp(-(df * df) / (2.0 * wdt * wdt));
            double common = rc.accel_strength * rf[q] * win * 1.0;
            fx = common * gx * rc.weight_x;
            fy = common * gy * 1.0;
            fz = common * gz * 1.0;
            double zp = k * DZ, ramp = 1.0;
            if (zp < rc.ramp_distance) ramp = zp / rc.ramp_distance;
            fx *= ramp; fy *= ramp; fz *= ramp;
        }
        const double Fm = rc.max_force;
        if (fx > Fm) fx = Fm;
        if (fx < -Fm) fx = -Fm;
        if (fy > Fm) fy = Fm;
        if (fy < -Fm) fy = -Fm;
        if (fz > Fm) fz = Fm;
        if (fz < -Fm) fz = -Fm;
        if (want & 1u) p.out[0][q] = -conv[0] + diff[0] + grav + fx;
        if (want & 2u) p.out[1][q] = -conv[1] + diff[1] + fy;
        if (want & 4u) p.out[2][q] = -conv[2] + diff[2] + fz;
    }
    if (!(want & ~7u)) return;

    const double sct = rc.sc_t;
    double nu0 = nut[q];
    double nxp = nut[q + X], nxm = nut[q - X], nyp = nut[q + Y], nym = nut[q - Y], nzp = nut[q + 1], nzm = nut[q - 1];
    // (b) G (:2710-2818)
    if (do_G) {
        double cg = conv_teno(Gf, q, near2, near3, S, vel, d);
        double sgs = fluxdiv((nxp + nu0) / (2.0 * sct), (nu0 + nxm) / (2.0 * sct), Gf[q + X], Gf[q], Gf[q - X], DX) +
                     fluxdiv((nyp + nu0) / (2.0 * sct), (nu0 + nym) / (2.0 * sct), Gf[q + Y], Gf[q], Gf[q - Y], DY) +
                     fluxdiv((nzp + nu0) / (2.0 * sct), (nu0 + nzm) / (2.0 * sct), Gf[q + 1], Gf[q], Gf[q - 1], DZ);
        double gx, gy, gz;
        if (t_loop >= rc.transition_step && !near3) {
            gx = weno5_grad(Gf, q, X, DX);
            gy = weno5_grad(Gf, q, Y, DY);
            gz = weno5_grad(Gf, q, 1, DZ);
        } else {
            gx = (Gf[q + X] - Gf[q - X]) / (2.0 * DX);
            gy = (Gf[q + Y] - Gf[q - Y]) / (2.0 * DY);
            gz = (Gf[q + 1] - Gf[q - 1]) / (2.0 * DZ);
        }
        double gmag = sqrt(gx * gx + gy * gy + gz * gz);
        p.gradG[q] = gmag;
        double prop = p.ST[q] * gmag;
        double art = 0.0;
        if (Gl > 0.1 && Gl < 0.99) {
            double df = Gl - ctr;
            double win = exp(-(df * df) / (2.0 * wdt * wdt));
            double lx = (Gf[q + X] - 2.0 * Gl + Gf[q - X]) / (DX * DX);
            double ly = (Gf[q + Y] - 2.0 * Gl + Gf[q - Y]) / (DY * DY);
            double lz = (Gf[q + 1] - 2.0 * Gl + Gf[q - 1]) / (DZ * DZ);
            double rp = fabs(j * DY);
            double tr = rc.tip_radius;
            double tip = 1.0 - exp(-(rp * rp) / (tr * tr));
            double coef = 5.0 * (nu0 + 1.0e-5) * win * tip;
            art = coef * (lx + ly + lz);
        }
        p.out[4][q] = -cg + prop + sgs + art;
    }
    if (!(want & ~31u) && !do_T) return;

    // (c) species (:2842-2896)
    double gv = Gl;
    if (gv < 0.0) gv = 0.0;
    if (gv > 1.0) gv = 1.0;
    double dxp = (nxp + nu0) / (2.0 * sct), dxm = (nu0 + nxm) / (2.0 * sct);
    double dyp = (nyp + nu0) / (2.0 * sct), dym = (nu0 + nym) / (2.0 * sct);
    double dzp = (nzp + nu0) / (2.0 * sct), dzm = (nu0 + nzm) / (2.0 * sct);
    for (int s = 0; s < LES_NSP; ++s) {
        if (!(want & (1u << (5 + s)))) continue;
        const double* Yf = p.Yf[s];
        double cy = conv_muscl(Yf, q, near2, S, vel, d);
        double lx = (Yf[q + X] - 2.0 * Yf[q] + Yf[q - X]) / (DX * DX);
        double ly = (Yf[q + Y] - 2.0 * Yf[q] + Yf[q - Y]) / (DY * DY);
        double lz = (Yf[q + 1] - 2.0 * Yf[q] + Yf[q - 1]) / (DZ * DZ);
        double mol = p.Dk[s][q] * (lx + ly + lz);
        double sg = fluxdiv(dxp, dxm, Yf[q + X], Yf[q], Yf[q - X], DX) +
                    fluxdiv(dyp, dym, Yf[q + Y], Yf[q], Yf[q - Y], DY) +
                    fluxdiv(dzp, dzm, Yf[q + 1], Yf[q], Yf[q - 1], DZ);
        double tgt;
        if (s == 0) tgt = rc.inletY[0] * (1.0 - gv);
        else if (s == 1) { double cons = rc.inletY[0] * rc.st_o2; tgt = fmax(0.0, rc.inletY[1] - cons * gv); }
        else if (s == 3) { double pr = rc.inletY[0] * rc.st_co2; tgt = rc.inletY[3] + pr * gv; }
        else if (s == 4) { double pr = rc.inletY[0] * rc.st_h2o; tgt = rc.inletY[4] + pr * gv; }
        else tgt = rc.inletY[2];
        p.out[5 + s][q] = -cy + mol + sg + rc.relaxation * (tgt - Yf[q]);
    }

    if (!do_T) return;

    // (d) temperature (:2916-2955)
    double ct = conv_muscl(Tf, q, near2, S, vel, d);
    double tx = (Tf[q + X] - 2.0 * Tf[q] + Tf[q - X]) / (DX * DX);
    double ty = (Tf[q + Y] - 2.0 * Tf[q] + Tf[q - Y]) / (DY * DY);
    double tz = (Tf[q + 1] - 2.0 * Tf[q] + Tf[q - 1]) / (DZ * DZ);
    double md = 0.0;
    if (rf[q] > 1e-6 && p.cp[q] > 1e-6) md = p.lam[q] / (rf[q] * p.cp[q]) * (tx + ty + tz);
    const double prt = rc.pr_t;
    double sd = fluxdiv((nxp + nu0) / (2.0 * prt), (nu0 + nxm) / (2.0 * prt), Tf[q + X], Tf[q], Tf[q - X], DX) +
                fluxdiv((nyp + nu0) / (2.0 * prt), (nu0 + nym) / (2.0 * prt), Tf[q + Y], Tf[q], Tf[q - Y], DY) +
                fluxdiv((nzp + nu0) / (2.0 * prt), (nu0 + nzm) / (2.0 * prt), Tf[q + 1], Tf[q], Tf[q - 1], DZ);
    double Tt = rc.inlet_temperature + gv * (rc.burnt_temperature - rc.inlet_temperature);
    p.out[3][q] = -ct + md + sd + rc.relaxation * (Tt - Tf[q]);
}

// ---------------------------------------------------------------------------
// K5/K6: RK4 stage combine + EOS (:4432-4455 ...) and final combine (:4517-4544)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void eos_cell(const RefConsts& rc, const double* Yc, double T, double fallback,
                                         double& rho)
{   // update_rho_from_EOS (:1893-1905)
    double s = 0.0;
    for (int b = 0; b < LES_NSP; ++b) s += Yc[b] / (rc.mw[b] / 1000.0);
    if (s > 1e-9 && T > 1.0) {
        double wtm = 1.0 / s;
        double rmix = LES_R_UNIV / wtm;
        rho = LES_P_CONST / (rmix * T);
    } else {
        rho = fallback;
    }
}

// DT of the step into device memory, so the RK stages need not wait for the
// host (which reads the same maxima back for the eddies)
__global__ void k_ref_dt_final(const unsigned long long* maxbits, double inlet_velocity, double courant,
                               double courant_diffusion, double dx, double* dt)
{
    if (threadIdx.x != 0) return;
    const double mv = __longlong_as_double((long long)maxbits[0]);
    const double mn = __longlong_as_double((long long)maxbits[1]);
    const double ma = __longlong_as_double((long long)maxbits[2]);
    dt[0] = ref_dt_rule(mv, mn, ma, inlet_velocity, courant, courant_diffusion, dx);
}

// stage < 3: the RK4 stage update with coefficient 0.5 DT, 0.5 DT, DT (:4432-4455)
__global__ void __launch_bounds__(256) k_ref_stage(GridDims g, RefConsts rc, RefStagePtrs s, int stage,
                                                   const double* dtp)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.n) return;
    const double DT = dtp[0];
    const double cdt = stage < 2 ? 0.5 * DT : 1.0 * DT;
    int a = s.attr[q];
    double Tn, Gn, Yn[LES_NSP];
    if (a == 3) {
        Tn = s.T0[q] + cdt * s.k[stage][3][q];
        Gn = s.G0[q] + cdt * s.k[stage][4][q];
        for (int b = 0; b < LES_NSP; ++b) Yn[b] = s.Y0[b][q] + cdt * s.k[stage][5 + b][q];
        Tn = fmax(200.0, fmin(3000.0, Tn));
        Gn = fmax(0.0, fmin(1.0, Gn));
        double sum = 0;
        for (int b = 0; b < LES_NSP; ++b) { Yn[b] = fmax(0.0, fmin(1.0, Yn[b])); sum += Yn[b]; }
        if (sum > LES_SMALL) for (int b = 0; b < LES_NSP; ++b) Yn[b] /= sum;
    } else {
        Tn = s.T0[q]; Gn = s.G0[q];
        for (int b = 0; b < LES_NSP; ++b) Yn[b] = s.Y0[b][q];
    }
    s.Tt[q] = Tn; s.Gt[q] = Gn;
    for (int b = 0; b < LES_NSP; ++b) s.Yt[b][q] = Yn[b];
    double rt = s.rhot[q];
    if (a == 3 || a == 2) { eos_cell(rc, Yn, Tn, s.rho[q], rt); s.rhot[q] = rt; }
    if (a == 3) {
        double ru = s.rhou0[q] + cdt * s.k[stage][0][q];
        double rv = s.rhov0[q] + cdt * s.k[stage][1][q];
        double rw = s.rhow0[q] + cdt * s.k[stage][2][q];
        if (rt > LES_SMALL) { s.ut[q] = ru / rt; s.vt[q] = rv / rt; s.wt[q] = rw / rt; }
        else { s.ut[q] = 0; s.vt[q] = 0; s.wt[q] = 0; }
    } else {
        s.ut[q] = s.u[q]; s.vt[q] = s.v[q]; s.wt[q] = s.w[q];
    }
}

__global__ void __launch_bounds__(256) k_ref_final(GridDims g, RefConsts rc, RefStagePtrs s, const double* dtp)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.n) return;
    const double h = dtp[0] / 6.0;
    int a = s.attr[q];
    if (a != 3 && a != 2) return;
    double* const* k1 = s.k[0];
    double* const* k2 = s.k[1];
    double* const* k3 = s.k[2];
    double* const* k4 = s.k[3];
    double Tn, Yn[LES_NSP];
    if (a == 3) {
        Tn = s.T0[q] + h * (k1[3][q] + 2.0 * k2[3][q] + 2.0 * k3[3][q] + k4[3][q]);
        double Gn = s.G0[q] + h * (k1[4][q] + 2.0 * k2[4][q] + 2.0 * k3[4][q] + k4[4][q]);
        for (int b = 0; b < LES_NSP; ++b)
            Yn[b] = s.Y0[b][q] + h * (k1[5 + b][q] + 2.0 * k2[5 + b][q] + 2.0 * k3[5 + b][q] + k4[5 + b][q]);
        Tn = fmax(200.0, fmin(3000.0, Tn));
        Gn = fmax(0.0, fmin(1.0, Gn));
        double sum = 0;
        for (int b = 0; b < LES_NSP; ++b) { Yn[b] = fmax(0.0, fmin(1.0, Yn[b])); sum += Yn[b]; }
        if (sum > LES_SMALL) for (int b = 0; b < LES_NSP; ++b) Yn[b] /= sum;
        s.T[q] = Tn; s.G[q] = Gn;
        for (int b = 0; b < LES_NSP; ++b) s.Y[b][q] = Yn[b];
    } else {
        Tn = s.T[q];
        for (int b = 0; b < LES_NSP; ++b) Yn[b] = s.Y[b][q];
    }
    double r = s.rho[q];
    eos_cell(rc, Yn, Tn, r, r);
    s.rho[q] = r;
    if (a == 3) {
        double ru = s.rhou0[q] + h * (k1[0][q] + 2.0 * k2[0][q] + 2.0 * k3[0][q] + k4[0][q]);
        double rv = s.rhov0[q] + h * (k1[1][q] + 2.0 * k2[1][q] + 2.0 * k3[1][q] + k4[1][q]);
        double rw = s.rhow0[q] + h * (k1[2][q] + 2.0 * k2[2][q] + 2.0 * k3[2][q] + k4[2][q]);
        s.rhou[q] = ru; s.rhov[q] = rv; s.rhow[q] = rw;
        if (r > LES_SMALL) { s.u[q] = ru / r; s.v[q] = rv / r; s.w[q] = rw / r; }
        else { s.u[q] = 0; s.v[q] = 0; s.w[q] = 0; }
    }
}

// save Phi_n (:4421-4426)
__global__ void __launch_bounds__(256) k_ref_save(GridDims g, RefStagePtrs s)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.n) return;
    s.rhou0[q] = s.rhou[q]; s.rhov0[q] = s.rhov[q]; s.rhow0[q] = s.rhow[q];
    s.T0[q] = s.T[q]; s.G0[q] = s.G[q];
    for (int b = 0; b < LES_NSP; ++b) s.Y0[b][q] = s.Y[b][q];
}

// ---------------------------------------------------------------------------
// boundary conditions (:2967-3161)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bc_copy(const RefStagePtrs& s, long long dst, long long src, bool vel)
{
    if (vel) { s.u[dst] = s.u[src]; s.v[dst] = s.v[src]; s.w[dst] = s.w[src]; }
    s.T[dst] = s.T[src]; s.G[dst] = s.G[src];
    for (int b = 0; b < LES_NSP; ++b) s.Y[b][dst] = s.Y[b][src];
}

__device__ __forceinline__ void bc_ambient(const RefConsts& rc, const RefStagePtrs& s, long long q)
{
    s.u[q] = 0.0; s.v[q] = 0.0; s.w[q] = 0.0;
    s.T[q] = rc.ambient_temperature; s.G[q] = 0.0;
    for (int b = 0; b < LES_NSP; ++b) s.Y[b][q] = rc.ambY[b];
}

// phase 1: inflow face i=0 and outflow face i=nx-1 (:2978-3043)
__global__ void __launch_bounds__(256) k_ref_bc_x(GridDims g, RefConsts rc, RefStagePtrs s, const double* inlet,
                                                  const double* noise)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (k >= g.nz) return;
    long long pl = (long long)j * g.nz + k;
    long long q0 = pl;
    int a = s.attr[q0];
    if (a == 1) {
        const double U = rc.inlet_velocity, amp = rc.noise_amp;
        double r1 = noise[3 * pl], r2 = noise[3 * pl + 1], r3 = noise[3 * pl + 2];
        s.u[q0] = fmax(0.0, U + inlet[pl] + amp * r1);
        s.v[q0] = inlet[g.ny * (long long)g.nz + pl] + amp * r2;
        s.w[q0] = inlet[2 * g.ny * (long long)g.nz + pl] + amp * r3;
        s.T[q0] = rc.inlet_temperature; s.G[q0] = 0.0;
        for (int b = 0; b < LES_NSP; ++b) s.Y[b][q0] = rc.inletY[b];
    } else if (a == 4) {
        if (s.u[q0 + g.sx] < 0) bc_ambient(rc, s, q0);
        else bc_copy(s, q0, q0 + g.sx, true);
    } else if (a == 6) {
        s.u[q0] = 0.0; s.v[q0] = 0.0; s.w[q0] = 0.0;
        bc_copy(s, q0, q0 + g.sx, false);
    }
    long long qn = (long long)(g.nx - 1) * g.sx + pl;
    if (s.attr[qn] == 2) bc_copy(s, qn, qn - g.sx, true);
}

// phase 2a: y faces j=0 (symmetry) and j=ny-1 (free) for 0<i<nx-1 (:3052-3088)
__global__ void __launch_bounds__(256) k_ref_bc_y(GridDims g, RefConsts rc, RefStagePtrs s)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y + 1;
    if (k >= g.nz || i >= g.nx - 1) return;
    long long q = lidx(g, i, 0, k);
    if (s.attr[q] == 5) {
        long long r = q + g.sy;
        s.T[q] = s.T[r]; s.G[q] = s.G[r];
        for (int b = 0; b < LES_NSP; ++b) s.Y[b][q] = s.Y[b][r];
        s.u[q] = s.u[r]; s.v[q] = 0.0; s.w[q] = s.w[r];
    }
    q = lidx(g, i, g.ny - 1, k);
    if (s.attr[q] == 4) {
        if (s.v[q - g.sy] < 0) bc_ambient(rc, s, q);
        else bc_copy(s, q, q - g.sy, true);
    }
}

// phase 2b: z faces k=0 and k=nz-1 for 0<i<nx-1, all j (:3092-3122)
__global__ void __launch_bounds__(256) k_ref_bc_z(GridDims g, RefConsts rc, RefStagePtrs s)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y + 1;
    if (j >= g.ny || i >= g.nx - 1) return;
    long long q = lidx(g, i, j, 0);
    if (s.attr[q] == 4) {
        if (s.w[q + 1] > 0) bc_ambient(rc, s, q);
        else bc_copy(s, q, q + 1, true);
    }
    q = lidx(g, i, j, g.nz - 1);
    if (s.attr[q] == 4) {
        if (s.w[q - 1] < 0) bc_ambient(rc, s, q);
        else bc_copy(s, q, q - 1, true);
    }
}

// phase 3: EOS, momentum, mass-fraction normalisation on every cell (:3130-3160)
__global__ void __launch_bounds__(256) k_ref_bc_eos(GridDims g, RefConsts rc, RefStagePtrs s)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.n) return;
    double sm = 0;
    for (int b = 0; b < LES_NSP; ++b) sm += s.Y[b][q] / (rc.mw[b] / 1000.0);
    double T = s.T[q], r;
    if (sm > 1e-9 && T > 1.0) r = LES_P_CONST / (LES_R_UNIV / (1.0 / sm) * T);
    else r = 1.2;
    s.rho[q] = r;
    s.rhou[q] = r * s.u[q]; s.rhov[q] = r * s.v[q]; s.rhow[q] = r * s.w[q];
    int a = s.attr[q];
    if (a == 3 || a == 2 || a == 5) {
        double sy = 0.0;
        for (int b = 0; b < LES_NSP; ++b) sy += s.Y[b][q];
        if (sy > 1e-6) for (int b = 0; b < LES_NSP; ++b) s.Y[b][q] /= sy;
    }
}

// ---------------------------------------------------------------------------
// DT maxima (:1472-1485, :4401-4408) - exact (order-independent) max
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ref_dt_max(GridDims g, const int* attr, const double* u, const double* v,
                                                    const double* w, const double* nu, const double* alpha,
                                                    unsigned long long* out3)
{
    double mv = 0.0, mn = 0.0, ma = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < g.n; q += (long long)gridDim.x * blockDim.x) {
        int a = attr[q];
        if (a == 1 || a == 3) {
            double vel = sqrt(u[q] * u[q] + v[q] * v[q] + w[q] * w[q]);
            if (vel > mv) mv = vel;
        }
        if (a == 3) {
            if (nu[q] > mn) mn = nu[q];
            if (alpha[q] > ma) ma = alpha[q];
        }
    }
    // wavefront max via shuffles, the block's four wave maxima through LDS,
    // then one atomic per block and value: the atomics on the three words are
    // serialised, so their count (not the cell loop) sets this kernel's time
    // when every wave issues its own
    __shared__ double red[3][4];
    for (int o = 32; o > 0; o >>= 1) {
        mv = fmax(mv, __shfl_xor(mv, o));
        mn = fmax(mn, __shfl_xor(mn, o));
        ma = fmax(ma, __shfl_xor(ma, o));
    }
    const int wv = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][wv] = mv; red[1][wv] = mn; red[2][wv] = ma; }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c3 = (int)threadIdx.x;
        double m = red[c3][0];
        for (int b = 1; b < (int)(blockDim.x >> 6); ++b) m = fmax(m, red[c3][b]);
        atomic_max_pos(&out3[c3], m);
    }
}

// ---------------------------------------------------------------------------
// SEM inflow plane (:964-1003) and its normalisation (:1005-1074)
// ---------------------------------------------------------------------------
// One wave per inlet-plane cell, 64 eddies at a time, one per lane.  The
// reference sums every eddy's strength x influence into the cell in eddy
// order (:965-1000); an eddy whose distance is not below its radius adds
// +0.0 or -0.0, which leaves the sum (never -0.0: it starts at +0.0 and a
// rounded sum of nonzero terms is never -0.0) unchanged, so only the eddies
// within their radius are added, in eddy order (lowest lane of each chunk
// first): bit for bit the reference's sums with one distance test per lane.
// Blocks of 256 threads: four cells per block.
__global__ void __launch_bounds__(256) k_ref_sem_plane(GridDims g, const int* attr, const double* vort, int nvort,
                                                       double* inlet)
{
    const int lane = (int)(threadIdx.x & 63);
    const long long P = (long long)g.ny * g.nz;
    const long long pl = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pl >= P) return;
    const int j = (int)(pl / g.nz), k = (int)(pl - (long long)j * g.nz);
    double su = 0.0, sv = 0.0, sw = 0.0;
    if (attr[pl] == 1) {
        const double yp = j * g.dy, zp = k * g.dz;
        for (int e0 = 0; e0 < nvort; e0 += 64) {
            const int e = e0 + lane;
            double cu = 0.0, cv = 0.0, cw = 0.0;
            bool hit = false;
            if (e < nvort) {
                const double* V = vort + 7 * e;
                double ddx = 0.0 - V[0], ddy = yp - V[1], ddz = zp - V[2];
                double dist = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                if (dist < V[3]) {
                    double inf = 1.0 - (dist / V[3]);
                    cu = V[4] * inf; cv = V[5] * inf; cw = V[6] * inf;
                    hit = true;
                }
            }
            unsigned long long m = __ballot(hit);
            while (m) {
                const int l = __ffsll((long long)m) - 1;
                m &= m - 1;
                su += __shfl(cu, l);
                sv += __shfl(cv, l);
                sw += __shfl(cw, l);
            }
        }
    }
    if (lane == 0) { inlet[pl] = su; inlet[P + pl] = sv; inlet[2 * P + pl] = sw; }
}

// The reference's serial sums over the inlet cells (:1005-1062), bit for bit:
// each component's sum runs in plane order over the inlet cells only (idx:
// their plane indices in plane order, n of them; the other cells add nothing
// and are skipped), the block staging the values in LDS chunks and threads
// 0..2 each running one component's sum; the mean subtraction and the rms
// scaling, elementwise, are spread over the block.  One block of
// SEM_NORM_THREADS threads.
#define SEM_NORM_THREADS 256
#define SEM_NORM_CHUNK 2048
__global__ void __launch_bounds__(SEM_NORM_THREADS) k_ref_sem_normalise(GridDims g, const int* idx, int n,
                                                                       double* inlet, double target)
{
    __shared__ double sv[3][SEM_NORM_CHUNK];
    __shared__ double tot[3];
    const long long P = (long long)g.ny * g.nz;
    double* const C[3] = {inlet, inlet + P, inlet + 2 * P};
    const int t = (int)threadIdx.x;
    if (n <= 0) return;
    for (int pass = 0; pass < 2; ++pass) {
        // pass 0: sums of u', v', w'; pass 1: sums of squares of the centred values
        double acc = 0.0;
        for (int c0 = 0; c0 < n; c0 += SEM_NORM_CHUNK) {
            const int m = n - c0 < SEM_NORM_CHUNK ? n - c0 : SEM_NORM_CHUNK;
            for (int e = t; e < m; e += SEM_NORM_THREADS) {
                const int pl = idx[c0 + e];
                sv[0][e] = C[0][pl];
                sv[1][e] = C[1][pl];
                sv[2][e] = C[2][pl];
            }
            __syncthreads();
            if (t < 3) {
                const double* x = sv[t];
                if (pass == 0) {
#pragma unroll 8
                    for (int e = 0; e < m; ++e) acc += x[e];
                } else {
#pragma unroll 8
                    for (int e = 0; e < m; ++e) acc += x[e] * x[e];
                }
            }
            __syncthreads();
        }
        if (t < 3) tot[t] = acc;
        __syncthreads();
        double f[3];
        if (pass == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) f[c] = tot[c] / n;                      // means
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double r = sqrt(tot[c] / n);
                f[c] = (r > LES_SMALL) ? target / r : 0.0;                   // rms scale factors
            }
        }
        for (int e = t; e < n; e += SEM_NORM_THREADS) {
            const int pl = idx[e];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (pass == 0) C[c][pl] -= f[c];
                else C[c][pl] *= f[c];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// probes (:3403-3471) - one thread per x-station; rows land in a device table
// ---------------------------------------------------------------------------
__global__ void k_ref_probe(GridDims g, const ProbeArgs pa, const double* u, const double* v, const double* w,
                            const double* G, const double* nu_t, const double* S_mag, double* rows, int* flags)
{
    int n = threadIdx.x;
    if (n >= pa.n_station) return;
    int ip = pa.i_probe[n];
    flags[n] = 0;
    if (ip <= 0 || ip >= g.nx - 1) return;
    int kp = g.nz / 2;
    double yf = -1.0;
    for (int j = 0; j < g.ny - 1; ++j) {
        double g1 = G[lidx(g, ip, j, kp)], g2 = G[lidx(g, ip, j + 1, kp)];
        if ((g1 - 0.5) * (g2 - 0.5) <= 0.0) {
            double y1 = j * g.dy, y2 = (j + 1) * g.dy;
            if (fabs(g2 - g1) < 1e-9) yf = y1;
            else yf = y1 + (0.5 - g1) / (g2 - g1) * (y2 - y1);
            break;
        }
    }
    if (!(yf > 0.0)) return;
    flags[n] = 1;
    for (int j = 0; j < g.ny; ++j) {
        long long q = lidx(g, ip, j, kp);
        double* r = rows + ((long long)n * g.ny + j) * 8;
        r[0] = yf; r[1] = u[q]; r[2] = v[q]; r[3] = w[q]; r[4] = G[q]; r[5] = nu_t[q]; r[6] = S_mag[q];
        r[7] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// statistics accumulators (:3167-3192), FP64
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_stats_accum(long long n, const double* u, const double* v, const double* w,
                                                     const double* G, StatPtrs st)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double a = u[q], b = v[q], c = w[q];
    st.acc[0][q] += a; st.acc[1][q] += b; st.acc[2][q] += c; st.acc[3][q] += G[q];
    st.acc[4][q] += a * a; st.acc[5][q] += b * b; st.acc[6][q] += c * c;
    st.acc[7][q] += a * b; st.acc[8][q] += a * c; st.acc[9][q] += b * c;
}

// ---------------------------------------------------------------------------
// monitor reductions (:3880-3897) + non-finite check
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_monitor(long long n, double dV, const double* rho, const double* T,
                                                 const double* G, const double* Y0, const double* u, double* partial)
{
    __shared__ double sh[6][256];
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, mnY = 1e9;
    double bad = 0.0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        tm += rho[q] * dV;
        mnT = fmin(mnT, T[q]); mxT = fmax(mxT, T[q]);
        mnG = fmin(mnG, G[q]); mxG = fmax(mxG, G[q]);
        mnY = fmin(mnY, Y0[q]);
        if (!isfinite(u[q]) || !isfinite(T[q]) || !isfinite(G[q]) || !isfinite(rho[q])) bad = 1.0;
    }
    sh[0][threadIdx.x] = tm; sh[1][threadIdx.x] = mnT; sh[2][threadIdx.x] = mxT;
    sh[3][threadIdx.x] = mnG; sh[4][threadIdx.x] = mxG; sh[5][threadIdx.x] = mnY;
    __shared__ double sbad[256];
    sbad[threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        int t = threadIdx.x;
        if (t < s) {
            sh[0][t] += sh[0][t + s];
            sh[1][t] = fmin(sh[1][t], sh[1][t + s]);
            sh[2][t] = fmax(sh[2][t], sh[2][t + s]);
            sh[3][t] = fmin(sh[3][t], sh[3][t + s]);
            sh[4][t] = fmax(sh[4][t], sh[4][t + s]);
            sh[5][t] = fmin(sh[5][t], sh[5][t + s]);
            sbad[t] = fmax(sbad[t], sbad[t + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) partial[7 * blockIdx.x + c] = sh[c][0];
    if (threadIdx.x == 0) partial[7 * blockIdx.x + 6] = sbad[0];
}
