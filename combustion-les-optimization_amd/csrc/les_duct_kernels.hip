// les_duct_kernels.hip - gfx950 kernels of the duct-mode scheme: periodic y/z
// duct, SSP-RK3, WENO5 or MUSCL-minmod upwind reconstruction, cp/mu tables,
// Gaussian SEM eddies from a counter-based RNG, level-set G with Sussman
// reinitialisation, FP64 statistics.  Templated on the storage/arithmetic type
// (double: the FP64 path checked against the oracle; float: the FP32 form).
// Reductions (domain-mean density, SEM plane moments, statistics, DT maxima)
// are FP64 in both forms.
//
// Per RK stage three kernels run on one stream:
//   k_duct_filter  state -> filtered primitives u, v, w, T, phi (+ block sums of rho)
//   k_duct_sgs     filtered velocity -> nu_t, S_T (+ exact max nu / alpha for DT)
//   k_duct_rhs     RHS of the 5 state fields fused with the SSP-RK3 combine
// A step adds: DT (1 thread), eddy update (1 block), inflow plane, its
// normalisation (1 block), inflow/outflow planes, optional reinitialisation
// and statistics.  No step needs the host: DT and every reduction stay on the
// device, so a whole step is captured once into a hipGraph and replayed.
#include "les_engine.hpp"

// The duct scheme is this engine's own specification: its kernels may fuse
// multiply-adds (the reference-scheme kernels above keep strict IEEE order).
#pragma clang fp contract(fast)

#define DNF 5

// ---------------------------------------------------------------------------
// device-side parameters of the duct scheme
// ---------------------------------------------------------------------------
struct DuctConsts {
    double rho_u_Tu;        // rho_unburnt * T_u  (rho = rho_u_Tu / T)
    double Tu, Tb, U, intensity;
    double wc;              // filter centre weight
    double cs_u, cs_b, cs_f;
    double s_l, nu_unburnt, d3, tau;
    double pr_t, sc_t, prandtl;
    double accel_strength, accel_width, accel_center, weight_x, max_force;
    double relaxation;
    double courant, courant_diff;
    double table_t0, table_dt;
    int table_n;
    int scheme;             // LES_ADV_WENO5 or LES_ADV_MUSCL_MINMOD
    int nvort;
    double L;               // eddy length scale
    double ly, lz;
    unsigned long long seed;
};

// Device buffers of one duct run.  Fields are stored as `real`.
struct DuctBufs {
    void* A[DNF];      // q^n, overwritten by q^{n+1} in the last stage
    void* B[DNF];      // RK stage state
    void* F[DNF];      // filtered u, v, w, T, phi
    void* nut;         // nu_t
    void* st;          // S_T
    double* cp_tab;
    double* mu_tab;
    double* eddy;      // nvort x 7
    double* inlet;     // 3 planes
    double* rho_part;  // block partial sums of rho(T_f)
    double* scal;      // [0] dt, [1] rho_avg, [2] maxvel, [3] last dt, [4] proj residual
    unsigned long long* maxbits;   // unused by the duct step (kept for upload paths)
    double* vpart;     // per-block max |u| of the new state (RHS stage 3, then inflow plane)
    double* dpart;     // per-block max nu, alpha: [3 stages][blocks][2]
    double* spart;     // statistics plane partial sums: [nx][chunks][10]
    double* stat[10];  // per-cell FP64 accumulators
    double* prof;      // nx x 10 FP64 plane accumulators
    double* points;    // point time series ring: cap x npts x 4
    int* point_idx;    // flat cell index of each point
    double* mon_part;
};

struct DuctState {
    DuctConsts k;
    DuctBufs b;
    int precision;
    int npts = 0;
    int point_cap = 0;
    int point_count = 0;
    int filter_blocks = 0;
    int bc_blocks = 0;
    int stat_chunks = 0;
    int prof_count = 0;
    size_t esize = 8;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool graph_ok = false;
    bool graph_stats = false;
    int use_graph = 1;
    double* stage64 = nullptr;   // FP64 staging for FP32 state transfers
};

// ---------------------------------------------------------------------------
// helpers shared by host-side checks and kernels
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ unsigned long long dmix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ double duniform(unsigned long long seed, unsigned long long a,
                                                    unsigned long long b, unsigned long long slot)
{
    unsigned long long h = dmix64(seed ^ dmix64(a * 0x9E3779B97F4A7C15ull ^ dmix64(b * 0xC2B2AE3D27D4EB4Full + slot)));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

template <typename real>
__device__ __forceinline__ real progress(real phi, real dx)
{
    return (real)0.5 * ((real)1.0 + tanh(phi / ((real)2.0 * dx)));
}

template <typename real>
__device__ __forceinline__ real table_lookup(const double* tab, const DuctConsts& k, real T)
{
    real s = (T - (real)k.table_t0) / (real)k.table_dt;
    if (s < (real)0.0) s = (real)0.0;
    if (s > (real)(k.table_n - 1)) s = (real)(k.table_n - 1);
    int i0 = (int)s;
    if (i0 > k.table_n - 2) i0 = k.table_n - 2;
    real f = s - (real)i0;
    real t0 = (real)tab[i0], t1 = (real)tab[i0 + 1];
    return t0 + f * (t1 - t0);
}

// WENO5 face value, division-lean form (one division per face).
template <typename real>
__device__ __forceinline__ real weno5_face(real v0, real v1, real v2, real v3, real v4)
{
    const real c13 = (real)(13.0 / 12.0), eps = (real)1.0e-6;
    real a0 = v0 - (real)2.0 * v1 + v2, b0 = v0 - (real)4.0 * v1 + (real)3.0 * v2;
    real a1 = v1 - (real)2.0 * v2 + v3, b1 = v1 - v3;
    real a2 = v2 - (real)2.0 * v3 + v4, b2 = (real)3.0 * v2 - (real)4.0 * v3 + v4;
    real e0 = eps + (c13 * (a0 * a0) + (real)0.25 * (b0 * b0));
    real e1 = eps + (c13 * (a1 * a1) + (real)0.25 * (b1 * b1));
    real e2 = eps + (c13 * (a2 * a2) + (real)0.25 * (b2 * b2));
    real s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
    real w0 = (real)0.1 * (s1 * s2), w1 = (real)0.6 * (s0 * s2), w2 = (real)0.3 * (s0 * s1);
    real p0 = (real)2.0 * v0 - (real)7.0 * v1 + (real)11.0 * v2;
    real p1 = -v1 + (real)5.0 * v2 + (real)2.0 * v3;
    real p2 = (real)2.0 * v2 + (real)5.0 * v3 - v4;
    return (w0 * p0 + w1 * p1 + w2 * p2) / ((real)6.0 * (w0 + w1 + w2));
}

template <typename real>
__device__ __forceinline__ real minmod(real a, real b)
{
    if (a * b <= (real)0.0) return (real)0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}

template <typename real>
__device__ __forceinline__ real face_left(int scheme, real v0, real v1, real v2, real v3, real v4)
{
    if (scheme == LES_ADV_WENO5) return weno5_face<real>(v0, v1, v2, v3, v4);
    return v2 + (real)0.5 * minmod<real>(v2 - v1, v3 - v2);
}

// upwind face values of the cell from the 7-point line v[0..6] (offset -3..3)
template <typename real>
__device__ __forceinline__ real face_plus(int scheme, const real* v, bool positive)
{
    return positive ? face_left<real>(scheme, v[1], v[2], v[3], v[4], v[5])
                    : face_left<real>(scheme, v[6], v[5], v[4], v[3], v[2]);
}

template <typename real>
__device__ __forceinline__ real face_minus(int scheme, const real* v, bool positive)
{
    return positive ? face_left<real>(scheme, v[0], v[1], v[2], v[3], v[4])
                    : face_left<real>(scheme, v[5], v[4], v[3], v[2], v[1]);
}

// ---------------------------------------------------------------------------
// K1: primitives + box filter (:1307-1470 structure), block sums of rho(T_f)
// ---------------------------------------------------------------------------
template <typename real>
__global__ void __launch_bounds__(256) k_duct_filter(GridDims g, DuctConsts k, const real* __restrict__ q0,
                                                     const real* __restrict__ q1, const real* __restrict__ q2,
                                                     const real* __restrict__ q3, const real* __restrict__ q4,
                                                     real* __restrict__ uf, real* __restrict__ vf,
                                                     real* __restrict__ wf, real* __restrict__ Tf,
                                                     real* __restrict__ pf, double* __restrict__ rho_part)
{
    __shared__ double red[256];
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double rsum = 0.0;
    if (kk < g.nz && j < g.ny) {
        const long long q = lidx(g, i, j, kk);
        const real rut = (real)k.rho_u_Tu, irut = (real)(1.0 / k.rho_u_Tu);
        real fu, fv, fw, fT, fp;
        if (i == 0 || i == g.nx - 1) {
            real T = q4[q], ri = T * irut;          // 1/rho
            fu = q0[q] * ri; fv = q1[q] * ri; fw = q2[q] * ri; fT = T; fp = q3[q];
        } else {
            // neighbour order (-x, -y, -z, centre, +z, +y, +x) as :1329-1331
            long long r[7];
            r[0] = q - g.sx;
            r[1] = lidx(g, i, wrapi(j - 1, g.ny), kk);
            r[2] = lidx(g, i, j, wrapi(kk - 1, g.nz));
            r[3] = q;
            r[4] = lidx(g, i, j, wrapi(kk + 1, g.nz));
            r[5] = lidx(g, i, wrapi(j + 1, g.ny), kk);
            r[6] = q + g.sx;
            real pu[7], pv[7], pw[7], pT[7], pp[7];
#pragma unroll
            for (int t = 0; t < 7; ++t) {
                real T = q4[r[t]], ri = T * irut;
                pu[t] = q0[r[t]] * ri; pv[t] = q1[r[t]] * ri; pw[t] = q2[r[t]] * ri;
                pT[t] = T; pp[t] = q3[r[t]];
            }
            const real wc = (real)k.wc, iden = (real)(1.0 / (6.0 + k.wc));
            fu = (pu[0] + pu[1] + pu[2] + wc * pu[3] + pu[4] + pu[5] + pu[6]) * iden;
            fv = (pv[0] + pv[1] + pv[2] + wc * pv[3] + pv[4] + pv[5] + pv[6]) * iden;
            fw = (pw[0] + pw[1] + pw[2] + wc * pw[3] + pw[4] + pw[5] + pw[6]) * iden;
            fT = (pT[0] + pT[1] + pT[2] + wc * pT[3] + pT[4] + pT[5] + pT[6]) * iden;
            fp = (pp[0] + pp[1] + pp[2] + wc * pp[3] + pp[4] + pp[5] + pp[6]) * iden;
        }
        uf[q] = fu; vf[q] = fv; wf[q] = fw; Tf[q] = fT; pf[q] = fp;
        rsum = (double)(rut / fT);
        (void)irut;
    }
    red[tid] = rsum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) {
        long long b = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        rho_part[b] = red[0];
    }
}

// deterministic finalisation of the block sums: rho_avg = sum / n
__global__ void __launch_bounds__(1024) k_duct_rho_final(long long n, int nparts, const double* __restrict__ part,
                                                         double* __restrict__ scal)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 1024) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) scal[1] = red[0] / (double)n;
}

// ---------------------------------------------------------------------------
// K2: properties, nu_t (:1569-1633), S_T (:1913-1965); exact DT maxima
// ---------------------------------------------------------------------------
template <typename real>
__global__ void __launch_bounds__(256) k_duct_sgs(GridDims g, DuctConsts k, const double* __restrict__ mu_tab,
                                                  const double* __restrict__ cp_tab, const real* __restrict__ uf,
                                                  const real* __restrict__ vf, const real* __restrict__ wf,
                                                  const real* __restrict__ Tf, const real* __restrict__ pf,
                                                  real* __restrict__ nut, real* __restrict__ st,
                                                  double* __restrict__ dpart)
{
    __shared__ double red[2][4];
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    double mn = 0.0, ma = 0.0;
    if (kk < g.nz && j < g.ny) {
        const long long q = lidx(g, i, j, kk);
        const real SL = (real)k.s_l;
        if (i == 0 || i == g.nx - 1) {
            nut[q] = (real)0.0; st[q] = SL;
        } else {
            const real DX = (real)g.dx, DY = (real)g.dy, DZ = (real)g.dz;
            real T = Tf[q], rf = (real)k.rho_u_Tu / T;
            real mu = table_lookup<real>(mu_tab, k, T);
            real cp = table_lookup<real>(cp_tab, k, T);
            real nu = mu / rf;
            real lam = mu * cp / (real)k.prandtl;
            real al = lam / (rf * cp);
            mn = (double)nu; ma = (double)al;
            long long xp = q + g.sx, xm = q - g.sx;
            long long yp = lidx(g, i, wrapi(j + 1, g.ny), kk), ym = lidx(g, i, wrapi(j - 1, g.ny), kk);
            long long zp = lidx(g, i, j, wrapi(kk + 1, g.nz)), zm = lidx(g, i, j, wrapi(kk - 1, g.nz));
            const real hx = (real)(0.5 / g.dx), hy = (real)(0.5 / g.dy), hz = (real)(0.5 / g.dz);
            real du_dx = (uf[xp] - uf[xm]) * hx;
            real dv_dy = (vf[yp] - vf[ym]) * hy;
            real dw_dz = (wf[zp] - wf[zm]) * hz;
            real du_dy = (uf[yp] - uf[ym]) * hy;
            real du_dz = (uf[zp] - uf[zm]) * hz;
            real dv_dx = (vf[xp] - vf[xm]) * hx;
            real dv_dz = (vf[zp] - vf[zm]) * hz;
            real dw_dx = (wf[xp] - wf[xm]) * hx;
            real dw_dy = (wf[yp] - wf[ym]) * hy;
            real S12 = (real)0.5 * (du_dy + dv_dx), S13 = (real)0.5 * (du_dz + dw_dx), S23 = (real)0.5 * (dv_dz + dw_dy);
            real ss = (real)2.0 * (S12 * S12 + S13 * S13 + S23 * S23) + du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz;
            real smag = sqrt((real)2.0 * ss);
            real div = du_dx + dv_dy + dw_dz;
            real gg = progress<real>(pf[q], DX);
            real cs = (real)k.cs_u + ((real)k.cs_b - (real)k.cs_u) * gg;
            cs += (real)k.cs_f * (real)4.0 * gg * ((real)1.0 - gg);
            nut[q] = (cs * DX) * (cs * DX) * smag;
            // SSR-SGS flame speed
            real num = ss - (div * div) * ((real)1.0 / (real)3.0);
            real X = (pow(DX, (real)4.0) / (nu * nu)) * num;
            real d_eta;
            if (X < (real)LES_SMALL) d_eta = (real)1.0;
            else {
                real px = pow(X, (real)0.75);
                real alg = (real)0.2187 * px;
                d_eta = (real)0.2340 * px * (alg > (real)1.0 ? log(alg) : (real)0.0);
            }
            real eta = DX / (d_eta + (real)LES_SMALL);
            real dF = (real)k.nu_unburnt / (SL + (real)LES_SMALL);
            real asf = (real)8.0 * exp((real)10.0 * dF / ((real)8.0 * eta + (real)LES_SMALL));
            real wr = pow(asf * eta / DX, (real)2.0 - (real)k.d3);
            real te = ((real)k.tau * SL > (real)LES_SMALL) ? DX * fmax((real)0.0, div) / ((real)k.tau * SL) : (real)0.0;
            st[q] = fmax(SL, SL * fmax((real)0.0, wr + te));
        }
    }
    // block max (wavefront shuffles, then LDS across the 4 wavefronts); one
    // slot per block, reduced by k_duct_dt: no atomics, order-independent
    for (int o = 32; o > 0; o >>= 1) {
        mn = fmax(mn, __shfl_xor(mn, o));
        ma = fmax(ma, __shfl_xor(ma, o));
    }
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if ((tid & 63) == 0) { red[0][tid >> 6] = mn; red[1][tid >> 6] = ma; }
    __syncthreads();
    if (tid == 0) {
        long long b = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        dpart[2 * b] = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
        dpart[2 * b + 1] = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
    }
}

// ---------------------------------------------------------------------------
// K3: RHS (:2552-2955 structure) fused with the SSP-RK3 combine
// ---------------------------------------------------------------------------
template <typename real>
struct RhsArgs {
    const real *uf, *vf, *wf, *Tf, *pf, *nut, *st;
    const real* qin[DNF];
    const real* q0[DNF];
    real* qout[DNF];
    const double* mu_tab;
    const double* cp_tab;
    const double* scal;
    double* vpart;
    int stage;
    double ca, cb;            // q_out = ca*q0 + cb*(q_in + dt*L) for stages 1, 2
    int track_vel;            // last stage: max |u| of the new state
    real* Lout[DNF];          // optional: the RHS itself (testing hook)
};

// Flat index of the point at offset o (-3..3) from cell (i, j, k) along axis d;
// y and z wrap (periodic duct), x never leaves the domain for the offsets used.
__device__ __forceinline__ int line_idx(const GridDims& g, int base_i, int j, int kk, int d, int o)
{
    if (d == 0) return base_i + o * (int)g.sx + j * g.nz + kk;
    if (d == 1) return base_i + wrapi(j + o, g.ny) * g.nz + kk;
    return base_i + j * g.nz + wrapi(kk + o, g.nz);
}

// K3a: scalar RHS (phi: G equation :2706-2818; T :2912-2955) + SSP-RK3 combine
template <typename real>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_duct_rhs_scalar(GridDims g, DuctConsts k, RhsArgs<real> a)
{
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (kk >= g.nz || j >= g.ny) return;
    const int bi = i * (int)g.sx;
    const int q = bi + j * g.nz + kk;
    if (i == 0 || i == g.nx - 1) {
        a.qout[3][q] = a.qin[3][q];
        a.qout[4][q] = a.qin[4][q];
        return;
    }
    const real DX = (real)g.dx, DY = (real)g.dy, DZ = (real)g.dz;
    const real rut = (real)k.rho_u_Tu;
    const bool xband = (i < 3 || i > g.nx - 4);
    const int scheme = k.scheme;
    const real* vel[3] = {a.uf, a.vf, a.wf};
    const real T0 = a.Tf[q];
    const real nu0 = a.nut[q];
    const int xp = q + (int)g.sx, xm = q - (int)g.sx;
    const int yp = line_idx(g, bi, j, kk, 1, 1), ym = line_idx(g, bi, j, kk, 1, -1);
    const int zp = line_idx(g, bi, j, kk, 2, 1), zm = line_idx(g, bi, j, kk, 2, -1);
    const real* S2[2] = {a.pf, a.Tf};
    real cv[2], sg[2], lap[2], grad[3];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const real* F = S2[s];
        const real isc2 = (real)(0.5 / ((s == 0) ? k.sc_t : k.pr_t));
        real acc = (real)0, sacc = (real)0, lacc = (real)0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double dd64 = d == 0 ? g.dx : (d == 1 ? g.dy : g.dz);
            const real id = (real)(1.0 / dd64), id2 = (real)(1.0 / (dd64 * dd64)), hid = (real)(0.5 / dd64);
            const int qp = d == 0 ? xp : (d == 1 ? yp : zp);
            const int qm = d == 0 ? xm : (d == 1 ? ym : zm);
            real ua = vel[d][q];
            real vm1 = F[qm], v0 = F[q], vp1 = F[qp];
            real t;
            if (d == 0 && xband) {
                t = (ua >= (real)0.0) ? ua * (v0 - vm1) * id : ua * (vp1 - v0) * id;
            } else {
                real v[7];
#pragma unroll
                for (int o = 0; o < 7; ++o) v[o] = F[line_idx(g, bi, j, kk, d, o - 3)];
                bool pos = ua >= (real)0.0;
                real fp = face_plus<real>(scheme, v, pos);
                real fm = face_minus<real>(scheme, v, pos);
                t = ua * (fp - fm) * id;
            }
            acc = (d == 0) ? t : acc + t;
            real ntp = a.nut[qp], ntm = a.nut[qm];
            real gp = (ntp + nu0) * isc2, gm = (nu0 + ntm) * isc2;
            real t2 = (gp * (vp1 - v0) - gm * (v0 - vm1)) * id2;
            sacc = (d == 0) ? t2 : sacc + t2;
            real t3 = (vp1 - (real)2.0 * v0 + vm1) * id2;
            lacc = (d == 0) ? t3 : lacc + t3;
            if (s == 0) grad[d] = (vp1 - vm1) * hid;
        }
        cv[s] = acc; sg[s] = sacc; lap[s] = lacc;
    }
    const real g0 = progress<real>(a.pf[q], DX);
    const real wdt = (real)k.accel_width, ctr = (real)k.accel_center;
    real gmag = sqrt(grad[0] * grad[0] + grad[1] * grad[1] + grad[2] * grad[2]);
    real art = (real)0;
    if (g0 > (real)0.1 && g0 < (real)0.99) {
        real df = g0 - ctr;
        real win = exp(-(df * df) / ((real)2.0 * wdt * wdt));
        art = (real)5.0 * (nu0 + (real)1.0e-5) * win * lap[0];
    }
    real L3 = -cv[0] + a.st[q] * gmag + sg[0] + art;
    const real rf0 = rut / T0;
    real mu = table_lookup<real>(a.mu_tab, k, T0);
    real cp = table_lookup<real>(a.cp_tab, k, T0);
    real lam = mu * cp / (real)k.prandtl;
    real al = lam / (rf0 * cp);
    const real Tu = (real)k.Tu, Tb = (real)k.Tb;
    real Ttg = Tu + g0 * (Tb - Tu);
    real L4 = -cv[1] + al * lap[1] + sg[1] + (real)k.relaxation * (Ttg - T0);
    if (a.Lout[0]) { a.Lout[3][q] = L3; a.Lout[4][q] = L4; }
    const real DT = (real)a.scal[0];
    real y3 = a.qin[3][q] + DT * L3, y4 = a.qin[4][q] + DT * L4;
    real v3, v4;
    if (a.stage == 0) { v3 = y3; v4 = y4; }
    else { v3 = (real)a.ca * a.q0[3][q] + (real)a.cb * y3; v4 = (real)a.ca * a.q0[4][q] + (real)a.cb * y4; }
    v4 = fmax((real)200.0, fmin((real)3000.0, v4));
    a.qout[3][q] = v3;
    a.qout[4][q] = v4;
}

// K3b: momentum RHS (:2552-2691 structure) + SSP-RK3 combine; in the last stage
// also the per-block max |u| of the new state (its T is written by K3a first)
template <typename real>
__global__ void __launch_bounds__(256) k_duct_rhs_momentum(GridDims g, DuctConsts k, RhsArgs<real> a)
{
    __shared__ double vred[4];
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    double vmax = 0.0;
    if (kk < g.nz && j < g.ny) {
        const int bi = i * (int)g.sx;
        const int q = bi + j * g.nz + kk;
        if (i == 0 || i == g.nx - 1) {
#pragma unroll
            for (int f = 0; f < 3; ++f) a.qout[f][q] = a.qin[f][q];
        } else {
            const real DX = (real)g.dx, DY = (real)g.dy, DZ = (real)g.dz;
            const real rut = (real)k.rho_u_Tu;
            const bool xband = (i < 3 || i > g.nx - 4);
            const int scheme = k.scheme;
            const real* vel[3] = {a.uf, a.vf, a.wf};
            const real T0 = a.Tf[q];
            const real rf0 = rut / T0;
            real conv[3] = {(real)0, (real)0, (real)0};
#pragma unroll 1
            for (int d = 0; d < 3; ++d) {
                const real id = (real)(1.0 / (d == 0 ? g.dx : (d == 1 ? g.dy : g.dz)));
                const int qp = line_idx(g, bi, j, kk, d, 1), qm = line_idx(g, bi, j, kk, d, -1);
                const real* U = vel[d];
                real up = (real)0.5 * (U[qp] + U[q]);
                real um = (real)0.5 * (U[q] + U[qm]);
                real rp = (real)0.5 * (rut / a.Tf[qp] + rf0);
                real rm = (real)0.5 * (rf0 + rut / a.Tf[qm]);
                real mp = rp * up, mm = rm * um;
                const bool band = (d == 0 && xband);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const real* V = vel[m];
                    real fp, fm;
                    if (band) {
                        real v2 = V[qm], v3 = V[q], v4 = V[qp];
                        fp = (up >= (real)0.0) ? v3 : v4;
                        fm = (um >= (real)0.0) ? v2 : v3;
                    } else {
                        real v[7];
#pragma unroll
                        for (int o = 0; o < 7; ++o) v[o] = V[line_idx(g, bi, j, kk, d, o - 3)];
                        fp = face_plus<real>(scheme, v, up >= (real)0.0);
                        fm = face_minus<real>(scheme, v, um >= (real)0.0);
                    }
                    real term = (mp * fp - mm * fm) * id;
                    conv[m] = (d == 0) ? term : conv[m] + term;
                }
            }
            real mu = table_lookup<real>(a.mu_tab, k, T0);
            real nu0 = a.nut[q];
            real mu_eff = mu + rf0 * nu0;
            const int xp = q + (int)g.sx, xm = q - (int)g.sx;
            const int yp = line_idx(g, bi, j, kk, 1, 1), ym = line_idx(g, bi, j, kk, 1, -1);
            const int zp = line_idx(g, bi, j, kk, 2, 1), zm = line_idx(g, bi, j, kk, 2, -1);
            const real ix2 = (real)(1.0 / (g.dx * g.dx)), iy2 = (real)(1.0 / (g.dy * g.dy)), iz2 = (real)(1.0 / (g.dz * g.dz));
            real diff[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const real* U = vel[m];
                real uc = U[q];
                real lx = (U[xp] - (real)2.0 * uc + U[xm]) * ix2;
                real ly = (U[yp] - (real)2.0 * uc + U[ym]) * iy2;
                real lz = (U[zp] - (real)2.0 * uc + U[zm]) * iz2;
                diff[m] = mu_eff * (lx + ly + lz);
            }
            const real rho_avg = (real)a.scal[1];
            real grav = (rf0 - rho_avg) * (real)LES_G_GRAV;
            const real* P = a.pf;
            real g0 = progress<real>(P[q], DX);
            real fx = (real)0, fy = (real)0, fz = (real)0;
            const real wdt = (real)k.accel_width, ctr = (real)k.accel_center, Fm = (real)k.max_force;
            if (g0 > (real)0.01 && g0 < (real)0.99) {
                real gx = (progress<real>(P[xp], DX) - progress<real>(P[xm], DX)) * (real)(0.5 / g.dx);
                real gy = (progress<real>(P[yp], DX) - progress<real>(P[ym], DX)) * (real)(0.5 / g.dy);
                real gz = (progress<real>(P[zp], DX) - progress<real>(P[zm], DX)) * (real)(0.5 / g.dz);
                real df = g0 - ctr;
                real win = exp(-(df * df) / ((real)2.0 * wdt * wdt));
                real common = (real)k.accel_strength * rf0 * win;
                fx = common * gx * (real)k.weight_x;
                fy = common * gy;
                fz = common * gz;
            }
            fx = fmax(-Fm, fmin(Fm, fx));
            fy = fmax(-Fm, fmin(Fm, fy));
            fz = fmax(-Fm, fmin(Fm, fz));
            real Lq[3];
            Lq[0] = -conv[0] + diff[0] + grav + fx;
            Lq[1] = -conv[1] + diff[1] + fy;
            Lq[2] = -conv[2] + diff[2] + fz;
            if (a.Lout[0]) {
#pragma unroll
                for (int f = 0; f < 3; ++f) a.Lout[f][q] = Lq[f];
            }
            const real DT = (real)a.scal[0];
            real out[3];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                real y = a.qin[f][q] + DT * Lq[f];
                out[f] = (a.stage == 0) ? y : (real)a.ca * a.q0[f][q] + (real)a.cb * y;
                a.qout[f][q] = out[f];
            }
            if (a.track_vel) {
                real rho = rut / a.qout[4][q];
                real uu = out[0] / rho, vv = out[1] / rho, ww = out[2] / rho;
                vmax = (double)sqrt(uu * uu + vv * vv + ww * ww);
            }
        }
    }
    if (a.track_vel) {
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o));
        const int tid = threadIdx.y * blockDim.x + threadIdx.x;
        if ((tid & 63) == 0) vred[tid >> 6] = vmax;
        __syncthreads();
        if (tid == 0) {
            long long b = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
            a.vpart[b] = fmax(fmax(vred[0], vred[1]), fmax(vred[2], vred[3]));
        }
    }
}

// ---------------------------------------------------------------------------
// DT (:4399-4411 structure): one block reduces the per-block maxima of the
// previous step (max |u| of the new state, max nu / alpha over its 3 stages)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_duct_dt(double dx, DuctConsts k, double* scal, const double* vpart,
                                                  int nv, const double* dpart, int nd)
{
    __shared__ double r[3][1024];
    double mv = 0.0, mn = 0.0, ma = 0.0;
    for (int b = threadIdx.x; b < nv; b += 1024) mv = fmax(mv, vpart[b]);
    for (int b = threadIdx.x; b < nd; b += 1024) { mn = fmax(mn, dpart[2 * b]); ma = fmax(ma, dpart[2 * b + 1]); }
    r[0][threadIdx.x] = mv; r[1][threadIdx.x] = mn; r[2][threadIdx.x] = ma;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; ++a) r[a][threadIdx.x] = fmax(r[a][threadIdx.x], r[a][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    mv = r[0][0]; mn = r[1][0]; ma = r[2][0];
    double maxvel = mv > LES_SMALL ? mv : k.U;
    double dtc = k.courant * dx / (maxvel + LES_SMALL);
    double dtn = k.courant_diff * (dx * dx) / (mn + LES_SMALL);
    double dta = k.courant_diff * (dx * dx) / (ma + LES_SMALL);
    scal[0] = fmin(dtc, fmin(dtn, dta));
    scal[2] = maxvel;
}

// max |u| over planes 0..nx-2 of a state (used after upload / init)
template <typename real>
__global__ void __launch_bounds__(256) k_duct_velmax(GridDims g, double rho_u_Tu, const real* q0, const real* q1,
                                                     const real* q2, const real* q4, double* vpart)
{
    __shared__ double vred[4];
    double vm = 0.0;
    long long lim = (long long)(g.nx - 1) * g.sx;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < lim; q += (long long)gridDim.x * blockDim.x) {
        real rho = (real)rho_u_Tu / q4[q];
        real u = q0[q] / rho, v = q1[q] / rho, w = q2[q] / rho;
        vm = fmax(vm, (double)sqrt(u * u + v * v + w * w));
    }
    for (int o = 32; o > 0; o >>= 1) vm = fmax(vm, __shfl_xor(vm, o));
    if ((threadIdx.x & 63) == 0) vred[threadIdx.x >> 6] = vm;
    __syncthreads();
    if (threadIdx.x == 0) vpart[blockIdx.x] = fmax(fmax(vred[0], vred[1]), fmax(vred[2], vred[3]));
}

// ---------------------------------------------------------------------------
// SEM: eddy advection/recycling (1 block), Gaussian plane, normalisation
// ---------------------------------------------------------------------------
__device__ __forceinline__ void eddy_draw_dev(const DuctConsts& k, double* e, int idx, unsigned long long gen, bool initial)
{
    if (initial) e[0] = (2.0 * duniform(k.seed, idx, gen, 0) - 1.0) * k.L;
    e[1] = duniform(k.seed, idx, gen, 1) * k.ly;
    e[2] = duniform(k.seed, idx, gen, 2) * k.lz;
    e[3] = 2.0 * duniform(k.seed, idx, gen, 3) - 1.0;
    e[4] = 2.0 * duniform(k.seed, idx, gen, 4) - 1.0;
    e[5] = 2.0 * duniform(k.seed, idx, gen, 5) - 1.0;
    e[6] = (double)gen;
}

__global__ void k_duct_eddy_init(DuctConsts k, double* eddy)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= k.nvort) return;
    eddy_draw_dev(k, eddy + 7 * (long long)e, e, 0ull, true);
}

__global__ void k_duct_eddy_update(DuctConsts k, const double* scal, double* eddy)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= k.nvort) return;
    double* E = eddy + 7 * (long long)e;
    E[0] += k.U * scal[0];
    if (E[0] > k.L) {
        E[0] -= 2.0 * k.L;
        eddy_draw_dev(k, E, e, (unsigned long long)E[6] + 1ull, false);
    }
}

// One thread per inflow cell; the eddy table is staged through LDS in chunks.
__global__ void __launch_bounds__(256) k_duct_sem_plane(GridDims g, DuctConsts k, const double* __restrict__ eddy,
                                                        double* __restrict__ inlet)
{
    __shared__ double se[256 * 7];
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const long long P = (long long)g.ny * g.nz;
    const double y = j * g.dy, z = kk * g.dz, inv_l2 = 1.0 / (k.L * k.L);
    double su = 0.0, sv = 0.0, sw = 0.0;
    for (int base = 0; base < k.nvort; base += 256) {
        int cnt = min(256, k.nvort - base);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt * 7; t += blockDim.x) se[t] = eddy[7 * (long long)base + t];
        __syncthreads();
        if (kk < g.nz) {
            for (int e = 0; e < cnt; ++e) {
                const double* E = se + 7 * e;
                double ddx = 0.0 - E[0];
                double ddy = y - E[1];
                ddy -= k.ly * rint(ddy / k.ly);
                double ddz = z - E[2];
                ddz -= k.lz * rint(ddz / k.lz);
                double f = exp(-M_PI * (ddx * ddx + ddy * ddy + ddz * ddz) * inv_l2);
                su += E[3] * f; sv += E[4] * f; sw += E[5] * f;
            }
        }
    }
    if (kk < g.nz) {
        long long p = (long long)j * g.nz + kk;
        inlet[p] = su; inlet[P + p] = sv; inlet[2 * P + p] = sw;
    }
}

// mean removal and rms rescale of the three planes, one 1024-thread block each
__global__ void __launch_bounds__(1024) k_duct_sem_norm(long long P, double target, double* __restrict__ inlet)
{
    __shared__ double red[1024];
    double* f = inlet + blockIdx.x * P;
    double s = 0.0;
    for (long long p = threadIdx.x; p < P; p += 1024) s += f[p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    double mean = red[0] / (double)P;
    __syncthreads();
    double qq = 0.0;
    for (long long p = threadIdx.x; p < P; p += 1024) {
        double v = f[p] - mean;
        f[p] = v;
        qq += v * v;
    }
    red[threadIdx.x] = qq;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    double rms = sqrt(red[0] / (double)P);
    double sc = (rms > LES_SMALL) ? target / rms : 0.0;
    for (long long p = threadIdx.x; p < P; p += 1024) f[p] *= sc;
}

// inflow plane from the SEM planes, zero-gradient outflow plane
template <typename real>
__global__ void __launch_bounds__(256) k_duct_bc(GridDims g, DuctConsts k, const double* __restrict__ inlet,
                                                 real* q0, real* q1, real* q2, real* q3, real* q4,
                                                 double* vpart)
{
    __shared__ double vred[4];
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    double vm = 0.0;
    if (kk < g.nz) {
        const long long P = (long long)g.ny * g.nz;
        const long long p = (long long)j * g.nz + kk;
        const double rho_u = k.rho_u_Tu / k.Tu;
        const double u = k.U + inlet[p], v = inlet[P + p], w = inlet[2 * P + p];
        q0[p] = (real)(rho_u * u);
        q1[p] = (real)(rho_u * v);
        q2[p] = (real)(rho_u * w);
        q3[p] = q3[p + g.sx] - (real)g.dx;
        q4[p] = (real)k.Tu;
        real rr = (real)k.rho_u_Tu / q4[p];
        real uu = q0[p] / rr, vv = q1[p] / rr, ww = q2[p] / rr;
        vm = (double)sqrt(uu * uu + vv * vv + ww * ww);
        const long long qn = (long long)(g.nx - 1) * g.sx + p, qm = qn - g.sx;
        q0[qn] = q0[qm]; q1[qn] = q1[qm]; q2[qn] = q2[qm]; q3[qn] = q3[qm]; q4[qn] = q4[qm];
    }
    for (int o = 32; o > 0; o >>= 1) vm = fmax(vm, __shfl_xor(vm, o));
    if ((threadIdx.x & 63) == 0) vred[threadIdx.x >> 6] = vm;
    __syncthreads();
    if (threadIdx.x == 0) vpart[(long long)blockIdx.y * gridDim.x + blockIdx.x] = fmax(fmax(vred[0], vred[1]), fmax(vred[2], vred[3]));
}

// ---------------------------------------------------------------------------
// level-set reinitialisation (Sussman-Smereka-Osher, first-order Godunov)
// ---------------------------------------------------------------------------
template <typename real>
__global__ void __launch_bounds__(256) k_duct_reinit(GridDims g, const real* __restrict__ phi0,
                                                     const real* __restrict__ phi, real* __restrict__ out)
{
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (kk >= g.nz || j >= g.ny) return;
    const long long q = lidx(g, i, j, kk);
    if (i == 0 || i == g.nx - 1) { out[q] = phi[q]; return; }
    const real dx = (real)g.dx, dy = (real)g.dy, dz = (real)g.dz, dtau = (real)0.3 * dx;
    real p0 = phi0[q];
    real s = p0 / sqrt(p0 * p0 + dx * dx);
    real c = phi[q];
    real ax = (c - phi[q - g.sx]) / dx, bx = (phi[q + g.sx] - c) / dx;
    real ay = (c - phi[lidx(g, i, wrapi(j - 1, g.ny), kk)]) / dy, by = (phi[lidx(g, i, wrapi(j + 1, g.ny), kk)] - c) / dy;
    real az = (c - phi[lidx(g, i, j, wrapi(kk - 1, g.nz))]) / dz, bz = (phi[lidx(g, i, j, wrapi(kk + 1, g.nz))] - c) / dz;
    real gx, gy, gz;
    if (p0 > (real)0.0) {
        real a1 = fmax(ax, (real)0), b1 = fmin(bx, (real)0), a2 = fmax(ay, (real)0), b2 = fmin(by, (real)0);
        real a3 = fmax(az, (real)0), b3 = fmin(bz, (real)0);
        gx = fmax(a1 * a1, b1 * b1); gy = fmax(a2 * a2, b2 * b2); gz = fmax(a3 * a3, b3 * b3);
    } else {
        real a1 = fmin(ax, (real)0), b1 = fmax(bx, (real)0), a2 = fmin(ay, (real)0), b2 = fmax(by, (real)0);
        real a3 = fmin(az, (real)0), b3 = fmax(bz, (real)0);
        gx = fmax(a1 * a1, b1 * b1); gy = fmax(a2 * a2, b2 * b2); gz = fmax(a3 * a3, b3 * b3);
    }
    real gm = sqrt(gx + gy + gz);
    out[q] = c - dtau * s * (gm - (real)1.0);
}

template <typename real>
__global__ void __launch_bounds__(256) k_copy(long long n, const real* __restrict__ a, real* __restrict__ b)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) b[q] = a[q];
}

// ---------------------------------------------------------------------------
// statistics: per-cell FP64 accumulators (:3167-3192) + y-z plane profiles.
// Grid (chunks, nx): each block sums one chunk of one x-plane; the chunk sums
// are added to the plane profile in a fixed order by k_duct_stats_final.
// ---------------------------------------------------------------------------
template <typename real>
__global__ void __launch_bounds__(256) k_duct_stats(GridDims g, DuctConsts k, const real* q0, const real* q1,
                                                    const real* q2, const real* q3, const real* q4, StatPtrs st,
                                                    int per_cell, double* spart)
{
    __shared__ double red[10][256];
    const int i = blockIdx.y;
    const long long P = (long long)g.ny * g.nz;
    const long long chunk = (P + gridDim.x - 1) / gridDim.x;
    const long long p0 = (long long)blockIdx.x * chunk;
    const long long p1 = p0 + chunk < P ? p0 + chunk : P;
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const real dx = (real)g.dx;
    for (long long p = p0 + threadIdx.x; p < p1; p += 256) {
        long long q = (long long)i * P + p;
        real rho = (real)k.rho_u_Tu / q4[q];
        double u = (double)(q0[q] / rho), v = (double)(q1[q] / rho), w = (double)(q2[q] / rho);
        double c = (double)progress<real>(q3[q], dx);
        if (per_cell) {
            st.acc[0][q] += u; st.acc[1][q] += v; st.acc[2][q] += w; st.acc[3][q] += c;
            st.acc[4][q] += u * u; st.acc[5][q] += v * v; st.acc[6][q] += w * w;
            st.acc[7][q] += u * v; st.acc[8][q] += u * w; st.acc[9][q] += v * w;
        }
        acc[0] += u; acc[1] += v; acc[2] += w; acc[3] += c;
        acc[4] += u * u; acc[5] += v * v; acc[6] += w * w; acc[7] += u * v; acc[8] += u * w; acc[9] += v * w;
    }
    for (int a = 0; a < 10; ++a) red[a][threadIdx.x] = acc[a];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 10; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 10) spart[((long long)i * gridDim.x + blockIdx.x) * 10 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_duct_stats_final(int nx, int chunks, const double* spart, double* prof)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * 10) return;
    int i = t / 10, a = t % 10;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += spart[((long long)i * chunks + c) * 10 + a];
    prof[t] += s;
}

// fixed-point time series (u, v, w, G) at npts cells into a ring of cap steps
template <typename real>
__global__ void k_duct_points(DuctConsts k, double dx, const int* idx, int npts, const real* q0, const real* q1,
                              const real* q2, const real* q3, const real* q4, double* ring, int slot)
{
    int t = threadIdx.x;
    if (t >= npts) return;
    long long q = idx[t];
    real rho = (real)k.rho_u_Tu / q4[q];
    double* o = ring + ((long long)slot * npts + t) * 4;
    o[0] = (double)(q0[q] / rho); o[1] = (double)(q1[q] / rho); o[2] = (double)(q2[q] / rho);
    o[3] = (double)progress<real>(q3[q], (real)dx);
}

// monitor reductions of the duct state (total mass, T and G range, non-finite)
template <typename real>
__global__ void __launch_bounds__(256) k_duct_monitor(GridDims g, DuctConsts k, const real* q0, const real* q3,
                                                      const real* q4, double* part)
{
    __shared__ double sh[6][256];
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, bad = 0;
    const double dV = g.dx * g.dy * g.dz;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < g.n; q += (long long)gridDim.x * 256) {
        double T = (double)q4[q];
        tm += k.rho_u_Tu / T * dV;
        mnT = fmin(mnT, T); mxT = fmax(mxT, T);
        double c = (double)progress<real>(q3[q], (real)g.dx);
        mnG = fmin(mnG, c); mxG = fmax(mxG, c);
        if (!isfinite(T) || !isfinite((double)q0[q]) || !isfinite((double)q3[q])) bad = 1.0;
    }
    sh[0][threadIdx.x] = tm; sh[1][threadIdx.x] = mnT; sh[2][threadIdx.x] = mxT;
    sh[3][threadIdx.x] = mnG; sh[4][threadIdx.x] = mxG; sh[5][threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        int t = threadIdx.x;
        if (t < s) {
            sh[0][t] += sh[0][t + s];
            sh[1][t] = fmin(sh[1][t], sh[1][t + s]); sh[2][t] = fmax(sh[2][t], sh[2][t + s]);
            sh[3][t] = fmin(sh[3][t], sh[3][t + s]); sh[4][t] = fmax(sh[4][t], sh[4][t + s]);
            sh[5][t] = fmax(sh[5][t], sh[5][t + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) part[6 * blockIdx.x + c] = sh[c][0];
}

// ---------------------------------------------------------------------------
// host side of the duct mode
// ---------------------------------------------------------------------------
#define DHIPC(call)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            snprintf(c->err, sizeof(c->err), "%s failed: %s (%s:%d)", #call,               \
                     hipGetErrorString(e_), __FILE__, __LINE__);                           \
            return LES_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

#define DLAUNCH_CHECK() DHIPC(hipGetLastError())

static DuctConsts make_duct_consts(const les_gpu* c)
{
    DuctConsts k;
    memset(&k, 0, sizeof(k));
    const les_params& p = c->p;
    k.rho_u_Tu = p.rho_unburnt * p.inlet_temperature;
    k.Tu = p.inlet_temperature; k.Tb = p.burnt_temperature;
    k.U = p.inlet_velocity; k.intensity = p.turbulence_intensity;
    k.wc = p.filter_weight_center;
    k.cs_u = p.cs_unburned; k.cs_b = p.cs_burnt; k.cs_f = p.cs_flame_boost;
    k.s_l = p.s_l; k.nu_unburnt = p.nu_unburnt_approx; k.d3 = p.fractal_d3;
    k.tau = p.burnt_temperature / p.inlet_temperature - 1.0;
    k.pr_t = p.pr_t; k.sc_t = p.sc_t; k.prandtl = p.prandtl;
    k.accel_strength = p.flame_accel_strength; k.accel_width = p.flame_accel_width;
    k.accel_center = p.flame_accel_center; k.weight_x = p.weight_x; k.max_force = p.max_flame_force;
    k.relaxation = p.relaxation_coeff;
    k.courant = p.courant; k.courant_diff = p.courant_diffusion;
    k.table_t0 = p.table_t0; k.table_dt = p.table_dt; k.table_n = p.table_n;
    k.scheme = p.advection;
    k.nvort = p.num_vortices;
    k.L = p.turbulent_length_scale;
    k.ly = p.ly; k.lz = p.lz;
    k.seed = (unsigned long long)p.seed;
    return k;
}

void duct_params_changed(les_gpu* c)
{
    if (c->d) {
        c->d->k = make_duct_consts(c);
        c->d->graph_ok = false;   // constants are baked into the captured graph
    }
}

static void* dalloc_bytes(les_gpu* c, size_t bytes)
{
    void* p = nullptr;
    if (hipMalloc(&p, bytes + 256) != hipSuccess) return nullptr;
    c->allocs.push_back(p);
    (void)hipMemset(p, 0, bytes + 256);
    return p;
}

int duct_create(les_gpu* c)
{
    if (c->p.integrator != LES_INT_SSPRK3 ||
        (c->p.advection != LES_ADV_WENO5 && c->p.advection != LES_ADV_MUSCL_MINMOD))
        return LES_ERR_PARAM;
    DuctState* d = new (std::nothrow) DuctState();
    if (!d) return LES_ERR_HIP;
    c->d = d;
    d->precision = c->p.precision;
    d->esize = (d->precision == LES_PREC_FP32) ? 4 : 8;
    d->k = make_duct_consts(c);
    const size_t n = (size_t)c->g.n, nb = n * d->esize;
    for (int f = 0; f < DNF; ++f) {
        if (!(d->b.A[f] = dalloc_bytes(c, nb)) || !(d->b.B[f] = dalloc_bytes(c, nb)) || !(d->b.F[f] = dalloc_bytes(c, nb)))
            return LES_ERR_HIP;
    }
    if (!(d->b.nut = dalloc_bytes(c, nb)) || !(d->b.st = dalloc_bytes(c, nb))) return LES_ERR_HIP;
    const int tn = c->p.table_n;
    std::vector<double> cp(tn), mu(tn);
    for (int t = 0; t < tn; ++t) {
        double T = c->p.table_t0 + t * c->p.table_dt;
        cp[t] = 1000.0 + 0.2 * (T - 300.0);
        mu[t] = 1.8e-5 * pow(T / 300.0, 0.7);
    }
    d->b.cp_tab = (double*)dalloc_bytes(c, tn * 8);
    d->b.mu_tab = (double*)dalloc_bytes(c, tn * 8);
    if (!d->b.cp_tab || !d->b.mu_tab) return LES_ERR_HIP;
    DHIPC(hipMemcpy(d->b.cp_tab, cp.data(), tn * 8, hipMemcpyHostToDevice));
    DHIPC(hipMemcpy(d->b.mu_tab, mu.data(), tn * 8, hipMemcpyHostToDevice));
    const size_t P = (size_t)c->p.ny * c->p.nz;
    d->b.eddy = (double*)dalloc_bytes(c, 7 * 8 * (size_t)(c->p.num_vortices > 0 ? c->p.num_vortices : 1));
    d->b.inlet = (double*)dalloc_bytes(c, 3 * 8 * P);
    Launch3 L = launch3(c->g.nx, c->g.ny, c->g.nz);
    d->filter_blocks = (int)(L.grid.x * L.grid.y * L.grid.z);
    d->b.rho_part = (double*)dalloc_bytes(c, 8 * (size_t)d->filter_blocks);
    d->b.scal = (double*)dalloc_bytes(c, 8 * 8);
    d->b.maxbits = (unsigned long long*)dalloc_bytes(c, 8 * 4);
    d->bc_blocks = (int)(((c->g.nz + 255) / 256) * c->g.ny);
    d->b.vpart = (double*)dalloc_bytes(c, 8 * (size_t)(d->filter_blocks + d->bc_blocks + 1024));
    d->b.dpart = (double*)dalloc_bytes(c, 8 * 2 * 3 * (size_t)d->filter_blocks);
    d->stat_chunks = 16;
    d->b.spart = (double*)dalloc_bytes(c, 8 * 10 * (size_t)d->stat_chunks * c->p.nx);
    if (!d->b.vpart || !d->b.dpart || !d->b.spart) return LES_ERR_HIP;
    d->b.prof = (double*)dalloc_bytes(c, 8 * 10 * (size_t)c->p.nx);
    d->b.mon_part = (double*)dalloc_bytes(c, 8 * 6 * 1024);
    if (!d->b.eddy || !d->b.inlet || !d->b.rho_part || !d->b.scal || !d->b.maxbits || !d->b.prof || !d->b.mon_part)
        return LES_ERR_HIP;
    if (c->p.stats_start_step >= 0 && c->p.stats_per_cell) {
        for (int a = 0; a < 10; ++a)
            if (!(d->b.stat[a] = (double*)dalloc_bytes(c, 8 * n))) return LES_ERR_HIP;
    }
    // 64 fixed probe points: 4 x-stations x 4 y x 4 z, evenly spaced
    d->npts = 64;
    d->point_cap = 4096;
    std::vector<int> pidx(64);
    int t = 0;
    for (int a = 0; a < 4; ++a)
        for (int b2 = 0; b2 < 4; ++b2)
            for (int e = 0; e < 4; ++e) {
                int ii = (a + 1) * c->p.nx / 5, jj = (2 * b2 + 1) * c->p.ny / 8, kk = (2 * e + 1) * c->p.nz / 8;
                pidx[t++] = (int)(((long long)ii * c->p.ny + jj) * c->p.nz + kk);
            }
    d->b.point_idx = (int*)dalloc_bytes(c, 64 * 4);
    d->b.points = (double*)dalloc_bytes(c, 8 * 4 * 64 * (size_t)d->point_cap);
    if (!d->b.point_idx || !d->b.points) return LES_ERR_HIP;
    DHIPC(hipMemcpy(d->b.point_idx, pidx.data(), 64 * 4, hipMemcpyHostToDevice));
    if (d->precision == LES_PREC_FP32 && !(d->stage64 = (double*)dalloc_bytes(c, 8 * n))) return LES_ERR_HIP;
    c->attr_h.assign(n, 3);
    return LES_OK;
}

void duct_destroy(les_gpu* c)
{
    if (!c->d) return;
    if (c->d->gexec) (void)hipGraphExecDestroy(c->d->gexec);
    if (c->d->graph) (void)hipGraphDestroy(c->d->graph);
    delete c->d;
    c->d = nullptr;
}

// ---- typed launches ----
template <typename real>
static int duct_velmax_t(les_gpu* c)
{
    DuctState* d = c->d;
    const GridDims& g = c->g;
    hipLaunchKernelGGL(k_duct_velmax<real>, dim3(1024), dim3(256), 0, c->stream, g, d->k.rho_u_Tu,
                       (const real*)d->b.A[0], (const real*)d->b.A[1], (const real*)d->b.A[2], (const real*)d->b.A[4],
                       d->b.vpart);
    DLAUNCH_CHECK();
    return LES_OK;
}

template <typename real>
static int duct_stage_t(les_gpu* c, int stage, const real* const* qin, real* const* qout, real* const* Lout = nullptr)
{
    DuctState* d = c->d;
    const GridDims& g = c->g;
    Launch3 L = launch3(g.nx, g.ny, g.nz);
    real** F = (real**)d->b.F;
    hipLaunchKernelGGL(k_duct_filter<real>, L.grid, L.block, 0, c->stream, g, d->k, qin[0], qin[1], qin[2], qin[3],
                       qin[4], F[0], F[1], F[2], F[3], F[4], d->b.rho_part);
    DLAUNCH_CHECK();
    hipLaunchKernelGGL(k_duct_rho_final, dim3(1), dim3(1024), 0, c->stream, g.n, d->filter_blocks, d->b.rho_part,
                       d->b.scal);
    DLAUNCH_CHECK();
    hipLaunchKernelGGL(k_duct_sgs<real>, L.grid, L.block, 0, c->stream, g, d->k, d->b.mu_tab, d->b.cp_tab, F[0], F[1],
                       F[2], F[3], F[4], (real*)d->b.nut, (real*)d->b.st,
                       d->b.dpart + 2 * (size_t)stage * d->filter_blocks);
    DLAUNCH_CHECK();
    RhsArgs<real> a;
    a.uf = F[0]; a.vf = F[1]; a.wf = F[2]; a.Tf = F[3]; a.pf = F[4];
    a.nut = (const real*)d->b.nut; a.st = (const real*)d->b.st;
    for (int f = 0; f < DNF; ++f) { a.qin[f] = qin[f]; a.q0[f] = (const real*)d->b.A[f]; a.qout[f] = qout[f]; }
    a.mu_tab = d->b.mu_tab; a.cp_tab = d->b.cp_tab; a.scal = d->b.scal; a.vpart = d->b.vpart;
    a.stage = stage;
    a.ca = (stage == 1) ? 0.75 : (1.0 / 3.0);
    a.cb = (stage == 1) ? 0.25 : (2.0 / 3.0);
    a.track_vel = (stage == 2);
    for (int f = 0; f < DNF; ++f) a.Lout[f] = Lout ? Lout[f] : nullptr;
    if (c->timing) DHIPC(hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_duct_rhs_scalar<real>, L.grid, L.block, 0, c->stream, g, d->k, a);
    hipLaunchKernelGGL(k_duct_rhs_momentum<real>, L.grid, L.block, 0, c->stream, g, d->k, a);
    DLAUNCH_CHECK();
    if (c->timing) {
        DHIPC(hipEventRecord(c->ev[1], c->stream));
        DHIPC(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        DHIPC(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->rhs_ms_sum += ms;
        c->rhs_launches++;
    }
    return LES_OK;
}

// One full time step, all launches on c->stream (capturable).
template <typename real>
static int duct_step_launches(les_gpu* c, bool reinit, bool stats, bool points)
{
    DuctState* d = c->d;
    const GridDims& g = c->g;
    const long long P = (long long)g.ny * g.nz;
    hipLaunchKernelGGL(k_duct_dt, dim3(1), dim3(1024), 0, c->stream, g.dx, d->k, d->b.scal, d->b.vpart,
                       d->filter_blocks + d->bc_blocks, d->b.dpart, 3 * d->filter_blocks);
    DLAUNCH_CHECK();
    if (d->k.nvort > 0) {
        hipLaunchKernelGGL(k_duct_eddy_update, dim3((d->k.nvort + 255) / 256), dim3(256), 0, c->stream, d->k, d->b.scal,
                           d->b.eddy);
        DLAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_duct_sem_plane, dim3((g.nz + 255) / 256, g.ny), dim3(256), 0, c->stream, g, d->k, d->b.eddy,
                       d->b.inlet);
    hipLaunchKernelGGL(k_duct_sem_norm, dim3(3), dim3(1024), 0, c->stream, P, d->k.U * d->k.intensity, d->b.inlet);
    DLAUNCH_CHECK();
    real** A = (real**)d->b.A;
    real** B = (real**)d->b.B;
    int e;
    if ((e = duct_stage_t<real>(c, 0, (const real* const*)A, B)) != LES_OK) return e;
    if ((e = duct_stage_t<real>(c, 1, (const real* const*)B, B)) != LES_OK) return e;
    if ((e = duct_stage_t<real>(c, 2, (const real* const*)B, A)) != LES_OK) return e;
    hipLaunchKernelGGL(k_duct_bc<real>, dim3((g.nz + 255) / 256, g.ny), dim3(256), 0, c->stream, g, d->k, d->b.inlet,
                       A[0], A[1], A[2], A[3], A[4], d->b.vpart + d->filter_blocks);
    DLAUNCH_CHECK();
    if (reinit) {
        Launch3 L = launch3(g.nx, g.ny, g.nz);
        int nbk = (int)((g.n + 255) / 256);
        // phi0 kept in B[3]; ping-pong between A[3] and B[4] (B is dead after the step)
        hipLaunchKernelGGL(k_copy<real>, dim3(nbk), dim3(256), 0, c->stream, g.n, (const real*)A[3], B[3]);
        for (int it = 0; it < c->p.reinit_iters; ++it) {
            hipLaunchKernelGGL(k_duct_reinit<real>, L.grid, L.block, 0, c->stream, g, (const real*)B[3],
                               (const real*)A[3], B[4]);
            hipLaunchKernelGGL(k_copy<real>, dim3(nbk), dim3(256), 0, c->stream, g.n, (const real*)B[4], A[3]);
        }
        DLAUNCH_CHECK();
    }
    if (stats) {
        StatPtrs st;
        for (int a = 0; a < 10; ++a) st.acc[a] = d->b.stat[a];
        hipLaunchKernelGGL(k_duct_stats<real>, dim3(d->stat_chunks, g.nx), dim3(256), 0, c->stream, g, d->k,
                           (const real*)A[0], (const real*)A[1], (const real*)A[2], (const real*)A[3],
                           (const real*)A[4], st, d->b.stat[0] != nullptr ? 1 : 0, d->b.spart);
        hipLaunchKernelGGL(k_duct_stats_final, dim3((g.nx * 10 + 255) / 256), dim3(256), 0, c->stream, g.nx,
                           d->stat_chunks, d->b.spart, d->b.prof);
        DLAUNCH_CHECK();
    }
    if (points) {
        int slot = d->point_count % d->point_cap;
        hipLaunchKernelGGL(k_duct_points<real>, dim3(1), dim3(64), 0, c->stream, d->k, g.dx, d->b.point_idx, d->npts,
                           (const real*)A[0], (const real*)A[1], (const real*)A[2], (const real*)A[3],
                           (const real*)A[4], d->b.points, slot);
        DLAUNCH_CHECK();
    }
    return LES_OK;
}

template <typename real>
static int duct_monitor_t(les_gpu* c, les_step_report* rep)
{
    DuctState* d = c->d;
    const GridDims& g = c->g;
    int nb = (int)((g.n + 255) / 256);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_duct_monitor<real>, dim3(nb), dim3(256), 0, c->stream, g, d->k, (const real*)d->b.A[0],
                       (const real*)d->b.A[3], (const real*)d->b.A[4], d->b.mon_part);
    DLAUNCH_CHECK();
    std::vector<double> h((size_t)nb * 6);
    double sc[4];
    DHIPC(hipMemcpyAsync(h.data(), d->b.mon_part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipMemcpyAsync(sc, d->b.scal, 4 * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, bad = 0;
    for (int b = 0; b < nb; ++b) {
        const double* q = &h[6 * (size_t)b];
        tm += q[0]; mnT = fmin(mnT, q[1]); mxT = fmax(mxT, q[2]); mnG = fmin(mnG, q[3]); mxG = fmax(mxG, q[4]);
        bad = fmax(bad, q[5]);
    }
    rep->dt = sc[0]; rep->max_velocity = sc[2];
    rep->total_mass = tm; rep->min_T = mnT; rep->max_T = mxT; rep->min_G = mnG; rep->max_G = mxG;
    rep->min_Y_fuel = 0.0; rep->nonfinite = bad > 0.0; rep->proj_residual = 0.0;
    return LES_OK;
}

template <typename real>
static int duct_steps_t(les_gpu* c, int n_steps, les_step_report* rep)
{
    DuctState* d = c->d;
    const bool stats_on = c->p.stats_start_step >= 0;
    for (int s = 0; s < n_steps; ++s) {
        const int t = c->step;
        const bool reinit = c->p.reinit_interval > 0 && (t + 1) % c->p.reinit_interval == 0;
        const bool stats = stats_on && t >= c->p.stats_start_step;
        // a plain step (no reinitialisation) replays a captured graph
        bool plain = !reinit && !c->timing && d->use_graph;
        if (plain && d->graph_ok && (stats != d->graph_stats)) plain = false;
        if (plain && !d->graph_ok) {
            d->graph_stats = stats;
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                int e = duct_step_launches<real>(c, false, stats, false);
                hipGraph_t gr = nullptr;
                hipError_t ce = hipStreamEndCapture(c->stream, &gr);
                if (e == LES_OK && ce == hipSuccess && gr) {
                    if (d->gexec) (void)hipGraphExecDestroy(d->gexec);
                    if (d->graph) (void)hipGraphDestroy(d->graph);
                    d->graph = gr;
                    if (hipGraphInstantiate(&d->gexec, gr, nullptr, nullptr, 0) == hipSuccess) d->graph_ok = true;
                } else if (gr) {
                    (void)hipGraphDestroy(gr);
                }
                (void)hipGetLastError();
            }
            if (!d->graph_ok) { d->use_graph = 0; plain = false; }
        }
        if (plain && d->graph_ok) {
            DHIPC(hipGraphLaunch(d->gexec, c->stream));
            int slot = d->point_count % d->point_cap;
            hipLaunchKernelGGL(k_duct_points<real>, dim3(1), dim3(64), 0, c->stream, d->k, c->g.dx, d->b.point_idx,
                               d->npts, (const real*)d->b.A[0], (const real*)d->b.A[1], (const real*)d->b.A[2],
                               (const real*)d->b.A[3], (const real*)d->b.A[4], d->b.points, slot);
            DLAUNCH_CHECK();
        } else {
            int e = duct_step_launches<real>(c, reinit, stats, true);
            if (e != LES_OK) return e;
        }
        d->point_count++;
        if (stats) { c->stat_count++; d->prof_count++; }
        c->step++;
    }
    if (rep) {
        rep->step = c->step - 1;
        return duct_monitor_t<real>(c, rep);
    }
    return LES_OK;
}

int duct_step(les_gpu* c, int n_steps, les_step_report* rep)
{
    DHIPC(hipEventRecord(c->ev[2], c->stream));
    int e = (c->d->precision == LES_PREC_FP32) ? duct_steps_t<float>(c, n_steps, nullptr)
                                               : duct_steps_t<double>(c, n_steps, nullptr);
    if (e != LES_OK) return e;
    DHIPC(hipEventRecord(c->ev[3], c->stream));
    DHIPC(hipEventSynchronize(c->ev[3]));
    float ms = 0.f;
    DHIPC(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->step_ms = n_steps > 0 ? ms / n_steps : 0.0;
    if (rep) {
        rep->step = c->step - 1;
        e = (c->d->precision == LES_PREC_FP32) ? duct_monitor_t<float>(c, rep) : duct_monitor_t<double>(c, rep);
        if (e != LES_OK) return e;
        if (rep->nonfinite) {
            snprintf(c->err, sizeof(c->err), "non-finite value in the state at step %d", rep->step);
            return LES_ERR_NAN;
        }
    }
    return LES_OK;
}

// ---- state exchange (fields rho*u, rho*v, rho*w, phi, T as double) ----
template <typename real>
__global__ void k_to_real(long long n, const double* src, real* dst)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = (real)src[q];
}

template <typename real>
__global__ void k_to_double(long long n, const real* src, double* dst)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = (double)src[q];
}

static int duct_xfer(les_gpu* c, void* dev, double* host, bool upload)
{
    DuctState* d = c->d;
    const size_t n = (size_t)c->g.n;
    if (d->precision == LES_PREC_FP64) {
        if (upload) DHIPC(hipMemcpyAsync(dev, host, n * 8, hipMemcpyHostToDevice, c->stream));
        else DHIPC(hipMemcpyAsync(host, dev, n * 8, hipMemcpyDeviceToHost, c->stream));
        return LES_OK;
    }
    // FP32 storage: convert through the context's FP64 staging buffer
    double* tmp = d->stage64;
    int nb = (int)((n + 255) / 256);
    if (upload) {
        DHIPC(hipMemcpyAsync(tmp, host, n * 8, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_to_real<float>, dim3(nb), dim3(256), 0, c->stream, (long long)n, tmp, (float*)dev);
    } else {
        hipLaunchKernelGGL(k_to_double<float>, dim3(nb), dim3(256), 0, c->stream, (long long)n, (const float*)dev, tmp);
        DHIPC(hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    return LES_OK;
}

static int duct_after_state_change(les_gpu* c)
{
    DuctState* d = c->d;
    DHIPC(hipMemsetAsync(d->b.vpart, 0, 8 * (size_t)(d->filter_blocks + d->bc_blocks), c->stream));
    DHIPC(hipMemsetAsync(d->b.dpart, 0, 8 * 2 * 3 * (size_t)d->filter_blocks, c->stream));
    int e = (d->precision == LES_PREC_FP32) ? duct_velmax_t<float>(c) : duct_velmax_t<double>(c);
    if (e != LES_OK) return e;
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

int duct_upload_state(les_gpu* c, const double* const* f, int nf)
{
    if (nf < DNF) { snprintf(c->err, sizeof(c->err), "duct mode state has 5 fields"); return LES_ERR_PARAM; }
    for (int a = 0; a < DNF; ++a) {
        int e = duct_xfer(c, c->d->b.A[a], (double*)f[a], true);
        if (e != LES_OK) return e;
    }
    return duct_after_state_change(c);
}

int duct_download_state(les_gpu* c, double* const* f, int nf)
{
    for (int a = 0; a < nf && a < DNF; ++a) {
        int e = duct_xfer(c, c->d->b.A[a], f[a], false);
        if (e != LES_OK) return e;
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

int duct_init_state(les_gpu* c)
{
    DuctState* d = c->d;
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const size_t n = (size_t)c->g.n;
    std::vector<double> h(5 * n);
    const double U = c->p.inlet_velocity, Tu = c->p.inlet_temperature, Tb = c->p.burnt_temperature;
    for (int i = 0; i < nx; ++i) {
        double phi = i * c->g.dx - c->p.flame_x0;
        double cc = 0.5 * (1.0 + tanh(phi / (2.0 * c->g.dx)));
        double T = Tu + cc * (Tb - Tu);
        double rho = c->p.rho_unburnt * Tu / T;
        for (size_t p = 0; p < (size_t)ny * nz; ++p) {
            size_t q = (size_t)i * ny * nz + p;
            h[q] = rho * U; h[n + q] = 0.0; h[2 * n + q] = 0.0; h[3 * n + q] = phi; h[4 * n + q] = T;
        }
    }
    const double* f[5] = {&h[0], &h[n], &h[2 * n], &h[3 * n], &h[4 * n]};
    int e = duct_upload_state(c, f, 5);
    if (e != LES_OK) return e;
    if (d->k.nvort > 0) {
        hipLaunchKernelGGL(k_duct_eddy_init, dim3((d->k.nvort + 255) / 256), dim3(256), 0, c->stream, d->k, d->b.eddy);
        DLAUNCH_CHECK();
    }
    DHIPC(hipMemsetAsync(d->b.inlet, 0, 3 * 8 * (size_t)ny * nz, c->stream));
    DHIPC(hipMemsetAsync(d->b.prof, 0, 10 * 8 * (size_t)nx, c->stream));
    if (d->b.stat[0])
        for (int a = 0; a < 10; ++a) DHIPC(hipMemsetAsync(d->b.stat[a], 0, n * 8, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    c->step = 0;
    c->stat_count = 0;
    d->prof_count = 0;
    d->point_count = 0;
    c->probe.clear();
    return LES_OK;
}

int duct_download_field(les_gpu* c, const char* name, double* out)
{
    DuctState* d = c->d;
    const size_t n = (size_t)c->g.n;
    struct { const char* nm; void* p; } tab[] = {
        {"u_f", d->b.F[0]}, {"v_f", d->b.F[1]}, {"w_f", d->b.F[2]}, {"T_f", d->b.F[3]}, {"phi_f", d->b.F[4]},
        {"nu_t", d->b.nut}, {"ST", d->b.st},
        {"rhou", d->b.A[0]}, {"rhov", d->b.A[1]}, {"rhow", d->b.A[2]}, {"phi", d->b.A[3]}, {"T", d->b.A[4]}};
    for (auto& t : tab)
        if (strcmp(t.nm, name) == 0) {
            int e = duct_xfer(c, t.p, out, false);
            if (e != LES_OK) return e;
            DHIPC(hipStreamSynchronize(c->stream));
            return LES_OK;
        }
    const size_t P = (size_t)c->p.ny * c->p.nz;
    struct { const char* nm; const double* p; size_t cnt; } dt[] = {
        {"inlet", d->b.inlet, 3 * P}, {"eddies", d->b.eddy, 7 * (size_t)c->p.num_vortices},
        {"profiles", d->b.prof, 10 * (size_t)c->p.nx}, {"scalars", d->b.scal, 8}};
    for (auto& t : dt)
        if (strcmp(t.nm, name) == 0) {
            DHIPC(hipMemcpyAsync(out, t.p, t.cnt * 8, hipMemcpyDeviceToHost, c->stream));
            DHIPC(hipStreamSynchronize(c->stream));
            return LES_OK;
        }
    if (strncmp(name, "stat", 4) == 0 && d->b.stat[0]) {
        int a = atoi(name + 4);
        if (a < 0 || a > 9) return LES_ERR_PARAM;
        DHIPC(hipMemcpyAsync(out, d->b.stat[a], n * 8, hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        return LES_OK;
    }
    snprintf(c->err, sizeof(c->err), "unknown duct field %s", name);
    return LES_ERR_PARAM;
}

// Testing hook: the RHS L(q) of the current state (stage-0 evaluation; its
// combined output goes to the scratch stage buffers B, L itself to F-sized
// scratch taken from B after the stage: B is dead between steps).
int duct_rhs(les_gpu* c, double* const* out, int n_out)
{
    if (n_out < DNF) return LES_ERR_PARAM;
    DuctState* d = c->d;
    const size_t n = (size_t)c->g.n;
    std::vector<void*> Lbuf(DNF, nullptr);
    for (int f = 0; f < DNF; ++f) {
        DHIPC(hipMalloc(&Lbuf[f], n * d->esize));
        DHIPC(hipMemsetAsync(Lbuf[f], 0, n * d->esize, c->stream));   // boundary planes read as 0
    }
    int e;
    if (d->precision == LES_PREC_FP32)
        e = duct_stage_t<float>(c, 0, (const float* const*)d->b.A, (float**)d->b.B, (float**)Lbuf.data());
    else
        e = duct_stage_t<double>(c, 0, (const double* const*)d->b.A, (double**)d->b.B, (double**)Lbuf.data());
    for (int f = 0; f < DNF && e == LES_OK; ++f) {
        e = duct_xfer(c, Lbuf[f], out[f], false);
    }
    DHIPC(hipStreamSynchronize(c->stream));
    for (int f = 0; f < DNF; ++f) (void)hipFree(Lbuf[f]);
    return e;
}

// fixed-point time series since the last fetch: out[step][point][u, v, w, G]
int duct_point_fetch(les_gpu* c, double* out, int max_steps, int* n_steps)
{
    DuctState* d = c->d;
    int avail = d->point_count < d->point_cap ? d->point_count : d->point_cap;
    int n = avail < max_steps ? avail : max_steps;
    int first = d->point_count - avail;    // oldest kept step
    for (int s = 0; s < n; ++s) {
        int slot = (first + s) % d->point_cap;
        DHIPC(hipMemcpyAsync(out + (size_t)s * d->npts * 4, d->b.points + (size_t)slot * d->npts * 4,
                             (size_t)d->npts * 4 * 8, hipMemcpyDeviceToHost, c->stream));
    }
    DHIPC(hipStreamSynchronize(c->stream));
    d->point_count = 0;
    *n_steps = n;
    return LES_OK;
}
