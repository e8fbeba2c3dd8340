// les_duct_kernels.hip - gfx950 kernels of the duct-mode scheme: periodic y/z
// duct, SSP-RK3, WENO5 or MUSCL-minmod upwind reconstruction, cp/mu tables,
// Gaussian SEM eddies from a counter-based RNG, level-set G with Sussman
// reinitialisation, FP64 statistics.  Templated on the storage/arithmetic type
// (double: the FP64 path checked against the oracle; float: the FP32 form).
// Reductions (domain-mean density, SEM plane moments, statistics, DT maxima)
// are FP64 in both forms.
//
// One engine context holds a pool of SLOTS; each slot is one run with its own
// grid, parameters, time step and step counter.  Field storage is shared
// arenas in HBM (first-fit cell ranges per slot) and every stencil kernel runs
// one block per entry of a device-side TILE LIST (first-fit tile ranges per
// slot), so one launch advances every active run: the ensemble and the served
// request stream use the same kernels as a single run (a pool of one).  All
// per-run state the kernels need (descriptor, dt, counters) lives on the
// device, so one time step of the whole pool is captured once into a hipGraph
// and replayed; admitting or retiring a run only rewrites device tables.
//
// Per RK stage:
//   k_duct_filter      state -> filtered primitives u, v, w, T, phi (+ per-tile sums of rho)
//   k_duct_sgs         nu_t, S_T (+ per-tile max nu / alpha for the next DT; the
//                      run's first tile reduces its domain-mean density)
//   k_duct_rhs_all     rho*u, rho*v, rho*w, phi, T: RHS fused with the SSP-RK3
//                      combine (+ per-tile max |u| in the last stage)
// Per step: DT, eddy update, inflow plane, its normalisation (on a side
// stream, beside the stages), the 3 stages, projection relaxation (runs with
// proj_iters > 0), inflow/outflow planes, reinitialisation (runs due),
// point series, probe rows, statistics, step end.  The host side
// (pool, launch sequence, graphs, ABI) is in les_duct_host.hip.
#include "les_engine.hpp"

// The duct scheme is this engine's own specification: its kernels may fuse
// multiply-adds (the reference-scheme kernels keep strict IEEE order).
#pragma clang fp contract(fast)

#define DNF 5
#define DUCT_NPTS 64
#define STAT_CHUNKS 16            // most chunks of one x-plane in the statistics pass
#define DT_BLOCKS_MAX 16          // most blocks per slot of the DT reduction
#define STAT_CHUNK_CELLS 2048     // target cells per chunk (8 per thread)

// ---------------------------------------------------------------------------
// device-side parameters of the duct scheme (one set per slot)
// ---------------------------------------------------------------------------
#define DUCT_NRC 21          // host-computed reciprocal constants per run

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
    double fixed_dt;        // > 0: prescribed DT_sim
    double table_t0, table_dt;
    int table_n;
    int scheme;             // LES_ADV_WENO5 or LES_ADV_MUSCL_MINMOD
    int nvort;
    int reinit_interval, reinit_iters;
    int proj_iters;         // projection relaxation sweeps per step (0: off)
    int stats_start, stats_per_cell;
    double dx4;             // dx^4 (host pow, as the oracle's)
    double L;               // eddy length scale
    double ly, lz;
    unsigned long long seed;
    // reciprocals of run constants, computed on the host in FP64 (the value a
    // device FP64 division gives) and rounded to FP32, so the kernels issue no
    // uniform divisions or conversions; indices DUCT_RC_*
    double rc[DUCT_NRC];
    float rcf[DUCT_NRC];
};

enum {
    DUCT_RC_ID = 0,     // 1/dx, 1/dy, 1/dz
    DUCT_RC_HID = 3,    // 0.5/dx, 0.5/dy, 0.5/dz
    DUCT_RC_ID2 = 6,    // 1/dx^2, 1/dy^2, 1/dz^2
    DUCT_RC_IRUT = 9,   // 1/(rho_u T_u)
    DUCT_RC_IDEN = 10,  // 1/(6 + filter centre weight)
    DUCT_RC_ISC = 11,   // 0.5/Sc_t
    DUCT_RC_IPR = 12,   // 0.5/Pr_t
    DUCT_RC_INVD = 13,  // 1/(2/dx^2 + 2/dy^2 + 2/dz^2) (Jacobi diagonal)
    DUCT_RC_ITDT = 14,  // 1/(property table step)
    DUCT_RC_IPRAN = 15, // 1/Pr
    DUCT_RC_IW2 = 16,   // 1/(2 w^2), w the flame-acceleration window width
    DUCT_RC_DF = 17,    // nu_u/(S_L + small): flame thickness of the SSR model
    DUCT_RC_ITSL = 18,  // 1/(tau S_L) (0 when tau S_L <= small)
    DUCT_RC_ILY = 19,   // 1/ly, 1/lz: periodic images of the SEM eddies
    DUCT_RC_ILZ = 20
};

template <typename real>
__device__ __forceinline__ real rcp_const(const DuctConsts& k, int i);
template <>
__device__ __forceinline__ float rcp_const<float>(const DuctConsts& k, int i) { return k.rcf[i]; }
template <>
__device__ __forceinline__ double rcp_const<double>(const DuctConsts& k, int i) { return k.rc[i]; }

// per-direction entry base + d for a d known only at run time: a select, not
// an indexed load (a dynamic index would move the constants to scratch)
template <typename real>
__device__ __forceinline__ real rcp_dir(const DuctConsts& k, int base, int d)
{
    const real a = rcp_const<real>(k, base), b = rcp_const<real>(k, base + 1), c = rcp_const<real>(k, base + 2);
    return d == 0 ? a : (d == 1 ? b : c);
}

// |u| of a cell from its conservative state (1/rho = T/(rho_u T_u)).  Every
// kernel that feeds the dt rule uses this one form, with explicit FMAs, so a
// run resumed from a checkpoint (whose first dt comes from a separate pass)
// takes bit for bit the dt of a run that never stopped.
template <typename real>
__device__ __forceinline__ double cell_speed(real m0, real m1, real m2, real T, const DuctConsts& k)
{
    const real ri = T * rcp_const<real>(k, DUCT_RC_IRUT);
    const real u = m0 * ri, v = m1 * ri, w = m2 * ri;
    return (double)sqrt(fma(w, w, fma(v, v, u * u)));
}

// Run descriptor: written by the host when a run is admitted or retired.
struct SlotDesc {
    GridDims g;
    long long off;          // first cell in the field arenas
    long long poff;         // first value in the plane arena (3 planes)
    long long xoff;         // first value in the profile arena (10 per x-plane)
    long long eoff;         // first eddy in the eddy arena
    long long ptoff;        // first value in the point-series arena
    int tile0, ntiles;      // tile range in the tile list
    int tiles_j, tiles_k;   // tiles of one x-plane
    int active;             // 1 while admitted
    int n_steps;            // steps requested
    int point_idx[DUCT_NPTS];
    int npts;               // points in use (<= DUCT_NPTS)
    int nprobe;             // flame-region probe stations
    int probe_i[LES_NPROBE_MAX];
    long long proff;        // first value of the run's probe ring
    DuctConsts k;
};

// Run state the kernels update (never written by the host while a step runs).
struct SlotDyn {
    double dt, rho_avg, maxvel, time;
    double proj_res0, proj_res;   // FP64 RMS residuals of the last projection
    int step;               // steps done
    int done;               // step >= n_steps
    int pad[2];
};

// Pool-wide device pointers handed to every kernel.
struct Pool {
    void* A[DNF];           // q^n, overwritten by q^{n+1} in the last stage
    void* B[DNF];           // RK stage state / scratch
    void* F[DNF];           // filtered u, v, w, T, phi
    void* nut;
    void* st;
    double* stat[10];       // per-cell FP64 accumulators (optional arena)
    const SlotDesc* desc;
    SlotDyn* dyn;
    const int* tile_slot;   // tile -> slot (-1: free)
    const int* ntile;       // tile high-water mark
    int tb, tn;             // tn > 0: the launch covers tiles [tb, tb + tn) only (a group of runs)
    double* mu_tab;         // viscosity against T (tables: t0, step, n in DuctConsts)
    double* eddy;           // 7 doubles per eddy
    double* inlet;          // 3 planes per slot
    double* prof;           // 10 FP64 per x-plane per slot
    double* points;         // ring: cap x 64 x 4 per slot
    double* rho_part;       // per tile
    double* vpart;          // per tile: max |u| of the new state
    double* dpart;          // [3][tile][2]: max nu, alpha per stage
    double* bcpart;         // [slot][ny_max][nzb]: max |u| of the inflow plane
    double* spart;          // [slot][nx_max][STAT_CHUNKS][10]
    double* ppart;          // per tile: FP64 sums of b^2 and r^2 of the projection
    int tile_cap;
    int ny_max, nz_max, nx_max, nzb;
    int point_cap;          // steps per point-series ring
    double* sem_part;       // SEM partial planes: [slot][sem_split][3][ny_max*nz_max]
    int sem_split;          // eddy chunks of SEM_CHUNK per run
    double* probe;          // probe ring: [slot][probe_cap][station][2 + 6*ny_max]
    int probe_cap;          // steps per probe ring (0: no probes)
    int* sdue;              // per slot: 1 if the run sampled statistics this step
    double* dtpart;         // [slot][DT_BLOCKS_MAX][3]: block maxima of the DT reduction
    unsigned* dtcnt;        // per slot: blocks of the DT reduction counted in (0 between launches)
};

#define PROBE_STRIDE(P) (2 + 6 * (P).ny_max)
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

// Arithmetic of the two storage types.  double (the FP64 path, compared with
// the oracle to 1e-10): the device libm's exp and log; x^y as exp(y log x);
// x^(3/4) as two square roots; divisions by Newton-Raphson from the hardware
// reciprocal (v_rcp_f64, one refinement step e = 1 - b r, r += r e, then one
// residual correction of the quotient q += r (a - b q), Markstein's scheme:
// the correction squares the refined reciprocal's relative error, so the
// quotient is within about 1 ulp of the IEEE one; measured on 4.2 million
// random operand pairs of magnitudes 1e-30..1e30 it equalled the IEEE
// quotient every time, as it did with two refinement steps, while with none
// it was up to 18 ulp off) for the finite, normal, moderate operands these
// kernels divide (temperatures, positive sums of squares, eddy viscosities)
// and without the scaling and fix-up steps the IEEE sequence adds.  float (the
// FP32 form): the hardware's native reciprocal, exp2 and log2 (v_rcp_f32,
// v_exp_f32, v_log_f32; about 1 ulp), which the FP32 tolerances cover.
template <typename real>
struct M {
    static __device__ __forceinline__ real div(real a, real b)
    {
        real r = __builtin_amdgcn_rcp(b);
        const real e = fma(-b, r, (real)1.0);
        r = fma(r, e, r);
        const real q = a * r;
        return fma(r, fma(-b, q, a), q);
    }
    static __device__ __forceinline__ real exp(real x) { return ::exp(x); }
    static __device__ __forceinline__ real log(real x) { return ::log(x); }
    // x^y of positive x as exp(y log x): relative error about |y log x| ulp,
    // without the special-case handling of the libm pow
    static __device__ __forceinline__ real pow(real x, real y) { return ::exp(y * ::log(x)); }
    static __device__ __forceinline__ real pow34(real x) { return sqrt(x * sqrt(x)); }
};

template <>
struct M<float> {
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504089f); }
    static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.69314718056f; }
    static __device__ __forceinline__ float pow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
    static __device__ __forceinline__ float pow34(float x) { return pow(x, 0.75f); }
};

// progress variable c = (1 + tanh(phi / (2 dx))) / 2, evaluated as the equal
// logistic 1 / (1 + exp(-phi / dx)) (one exponential and one division; no
// cancellation where c is small); i2dx = 0.5/dx of the run
template <typename real>
__device__ __forceinline__ real progress(real phi, real i2dx)
{
    return M<real>::div((real)1.0, (real)1.0 + M<real>::exp((real)-2.0 * (phi * i2dx)));
}

template <typename real>
__device__ __forceinline__ real table_lookup(const double* tab, const DuctConsts& k, real T)
{
    real s = (T - (real)k.table_t0) * rcp_const<real>(k, DUCT_RC_ITDT);
    if (s < (real)0.0) s = (real)0.0;
    if (s > (real)(k.table_n - 1)) s = (real)(k.table_n - 1);
    int i0 = (int)s;
    if (i0 > k.table_n - 2) i0 = k.table_n - 2;
    real f = s - (real)i0;
    real t0 = (real)tab[i0], t1 = (real)tab[i0 + 1];
    return t0 + f * (t1 - t0);
}

// WENO5 face value (Jiang-Shu, linear weights 0.1/0.6/0.3, eps 1e-6), one
// division per face: the nonlinear weights are scaled by the product of the
// three (eps + IS_k)^2.  Constant factors are folded for the instruction
// count: the three eps + IS_k are taken times 12/13 (a common factor of the
// weights' numerators and denominator), so the second differences enter
// unscaled and sqrt(12/13) x 1/2 goes into the first differences' coefficients,
// each IS_k being two FMAs; the linear weights and the common 1/6 go into the
// candidate values and the denominator.  Algebraically the textbook formula;
// only the rounding differs from it.  weno5_nd returns the quotient's
// numerator and denominator.
template <typename real>
__device__ __forceinline__ void weno5_nd(real v0, real v1, real v2, real v3, real v4, real& num_out, real& den_out)
{
    // the three IS_k + eps scaled by 12/13 (the weights only see their ratios):
    // the second differences unscaled, c = sqrt(12/13) folded into the first
    constexpr double c = 0.96076892283052284;
    const real eps = (real)(1.0e-6 * 12.0 / 13.0);
    const real a0 = v0 - (real)2.0 * v1 + v2, h0 = (real)(0.5 * c) * v0 - (real)(2.0 * c) * v1 + (real)(1.5 * c) * v2;
    const real a1 = v1 - (real)2.0 * v2 + v3, h1 = (real)(0.5 * c) * (v1 - v3);
    const real a2 = v2 - (real)2.0 * v3 + v4, h2 = (real)(1.5 * c) * v2 - (real)(2.0 * c) * v3 + (real)(0.5 * c) * v4;
    const real e0 = a0 * a0 + (h0 * h0 + eps);
    const real e1 = a1 * a1 + (h1 * h1 + eps);
    const real e2 = a2 * a2 + (h2 * h2 + eps);
    const real s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
    const real t12 = s1 * s2, t02 = s0 * s2, t01 = s0 * s1;
    const real q0 = (real)0.2 * v0 - (real)0.7 * v1 + (real)1.1 * v2;       // 0.1 x (2, -7, 11)
    const real q1 = (real)-0.6 * v1 + (real)3.0 * v2 + (real)1.2 * v3;      // 0.6 x (-1, 5, 2)
    const real q2 = (real)0.6 * v2 + (real)1.5 * v3 - (real)0.3 * v4;       // 0.3 x (2, 5, -1)
    const real num = t12 * q0 + t02 * q1 + t01 * q2;
    const real den = (real)0.6 * t12 + (real)3.6 * t02 + (real)1.8 * t01;  // 6 x (0.1, 0.6, 0.3)
    num_out = num;
    den_out = den;
}

template <typename real>
__device__ __forceinline__ real weno5_face(real v0, real v1, real v2, real v3, real v4)
{
    real n, d;
    weno5_nd<real>(v0, v1, v2, v3, v4, n, d);
    return M<real>::div(n, d);
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

// upwind face values of the cell from the 7-point line v[0..6] (offset -3..3);
// the upwind-ordered stencil is selected first and reconstructed once, so a
// wave whose lanes disagree on the sign evaluates one reconstruction, not two
template <typename real>
__device__ __forceinline__ real face_plus(int scheme, const real* v, bool positive)
{
    return face_left<real>(scheme, positive ? v[1] : v[6], positive ? v[2] : v[5], positive ? v[3] : v[4],
                           positive ? v[4] : v[3], positive ? v[5] : v[2]);
}

template <typename real>
__device__ __forceinline__ real face_minus(int scheme, const real* v, bool positive)
{
    return face_left<real>(scheme, positive ? v[0] : v[5], positive ? v[1] : v[4], positive ? v[2] : v[3],
                           positive ? v[3] : v[2], positive ? v[4] : v[1]);
}

// p f(i+1/2) - m f(i-1/2) of the cell's two upwind faces (signs pos_p, pos_m).
// FP64 WENO5: the two quotients over one division, (p n+ d- - m n- d+) / (d+ d-);
// each denominator is a sum of products of two (eps + IS_k)^2 terms, between
// about 1e-24 (eps = 1e-6) and 1e35 (temperature jumps of 2800 K, the clip
// range), so their product stays far inside the FP64 range (not the FP32 one,
// which keeps one division per face).
template <typename real>
__device__ __forceinline__ real face_pair(int scheme, const real* v, bool pos_p, bool pos_m, real p, real m)
{
    if constexpr (sizeof(real) == 8) {
        if (scheme == LES_ADV_WENO5) {
            real np, dp, nm, dm;
            weno5_nd<real>(pos_p ? v[1] : v[6], pos_p ? v[2] : v[5], pos_p ? v[3] : v[4], pos_p ? v[4] : v[3],
                           pos_p ? v[5] : v[2], np, dp);
            weno5_nd<real>(pos_m ? v[0] : v[5], pos_m ? v[1] : v[4], pos_m ? v[2] : v[3], pos_m ? v[3] : v[2],
                           pos_m ? v[4] : v[1], nm, dm);
            return M<real>::div(p * np * dm - m * nm * dp, dp * dm);
        }
    }
    return p * face_plus<real>(scheme, v, pos_p) - m * face_minus<real>(scheme, v, pos_m);
}

// f(i+1/2) - f(i-1/2) of the cell's two upwind faces when both take the same
// side (the scalars: upwind by the cell's own velocity): the line is reversed
// once when the side is negative, so the two faces read fixed positions of it
// (A = f(r1..r5), B = f(r0..r4): the faces are (A, B) on the positive side
// and (B, A) on the negative one) instead of selecting ten stencil values
template <typename real>
__device__ __forceinline__ real face_diff_same(int scheme, const real* v, bool pos)
{
    real r[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) r[o] = (o == 3 || pos) ? v[o] : v[6 - o];
    real d;
    if constexpr (sizeof(real) == 8) {
        if (scheme == LES_ADV_WENO5) {
            real na, da, nb, db;
            weno5_nd<real>(r[1], r[2], r[3], r[4], r[5], na, da);
            weno5_nd<real>(r[0], r[1], r[2], r[3], r[4], nb, db);
            d = M<real>::div(na * db - nb * da, da * db);
            return pos ? d : -d;
        }
    }
    d = face_left<real>(scheme, r[1], r[2], r[3], r[4], r[5]) - face_left<real>(scheme, r[0], r[1], r[2], r[3], r[4]);
    return pos ? d : -d;
}

// FP64 WENO5 only: the quotients of face_pair and face_diff_same as numerator
// and denominator, so that two of them can share one division
__device__ __forceinline__ void face_pair_nd(const double* v, bool pos_p, bool pos_m, double p, double m, double& n,
                                             double& d)
{
    double np, dp, nm, dm;
    weno5_nd<double>(pos_p ? v[1] : v[6], pos_p ? v[2] : v[5], pos_p ? v[3] : v[4], pos_p ? v[4] : v[3],
                     pos_p ? v[5] : v[2], np, dp);
    weno5_nd<double>(pos_m ? v[0] : v[5], pos_m ? v[1] : v[4], pos_m ? v[2] : v[3], pos_m ? v[3] : v[2],
                     pos_m ? v[4] : v[1], nm, dm);
    n = p * np * dm - m * nm * dp;
    d = dp * dm;
}

__device__ __forceinline__ void face_diff_same_nd(const double* v, bool pos, double& n, double& d)
{
    double r[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) r[o] = (o == 3 || pos) ? v[o] : v[6 - o];
    double na, da, nb, db;
    weno5_nd<double>(r[1], r[2], r[3], r[4], r[5], na, da);
    weno5_nd<double>(r[0], r[1], r[2], r[3], r[4], nb, db);
    const double nn = na * db - nb * da;
    n = pos ? nn : -nn;
    d = da * db;
}


// ---------------------------------------------------------------------------
// tile-to-block mapping (XCD-aware) and tile decoding
// ---------------------------------------------------------------------------
#define DUCT_CAS __attribute__((address_space(4)))

#define XCD_REVERSE 8      // flag on a launch's xcd_map: its blocks walk their XCD's tiles backwards

// tiles per chunk of the chunked XCD mappings (xcd_map 2 and 3); 1 otherwise
__host__ __device__ inline int xcd_chunk(int xcd_map) { return (xcd_map & 7) == 2 ? 16 : ((xcd_map & 7) == 3 ? 64 : 1); }

// One tile per block.  Blocks are dispatched round-robin over the 8 XCDs
// (block b runs on XCD b % 8); with xcd_map the blocks of one XCD take one
// contiguous range of the tile list, so tiles that share stencil halos
// (neighbours in j, k and i) are read through the same L2.  The grid is the
// tile high-water mark rounded up to whole rounds (8 blocks, or 8 chunks).
// Tile of a launch-local tile index: the launch's group of runs (tiles
// [tb, tb + tn)) or the whole pool (tn = 0: up to the tile high-water mark)
__device__ __forceinline__ int tile_in_range(const Pool& P, int b)
{
    if (P.tn > 0) return b < P.tn ? P.tb + b : -1;
    return b < *(const DUCT_CAS int*)P.ntile ? b : -1;
}

__device__ __forceinline__ int tile_of_block(const Pool& P, int xcd_map)
{
    int b = blockIdx.x;
    if ((xcd_map & XCD_REVERSE) && (gridDim.x & 7) == 0) {
        // each XCD takes the same tiles in reverse order: a launch that
        // follows one in forward order starts on the tiles the other wrote last
        b = (((int)(gridDim.x >> 3) - 1 - (b >> 3)) << 3) | (b & 7);
    }
    xcd_map &= 7;
    if (xcd_map == 4) {
        // slabs: each XCD takes the same eighth of every x-plane's tiles (the
        // plane layout of the run in slot 0), plane after plane, so a tile's
        // x neighbours are read through the XCD's own L2; planes whose tile
        // count is not a multiple of 8 fall back to contiguous ranges
        const int tpp = *(const DUCT_CAS int*)&P.desc[0].tiles_j * *(const DUCT_CAS int*)&P.desc[0].tiles_k;
        if (tpp > 0 && (tpp & 7) == 0) {
            const int w = tpp >> 3, l = b >> 3;
            return tile_in_range(P, (l / w) * tpp + (b & 7) * w + l % w);
        }
        xcd_map = 1;
    }
    if (xcd_map == 1 && (gridDim.x & 7) == 0) {
        b = (b & 7) * (int)(gridDim.x >> 3) + (b >> 3);
    } else if (xcd_map >= 2) {
        // chunks of C consecutive tiles dealt round-robin over the XCDs (the
        // grid is a multiple of 8 C): the XCDs sweep the tile list together,
        // each keeping the y/z neighbours within its chunks in its own L2
        const int C = xcd_chunk(xcd_map), m = b >> 3;
        b = (((m / C) << 3) + (b & 7)) * C + m % C;
    }
    return tile_in_range(P, b);
}

// Loads through the constant address space: run descriptors, the tile list
// and the run state are not written while a stencil kernel runs, so loads
// with a block-uniform address become scalar loads into SGPRs.
template <typename T>
__device__ __forceinline__ T ldc(const T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const DUCT_CAS T*)p;
#else
    return *p;      // host pass of a device function: never executed
#endif
}

// A run-local array at a block-uniform base (an arena pointer plus the run's
// cell offset).  Element i is addressed through a 32-bit byte offset, so a
// load or store takes the SGPR-base + 32-bit VGPR-offset form instead of a
// 64-bit address per access; admission keeps one run's field below 4 GiB.
template <typename T>
struct RunArr {
    T* p;
    __device__ __forceinline__ T& operator[](int i) const
    {
        using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
        return *(T*)((B*)p + (unsigned)i * (unsigned)sizeof(T));
    }
};

template <typename T>
__device__ __forceinline__ RunArr<T> run_arr(T* p) { return RunArr<T>{p}; }

// A run-local array that the kernel only reads, behind a buffer resource (base
// and size in SGPRs): element i is one buffer load at a 32-bit byte offset,
// constant neighbour offsets fold into the instruction, and a read outside
// the run's n elements returns 0.  Only for arrays no thread of the kernel
// writes (buffer loads are not ordered against the kernel's global stores).
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
struct RunBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ T operator[](int i) const
    {
        static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit elements");
        const int off = (int)((unsigned)i * (unsigned)sizeof(T));
        if constexpr (sizeof(T) == 8)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    }
    // element at a byte offset (a multiple of sizeof(T))
    __device__ __forceinline__ T at(int boff) const
    {
        if constexpr (sizeof(T) == 8)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, boff, 0, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, boff, 0, 0));
    }
};

// 0x00020000: the gfx9-family buffer descriptor word 3 for raw 32-bit data
template <typename T>
__device__ __forceinline__ RunBuf<T> run_buf(const T* p, long long n)
{
    RunBuf<T> b;
    b.r = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (long long)((unsigned long long)n * sizeof(T)),
                                            0x00020000);
    return b;
}
#else
template <typename T>
struct RunBuf {                                      // host pass only, never executed
    const T* p;
    __device__ T operator[](int i) const { return p[i]; }
    __device__ T at(int boff) const { return *(const T*)((const char*)p + boff); }
};
template <typename T>
__device__ RunBuf<T> run_buf(const T* p, long long) { return RunBuf<T>{p}; }
#endif

// The parts of a run descriptor a stencil kernel uses, held in SGPRs.
struct SlotView {
    GridDims g;
    DuctConsts k;
    long long off;
};

__device__ __forceinline__ SlotView slot_view(const Pool& P, int s)
{
    const SlotDesc* D = P.desc + s;
    SlotView v;
    v.g = ldc(&D->g);
    v.k = ldc(&D->k);
    v.off = ldc(&D->off);
    return v;
}

struct TilePos {
    int s, i, j, kk;
    bool in;
};

// Decodes tile t for this thread; returns false (uniformly for the block) if
// the tile is free or its run is inactive or finished.  TK > 0: a tile is
// TK cells in k by 256/TK rows in j of one x-plane; TK == 0 (flat): 256
// consecutive cells of one x-plane (no idle lanes when nz is not a multiple
// of the tile width).
template <int TK>
__device__ __forceinline__ bool tile_decode(const Pool& P, int t, TilePos& tp)
{
    const int s = ldc(P.tile_slot + t);
    if (s < 0) return false;
    const SlotDesc* D = P.desc + s;
    if (!ldc(&D->active) || ldc(&P.dyn[s].done)) return false;
    const int lt = t - ldc(&D->tile0);
    const int tiles_k = ldc(&D->tiles_k);
    const int ny = ldc(&D->g.ny), nz = ldc(&D->g.nz);
    tp.s = s;
    if constexpr (TK == 0) {
        tp.i = lt / tiles_k;
        const int p = (lt - tp.i * tiles_k) * 256 + (int)threadIdx.x;
        tp.in = p < ny * nz;
        tp.j = p / nz;
        tp.kk = p - tp.j * nz;
    } else {
        constexpr int TJ = 256 / TK;
        const int tiles_j = ldc(&D->tiles_j);
        const int kb = lt % tiles_k;
        const int rest = lt / tiles_k;
        const int jb = rest % tiles_j;
        tp.i = rest / tiles_j;
        tp.kk = kb * TK + (int)(threadIdx.x % TK);
        tp.j = jb * TJ + (int)(threadIdx.x / TK);
        tp.in = tp.kk < nz && tp.j < ny;
    }
    return true;
}

__device__ __forceinline__ int line_idx(const GridDims& g, int bi, int j, int kk, int d, int o)
{
    if (d == 0) return bi + o * (int)g.sx + j * g.nz + kk;
    if (d == 1) return bi + wrapi(j + o, g.ny) * g.nz + kk;
    return bi + j * g.nz + wrapi(kk + o, g.nz);
}

// Fixed-order sum of one value per thread of a 256-thread block: butterfly
// within each wave (lane 0's association order is fixed), then the four wave
// sums in order; every thread gets the total.
__device__ __forceinline__ double block_sum256(double v, double* red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void block_max2_256(double& a, double& b, double* red)
{
    for (int o = 32; o > 0; o >>= 1) { a = fmax(a, __shfl_xor(a, o)); b = fmax(b, __shfl_xor(b, o)); }
    if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = a; red[4 + (threadIdx.x >> 6)] = b; }
    __syncthreads();
    a = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    b = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
    __syncthreads();
}

// ---------------------------------------------------------------------------
// K1: primitives + box filter (:1307-1470 structure), per-tile sums of rho
// ---------------------------------------------------------------------------
// Row tiles stage the primitives of the tile plus one cell each side in j and
// k (wrapped) in LDS; the y and z neighbours of the filter come from there, the
// x neighbours from memory.  FP64 on rows of 64, 32 or 16 cells (its filter is
// held by the texture-address path: 16 cycles per 64-bit load; 128^3 step
// -2.3 %), FP32 on rows of 32 or 16 (on 64-wide rows its filter already moves
// bytes at the copy rate and the staging's extra rows cost 1 %).
template <int TK, int ES>
struct FiltStage {
    static constexpr bool on = (TK == 16 || TK == 32 || (TK == 64 && ES == 8));
    static constexpr int TJ = on ? 256 / TK : 1, SK = on ? TK + 2 : 1, SN = on ? (TJ + 2) * SK : 1;
};

template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_filter(Pool P, int qin_is_B, int xcd_map)
{
    using FS = FiltStage<TK, (int)sizeof(real)>;
    __shared__ double red[256];
    __shared__ real sp[FS::on ? 5 : 1][FS::SN];
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        const SlotView D = slot_view(P, tp.s);
        const GridDims g = D.g;
        const DuctConsts& k = D.k;
        void* const* Q = qin_is_B ? P.B : P.A;
        const auto q0 = run_buf((const real*)Q[0] + D.off, g.n);
        const auto q1 = run_buf((const real*)Q[1] + D.off, g.n);
        const auto q2 = run_buf((const real*)Q[2] + D.off, g.n);
        const auto q3 = run_buf((const real*)Q[3] + D.off, g.n);
        const auto q4 = run_buf((const real*)Q[4] + D.off, g.n);
        const real rut = (real)k.rho_u_Tu, irut = rcp_const<real>(k, DUCT_RC_IRUT);
        const bool endplane = (tp.i == 0 || tp.i == g.nx - 1);      // uniform over the tile
        if constexpr (FS::on) {
            if (!endplane) {
                // the tile's neighbourhood (a row or column beyond one wrap only
                // occurs past the grid's edge, for cells outside the run, and reads
                // cell 0)
                constexpr int TKS = TK > 0 ? TK : 1;
                const int j0 = tp.j - (int)threadIdx.x / TKS, k0 = tp.kk - (int)threadIdx.x % TKS;
                const int bi = tp.i * (int)g.sx;
                for (int r = threadIdx.x; r < FS::SN; r += 256) {
                    const int jr = r / FS::SK, kr = r - jr * FS::SK;
                    int gj = wrapi(j0 + jr - 1, g.ny), gk = wrapi(k0 + kr - 1, g.nz);
                    gj = (unsigned)gj < (unsigned)g.ny ? gj : 0;
                    gk = (unsigned)gk < (unsigned)g.nz ? gk : 0;
                    const int c = bi + gj * g.nz + gk;
                    const real T = q4[c], ri = T * irut;          // 1/rho
                    sp[0][r] = q0[c] * ri; sp[1][r] = q1[c] * ri; sp[2][r] = q2[c] * ri;
                    sp[3][r] = T; sp[4][r] = q3[c];
                }
                __syncthreads();
            }
        }
        double rsum = 0.0;
        if (tp.in) {
            const int i = tp.i, j = tp.j, kk = tp.kk;
            const int q = (int)lidx(g, i, j, kk);
            real fu, fv, fw, fT, fp;
            if (endplane) {
                real T = q4[q], ri = T * irut;          // 1/rho
                fu = q0[q] * ri; fv = q1[q] * ri; fw = q2[q] * ri; fT = T; fp = q3[q];
            } else {
                // neighbour order (-x, -y, -z, centre, +z, +y, +x) as :1329-1331
                real pu[7], pv[7], pw[7], pT[7], pp[7];
                if constexpr (FS::on) {
                    constexpr int TKS = TK > 0 ? TK : 1;
                    const int lb = ((int)threadIdx.x / TKS + 1) * FS::SK + (int)threadIdx.x % TKS + 1;
                    const int lo[5] = {lb - FS::SK, lb - 1, lb, lb + 1, lb + FS::SK};
#pragma unroll
                    for (int u = 0; u < 5; ++u) {
                        pu[u + 1] = sp[0][lo[u]]; pv[u + 1] = sp[1][lo[u]]; pw[u + 1] = sp[2][lo[u]];
                        pT[u + 1] = sp[3][lo[u]]; pp[u + 1] = sp[4][lo[u]];
                    }
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const int r = u == 0 ? q - (int)g.sx : q + (int)g.sx;
                        const int o = u == 0 ? 0 : 6;
                        real T = q4[r], ri = T * irut;
                        pu[o] = q0[r] * ri; pv[o] = q1[r] * ri; pw[o] = q2[r] * ri;
                        pT[o] = T; pp[o] = q3[r];
                    }
                } else {
                    int r[7];
                    r[0] = q - (int)g.sx;
                    r[1] = (int)lidx(g, i, wrapi(j - 1, g.ny), kk);
                    r[2] = (int)lidx(g, i, j, wrapi(kk - 1, g.nz));
                    r[3] = q;
                    r[4] = (int)lidx(g, i, j, wrapi(kk + 1, g.nz));
                    r[5] = (int)lidx(g, i, wrapi(j + 1, g.ny), kk);
                    r[6] = q + (int)g.sx;
#pragma unroll
                    for (int u = 0; u < 7; ++u) {
                        real T = q4[r[u]], ri = T * irut;
                        pu[u] = q0[r[u]] * ri; pv[u] = q1[r[u]] * ri; pw[u] = q2[r[u]] * ri;
                        pT[u] = T; pp[u] = q3[r[u]];
                    }
                }
                const real wc = (real)k.wc, iden = rcp_const<real>(k, DUCT_RC_IDEN);
                fu = (pu[0] + pu[1] + pu[2] + wc * pu[3] + pu[4] + pu[5] + pu[6]) * iden;
                fv = (pv[0] + pv[1] + pv[2] + wc * pv[3] + pv[4] + pv[5] + pv[6]) * iden;
                fw = (pw[0] + pw[1] + pw[2] + wc * pw[3] + pw[4] + pw[5] + pw[6]) * iden;
                fT = (pT[0] + pT[1] + pT[2] + wc * pT[3] + pT[4] + pT[5] + pT[6]) * iden;
                fp = (pp[0] + pp[1] + pp[2] + wc * pp[3] + pp[4] + pp[5] + pp[6]) * iden;
            }
            const auto F0 = run_arr((real*)P.F[0] + D.off);
            const auto F1 = run_arr((real*)P.F[1] + D.off);
            const auto F2 = run_arr((real*)P.F[2] + D.off);
            const auto F3 = run_arr((real*)P.F[3] + D.off);
            const auto F4 = run_arr((real*)P.F[4] + D.off);
            F0[q] = fu; F1[q] = fv; F2[q] = fw; F3[q] = fT; F4[q] = fp;
            rsum = (double)(rut / fT);
        }
        double tot = block_sum256(rsum, red);
        if (threadIdx.x == 0) P.rho_part[t] = tot;
    }
}

// ---------------------------------------------------------------------------
// K2: properties, nu_t (:1569-1633), S_T (:1913-1965); per-tile DT maxima
// ---------------------------------------------------------------------------
// Properties (tables), nu_t (:1569-1633) and the SSR flame speed
// (:1913-1965) of one interior cell from its filtered T and phi and the
// central differences of the filtered velocities; mn/ma receive the cell's
// nu and alpha for the diffusive DT limit.
template <typename real>
__device__ __forceinline__ void sgs_cell(const Pool& P, const DuctConsts& k, const GridDims& g, real T, real pfq,
                                         real du_dx, real dv_dy, real dw_dz, real du_dy, real du_dz, real dv_dx,
                                         real dv_dz, real dw_dx, real dw_dy, real& nut_out, real& st_out, double& mn,
                                         double& ma)
{
    const real SL = (real)k.s_l;
    const real DX = (real)g.dx;
    // nu = mu / rho with rho = rho_u T_u / T, and alpha = lambda / (rho cp) with
    // lambda = mu cp / Pr, so alpha = nu / Pr: the cp table and the divisions
    // cancel (the same quantities up to rounding)
    real mu = table_lookup<real>(P.mu_tab, k, T);
    real nu = mu * T * rcp_const<real>(k, DUCT_RC_IRUT);
    real al = nu * rcp_const<real>(k, DUCT_RC_IPRAN);
    mn = (double)nu; ma = (double)al;
    real S12 = (real)0.5 * (du_dy + dv_dx), S13 = (real)0.5 * (du_dz + dw_dx), S23 = (real)0.5 * (dv_dz + dw_dy);
    real ss = (real)2.0 * (S12 * S12 + S13 * S13 + S23 * S23) + du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz;
    real smag = sqrt((real)2.0 * ss);
    real div = du_dx + dv_dy + dw_dz;
    real gg = progress<real>(pfq, rcp_const<real>(k, DUCT_RC_HID));
    real cs = (real)k.cs_u + ((real)k.cs_b - (real)k.cs_u) * gg;
    cs += (real)k.cs_f * (real)4.0 * gg * ((real)1.0 - gg);
    nut_out = (cs * DX) * (cs * DX) * smag;
    // SSR-SGS flame speed
    real num = ss - (div * div) * ((real)1.0 / (real)3.0);
    real X = M<real>::div((real)k.dx4, nu * nu) * num;
    real d_eta;
    if (X < (real)LES_SMALL) d_eta = (real)1.0;
    else {
        real px = M<real>::pow34(X);
        real alg = (real)0.2187 * px;
        d_eta = (real)0.2340 * px * (alg > (real)1.0 ? M<real>::log(alg) : (real)0.0);
    }
    const real dd = d_eta + (real)LES_SMALL;   // the Kolmogorov scale eta = dx / dd
    const real dF = rcp_const<real>(k, DUCT_RC_DF);
    // wrinkling (asf eta / dx)^(2 - D3) with asf = 8 exp(10 dF / (8 eta + small)),
    // taken as one exponential: exp((2 - D3) (ln 8 + 10 dF / (8 eta + small) + ln(eta / dx)))
    // with 10 dF / (8 eta + small) = 10 dF dd / (8 dx + small dd) and ln(eta / dx) = -ln dd
    const real lg = (real)2.0794415416798357 +
                    M<real>::div((real)10.0 * dF * dd, (real)8.0 * DX + (real)LES_SMALL * dd) - M<real>::log(dd);
    real wr = M<real>::exp(((real)2.0 - (real)k.d3) * lg);
    real te = DX * fmax((real)0.0, div) * rcp_const<real>(k, DUCT_RC_ITSL);
    st_out = fmax(SL, SL * fmax((real)0.0, wr + te));
}

template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_sgs(Pool P, int stage, int xcd_map)
{
    __shared__ double red[8];
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        const SlotView D = slot_view(P, tp.s);
        const GridDims g = D.g;
        const DuctConsts& k = D.k;
        double mn = 0.0, ma = 0.0;
        if (tp.in) {
            const int i = tp.i, j = tp.j, kk = tp.kk;
            const int q = (int)lidx(g, i, j, kk);
            const auto uf = run_buf((const real*)P.F[0] + D.off, g.n);
            const auto vf = run_buf((const real*)P.F[1] + D.off, g.n);
            const auto wf = run_buf((const real*)P.F[2] + D.off, g.n);
            const auto Tf = run_buf((const real*)P.F[3] + D.off, g.n);
            const auto pf = run_buf((const real*)P.F[4] + D.off, g.n);
            const auto nut = run_arr((real*)P.nut + D.off);
            const auto stv = run_arr((real*)P.st + D.off);
            const real SL = (real)k.s_l;
            if (i == 0 || i == g.nx - 1) {
                nut[q] = (real)0.0; stv[q] = SL;
            } else {
                const int xp = q + (int)g.sx, xm = q - (int)g.sx;
                const int yp = (int)lidx(g, i, wrapi(j + 1, g.ny), kk), ym = (int)lidx(g, i, wrapi(j - 1, g.ny), kk);
                const int zp = (int)lidx(g, i, j, wrapi(kk + 1, g.nz)), zm = (int)lidx(g, i, j, wrapi(kk - 1, g.nz));
                const real hx = rcp_const<real>(D.k, DUCT_RC_HID), hy = rcp_const<real>(D.k, DUCT_RC_HID + 1), hz = rcp_const<real>(D.k, DUCT_RC_HID + 2);
                real nt, stq;
                sgs_cell<real>(P, k, g, Tf[q], pf[q],
                               (uf[xp] - uf[xm]) * hx, (vf[yp] - vf[ym]) * hy, (wf[zp] - wf[zm]) * hz,
                               (uf[yp] - uf[ym]) * hy, (uf[zp] - uf[zm]) * hz, (vf[xp] - vf[xm]) * hx,
                               (vf[zp] - vf[zm]) * hz, (wf[xp] - wf[xm]) * hx, (wf[yp] - wf[ym]) * hy, nt, stq, mn, ma);
                nut[q] = nt;
                stv[q] = stq;
            }
        }
        block_max2_256(mn, ma, red);
        if (threadIdx.x == 0) {
            double* dp = P.dpart + 2 * ((long long)stage * P.tile_cap + t);
            dp[0] = mn;
            dp[1] = ma;
        }
        // the run's first tile also reduces the run's per-tile sums of rho
        // (written by the filter) into its domain-mean density, in a fixed
        // order (strided per thread, then the block's fixed-order sum); the
        // momentum kernel of this stage reads it
        if (t == ldc(&P.desc[tp.s].tile0)) {
            const int nt = ldc(&P.desc[tp.s].ntiles);
            double v = 0.0;
            for (int b = threadIdx.x; b < nt; b += 256) v += P.rho_part[t + b];
            const double tot = block_sum256(v, red);
            if (threadIdx.x == 0) P.dyn[tp.s].rho_avg = tot / (double)g.n;
        }
    }
}

// ---------------------------------------------------------------------------
// K3: RHS fused with the SSP-RK3 combine
// ---------------------------------------------------------------------------
template <typename real>
struct RhsArgs {
    int qin_is_B;        // stage input: B (stages 2, 3) or A (stage 1)
    int qout_is_B;       // stage output: B (stages 1, 2) or A (stage 3)
    int stage;
    double ca, cb;       // q_out = ca*q0 + cb*(q_in + dt*L) for stages 2, 3
    int track_vel;       // last stage: per-tile max |u| of the new state
    int xcd_map;
    real* Lout[DNF];     // optional: the RHS itself, slot-local indices (testing hook)
};

// The RHS of all five fields of one RK stage fused with the SSP-RK3 combine,
// one launch:
//   rho*u, rho*v, rho*w (:2552-2691 structure): conservative WENO5/minmod face
//     fluxes with face-averaged rho and velocity, mu_eff Laplacian, buoyancy
//     (rho - rho_avg) g, flame force in the progress band;
//   phi (G equation, :2706-2818) and T (:2912-2955): upwind convection by the
//     cell's velocity, SGS fluxes with face-averaged nu_t, S_T |grad phi|,
//     artificial diffusion in the flame band, conduction, nudging to the
//     G-weighted temperature;
// and, in the last stage, the per-tile max |u| of the new state.  Per cell the
// quantities the equations share are formed once: each direction's line
// offsets, the cell's velocity along it, the neighbours' T, rho and nu_t, the
// progress variable, the property tables; the direction loop stays rolled (one
// copy of its code and registers) and inside it the seven-point lines of u, v,
// w, phi and T are loaded one ahead of the reconstruction that uses them.
// FP64: 4 waves per SIMD (125 VGPRs); FP32: 6 waves (80 VGPRs; 7 spill).
// Sums of the RHS over the three directions: momentum convection (cv0..cv2);
// phi and T: SGS flux minus convection (ap, aT) and Laplacian (lp, lT);
// |grad phi|^2 (gg).
template <typename real>
struct RhsSums {
    real cv0, cv1, cv2, ap, aT, lp, lT, gg;
};

// One direction's terms of the five RHS from the seven-point lines through the
// cell.  ld(f, o) is field f (0..2 velocities u, v, w; 3 phi; 4 T; 5 nu_t) at
// offset o - 3 along the direction; vd is the direction's velocity field
// (0..2).  Line m + 1 is loaded while line m is reconstructed.  first: the
// first direction (starts the sums); band: first-order faces (the x band next
// to the inflow and outflow planes).
template <typename real, bool PAIR, typename LD>
__device__ __forceinline__ void rhs_line(RhsSums<real>& A, const LD& ld, int vd, bool first, bool band, int scheme,
                                         real rut, real rf0, real nu0, real isc, real ipr, real id, real id2,
                                         real hid)
{
    const real u0 = ld(vd, 3);
    const real up = (real)0.5 * (ld(vd, 4) + u0);
    const real um = (real)0.5 * (u0 + ld(vd, 2));
    // face densities from the neighbours' rho = rho_u T_u / T, the two
    // neighbours over one division: rho_u T_u / (T+ T-) times T- and T+
    const real Tp = ld(4, 4), Tm = ld(4, 2);
    const real rq = M<real>::div(rut, Tp * Tm);
    const real rp = (real)0.5 * (rq * Tm + rf0);
    const real rm = (real)0.5 * (rf0 + rq * Tp);
    const real mp = rp * up, mm = rm * um;
    real ntp = (real)0, ntm = (real)0;     // nu_t at +-1, loaded with the phi line
    // FP64 WENO5 away from the end-plane band (PAIR: the LDS-staged kernels,
    // whose registers leave room for the held pair): the quotients of fields 0 and 1,
    // and of fields 2 and 3, share one division each (1/(a b) times b and a;
    // the denominators are products of the smoothness weights, far inside
    // the FP64 range even in pairs); (pn, pd) holds the first of a pair
    const bool pairdiv = PAIR && sizeof(real) == 8 && scheme == LES_ADV_WENO5 && !band;
    real pn = (real)0, pd = (real)1;
    real nxt[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) nxt[o] = ld(0, o);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        real v[7];
#pragma unroll
        for (int o = 0; o < 7; ++o) v[o] = nxt[o];
        if (m < 4) {
#pragma unroll
            for (int o = 0; o < 7; ++o) nxt[o] = ld(m + 1, o);
        }
        if (m == 2) { ntp = ld(5, 4); ntm = ld(5, 2); }
        if (m < 3) {
            real term;
            if constexpr (sizeof(real) == 8) {
                if (pairdiv) {
                    real n, d;
                    face_pair_nd(v, up >= (real)0.0, um >= (real)0.0, mp, mm, n, d);
                    if (m == 1) {
                        // fields 0 and 1 over one division
                        const real R = M<real>::div((real)1.0, pd * d);
                        const real t0 = pn * (d * R) * id;
                        A.cv0 = first ? t0 : A.cv0 + t0;
                        term = n * (pd * R) * id;
                    } else {
                        pn = n; pd = d;            // fields 0 and 2: completed by the next field
                        continue;
                    }
                }
            }
            if (!pairdiv) {
                if (band) {
                    const real fp = (up >= (real)0.0) ? v[3] : v[4];
                    const real fm = (um >= (real)0.0) ? v[2] : v[3];
                    term = (mp * fp - mm * fm) * id;
                } else {
                    term = face_pair<real>(scheme, v, up >= (real)0.0, um >= (real)0.0, mp, mm) * id;
                }
            }
            const real c = m == 0 ? A.cv0 : (m == 1 ? A.cv1 : A.cv2);
            const real nc = first ? term : c + term;
            A.cv0 = m == 0 ? nc : A.cv0;
            A.cv1 = m == 1 ? nc : A.cv1;
            A.cv2 = m == 2 ? nc : A.cv2;
        } else {
            // scalar phi (m = 3) or T (m = 4): upwind by the cell's velocity
            const real vm1 = v[2], v0 = v[3], vp1 = v[4];
            real tt;
            if (band) {
                tt = (u0 >= (real)0.0) ? u0 * (v0 - vm1) * id : u0 * (vp1 - v0) * id;
            } else {
                const bool pos = u0 >= (real)0.0;
                bool done = false;
                if constexpr (sizeof(real) == 8) {
                    if (pairdiv && m == 3) {
                        // field 2 (held) and phi over one division
                        real n, d;
                        face_diff_same_nd(v, pos, n, d);
                        const real R = M<real>::div((real)1.0, pd * d);
                        const real t2 = pn * (d * R) * id;
                        A.cv2 = first ? t2 : A.cv2 + t2;
                        tt = u0 * (n * (pd * R)) * id;
                        done = true;
                    }
                }
                if (!done) tt = u0 * face_diff_same<real>(scheme, v, pos) * id;
            }
            const real isc2 = m == 3 ? isc : ipr;
            const real gp = (ntp + nu0) * isc2, gm = (nu0 + ntm) * isc2;
            const real t2 = (gp * (vp1 - v0) - gm * (v0 - vm1)) * id2;
            const real t3 = ((vp1 - v0) - (v0 - vm1)) * id2;
            if (m == 3) {
                A.ap = first ? t2 - tt : A.ap + (t2 - tt);
                A.lp = first ? t3 : A.lp + t3;
                const real gd = (vp1 - vm1) * hid;
                A.gg = first ? gd * gd : A.gg + gd * gd;
            } else {
                A.aT = first ? t2 - tt : A.aT + (t2 - tt);
                A.lT = first ? t3 : A.lT + t3;
            }
        }
    }
}

// Row tiles of 64, 32 or 16 cells in k stage the y/z neighbourhood of the tile
// (the tile plus 3 cells each side in j and k, wrapped) of the filtered u, v,
// w, phi, T and nu_t in LDS: the y and z lines come from there, the x lines
// from memory.
template <int TK>
struct RhsStage {
    static constexpr bool on = (TK == 16 || TK == 32 || TK == 64);
    static constexpr int TJ = on ? 256 / TK : 1, SJ = TJ + 6, SK = on ? TK + 6 : 1, SN = on ? SJ * SK : 1;
    // occupancy hint per element size: FP64 4 waves per SIMD (64-wide rows are
    // held there by their 33.7 KB of LDS; compiled for 5 they schedule better:
    // 128^3 step 0.88 -> 0.84 ms, while 32- and 16-wide rows spill at 5), FP32 6
    // (80 VGPRs unstaged; the staged form needs 58, so runs 8 waves)
    static constexpr int waves(int esize) { return esize == 8 ? (TK == 64 ? 5 : 4) : 6; }
};

// The RHS of all five fields of one RK stage fused with the SSP-RK3 combine,
// one launch:
//   rho*u, rho*v, rho*w (:2552-2691 structure): conservative WENO5/minmod face
//     fluxes with face-averaged rho and velocity, mu_eff Laplacian, buoyancy
//     (rho - rho_avg) g, flame force in the progress band;
//   phi (G equation, :2706-2818) and T (:2912-2955): upwind convection by the
//     cell's velocity, SGS fluxes with face-averaged nu_t, S_T |grad phi|,
//     artificial diffusion in the flame band, conduction, nudging to the
//     G-weighted temperature;
// and, in the last stage, the per-tile max |u| of the new state.  Per cell the
// quantities the equations share are formed once: each direction's line
// offsets, the cell's velocity along it, the neighbours' T, rho and nu_t, the
// progress variable, the property tables.  Flat tiles and 128-wide rows walk
// the directions in a rolled loop over memory; 64-, 32- and 16-wide rows take
// their y and z lines from the LDS-staged neighbourhood (RhsStage).
// Occupancy: RhsStage::waves.
template <typename real, int TK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RhsStage<TK>::waves(sizeof(real)))))
k_duct_rhs_all(Pool P, RhsArgs<real> a)
{
    using ST = RhsStage<TK>;
    __shared__ double vred[8];
    __shared__ real sF[ST::on ? 6 : 1][ST::SN];
    const int t = tile_of_block(P, a.xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        const SlotView D = slot_view(P, tp.s);
        const GridDims g = D.g;
        const DuctConsts& k = D.k;
        const bool endplane = (tp.i == 0 || tp.i == g.nx - 1);      // uniform over the tile
        if constexpr (ST::on) {
            if (!endplane) {
                // the block's tile origin, then its neighbourhood: SJ x SK cells of
                // each of the six fields, rows and columns wrapped (a row or column
                // beyond one wrap only occurs past the grid's edge, for cells outside
                // the run, and reads cell 0)
                const int j0 = tp.j - (int)threadIdx.x / TK, k0 = tp.kk - (int)threadIdx.x % TK;
                const int bi = tp.i * (int)g.sx;
                const auto b0 = run_buf((const real*)P.F[0] + D.off, g.n), b1 = run_buf((const real*)P.F[1] + D.off, g.n);
                const auto b2 = run_buf((const real*)P.F[2] + D.off, g.n), b3 = run_buf((const real*)P.F[4] + D.off, g.n);
                const auto b4 = run_buf((const real*)P.F[3] + D.off, g.n), b5 = run_buf((const real*)P.nut + D.off, g.n);
                for (int r = threadIdx.x; r < ST::SN; r += 256) {
                    const int jr = r / ST::SK, kr = r - jr * ST::SK;
                    int gj = wrapi(j0 + jr - 3, g.ny), gk = wrapi(k0 + kr - 3, g.nz);
                    gj = (unsigned)gj < (unsigned)g.ny ? gj : 0;
                    gk = (unsigned)gk < (unsigned)g.nz ? gk : 0;
                    const int c = bi + gj * g.nz + gk;
                    sF[0][r] = b0[c]; sF[1][r] = b1[c]; sF[2][r] = b2[c];
                    sF[3][r] = b3[c]; sF[4][r] = b4[c]; sF[5][r] = b5[c];
                }
                __syncthreads();
            }
        }
        double vmax = 0.0;
        if (tp.in) {
            const int i = tp.i, j = tp.j, kk = tp.kk;
            const int bi = i * (int)g.sx;
            const int q = bi + j * g.nz + kk;
            void* const* QI = a.qin_is_B ? P.B : P.A;
            void* const* QO = a.qout_is_B ? P.B : P.A;
            if (endplane) {
#pragma unroll
                for (int f = 0; f < DNF; ++f) run_arr((real*)QO[f] + D.off)[q] = run_arr((const real*)QI[f] + D.off)[q];
            } else {
                const auto uf = run_buf((const real*)P.F[0] + D.off, g.n);
                const auto vf = run_buf((const real*)P.F[1] + D.off, g.n);
                const auto wf = run_buf((const real*)P.F[2] + D.off, g.n);
                const auto Tf = run_buf((const real*)P.F[3] + D.off, g.n);
                const auto pf = run_buf((const real*)P.F[4] + D.off, g.n);
                const auto nut = run_buf((const real*)P.nut + D.off, g.n);
                const real rut = (real)k.rho_u_Tu;
                const bool xband = (i < 3 || i > g.nx - 4);
                const int scheme = k.scheme;
                const RunBuf<real> vel[3] = {uf, vf, wf};
                const RunBuf<real> fld[6] = {uf, vf, wf, pf, Tf, nut};
                const real T0 = Tf[q];
                const real rf0 = M<real>::div(rut, T0);
                const real nu0 = nut[q];
                const real isc = rcp_const<real>(k, DUCT_RC_ISC), ipr = rcp_const<real>(k, DUCT_RC_IPR);
                RhsSums<real> S;
                S.cv0 = S.cv1 = S.cv2 = S.ap = S.aT = S.lp = S.lT = S.gg = (real)0;
                // the cell in the staged neighbourhood (row tiles of 64, 32 or 16)
                constexpr int TKS = TK > 0 ? TK : 1;
                const int lb = ST::on ? ((int)threadIdx.x / TKS + 3) * ST::SK + (int)threadIdx.x % TKS + 3 : 0;
                if constexpr (ST::on) {
                    // x lines from memory (the band next to the end planes reads
                    // offsets -1..1 only: there the outer offsets repeat them)
                    int lo[7];
#pragma unroll
                    for (int o = 0; o < 7; ++o) lo[o] = (q + (o - 3) * (int)g.sx) * (int)sizeof(real);
                    if (xband) {
                        lo[0] = lo[1] = lo[2];
                        lo[5] = lo[6] = lo[4];
                    }
                    auto ldx = [&](int f, int o) {
                        return (f == 0 ? fld[0] : (f == 1 ? fld[1] : (f == 2 ? fld[2] : (f == 3 ? fld[3] : (f == 4 ? fld[4] : fld[5]))))).at(lo[o]);
                    };
                    rhs_line<real, ST::on>(S, ldx, 0, true, xband, scheme, rut, rf0, nu0, isc, ipr,
                                   rcp_const<real>(k, DUCT_RC_ID), rcp_const<real>(k, DUCT_RC_ID2),
                                   rcp_const<real>(k, DUCT_RC_HID));
                    // y and z lines from LDS
#ifdef EXP_LDS_UNROLLED
                    auto ldy = [&](int f, int o) { return sF[f][lb + (o - 3) * ST::SK]; };
                    rhs_line<real, ST::on>(S, ldy, 1, false, false, scheme, rut, rf0, nu0, isc, ipr,
                                   rcp_const<real>(k, DUCT_RC_ID + 1), rcp_const<real>(k, DUCT_RC_ID2 + 1),
                                   rcp_const<real>(k, DUCT_RC_HID + 1));
                    __builtin_amdgcn_sched_barrier(0);
                    auto ldz = [&](int f, int o) { return sF[f][lb + (o - 3)]; };
                    rhs_line<real, ST::on>(S, ldz, 2, false, false, scheme, rut, rf0, nu0, isc, ipr,
                                   rcp_const<real>(k, DUCT_RC_ID + 2), rcp_const<real>(k, DUCT_RC_ID2 + 2),
                                   rcp_const<real>(k, DUCT_RC_HID + 2));
#else
#pragma unroll 1
                    for (int d = 1; d < 3; ++d) {
                        const int st = d == 1 ? ST::SK : 1;
                        const int l0 = lb - 3 * st;
                        auto ldl = [&](int f, int o) { return sF[f == 6 ? d : f][l0 + o * st]; };
                        rhs_line<real, ST::on>(S, ldl, 6, false, false, scheme, rut, rf0, nu0, isc, ipr,
                                       rcp_dir<real>(k, DUCT_RC_ID, d), rcp_dir<real>(k, DUCT_RC_ID2, d),
                                       rcp_dir<real>(k, DUCT_RC_HID, d));
                    }
#endif
                } else {
                    // the direction loop stays rolled (one copy of its code and
                    // registers); each direction's line offsets are selected once
#pragma unroll 1
                    for (int d = 0; d < 3; ++d) {
                        int lo[7];
#pragma unroll
                        for (int o = 0; o < 7; ++o) {
                            lo[o] = line_idx(g, bi, j, kk, d, o - 3) * (int)sizeof(real);
                            asm volatile("" : "+v"(lo[o]));
                        }
                        const bool band = (d == 0 && xband);
                        if (band) {
                            lo[0] = lo[1] = lo[2];
                            lo[5] = lo[6] = lo[4];
                        }
                        const RunBuf<real> U = d == 0 ? vel[0] : (d == 1 ? vel[1] : vel[2]);
                        auto ld = [&](int f, int o) {
                            // f == vd (the direction's velocity) is read through U
                            return (f == 6 ? U : (f == 0 ? fld[0] : (f == 1 ? fld[1] : (f == 2 ? fld[2] : (f == 3 ? fld[3] : (f == 4 ? fld[4] : fld[5])))))).at(lo[o]);
                        };
                        rhs_line<real, false>(S, ld, 6, d == 0, band, scheme, rut, rf0, nu0, isc, ipr,
                                       rcp_dir<real>(k, DUCT_RC_ID, d), rcp_dir<real>(k, DUCT_RC_ID2, d),
                                       rcp_dir<real>(k, DUCT_RC_HID, d));
                    }
                }
                const real cv0 = S.cv0, cv1 = S.cv1, cv2 = S.cv2, ap = S.ap, aT = S.aT, lp = S.lp, lT = S.lT, gg = S.gg;
                const real conv[3] = {cv0, cv1, cv2};
                const real mu = table_lookup<real>(P.mu_tab, k, T0);
                const real mu_eff = mu + rf0 * nu0;
                const int xp = q + (int)g.sx, xm = q - (int)g.sx;
                const int yp = line_idx(g, bi, j, kk, 1, 1), ym = line_idx(g, bi, j, kk, 1, -1);
                const int zp = line_idx(g, bi, j, kk, 2, 1), zm = line_idx(g, bi, j, kk, 2, -1);
                const real ix2 = rcp_const<real>(k, DUCT_RC_ID2), iy2 = rcp_const<real>(k, DUCT_RC_ID2 + 1), iz2 = rcp_const<real>(k, DUCT_RC_ID2 + 2);
                real df0 = (real)0, df1 = (real)0, df2 = (real)0;
                // field f (0..2 velocities, 3 phi) at the cell (n = 0) or at its
                // +-y (n = 1, 2) and +-z (n = 3, 4) neighbours: from the staged
                // neighbourhood where there is one
                auto yz = [&](int f, int n) -> real {
                    if constexpr (ST::on) {
                        const int o = n == 0 ? 0 : (n == 1 ? ST::SK : (n == 2 ? -ST::SK : (n == 3 ? 1 : -1)));
                        return sF[f][lb + o];
                    } else {
                        const int c = n == 0 ? q : (n == 1 ? yp : (n == 2 ? ym : (n == 3 ? zp : zm)));
                        return fld[f][c];
                    }
                };
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const RunBuf<real> U = m == 0 ? vel[0] : (m == 1 ? vel[1] : vel[2]);
                    const real uc = yz(m, 0);
                    // second differences as differences of neighbour differences
                    const real lx = ((U[xp] - uc) - (uc - U[xm])) * ix2;
                    const real ly = ((yz(m, 1) - uc) - (uc - yz(m, 2))) * iy2;
                    const real lz = ((yz(m, 3) - uc) - (uc - yz(m, 4))) * iz2;
                    const real dm = mu_eff * (lx + ly + lz);
                    df0 = m == 0 ? dm : df0;
                    df1 = m == 1 ? dm : df1;
                    df2 = m == 2 ? dm : df2;
                }
                const real rho_avg = (real)ldc(&P.dyn[tp.s].rho_avg);
                const real grav = (rf0 - rho_avg) * (real)LES_G_GRAV;
                const real g0 = progress<real>(yz(3, 0), rcp_const<real>(k, DUCT_RC_HID));
                const real ctr = (real)k.accel_center, Fm = (real)k.max_force;
                // the flame window exp(-(c - c0)^2 / (2 w^2)), shared by the flame
                // force (0.01 < c < 0.99) and the artificial diffusion (0.1 < c < 0.99)
                real win = (real)0;
                if (g0 > (real)0.01 && g0 < (real)0.99) {
                    const real df = g0 - ctr;
                    win = M<real>::exp(-(df * df) * rcp_const<real>(k, DUCT_RC_IW2));
                }
                real fx = (real)0, fy = (real)0, fz = (real)0;
                if (g0 > (real)0.01 && g0 < (real)0.99) {
                    real gx = (real)0, gy = (real)0, gz = (real)0;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const real pa = d == 0 ? pf[xp] : yz(3, 2 * d - 1), pb = d == 0 ? pf[xm] : yz(3, 2 * d);
                        const real gd = (progress<real>(pa, rcp_const<real>(k, DUCT_RC_HID)) -
                                         progress<real>(pb, rcp_const<real>(k, DUCT_RC_HID))) * rcp_dir<real>(k, DUCT_RC_HID, d);
                        gx = d == 0 ? gd : gx;
                        gy = d == 1 ? gd : gy;
                        gz = d == 2 ? gd : gz;
                    }
                    const real common = (real)k.accel_strength * rf0 * win;
                    fx = common * gx * (real)k.weight_x;
                    fy = common * gy;
                    fz = common * gz;
                }
                fx = fmax(-Fm, fmin(Fm, fx));
                fy = fmax(-Fm, fmin(Fm, fy));
                fz = fmax(-Fm, fmin(Fm, fz));
                real L[DNF];
                L[0] = -conv[0] + df0 + grav + fx;
                L[1] = -conv[1] + df1 + fy;
                L[2] = -conv[2] + df2 + fz;
                // phi: S_T |grad phi|, SGS flux, artificial diffusion in the flame band
                real art = (real)0;
                if (g0 > (real)0.1 && g0 < (real)0.99) art = (real)5.0 * (nu0 + (real)1.0e-5) * win * lp;
                const auto stv = run_buf((const real*)P.st + D.off, g.n);
                L[3] = ap + stv[q] * sqrt(gg) + art;
                // T: conduction with alpha = lambda / (rho cp), lambda = mu cp / Pr, so
                // alpha = mu T / (rho_u T_u Pr) (no cp table, no division), SGS flux, nudging
                const real al = mu * T0 * rcp_const<real>(k, DUCT_RC_IRUT) * rcp_const<real>(k, DUCT_RC_IPRAN);
                const real Tu = (real)k.Tu, Tb = (real)k.Tb;
                const real Ttg = Tu + g0 * (Tb - Tu);
                L[4] = aT + al * lT + (real)k.relaxation * (Ttg - T0);
                if (a.Lout[0]) {
#pragma unroll
                    for (int f = 0; f < DNF; ++f) a.Lout[f][q] = L[f];
                }
                const real DT = (real)ldc(&P.dyn[tp.s].dt);
                real out[DNF];
#pragma unroll
                for (int f = 0; f < DNF; ++f) {
                    const real y = run_arr((const real*)QI[f] + D.off)[q] + DT * L[f];
                    out[f] = a.stage == 0 ? y : (real)a.ca * run_arr((const real*)P.A[f] + D.off)[q] + (real)a.cb * y;
                }
                out[4] = fmax((real)200.0, fmin((real)3000.0, out[4]));
#pragma unroll
                for (int f = 0; f < DNF; ++f) run_arr((real*)QO[f] + D.off)[q] = out[f];
                if (a.track_vel) vmax = cell_speed<real>(out[0], out[1], out[2], out[4], k);
            }
        }
        if (a.track_vel) {
            double dummy = 0.0;
            block_max2_256(vmax, dummy, vred);
            if (threadIdx.x == 0) P.vpart[t] = vmax;
        }
    }
}

// ---------------------------------------------------------------------------
// DT per slot (:4399-4411 structure): the per-tile maxima of the previous step
// (max |u| of the new state, max nu / alpha over its 3 stages) and the
// inflow-plane maxima, reduced by nb blocks per slot (grid nb x slots), each a
// strided share of the tiles, four per thread and pass with their loads in
// flight together.  The maxima are exact in any order.  The slot's last block
// to count itself in takes the nb block maxima and sets the DT; it also resets
// the slot's arrival counter, so the counter is zero at every launch.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_duct_dt(Pool P, int nb)
{
    __shared__ double r[3][4];
    __shared__ int last;
    const int s = blockIdx.y, bk = blockIdx.x;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    double mv = 0.0, mn = 0.0, ma = 0.0;
    const int nt = D.ntiles;
    for (int b0 = bk * 1024; b0 < nt; b0 += nb * 1024) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int b = b0 + u * 256 + (int)threadIdx.x;
            if (b < nt) {
                const int t = D.tile0 + b;
                mv = fmax(mv, P.vpart[t]);
#pragma unroll
                for (int stg = 0; stg < 3; ++stg) {
                    const double* dp = P.dpart + 2 * ((long long)stg * P.tile_cap + t);
                    mn = fmax(mn, dp[0]);
                    ma = fmax(ma, dp[1]);
                }
            }
        }
    }
    const int nbc = P.ny_max * P.nzb;
    for (int b = bk * 256 + (int)threadIdx.x; b < nbc; b += nb * 256)
        mv = fmax(mv, P.bcpart[(long long)s * nbc + b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mv = fmax(mv, __shfl_xor(mv, o));
        mn = fmax(mn, __shfl_xor(mn, o));
        ma = fmax(ma, __shfl_xor(ma, o));
    }
    if ((threadIdx.x & 63) == 0) {
        r[0][threadIdx.x >> 6] = mv; r[1][threadIdx.x >> 6] = mn; r[2][threadIdx.x >> 6] = ma;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            mv = fmax(mv, r[0][w]); mn = fmax(mn, r[1][w]); ma = fmax(ma, r[2][w]);
        }
        double* bp = P.dtpart + 3 * ((long long)s * DT_BLOCKS_MAX + bk);
        bp[0] = mv; bp[1] = mn; bp[2] = ma;
        // the block maxima are released before the block counts itself in
        __threadfence();
        last = atomicAdd(&P.dtcnt[s], 1u) == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __threadfence();
    P.dtcnt[s] = 0u;
    for (int b = 0; b < nb; ++b) {
        const double* bp = P.dtpart + 3 * ((long long)s * DT_BLOCKS_MAX + b);
        mv = fmax(mv, bp[0]); mn = fmax(mn, bp[1]); ma = fmax(ma, bp[2]);
    }
    const DuctConsts& k = D.k;
    const double dx = D.g.dx;
    double maxvel = mv > LES_SMALL ? mv : k.U;
    double dtc = k.courant * dx / (maxvel + LES_SMALL);
    double dtn = k.courant_diff * (dx * dx) / (mn + LES_SMALL);
    double dta = k.courant_diff * (dx * dx) / (ma + LES_SMALL);
    P.dyn[s].dt = k.fixed_dt > 0 ? k.fixed_dt : fmin(dtc, fmin(dtn, dta));
    P.dyn[s].maxvel = maxvel;
}

// max |u| per tile of a state in A (after an upload or init)
template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_velmax(Pool P, int only_slot, int xcd_map)
{
    __shared__ double red[8];
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        if (tp.s != only_slot) return;
        const SlotView D = slot_view(P, tp.s);
        double vm = 0.0;
        if (tp.in && tp.i <= D.g.nx - 2) {
            const int q = (int)lidx(D.g, tp.i, tp.j, tp.kk);
            vm = cell_speed<real>(((const real*)P.A[0] + D.off)[q], ((const real*)P.A[1] + D.off)[q],
                                  ((const real*)P.A[2] + D.off)[q], ((const real*)P.A[4] + D.off)[q], D.k);
        }
        double dummy = 0.0;
        block_max2_256(vm, dummy, red);
        if (threadIdx.x == 0) {
            P.vpart[t] = vm;
            for (int stg = 0; stg < 3; ++stg) {
                double* dp = P.dpart + 2 * ((long long)stg * P.tile_cap + t);
                dp[0] = 0.0;
                dp[1] = 0.0;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// SEM: eddy advection/recycling, Gaussian plane, normalisation
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

__global__ void k_duct_eddy_init(Pool P, int s)
{
    const SlotDesc& D = P.desc[s];
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= D.k.nvort) return;
    eddy_draw_dev(D.k, P.eddy + 7 * (D.eoff + e), e, 0ull, true);
}

__global__ void k_duct_eddy_update(Pool P)
{
    const int s = blockIdx.y;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= D.k.nvort) return;
    double* E = P.eddy + 7 * (D.eoff + e);
    E[0] += D.k.U * P.dyn[s].dt;
    if (E[0] > D.k.L) {
        E[0] -= 2.0 * D.k.L;
        eddy_draw_dev(D.k, E, e, (unsigned long long)E[6] + 1ull, false);
    }
}

// One block per 16 x 16 patch of the inflow plane (rows j, columns k), eddy
// chunk of SEM_CHUNK eddies and run, into an FP64 partial plane per chunk
// (k_duct_sem_norm adds the chunks in order).  An eddy's Gaussian
// exp(-pi |d|^2 / L^2) is the product of its x, y and z factors, so the block
// takes the chunk's y factors for its 16 rows and z factors for its 16 columns
// (32 exponentials per eddy instead of 256, staged in LDS with each eddy's
// amplitudes times its x factor), and each cell adds
// (amplitude x x factor) x (y factor x z factor) over the chunk's eddies.
#define SEM_CHUNK 64
#define SEM_PATCH 16
__global__ void __launch_bounds__(256) k_duct_sem_plane(Pool P)
{
    __shared__ double fy[SEM_CHUNK][SEM_PATCH], fz[SEM_CHUNK][SEM_PATCH], amp[SEM_CHUNK][3];
    const int pk = (P.nz_max + SEM_PATCH - 1) / SEM_PATCH;
    const int npc = (P.ny_max + SEM_PATCH - 1) / SEM_PATCH * pk;
    const int s = blockIdx.y, sp = (int)blockIdx.x / npc, pc = (int)blockIdx.x - sp * npc;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    const GridDims& g = D.g;
    const DuctConsts& k = D.k;
    const int J0 = (pc / pk) * SEM_PATCH, K0 = (pc % pk) * SEM_PATCH;
    if (J0 >= g.ny || K0 >= g.nz) return;
    const long long PLM = (long long)P.ny_max * P.nz_max;
    double* out = P.sem_part + ((long long)s * P.sem_split + sp) * 3 * PLM;
    const int e0 = sp * SEM_CHUNK;
    const int cnt = min(SEM_CHUNK, k.nvort - e0);
    const int jj = (int)threadIdx.x / SEM_PATCH, kk = (int)threadIdx.x % SEM_PATCH;
    double su = 0.0, sv = 0.0, sw = 0.0;
    if (cnt > 0) {
        const double* eddy = P.eddy + 7 * (D.eoff + e0);
        const double inv_l2 = 1.0 / (k.L * k.L);
        for (int u = threadIdx.x; u < cnt * 2 * SEM_PATCH; u += blockDim.x) {
            const int e = u / (2 * SEM_PATCH), r = u % (2 * SEM_PATCH);
            const double* E = eddy + 7 * e;
            if (r < SEM_PATCH) {
                double ddy = (J0 + r) * g.dy - E[1];
                ddy -= k.ly * rint(ddy * k.rc[DUCT_RC_ILY]);     // nearest periodic image
                fy[e][r] = exp(-M_PI * (ddy * ddy) * inv_l2);
            } else {
                double ddz = (K0 + r - SEM_PATCH) * g.dz - E[2];
                ddz -= k.lz * rint(ddz * k.rc[DUCT_RC_ILZ]);
                fz[e][r - SEM_PATCH] = exp(-M_PI * (ddz * ddz) * inv_l2);
            }
        }
        for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
            const double* E = eddy + 7 * e;
            const double fx = exp(-M_PI * (E[0] * E[0]) * inv_l2);
            amp[e][0] = E[3] * fx; amp[e][1] = E[4] * fx; amp[e][2] = E[5] * fx;
        }
        __syncthreads();
        for (int e = 0; e < cnt; ++e) {
            const double w = fy[e][jj] * fz[e][kk];
            su += amp[e][0] * w; sv += amp[e][1] * w; sw += amp[e][2] * w;
        }
    }
    const int j = J0 + jj, kc = K0 + kk;
    if (j < g.ny && kc < g.nz) {
        const long long p = (long long)j * g.nz + kc;
        out[p] = su; out[PLM + p] = sv; out[2 * PLM + p] = sw;
    }
}

// fixed-order sum over a block of NT threads: butterfly within each wave,
// then the wave sums in order
template <int NT>
__device__ __forceinline__ double block_sum_nt(double v, double* red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r += red[w];
    __syncthreads();
    return r;
}

// chunk sums -> plane (fixed order), mean removal and rms rescale; one block
// of NT threads per (component, run) (1024 for large planes, 256 where a pool
// of many small runs launches many of them beside the stages); the plane sums
// in a fixed order (strided per thread, then the block's fixed-order sum)
template <int NT>
__global__ void __launch_bounds__(NT) k_duct_sem_norm(Pool P)
{
    __shared__ double red[NT / 64];
    const int s = blockIdx.y;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    const long long PL = (long long)D.g.ny * D.g.nz;
    const long long PLM = (long long)P.ny_max * P.nz_max;
    const double target = D.k.U * D.k.intensity;
    double* f = P.inlet + D.poff + blockIdx.x * PL;
    const double* part = P.sem_part + (long long)s * P.sem_split * 3 * PLM + blockIdx.x * PLM;
    double sm = 0.0;
    for (long long p = threadIdx.x; p < PL; p += NT) {
        double v = 0.0;
        for (int sp = 0; sp < P.sem_split; ++sp) v += part[(long long)sp * 3 * PLM + p];
        f[p] = v;
        sm += v;
    }
    const double mean = block_sum_nt<NT>(sm, red) / (double)PL;
    double qq = 0.0;
    for (long long p = threadIdx.x; p < PL; p += NT) {
        double v = f[p] - mean;
        f[p] = v;
        qq += v * v;
    }
    const double rms = sqrt(block_sum_nt<NT>(qq, red) / (double)PL);
    const double sc = (rms > LES_SMALL) ? target / rms : 0.0;
    for (long long p = threadIdx.x; p < PL; p += NT) f[p] *= sc;
}

// inflow plane from the SEM planes, zero-gradient outflow plane
template <typename real>
__global__ void __launch_bounds__(256) k_duct_bc(Pool P)
{
    __shared__ double red[8];
    const int s = blockIdx.z;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    const GridDims& g = D.g;
    const DuctConsts& k = D.k;
    const int j = blockIdx.y;
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    double vm = 0.0;
    if (j < g.ny && kk < g.nz) {
        real* q0 = (real*)P.A[0] + D.off;
        real* q1 = (real*)P.A[1] + D.off;
        real* q2 = (real*)P.A[2] + D.off;
        real* q3 = (real*)P.A[3] + D.off;
        real* q4 = (real*)P.A[4] + D.off;
        const double* in = P.inlet + D.poff;
        const long long PL = (long long)g.ny * g.nz;
        const long long p = (long long)j * g.nz + kk;
        const double rho_u = k.rho_u_Tu / k.Tu;
        const double u = k.U + in[p], v = in[PL + p], w = in[2 * PL + p];
        q0[p] = (real)(rho_u * u);
        q1[p] = (real)(rho_u * v);
        q2[p] = (real)(rho_u * w);
        q3[p] = q3[p + g.sx] - (real)g.dx;
        q4[p] = (real)k.Tu;
        vm = cell_speed<real>(q0[p], q1[p], q2[p], q4[p], k);
        const long long qn = (long long)(g.nx - 1) * g.sx + p, qm = qn - g.sx;
        q0[qn] = q0[qm]; q1[qn] = q1[qm]; q2[qn] = q2[qm]; q3[qn] = q3[qm]; q4[qn] = q4[qm];
    }
    double dummy = 0.0;
    block_max2_256(vm, dummy, red);
    if (threadIdx.x == 0)
        P.bcpart[(long long)s * P.ny_max * P.nzb + (long long)blockIdx.y * P.nzb + blockIdx.x] = vm;
}

// ---------------------------------------------------------------------------
// level-set reinitialisation (Sussman-Smereka-Osher, first-order Godunov) of
// the slots due this step; phi0 in B[3], iterate A[3] -> B[4] -> A[3]
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool reinit_due(const SlotView& D, int step, int iter)
{
    return D.k.reinit_interval > 0 && (step + 1) % D.k.reinit_interval == 0 && iter < D.k.reinit_iters;
}

// Level-set reinitialisation sweeps, ping-ponged between phi (A[3]) and a
// stage buffer (B[4]): sweep `iter` reads the even sweeps' input from phi and
// the odd ones' from B[4] and writes the other (the first sweep also keeps
// phi0 = phi in B[3]), so each sweep is one launch; phase 2 moves the result of
// a run with an odd sweep count back into phi.  The same arithmetic as a sweep
// into a temporary followed by a copy, one pass over the field fewer.
template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_reinit(Pool P, int phase, int iter, int xcd_map)
{
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        const SlotView D = slot_view(P, tp.s);
        const int step = ldc(&P.dyn[tp.s].step);
        if (!tp.in) return;
        const GridDims g = D.g;
        const int i = tp.i, j = tp.j, kk = tp.kk;
        const int q = (int)lidx(g, i, j, kk);
        const auto phi = run_arr((real*)P.A[3] + D.off);
        const auto phi0 = run_arr((real*)P.B[3] + D.off);
        const auto tmp = run_arr((real*)P.B[4] + D.off);
        if (phase == 2) {             // odd sweep count: the result is in B[4]
            if (reinit_due(D, step, 0) && (D.k.reinit_iters & 1)) phi[q] = tmp[q];
            return;
        }
        if (!reinit_due(D, step, iter)) return;
        const auto src = (iter & 1) ? tmp : phi;
        const auto dst = (iter & 1) ? phi : tmp;
        const real c = src[q];
        if (iter == 0) phi0[q] = c;
        if (i == 0 || i == g.nx - 1) { dst[q] = c; return; }
        const real dx = (real)g.dx, dtau = (real)0.3 * dx;
        const real idx = rcp_const<real>(D.k, DUCT_RC_ID), idy = rcp_const<real>(D.k, DUCT_RC_ID + 1),
                   idz = rcp_const<real>(D.k, DUCT_RC_ID + 2);
        const real p0 = iter == 0 ? c : phi0[q];
        real sgn = M<real>::div(p0, sqrt(p0 * p0 + dx * dx));
        real ax = (c - src[q - (int)g.sx]) * idx, bx = (src[q + (int)g.sx] - c) * idx;
        real ay = (c - src[(int)lidx(g, i, wrapi(j - 1, g.ny), kk)]) * idy, by = (src[(int)lidx(g, i, wrapi(j + 1, g.ny), kk)] - c) * idy;
        real az = (c - src[(int)lidx(g, i, j, wrapi(kk - 1, g.nz))]) * idz, bz = (src[(int)lidx(g, i, j, wrapi(kk + 1, g.nz))] - c) * idz;
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
        dst[q] = c - dtau * sgn * (gm - (real)1.0);
    }
}

// ---------------------------------------------------------------------------
// projection relaxation of the mass flux m = rho*u on interior planes:
// b = div m (central, periodic y/z); proj_iters Richardson sweeps
// phi <- phi + (L phi - b)/d from phi = 0 on the compact Laplacian with ghost
// rules phi(0) = phi(1) (inflow, Neumann) and phi(nx-1) = 0 (outflow); then
// m <- m - grad phi.  b in B[0], phi ping-pongs B[1] -> B[2] -> B[1] ...
// (B is dead after the last stage).  FP64 per-tile sums of b^2 and of the
// final residual (b - L phi)^2, reduced per run in a fixed order.
// ---------------------------------------------------------------------------
template <typename real, typename Arr>
__device__ __forceinline__ real proj_phi(const Arr& phi, const GridDims& g, int i, int j, int kk)
{
    if (i <= 0) i = 1;
    if (i >= g.nx - 1) return (real)0.0;
    return phi[i * (int)g.sx + wrapi(j, g.ny) * g.nz + wrapi(kk, g.nz)];
}

template <typename real, typename Arr>
__device__ __forceinline__ real proj_lap(const Arr& phi, const GridDims& g, const DuctConsts& k, int i, int j, int kk)
{
    const real ix2 = rcp_const<real>(k, DUCT_RC_ID2), iy2 = rcp_const<real>(k, DUCT_RC_ID2 + 1), iz2 = rcp_const<real>(k, DUCT_RC_ID2 + 2);
    const real p0 = phi[i * (int)g.sx + j * g.nz + kk];
    real lx = ((proj_phi<real>(phi, g, i + 1, j, kk) - p0) - (p0 - proj_phi<real>(phi, g, i - 1, j, kk))) * ix2;
    real ly = ((proj_phi<real>(phi, g, i, j + 1, kk) - p0) - (p0 - proj_phi<real>(phi, g, i, j - 1, kk))) * iy2;
    real lz = ((proj_phi<real>(phi, g, i, j, kk + 1) - p0) - (p0 - proj_phi<real>(phi, g, i, j, kk - 1))) * iz2;
    return lx + ly + lz;
}

template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_proj_div(Pool P, int xcd_map)
{
    __shared__ double red[256];
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    const SlotView D = slot_view(P, tp.s);
    if (D.k.proj_iters <= 0) return;
    const GridDims g = D.g;
    const int i = tp.i, j = tp.j, kk = tp.kk;
    double sb = 0.0;
    if (tp.in && i > 0 && i < g.nx - 1) {
        const int q = i * (int)g.sx + j * g.nz + kk;
        const auto m0 = run_arr((const real*)P.A[0] + D.off);
        const auto m1 = run_arr((const real*)P.A[1] + D.off);
        const auto m2 = run_arr((const real*)P.A[2] + D.off);
        const real hx = rcp_const<real>(D.k, DUCT_RC_HID), hy = rcp_const<real>(D.k, DUCT_RC_HID + 1), hz = rcp_const<real>(D.k, DUCT_RC_HID + 2);
        real v = (m0[q + (int)g.sx] - m0[q - (int)g.sx]) * hx +
                 (m1[i * (int)g.sx + wrapi(j + 1, g.ny) * g.nz + kk] - m1[i * (int)g.sx + wrapi(j - 1, g.ny) * g.nz + kk]) * hy +
                 (m2[i * (int)g.sx + j * g.nz + wrapi(kk + 1, g.nz)] - m2[i * (int)g.sx + j * g.nz + wrapi(kk - 1, g.nz)]) * hz;
        ((real*)P.B[0] + D.off)[q] = v;
        sb = (double)v * (double)v;
    }
    double tot = block_sum256(sb, red);
    if (threadIdx.x == 0) {
        P.ppart[2 * (long long)t] = tot;
        P.ppart[2 * (long long)t + 1] = 0.0;
    }
}

template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_proj_sweep(Pool P, int it, int xcd_map)
{
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    const SlotView D = slot_view(P, tp.s);
    const GridDims g = D.g;
    const int i = tp.i;
    if (it >= D.k.proj_iters || !tp.in || i == 0 || i == g.nx - 1) return;
    const int q = i * (int)g.sx + tp.j * g.nz + tp.kk;
    const auto po = run_arr((const real*)P.B[1 + (it & 1)] + D.off);
    const auto pn = run_arr((real*)P.B[1 + ((it + 1) & 1)] + D.off);
    const real b = ((const real*)P.B[0] + D.off)[q];
    const real inv_d = rcp_const<real>(D.k, DUCT_RC_INVD);
    if (it == 0) pn[q] = ((real)0.0 - b) * inv_d;           // phi = 0 before the first sweep
    else pn[q] = po[q] + (proj_lap<real>(po, g, D.k, i, tp.j, tp.kk) - b) * inv_d;
}

template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_proj_correct(Pool P, int xcd_map)
{
    __shared__ double red[256];
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    const SlotView D = slot_view(P, tp.s);
    if (D.k.proj_iters <= 0) return;
    const GridDims g = D.g;
    const int i = tp.i, j = tp.j, kk = tp.kk;
    double sr = 0.0, vm = 0.0;
    if (tp.in && i > 0 && i < g.nx - 1) {
        const int q = i * (int)g.sx + j * g.nz + kk;
        const auto phi = run_arr((const real*)P.B[1 + (D.k.proj_iters & 1)] + D.off);
        const real b = ((const real*)P.B[0] + D.off)[q];
        const real r = b - proj_lap<real>(phi, g, D.k, i, j, kk);
        sr = (double)r * (double)r;
        const real hx = rcp_const<real>(D.k, DUCT_RC_HID), hy = rcp_const<real>(D.k, DUCT_RC_HID + 1), hz = rcp_const<real>(D.k, DUCT_RC_HID + 2);
        const auto m0 = run_arr((real*)P.A[0] + D.off);
        const auto m1 = run_arr((real*)P.A[1] + D.off);
        const auto m2 = run_arr((real*)P.A[2] + D.off);
        const real a0 = m0[q] - (proj_phi<real>(phi, g, i + 1, j, kk) - proj_phi<real>(phi, g, i - 1, j, kk)) * hx;
        const real a1 = m1[q] - (proj_phi<real>(phi, g, i, j + 1, kk) - proj_phi<real>(phi, g, i, j - 1, kk)) * hy;
        const real a2 = m2[q] - (proj_phi<real>(phi, g, i, j, kk + 1) - proj_phi<real>(phi, g, i, j, kk - 1)) * hz;
        m0[q] = a0; m1[q] = a1; m2[q] = a2;
        vm = cell_speed<real>(a0, a1, a2, ((const real*)P.A[4] + D.off)[q], D.k);
    }
    double tot = block_sum256(sr, red);
    double dummy = 0.0;
    block_max2_256(vm, dummy, red);
    if (threadIdx.x == 0) {
        P.ppart[2 * (long long)t + 1] = tot;
        P.vpart[t] = vm;           // max |u| of the corrected state (next step's DT)
    }
}

// per-run FP64 residual RMS before and after the sweeps
__global__ void __launch_bounds__(1024) k_duct_proj_final(Pool P)
{
    __shared__ double r0[1024], r1[1024];
    const int s = blockIdx.x;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done || D.k.proj_iters <= 0) return;
    double a = 0.0, b = 0.0;
    for (int u = threadIdx.x; u < D.ntiles; u += 1024) {
        a += P.ppart[2 * (long long)(D.tile0 + u)];
        b += P.ppart[2 * (long long)(D.tile0 + u) + 1];
    }
    r0[threadIdx.x] = a;
    r1[threadIdx.x] = b;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { r0[threadIdx.x] += r0[threadIdx.x + w]; r1[threadIdx.x] += r1[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n_int = (double)(D.g.nx - 2) * D.g.ny * D.g.nz;
        P.dyn[s].proj_res0 = sqrt(r0[0] / n_int);
        P.dyn[s].proj_res = sqrt(r1[0] / n_int);
    }
}

// ---------------------------------------------------------------------------
// statistics: per-cell FP64 accumulators (:3167-3192) + y-z plane profiles.
// Grid (chunks, nx_max, slots); chunk sums are added in a fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool stats_due(const SlotDesc& D, const SlotDyn& y)
{
    return D.k.stats_start >= 0 && y.step >= D.k.stats_start;
}

// Chunks of one x-plane of a run in the statistics pass: a function of the
// run's own plane size only, so its sums do not depend on the pool.
__device__ __forceinline__ int stat_chunks(long long PL)
{
    const long long c = (PL + STAT_CHUNK_CELLS - 1) / STAT_CHUNK_CELLS;
    return c < 1 ? 1 : (c > STAT_CHUNKS ? STAT_CHUNKS : (int)c);
}

// Statistics samples of chunk ch of plane i of run s: FP64 sums of u, v, w, c
// and their products into the chunk's partials (and the per-cell FP64
// accumulators when the run keeps them).  The block of chunk 0 of plane 0
// records whether the run sampled this step, for the plane sums that follow.
template <typename real>
__device__ __forceinline__ void stats_block(const Pool& P, int s, int i, int ch, double (*red)[10])
{
    const SlotDesc& D = P.desc[s];
    const bool due = D.active && !P.dyn[s].done && stats_due(D, P.dyn[s]);
    if (i == 0 && ch == 0 && threadIdx.x == 0) P.sdue[s] = due ? 1 : 0;
    if (!due) return;
    const GridDims& g = D.g;
    if (i >= g.nx) return;
    const DuctConsts& k = D.k;
    const long long PL = (long long)g.ny * g.nz;
    const int nch = stat_chunks(PL);
    if (ch >= nch) return;
    const long long chunk = (PL + nch - 1) / nch;
    const long long p0 = (long long)ch * chunk;
    const long long p1 = p0 + chunk < PL ? p0 + chunk : PL;
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const auto A0 = run_arr((const real*)P.A[0] + D.off);
    const auto A1 = run_arr((const real*)P.A[1] + D.off);
    const auto A2 = run_arr((const real*)P.A[2] + D.off);
    const auto A3 = run_arr((const real*)P.A[3] + D.off);
    const auto A4 = run_arr((const real*)P.A[4] + D.off);
    const bool per_cell = k.stats_per_cell && P.stat[0];
    const real irut = rcp_const<real>(k, DUCT_RC_IRUT);
    for (long long p = p0 + threadIdx.x; p < p1; p += 256) {
        long long q = (long long)i * PL + p;
        const real ri = A4[q] * irut;                  // 1 / rho = T / (rho_u T_u)
        double u = (double)(A0[q] * ri), v = (double)(A1[q] * ri), w = (double)(A2[q] * ri);
        double c = (double)progress<real>(A3[q], rcp_const<real>(D.k, DUCT_RC_HID));
        if (per_cell) {
            long long qa = D.off + q;
            P.stat[0][qa] += u; P.stat[1][qa] += v; P.stat[2][qa] += w; P.stat[3][qa] += c;
            P.stat[4][qa] += u * u; P.stat[5][qa] += v * v; P.stat[6][qa] += w * w;
            P.stat[7][qa] += u * v; P.stat[8][qa] += u * w; P.stat[9][qa] += v * w;
        }
        acc[0] += u; acc[1] += v; acc[2] += w; acc[3] += c;
        acc[4] += u * u; acc[5] += v * v; acc[6] += w * w; acc[7] += u * v; acc[8] += u * w; acc[9] += v * w;
    }
    // fixed-order reduction: butterfly within each wave, then the 4 waves in order
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int a = 0; a < 10; ++a) acc[a] += __shfl_xor(acc[a], o);
    const int wv = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int a = 0; a < 10; ++a) red[wv][a] = acc[a];
    __syncthreads();
    if (threadIdx.x < 10)
        P.spart[(((long long)s * P.nx_max + i) * STAT_CHUNKS + ch) * 10 + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// The chunk partials of plane values t = 10 i + a of run s added in order into
// the run's plane profiles, where the run sampled this step (the flag its
// statistics launch recorded: the step counter itself advances in the same
// launch).
__device__ __forceinline__ void stats_final_block(const Pool& P, int s, int bx)
{
    const SlotDesc& D = P.desc[s];
    if (!D.active || !P.sdue[s]) return;
    int t = bx * blockDim.x + threadIdx.x;
    if (t >= D.g.nx * 10) return;
    int i = t / 10, a = t % 10;
    double sm = 0.0;
    const int nch = stat_chunks((long long)D.g.ny * D.g.nz);
    for (int c = 0; c < nch; ++c) sm += P.spart[(((long long)s * P.nx_max + i) * STAT_CHUNKS + c) * 10 + a];
    P.prof[D.xoff + t] += sm;
}

// fixed-point series (u, v, w, G) of run s into its ring
template <typename real>
__device__ __forceinline__ void points_block(const Pool& P, int s)
{
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    const int t = threadIdx.x;
    if (t >= D.npts) return;
    const int q = D.point_idx[t];
    real rho = (real)D.k.rho_u_Tu / ((const real*)P.A[4] + D.off)[q];
    const int slot = P.dyn[s].step % P.point_cap;
    double* o = P.points + D.ptoff + ((long long)slot * DUCT_NPTS + t) * 4;
    o[0] = (double)(((const real*)P.A[0] + D.off)[q] / rho);
    o[1] = (double)(((const real*)P.A[1] + D.off)[q] / rho);
    o[2] = (double)(((const real*)P.A[2] + D.off)[q] / rho);
    o[3] = (double)progress<real>(((const real*)P.A[3] + D.off)[q], rcp_const<real>(D.k, DUCT_RC_HID));
}

// flame-region probe lines (record_flame_region_data :3403-3471 structure):
// one block per (station, run).  The first G = 0.5 crossing along y on the
// line (i_probe, ., nz/2) is the smallest j whose segment [j, j+1] crosses
// (block minimum in LDS), located by linear interpolation; then the line's u,
// v, w, G, nu_t and |S| (|S| recomputed from the last stage's filtered
// velocities, nu_t of the last stage) go into the run's ring, one thread per j.
template <typename real>
__device__ __forceinline__ void probe_block(const Pool& P, int s, int n, int& jmin)
{
    const SlotDesc& D = P.desc[s];
    if (P.probe_cap <= 0 || !D.active || P.dyn[s].done || n >= D.nprobe) return;
    const GridDims g = D.g;
    double* o = P.probe + D.proff + ((long long)(P.dyn[s].step % P.probe_cap) * LES_NPROBE_MAX + n) * PROBE_STRIDE(P);
    const int ip = D.probe_i[n];
    if (ip <= 0 || ip >= g.nx - 1) {
        if (threadIdx.x == 0) o[0] = 0.0;
        return;
    }
    const int kp = g.nz / 2;
    const real* A3 = (const real*)P.A[3] + D.off;
    if (threadIdx.x == 0) jmin = 0x7fffffff;
    __syncthreads();
    for (int j = threadIdx.x; j < g.ny - 1; j += blockDim.x) {
        double g1 = (double)progress<real>(A3[lidx(g, ip, j, kp)], rcp_const<real>(D.k, DUCT_RC_HID));
        double g2 = (double)progress<real>(A3[lidx(g, ip, j + 1, kp)], rcp_const<real>(D.k, DUCT_RC_HID));
        if ((g1 - 0.5) * (g2 - 0.5) <= 0.0) atomicMin(&jmin, j);
    }
    __syncthreads();
    const int jc = jmin;
    double yf = -1.0;
    if (jc != 0x7fffffff) {
        double g1 = (double)progress<real>(A3[lidx(g, ip, jc, kp)], rcp_const<real>(D.k, DUCT_RC_HID));
        double g2 = (double)progress<real>(A3[lidx(g, ip, jc + 1, kp)], rcp_const<real>(D.k, DUCT_RC_HID));
        double y1 = jc * g.dy, y2 = (jc + 1) * g.dy;
        if (fabs(g2 - g1) < 1e-9) yf = y1;
        else yf = y1 + (0.5 - g1) / (g2 - g1) * (y2 - y1);
    }
    if (!(yf > 0.0)) {
        if (threadIdx.x == 0) o[0] = 0.0;
        return;
    }
    if (threadIdx.x == 0) { o[0] = 1.0; o[1] = yf; }
    const real* uf = (const real*)P.F[0] + D.off;
    const real* vf = (const real*)P.F[1] + D.off;
    const real* wf = (const real*)P.F[2] + D.off;
    const real hx = rcp_const<real>(D.k, DUCT_RC_HID), hy = rcp_const<real>(D.k, DUCT_RC_HID + 1), hz = rcp_const<real>(D.k, DUCT_RC_HID + 2);
    for (int j = threadIdx.x; j < g.ny; j += blockDim.x) {
        const int q = (int)lidx(g, ip, j, kp);
        const real rho = (real)D.k.rho_u_Tu / ((const real*)P.A[4] + D.off)[q];
        double* r = o + 2 + 6 * j;
        r[0] = (double)(((const real*)P.A[0] + D.off)[q] / rho);
        r[1] = (double)(((const real*)P.A[1] + D.off)[q] / rho);
        r[2] = (double)(((const real*)P.A[2] + D.off)[q] / rho);
        r[3] = (double)progress<real>(A3[q], rcp_const<real>(D.k, DUCT_RC_HID));
        r[4] = (double)((const real*)P.nut + D.off)[q];
        const int xp = q + (int)g.sx, xm = q - (int)g.sx;
        const int yp = (int)lidx(g, ip, wrapi(j + 1, g.ny), kp), ym = (int)lidx(g, ip, wrapi(j - 1, g.ny), kp);
        const int zp = (int)lidx(g, ip, j, wrapi(kp + 1, g.nz)), zm = (int)lidx(g, ip, j, wrapi(kp - 1, g.nz));
        real du_dx = (uf[xp] - uf[xm]) * hx, dv_dy = (vf[yp] - vf[ym]) * hy, dw_dz = (wf[zp] - wf[zm]) * hz;
        real du_dy = (uf[yp] - uf[ym]) * hy, du_dz = (uf[zp] - uf[zm]) * hz;
        real dv_dx = (vf[xp] - vf[xm]) * hx, dv_dz = (vf[zp] - vf[zm]) * hz;
        real dw_dx = (wf[xp] - wf[xm]) * hx, dw_dy = (wf[yp] - wf[ym]) * hy;
        real S12 = (real)0.5 * (du_dy + dv_dx), S13 = (real)0.5 * (du_dz + dw_dx), S23 = (real)0.5 * (dv_dz + dw_dy);
        real ss = (real)2.0 * (S12 * S12 + S13 * S13 + S23 * S23) + du_dx * du_dx + dv_dy * dv_dy + dw_dz * dw_dz;
        r[5] = (double)sqrt((real)2.0 * ss);
    }
}

// The outputs of the new state in one launch (they only read it): blocks
// [0, nst) the statistics chunks (sch per plane, nx_max planes per run), then
// LES_NPROBE_MAX probe stations per run (when probes are on), then one
// point-series block per run.
template <typename real>
__global__ void __launch_bounds__(256) k_duct_outputs(Pool P, int sch, int nslots)
{
    __shared__ double red[4][10];
    __shared__ int jmin;
    int b = blockIdx.x;
    const int nst = sch * P.nx_max * nslots;
    if (b < nst) {
        stats_block<real>(P, b / (sch * P.nx_max), (b / sch) % P.nx_max, b % sch, red);
        return;
    }
    b -= nst;
    const int npr = P.probe_cap > 0 ? LES_NPROBE_MAX * nslots : 0;
    if (b < npr) {
        probe_block<real>(P, b / LES_NPROBE_MAX, b % LES_NPROBE_MAX, jmin);
        return;
    }
    b -= npr;
    if (b < nslots) points_block<real>(P, b);
}

// The plane sums of the statistics (blocks [0, nfin), nfin = fx per run), then
// the step end of every run (the remaining blocks, one thread per run): time,
// step counter, done flag.
__global__ void __launch_bounds__(256) k_duct_step_close(Pool P, int fx, int nslots)
{
    const int nfin = fx * nslots;
    if ((int)blockIdx.x < nfin) {
        stats_final_block(P, (int)blockIdx.x / fx, (int)blockIdx.x % fx);
        return;
    }
    const int s = ((int)blockIdx.x - nfin) * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const SlotDesc& D = P.desc[s];
    SlotDyn& y = P.dyn[s];
    if (!D.active || y.done) return;
    y.time += y.dt;
    y.step += 1;
    if (D.n_steps > 0 && y.step >= D.n_steps) y.done = 1;
}

// monitor reductions of one slot (total mass, T and G range, non-finite)
template <typename real>
__global__ void __launch_bounds__(256) k_duct_monitor(Pool P, int s, double* part)
{
    __shared__ double sh[6][256];
    const SlotDesc& D = P.desc[s];
    const GridDims& g = D.g;
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, bad = 0;
    const double dV = g.dx * g.dy * g.dz;
    const real* A0 = (const real*)P.A[0] + D.off;
    const real* A3 = (const real*)P.A[3] + D.off;
    const real* A4 = (const real*)P.A[4] + D.off;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < g.n; q += (long long)gridDim.x * 256) {
        double T = (double)A4[q];
        tm += D.k.rho_u_Tu / T * dV;
        mnT = fmin(mnT, T); mxT = fmax(mxT, T);
        double c = (double)progress<real>(A3[q], rcp_const<real>(D.k, DUCT_RC_HID));
        mnG = fmin(mnG, c); mxG = fmax(mxG, c);
        if (!isfinite(T) || !isfinite((double)A0[q]) || !isfinite((double)A3[q])) bad = 1.0;
    }
    sh[0][threadIdx.x] = tm; sh[1][threadIdx.x] = mnT; sh[2][threadIdx.x] = mxT;
    sh[3][threadIdx.x] = mnG; sh[4][threadIdx.x] = mxG; sh[5][threadIdx.x] = bad;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        int t = threadIdx.x;
        if (t < w) {
            sh[0][t] += sh[0][t + w];
            sh[1][t] = fmin(sh[1][t], sh[1][t + w]); sh[2][t] = fmax(sh[2][t], sh[2][t + w]);
            sh[3][t] = fmin(sh[3][t], sh[3][t + w]); sh[4][t] = fmax(sh[4][t], sh[4][t + w]);
            sh[5][t] = fmax(sh[5][t], sh[5][t + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 6; ++c) part[6 * blockIdx.x + c] = sh[c][0];
}

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

// x-planes of the five state fields (arena A) from one run's range to
// another's, as 32-bit words (both ranges start on an element boundary);
// one launch for the five fields
__global__ void __launch_bounds__(256) k_duct_xface_copy(Pool P, long long src_word, long long dst_word,
                                                         long long words)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (int f = 0; f < DNF; ++f) {
        const unsigned* s = (const unsigned*)P.A[f] + src_word;
        unsigned* d = (unsigned*)P.A[f] + dst_word;
        for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) d[w] = s[w];
    }
}

