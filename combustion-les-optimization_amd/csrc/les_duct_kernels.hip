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
// arenas in HBM (first-fit cell ranges per slot) and every stencil kernel is a
// persistent loop over a device-side TILE LIST (first-fit tile ranges per
// slot), so one launch advances every active run: the ensemble and the served
// request stream use the same kernels as a single run (a pool of one).  All
// per-run state the kernels need (descriptor, dt, counters) lives on the
// device, so one time step of the whole pool is captured once into a hipGraph
// and replayed; admitting or retiring a run only rewrites device tables.
//
// Per RK stage:
//   k_duct_filter      state -> filtered primitives u, v, w, T, phi (+ per-tile sums of rho)
//   k_duct_rho_final   per-slot domain-mean density (fixed-order reduction)
//   k_duct_sgs         nu_t, S_T (+ per-tile max nu / alpha for the next DT)
//   k_duct_rhs_scalar  phi and T: RHS fused with the SSP-RK3 combine
//   k_duct_rhs_momentum rho*u, rho*v, rho*w: RHS fused with the combine
// Per step: DT, eddy update, inflow plane, its normalisation, the 3 stages,
// projection relaxation (runs with proj_iters > 0), inflow/outflow planes,
// reinitialisation (runs due), statistics, point series, step end.
#include "les_engine.hpp"

// The duct scheme is this engine's own specification: its kernels may fuse
// multiply-adds (the reference-scheme kernels keep strict IEEE order).
#pragma clang fp contract(fast)

#define DNF 5
#define DUCT_NPTS 64
#define STAT_CHUNKS 16            // most chunks of one x-plane in the statistics pass
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
    double* cp_tab;
    double* mu_tab;
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

// Arithmetic of the two storage types.  double: IEEE operations and the
// device libm (the FP64 path is compared with the oracle).  float (the FP32
// form): the hardware's native reciprocal, exp2 and log2 (v_rcp_f32,
// v_exp_f32, v_log_f32; about 1 ulp), which the FP32 tolerances cover.
template <typename real>
struct M {
    static __device__ __forceinline__ real div(real a, real b) { return a / b; }
    static __device__ __forceinline__ real exp(real x) { return ::exp(x); }
    static __device__ __forceinline__ real log(real x) { return ::log(x); }
    static __device__ __forceinline__ real pow(real x, real y) { return ::pow(x, y); }
    static __device__ __forceinline__ real tanh(real x) { return ::tanh(x); }
};

template <>
struct M<float> {
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504089f); }
    static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.69314718056f; }
    static __device__ __forceinline__ float pow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
    static __device__ __forceinline__ float tanh(float x)
    {
        const float t = __builtin_amdgcn_exp2f(-2.88539008178f * fabsf(x));     // exp(-2|x|)
        return copysignf((1.0f - t) * __builtin_amdgcn_rcpf(1.0f + t), x);
    }
};

// progress variable c = (1 + tanh(phi / (2 dx))) / 2; i2dx = 0.5/dx of the run
template <typename real>
__device__ __forceinline__ real progress(real phi, real i2dx)
{
    return (real)0.5 * ((real)1.0 + M<real>::tanh(phi * i2dx));
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
// count: sqrt(13/12) into the second differences and 1/2 into the first, so
// each IS_k is two FMAs; the linear weights and the common 1/6 into the
// candidate values and the denominator.  Algebraically the textbook formula;
// only the rounding differs from it.
template <typename real>
__device__ __forceinline__ real weno5_face(real v0, real v1, real v2, real v3, real v4)
{
    const real k = (real)1.0408329997330663, eps = (real)1.0e-6;    // k = sqrt(13/12)
    const real a0 = k * (v0 - (real)2.0 * v1 + v2), h0 = (real)0.5 * v0 - (real)2.0 * v1 + (real)1.5 * v2;
    const real a1 = k * (v1 - (real)2.0 * v2 + v3), h1 = (real)0.5 * (v1 - v3);
    const real a2 = k * (v2 - (real)2.0 * v3 + v4), h2 = (real)1.5 * v2 - (real)2.0 * v3 + (real)0.5 * v4;
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
    return M<real>::div(num, den);
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


// ---------------------------------------------------------------------------
// persistent tile loop (XCD-aware) and tile decoding
// ---------------------------------------------------------------------------
#define DUCT_CAS __attribute__((address_space(4)))

// One tile per block.  Blocks are dispatched round-robin over the 8 XCDs
// (block b runs on XCD b % 8); with xcd_map the blocks of one XCD take one
// contiguous range of the tile list, so tiles that share stencil halos
// (neighbours in j, k and i) are read through the same L2.  The grid is the
// tile high-water mark rounded up to a multiple of 8.
__device__ __forceinline__ int tile_of_block(const Pool& P, int xcd_map)
{
    int b = blockIdx.x;
    if (xcd_map && (gridDim.x & 7) == 0) b = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    return b < *(const DUCT_CAS int*)P.ntile ? b : -1;
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

__device__ __forceinline__ double block_sum256(double v, double* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    double r = red[0];
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
template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_filter(Pool P, int qin_is_B, int xcd_map)
{
    __shared__ double red[256];
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
        double rsum = 0.0;
        if (tp.in) {
            const int i = tp.i, j = tp.j, kk = tp.kk;
            const int q = (int)lidx(g, i, j, kk);
            const real rut = (real)k.rho_u_Tu, irut = rcp_const<real>(k, DUCT_RC_IRUT);
            real fu, fv, fw, fT, fp;
            if (i == 0 || i == g.nx - 1) {
                real T = q4[q], ri = T * irut;          // 1/rho
                fu = q0[q] * ri; fv = q1[q] * ri; fw = q2[q] * ri; fT = T; fp = q3[q];
            } else {
                // neighbour order (-x, -y, -z, centre, +z, +y, +x) as :1329-1331
                int r[7];
                r[0] = q - (int)g.sx;
                r[1] = (int)lidx(g, i, wrapi(j - 1, g.ny), kk);
                r[2] = (int)lidx(g, i, j, wrapi(kk - 1, g.nz));
                r[3] = q;
                r[4] = (int)lidx(g, i, j, wrapi(kk + 1, g.nz));
                r[5] = (int)lidx(g, i, wrapi(j + 1, g.ny), kk);
                r[6] = q + (int)g.sx;
                real pu[7], pv[7], pw[7], pT[7], pp[7];
#pragma unroll
                for (int u = 0; u < 7; ++u) {
                    real T = q4[r[u]], ri = T * irut;
                    pu[u] = q0[r[u]] * ri; pv[u] = q1[r[u]] * ri; pw[u] = q2[r[u]] * ri;
                    pT[u] = T; pp[u] = q3[r[u]];
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

// per-slot domain-mean density: fixed-order reduction of the slot's tile sums
__global__ void __launch_bounds__(1024) k_duct_rho_final(Pool P)
{
    __shared__ double red[1024];
    const int s = blockIdx.x;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    double v = 0.0;
    for (int b = threadIdx.x; b < D.ntiles; b += 1024) v += P.rho_part[D.tile0 + b];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) P.dyn[s].rho_avg = red[0] / (double)D.g.n;
}

// ---------------------------------------------------------------------------
// K2: properties, nu_t (:1569-1633), S_T (:1913-1965); per-tile DT maxima
// ---------------------------------------------------------------------------
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
                const real DX = (real)g.dx;
                real T = Tf[q], rf = M<real>::div((real)k.rho_u_Tu, T);
                real mu = table_lookup<real>(P.mu_tab, k, T);
                real cp = table_lookup<real>(P.cp_tab, k, T);
                real nu = M<real>::div(mu, rf);
                real lam = mu * cp * rcp_const<real>(k, DUCT_RC_IPRAN);
                real al = M<real>::div(lam, rf * cp);
                mn = (double)nu; ma = (double)al;
                const int xp = q + (int)g.sx, xm = q - (int)g.sx;
                const int yp = (int)lidx(g, i, wrapi(j + 1, g.ny), kk), ym = (int)lidx(g, i, wrapi(j - 1, g.ny), kk);
                const int zp = (int)lidx(g, i, j, wrapi(kk + 1, g.nz)), zm = (int)lidx(g, i, j, wrapi(kk - 1, g.nz));
                const real hx = rcp_const<real>(D.k, DUCT_RC_HID), hy = rcp_const<real>(D.k, DUCT_RC_HID + 1), hz = rcp_const<real>(D.k, DUCT_RC_HID + 2);
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
                real gg = progress<real>(pf[q], rcp_const<real>(D.k, DUCT_RC_HID));
                real cs = (real)k.cs_u + ((real)k.cs_b - (real)k.cs_u) * gg;
                cs += (real)k.cs_f * (real)4.0 * gg * ((real)1.0 - gg);
                nut[q] = (cs * DX) * (cs * DX) * smag;
                // SSR-SGS flame speed
                real num = ss - (div * div) * ((real)1.0 / (real)3.0);
                real X = M<real>::div((real)k.dx4, nu * nu) * num;
                real d_eta;
                if (X < (real)LES_SMALL) d_eta = (real)1.0;
                else {
                    real px = M<real>::pow(X, (real)0.75);
                    real alg = (real)0.2187 * px;
                    d_eta = (real)0.2340 * px * (alg > (real)1.0 ? M<real>::log(alg) : (real)0.0);
                }
                real eta = M<real>::div(DX, d_eta + (real)LES_SMALL);
                const real dF = rcp_const<real>(k, DUCT_RC_DF);
                real asf = (real)8.0 * M<real>::exp(M<real>::div((real)10.0 * dF, (real)8.0 * eta + (real)LES_SMALL));
                real wr = M<real>::pow(asf * eta * rcp_const<real>(k, DUCT_RC_ID), (real)2.0 - (real)k.d3);
                real te = DX * fmax((real)0.0, div) * rcp_const<real>(k, DUCT_RC_ITSL);
                stv[q] = fmax(SL, SL * fmax((real)0.0, wr + te));
            }
        }
        block_max2_256(mn, ma, red);
        if (threadIdx.x == 0) {
            double* dp = P.dpart + 2 * ((long long)stage * P.tile_cap + t);
            dp[0] = mn;
            dp[1] = ma;
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

// scalar RHS (phi: G equation :2706-2818; T :2912-2955) + SSP-RK3 combine
template <typename real, int TK>
// FP64: at least 4 waves per SIMD (122 VGPRs; a 5-wave cap spills 108 bytes);
// FP32: 8 waves with flat tiles, 7 with row tiles (8 would spill 16 bytes)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(real) == 8 ? 4 : (TK == 0 ? 8 : 7))))
k_duct_rhs_scalar(Pool P, RhsArgs<real> a)
{
    const int t = tile_of_block(P, a.xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        if (!tp.in) return;
        const SlotView D = slot_view(P, tp.s);
        const GridDims g = D.g;
        const DuctConsts& k = D.k;
        const int i = tp.i, j = tp.j, kk = tp.kk;
        const int bi = i * (int)g.sx;
        const int q = bi + j * g.nz + kk;
        void* const* QI = a.qin_is_B ? P.B : P.A;
        void* const* QO = a.qout_is_B ? P.B : P.A;
        const auto qin3 = run_arr((const real*)QI[3] + D.off);
        const auto qin4 = run_arr((const real*)QI[4] + D.off);
        const auto qo3 = run_arr((real*)QO[3] + D.off);
        const auto qo4 = run_arr((real*)QO[4] + D.off);
        if (i == 0 || i == g.nx - 1) {
            qo3[q] = qin3[q];
            qo4[q] = qin4[q];
            return;
        }
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
        const real T0 = Tf[q];
        const real nu0 = nut[q];
        const int xp = q + (int)g.sx, xm = q - (int)g.sx;
        const int yp = line_idx(g, bi, j, kk, 1, 1), ym = line_idx(g, bi, j, kk, 1, -1);
        const int zp = line_idx(g, bi, j, kk, 2, 1), zm = line_idx(g, bi, j, kk, 2, -1);
        const RunBuf<real> S2[2] = {pf, Tf};
        real cv[2], sg[2], lap[2], grad[3];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const RunBuf<real> Fs = S2[s];
            const real isc2 = rcp_const<real>(k, (s == 0) ? DUCT_RC_ISC : DUCT_RC_IPR);
            real acc = (real)0, sacc = (real)0, lacc = (real)0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const real id = rcp_const<real>(k, DUCT_RC_ID + d), id2 = rcp_const<real>(k, DUCT_RC_ID2 + d);
                const real hid = rcp_const<real>(k, DUCT_RC_HID + d);
                const int qp = d == 0 ? xp : (d == 1 ? yp : zp);
                const int qm = d == 0 ? xm : (d == 1 ? ym : zm);
                real ua = vel[d][q];
                real vm1 = Fs[qm], v0 = Fs[q], vp1 = Fs[qp];
                real tt;
                if (d == 0 && xband) {
                    tt = (ua >= (real)0.0) ? ua * (v0 - vm1) * id : ua * (vp1 - v0) * id;
                } else {
                    real v[7];
#pragma unroll
                    for (int o = 0; o < 7; ++o) v[o] = Fs[line_idx(g, bi, j, kk, d, o - 3)];
                    bool pos = ua >= (real)0.0;
                    real fp = face_plus<real>(scheme, v, pos);
                    real fm = face_minus<real>(scheme, v, pos);
                    tt = ua * (fp - fm) * id;
                }
                acc = (d == 0) ? tt : acc + tt;
                real ntp = nut[qp], ntm = nut[qm];
                real gp = (ntp + nu0) * isc2, gm = (nu0 + ntm) * isc2;
                real t2 = (gp * (vp1 - v0) - gm * (v0 - vm1)) * id2;
                sacc = (d == 0) ? t2 : sacc + t2;
                real t3 = (vp1 - (real)2.0 * v0 + vm1) * id2;
                lacc = (d == 0) ? t3 : lacc + t3;
                if (s == 0) grad[d] = (vp1 - vm1) * hid;
            }
            cv[s] = acc; sg[s] = sacc; lap[s] = lacc;
        }
        const real g0 = progress<real>(pf[q], rcp_const<real>(D.k, DUCT_RC_HID));
        const real ctr = (real)k.accel_center;
        real gmag = sqrt(grad[0] * grad[0] + grad[1] * grad[1] + grad[2] * grad[2]);
        real art = (real)0;
        if (g0 > (real)0.1 && g0 < (real)0.99) {
            real df = g0 - ctr;
            real win = M<real>::exp(-(df * df) * rcp_const<real>(k, DUCT_RC_IW2));
            art = (real)5.0 * (nu0 + (real)1.0e-5) * win * lap[0];
        }
        const auto stv = run_buf((const real*)P.st + D.off, g.n);
        real L3 = -cv[0] + stv[q] * gmag + sg[0] + art;
        const real rf0 = M<real>::div(rut, T0);
        real mu = table_lookup<real>(P.mu_tab, k, T0);
        real cp = table_lookup<real>(P.cp_tab, k, T0);
        real lam = mu * cp * rcp_const<real>(k, DUCT_RC_IPRAN);
        real al = M<real>::div(lam, rf0 * cp);
        const real Tu = (real)k.Tu, Tb = (real)k.Tb;
        real Ttg = Tu + g0 * (Tb - Tu);
        real L4 = -cv[1] + al * lap[1] + sg[1] + (real)k.relaxation * (Ttg - T0);
        if (a.Lout[0]) { a.Lout[3][q] = L3; a.Lout[4][q] = L4; }
        const real DT = (real)ldc(&P.dyn[tp.s].dt);
        real y3 = qin3[q] + DT * L3, y4 = qin4[q] + DT * L4;
        real v3, v4;
        if (a.stage == 0) { v3 = y3; v4 = y4; }
        else {
            const auto qa3 = run_arr((const real*)P.A[3] + D.off);
            const auto qa4 = run_arr((const real*)P.A[4] + D.off);
            v3 = (real)a.ca * qa3[q] + (real)a.cb * y3;
            v4 = (real)a.ca * qa4[q] + (real)a.cb * y4;
        }
        v4 = fmax((real)200.0, fmin((real)3000.0, v4));
        qo3[q] = v3;
        qo4[q] = v4;
    }
}

// momentum RHS (:2552-2691 structure) + SSP-RK3 combine; in the last stage also
// the per-tile max |u| of the new state (its T was written by the scalar kernel)
template <typename real, int TK>
// FP64: at least 6 waves per SIMD with the component loops rolled (79 VGPRs;
// unrolled it needs 115 at 4 waves); FP32: 8 waves (64 VGPRs, no spill)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(real) == 8 ? 6 : 8)))
k_duct_rhs_momentum(Pool P, RhsArgs<real> a)
{
    __shared__ double vred[8];
    const int t = tile_of_block(P, a.xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        const SlotView D = slot_view(P, tp.s);
        const GridDims g = D.g;
        const DuctConsts& k = D.k;
        double vmax = 0.0;
        if (tp.in) {
            const int i = tp.i, j = tp.j, kk = tp.kk;
            const int bi = i * (int)g.sx;
            const int q = bi + j * g.nz + kk;
            void* const* QI = a.qin_is_B ? P.B : P.A;
            void* const* QO = a.qout_is_B ? P.B : P.A;
            const RunArr<const real> qi[3] = {run_arr((const real*)QI[0] + D.off), run_arr((const real*)QI[1] + D.off),
                                              run_arr((const real*)QI[2] + D.off)};
            const RunArr<real> qo[3] = {run_arr((real*)QO[0] + D.off), run_arr((real*)QO[1] + D.off),
                                        run_arr((real*)QO[2] + D.off)};
            if (i == 0 || i == g.nx - 1) {
#pragma unroll
                for (int f = 0; f < 3; ++f) qo[f][q] = qi[f][q];
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
                const real T0 = Tf[q];
                const real rf0 = M<real>::div(rut, T0);
                // FP64 keeps the component loops rolled (fewer live registers,
                // 6 waves per SIMD); FP32 unrolls them
                constexpr int MU = sizeof(real) == 8 ? 1 : 3;
                real cv0 = (real)0, cv1 = (real)0, cv2 = (real)0;
#pragma unroll 1
                for (int d = 0; d < 3; ++d) {
                    const real id = rcp_dir<real>(k, DUCT_RC_ID, d);
                    const int qp = line_idx(g, bi, j, kk, d, 1), qm = line_idx(g, bi, j, kk, d, -1);
                    const RunBuf<real> U = d == 0 ? vel[0] : (d == 1 ? vel[1] : vel[2]);
                    real up = (real)0.5 * (U[qp] + U[q]);
                    real um = (real)0.5 * (U[q] + U[qm]);
                    real rp = (real)0.5 * (M<real>::div(rut, Tf[qp]) + rf0);
                    real rm = (real)0.5 * (rf0 + M<real>::div(rut, Tf[qm]));
                    real mp = rp * up, mm = rm * um;
                    const bool band = (d == 0 && xband);
#pragma unroll MU
                    for (int m = 0; m < 3; ++m) {
                        const RunBuf<real> V = m == 0 ? vel[0] : (m == 1 ? vel[1] : vel[2]);
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
                        const real c = m == 0 ? cv0 : (m == 1 ? cv1 : cv2);
                        const real nc = (d == 0) ? term : c + term;
                        cv0 = m == 0 ? nc : cv0;
                        cv1 = m == 1 ? nc : cv1;
                        cv2 = m == 2 ? nc : cv2;
                    }
                }
                const real conv[3] = {cv0, cv1, cv2};
                real mu = table_lookup<real>(P.mu_tab, k, T0);
                real nu0 = nut[q];
                real mu_eff = mu + rf0 * nu0;
                const int xp = q + (int)g.sx, xm = q - (int)g.sx;
                const int yp = line_idx(g, bi, j, kk, 1, 1), ym = line_idx(g, bi, j, kk, 1, -1);
                const int zp = line_idx(g, bi, j, kk, 2, 1), zm = line_idx(g, bi, j, kk, 2, -1);
                const real ix2 = rcp_const<real>(k, DUCT_RC_ID2), iy2 = rcp_const<real>(k, DUCT_RC_ID2 + 1), iz2 = rcp_const<real>(k, DUCT_RC_ID2 + 2);
                real diff[3];
                real df0 = (real)0, df1 = (real)0, df2 = (real)0;
#pragma unroll MU
                for (int m = 0; m < 3; ++m) {
                    const RunBuf<real> U = m == 0 ? vel[0] : (m == 1 ? vel[1] : vel[2]);
                    real uc = U[q];
                    real lx = (U[xp] - (real)2.0 * uc + U[xm]) * ix2;
                    real ly = (U[yp] - (real)2.0 * uc + U[ym]) * iy2;
                    real lz = (U[zp] - (real)2.0 * uc + U[zm]) * iz2;
                    const real dm = mu_eff * (lx + ly + lz);
                    df0 = m == 0 ? dm : df0;
                    df1 = m == 1 ? dm : df1;
                    df2 = m == 2 ? dm : df2;
                }
                diff[0] = df0; diff[1] = df1; diff[2] = df2;
                const real rho_avg = (real)ldc(&P.dyn[tp.s].rho_avg);
                real grav = (rf0 - rho_avg) * (real)LES_G_GRAV;
                real g0 = progress<real>(pf[q], rcp_const<real>(D.k, DUCT_RC_HID));
                real fx = (real)0, fy = (real)0, fz = (real)0;
                const real ctr = (real)k.accel_center, Fm = (real)k.max_force;
                if (g0 > (real)0.01 && g0 < (real)0.99) {
                    real gx = (real)0, gy = (real)0, gz = (real)0;
#pragma unroll MU
                    for (int d = 0; d < 3; ++d) {
                        const int a = d == 0 ? xp : (d == 1 ? yp : zp), b = d == 0 ? xm : (d == 1 ? ym : zm);
                        const real gd = (progress<real>(pf[a], rcp_const<real>(D.k, DUCT_RC_HID)) -
                                         progress<real>(pf[b], rcp_const<real>(D.k, DUCT_RC_HID))) * rcp_dir<real>(k, DUCT_RC_HID, d);
                        gx = d == 0 ? gd : gx;
                        gy = d == 1 ? gd : gy;
                        gz = d == 2 ? gd : gz;
                    }
                    real df = g0 - ctr;
                    real win = M<real>::exp(-(df * df) * rcp_const<real>(k, DUCT_RC_IW2));
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
                const real DT = (real)ldc(&P.dyn[tp.s].dt);
                real out[3];
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    real y = qi[f][q] + DT * Lq[f];
                    if (a.stage == 0) out[f] = y;
                    else out[f] = (real)a.ca * run_arr((const real*)P.A[f] + D.off)[q] + (real)a.cb * y;
                    qo[f][q] = out[f];
                }
                if (a.track_vel) {
                    vmax = cell_speed<real>(out[0], out[1], out[2], run_arr((const real*)QO[4] + D.off)[q], k);
                }
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
// DT per slot (:4399-4411 structure): one block per slot reduces the per-tile
// maxima of the previous step (max |u| of the new state, max nu / alpha over
// its 3 stages) and the inflow-plane maxima.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_duct_dt(Pool P)
{
    __shared__ double r[3][1024];
    const int s = blockIdx.x;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    const DuctConsts& k = D.k;
    double mv = 0.0, mn = 0.0, ma = 0.0;
    for (int b = threadIdx.x; b < D.ntiles; b += 1024) {
        const int t = D.tile0 + b;
        mv = fmax(mv, P.vpart[t]);
        for (int stg = 0; stg < 3; ++stg) {
            const double* dp = P.dpart + 2 * ((long long)stg * P.tile_cap + t);
            mn = fmax(mn, dp[0]);
            ma = fmax(ma, dp[1]);
        }
    }
    const int nbc = P.ny_max * P.nzb;
    for (int b = threadIdx.x; b < nbc; b += 1024) mv = fmax(mv, P.bcpart[(long long)s * nbc + b]);
    r[0][threadIdx.x] = mv; r[1][threadIdx.x] = mn; r[2][threadIdx.x] = ma;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c) r[c][threadIdx.x] = fmax(r[c][threadIdx.x], r[c][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    mv = r[0][0]; mn = r[1][0]; ma = r[2][0];
    const double dx = D.g.dx;
    double maxvel = mv > LES_SMALL ? mv : k.U;
    double dtc = k.courant * dx / (maxvel + LES_SMALL);
    double dtn = k.courant_diff * (dx * dx) / (mn + LES_SMALL);
    double dta = k.courant_diff * (dx * dx) / (ma + LES_SMALL);
    P.dyn[s].dt = fmin(dtc, fmin(dtn, dta));
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

// One thread per inflow cell and eddy chunk: block (k-chunk of 64 x eddy
// chunk, j, run) sums the Gaussian contributions of one chunk of SEM_CHUNK eddies
// (staged through LDS) into an FP64 partial plane; k_duct_sem_norm adds the
// chunks in order.
#define SEM_CHUNK 64
__global__ void __launch_bounds__(64) k_duct_sem_plane(Pool P)
{
    __shared__ double se[SEM_CHUNK * 7];
    const int nkc = (P.nz_max + 63) / 64;
    const int s = blockIdx.z, sp = (int)blockIdx.x / nkc;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    const GridDims& g = D.g;
    const DuctConsts& k = D.k;
    const int j = blockIdx.y;
    if (j >= g.ny) return;
    const int kk = ((int)blockIdx.x - sp * nkc) * 64 + threadIdx.x;
    const long long PLM = (long long)P.ny_max * P.nz_max;
    double* out = P.sem_part + ((long long)s * P.sem_split + sp) * 3 * PLM;
    const int e0 = sp * SEM_CHUNK;
    const int cnt = min(SEM_CHUNK, k.nvort - e0);
    double su = 0.0, sv = 0.0, sw = 0.0;
    if (cnt > 0) {
        const double* eddy = P.eddy + 7 * (D.eoff + e0);
        for (int u = threadIdx.x; u < cnt * 7; u += blockDim.x) se[u] = eddy[u];
        __syncthreads();
        if (kk < g.nz) {
            const double y = j * g.dy, z = kk * g.dz, inv_l2 = 1.0 / (k.L * k.L);
            for (int e = 0; e < cnt; ++e) {
                const double* E = se + 7 * e;
                double ddx = 0.0 - E[0];
                double ddy = y - E[1];
                ddy -= k.ly * rint(ddy * k.rc[DUCT_RC_ILY]);     // nearest periodic image
                double ddz = z - E[2];
                ddz -= k.lz * rint(ddz * k.rc[DUCT_RC_ILZ]);
                double f = exp(-M_PI * (ddx * ddx + ddy * ddy + ddz * ddz) * inv_l2);
                su += E[3] * f; sv += E[4] * f; sw += E[5] * f;
            }
        }
    }
    if (kk < g.nz) {
        long long p = (long long)j * g.nz + kk;
        out[p] = su; out[PLM + p] = sv; out[2 * PLM + p] = sw;
    }
}

// chunk sums -> plane (fixed order), mean removal and rms rescale; one
// 1024-thread block per (component, run)
__global__ void __launch_bounds__(1024) k_duct_sem_norm(Pool P)
{
    __shared__ double red[1024];
    const int s = blockIdx.y;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done) return;
    const long long PL = (long long)D.g.ny * D.g.nz;
    const long long PLM = (long long)P.ny_max * P.nz_max;
    const double target = D.k.U * D.k.intensity;
    double* f = P.inlet + D.poff + blockIdx.x * PL;
    const double* part = P.sem_part + (long long)s * P.sem_split * 3 * PLM + blockIdx.x * PLM;
    double sm = 0.0;
    for (long long p = threadIdx.x; p < PL; p += 1024) {
        double v = 0.0;
        for (int sp = 0; sp < P.sem_split; ++sp) v += part[(long long)sp * 3 * PLM + p];
        f[p] = v;
        sm += v;
    }
    red[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    double mean = red[0] / (double)PL;
    __syncthreads();
    double qq = 0.0;
    for (long long p = threadIdx.x; p < PL; p += 1024) {
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
    double rms = sqrt(red[0] / (double)PL);
    double sc = (rms > LES_SMALL) ? target / rms : 0.0;
    for (long long p = threadIdx.x; p < PL; p += 1024) f[p] *= sc;
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

template <typename real, int TK>
__global__ void __launch_bounds__(256) k_duct_reinit(Pool P, int phase, int iter, int xcd_map)
{
    const int t = tile_of_block(P, xcd_map);
    TilePos tp;
    if (t < 0 || !tile_decode<TK>(P, t, tp)) return;
    {
        const SlotView D = slot_view(P, tp.s);
        if (!reinit_due(D, ldc(&P.dyn[tp.s].step), iter) || !tp.in) return;
        const GridDims g = D.g;
        const int i = tp.i, j = tp.j, kk = tp.kk;
        const int q = (int)lidx(g, i, j, kk);
        const auto phi = run_arr((real*)P.A[3] + D.off);
        const auto phi0 = run_arr((real*)P.B[3] + D.off);
        const auto tmp = run_arr((real*)P.B[4] + D.off);
        if (phase == 0) {             // save phi0 (first iteration only)
            if (iter == 0) phi0[q] = phi[q];
            return;
        }
        if (phase == 2) {             // copy back
            phi[q] = tmp[q];
            return;
        }
        if (i == 0 || i == g.nx - 1) { tmp[q] = phi[q]; return; }
        const real dx = (real)g.dx, dtau = (real)0.3 * dx;
        const real idx = rcp_const<real>(D.k, DUCT_RC_ID), idy = rcp_const<real>(D.k, DUCT_RC_ID + 1),
                   idz = rcp_const<real>(D.k, DUCT_RC_ID + 2);
        real p0 = phi0[q];
        real sgn = M<real>::div(p0, sqrt(p0 * p0 + dx * dx));
        real c = phi[q];
        real ax = (c - phi[q - (int)g.sx]) * idx, bx = (phi[q + (int)g.sx] - c) * idx;
        real ay = (c - phi[(int)lidx(g, i, wrapi(j - 1, g.ny), kk)]) * idy, by = (phi[(int)lidx(g, i, wrapi(j + 1, g.ny), kk)] - c) * idy;
        real az = (c - phi[(int)lidx(g, i, j, wrapi(kk - 1, g.nz))]) * idz, bz = (phi[(int)lidx(g, i, j, wrapi(kk + 1, g.nz))] - c) * idz;
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
        tmp[q] = c - dtau * sgn * (gm - (real)1.0);
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
    real lx = (proj_phi<real>(phi, g, i + 1, j, kk) - (real)2.0 * p0 + proj_phi<real>(phi, g, i - 1, j, kk)) * ix2;
    real ly = (proj_phi<real>(phi, g, i, j + 1, kk) - (real)2.0 * p0 + proj_phi<real>(phi, g, i, j - 1, kk)) * iy2;
    real lz = (proj_phi<real>(phi, g, i, j, kk + 1) - (real)2.0 * p0 + proj_phi<real>(phi, g, i, j, kk - 1)) * iz2;
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

template <typename real>
__global__ void __launch_bounds__(256) k_duct_stats(Pool P)
{
    __shared__ double red[4][10];
    const int s = blockIdx.z;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done || !stats_due(D, P.dyn[s])) return;
    const GridDims& g = D.g;
    const int i = blockIdx.y;
    if (i >= g.nx) return;
    const DuctConsts& k = D.k;
    const long long PL = (long long)g.ny * g.nz;
    const int nch = stat_chunks(PL);
    if ((int)blockIdx.x >= nch) return;
    const long long chunk = (PL + nch - 1) / nch;
    const long long p0 = (long long)blockIdx.x * chunk;
    const long long p1 = p0 + chunk < PL ? p0 + chunk : PL;
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const real rut = (real)k.rho_u_Tu;
    const auto A0 = run_arr((const real*)P.A[0] + D.off);
    const auto A1 = run_arr((const real*)P.A[1] + D.off);
    const auto A2 = run_arr((const real*)P.A[2] + D.off);
    const auto A3 = run_arr((const real*)P.A[3] + D.off);
    const auto A4 = run_arr((const real*)P.A[4] + D.off);
    const bool per_cell = k.stats_per_cell && P.stat[0];
    for (long long p = p0 + threadIdx.x; p < p1; p += 256) {
        long long q = (long long)i * PL + p;
        real rho = rut / A4[q];
        double u = (double)(A0[q] / rho), v = (double)(A1[q] / rho), w = (double)(A2[q] / rho);
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
        P.spart[(((long long)s * P.nx_max + i) * STAT_CHUNKS + blockIdx.x) * 10 + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

__global__ void k_duct_stats_final(Pool P)
{
    const int s = blockIdx.y;
    const SlotDesc& D = P.desc[s];
    if (!D.active || P.dyn[s].done || !stats_due(D, P.dyn[s])) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.g.nx * 10) return;
    int i = t / 10, a = t % 10;
    double sm = 0.0;
    const int nch = stat_chunks((long long)D.g.ny * D.g.nz);
    for (int c = 0; c < nch; ++c) sm += P.spart[(((long long)s * P.nx_max + i) * STAT_CHUNKS + c) * 10 + a];
    P.prof[D.xoff + t] += sm;
}

// fixed-point series (u, v, w, G) of each slot into its ring, then the step end
template <typename real>
__global__ void k_duct_points(Pool P)
{
    const int s = blockIdx.x;
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
__global__ void __launch_bounds__(256) k_duct_probe(Pool P)
{
    __shared__ int jmin;
    const int s = blockIdx.y, n = blockIdx.x;
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

__global__ void k_duct_step_end(Pool P, int nslots)

{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
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

// host code below keeps IEEE operation order
#pragma clang fp contract(off)

// ---------------------------------------------------------------------------
// host side
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

// Calls f with the tile width as a compile-time constant.
template <typename F>
static int tk_dispatch(int tk, F&& f)
{
    switch (tk) {
    case 0: return f(std::integral_constant<int, 0>());
    case 32: return f(std::integral_constant<int, 32>());
    case 128: return f(std::integral_constant<int, 128>());
    default: return f(std::integral_constant<int, 64>());
    }
}

static bool tk_valid(int tk) { return tk == 0 || tk == 32 || tk == 64 || tk == 128; }

// Tiles of one run: TK > 0 covers TK x (256/TK) cells of an x-plane, flat
// tiles 256 consecutive cells of an x-plane.
static void run_tiles(int tk, int nx, int ny, int nz, int& tiles_j, int& tiles_k, long long& ntiles)
{
    if (tk == 0) {
        tiles_j = 1;
        tiles_k = (int)(((long long)ny * nz + 255) / 256);
    } else {
        const int tj = 256 / tk;
        tiles_j = (ny + tj - 1) / tj;
        tiles_k = (nz + tk - 1) / tk;
    }
    ntiles = (long long)nx * tiles_j * tiles_k;
}

// First-fit allocator of ranges in [0, cap).
struct RangeAlloc {
    long long cap = 0;
    std::vector<std::pair<long long, long long>> free_;   // (start, length), sorted
    void reset(long long c) { cap = c; free_.clear(); free_.push_back({0, c}); }
    bool alloc(long long n, long long& start)
    {
        for (size_t a = 0; a < free_.size(); ++a)
            if (free_[a].second >= n) {
                start = free_[a].first;
                free_[a].first += n;
                free_[a].second -= n;
                if (free_[a].second == 0) free_.erase(free_.begin() + a);
                return true;
            }
        return false;
    }
    void release(long long start, long long n)
    {
        size_t a = 0;
        while (a < free_.size() && free_[a].first < start) ++a;
        free_.insert(free_.begin() + a, {start, n});
        // merge neighbours
        for (size_t b = 0; b + 1 < free_.size();) {
            if (free_[b].first + free_[b].second == free_[b + 1].first) {
                free_[b].second += free_[b + 1].second;
                free_.erase(free_.begin() + b + 1);
            } else ++b;
        }
    }
    long long largest() const
    {
        long long m = 0;
        for (auto& f : free_) m = f.second > m ? f.second : m;
        return m;
    }
};

struct DuctState {
    int precision = LES_PREC_FP64;
    size_t esize = 8;
    int max_slots = 1;
    long long pool_cells = 0;
    int tile_cap = 0;
    int nx_max = 0, ny_max = 0, nz_max = 0, nvort_max = 0, nzb = 0;
    int point_cap = 1024;
    int probe_cap = 0;
    int stats_cells = 0;
    int cus = 256;
    // launch configuration (autotuned)
    int tk = 64;
    int grid = 8;                  // tile high-water mark rounded up to a multiple of 8
    int xcd_map = 1;
    double tuned_ms = 0.0;
    Pool P;
    SlotDesc* d_desc = nullptr;
    SlotDyn* d_dyn = nullptr;
    int* d_tile_slot = nullptr;
    int* d_ntile = nullptr;
    std::vector<SlotDesc> desc;
    std::vector<int> tile_slot;
    std::vector<int> taken;        // slot admitted (its state may not be set yet)
    std::vector<int> host_step;
    std::vector<int> point_base;   // host step index of the oldest unfetched point sample
    std::vector<int> probe_base;   // same for the probe ring
    std::vector<std::vector<les_probe_row>> probe_rows;   // drained, not yet fetched (batch)
    std::vector<std::vector<double>> probe_x;             // station x per run
    RangeAlloc cells, tiles;
    int ntile_hwm = 0;
    double* stage64 = nullptr;
    double* mon_part = nullptr;
    int max_reinit_iters = 0;
    int max_proj_iters = 0;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};   // [0] plain step, [1] with reinitialisation
    bool graph_ok[2] = {false, false};
    int use_graph = 1;
    // side stream for the launches that run beside the main chain (inflow
    // turbulence beside the RK stages, point/probe output beside the
    // statistics), forked from and joined back into the main stream
    hipStream_t aux = nullptr;
    hipEvent_t evf[2] = {nullptr, nullptr}, evj[2] = {nullptr, nullptr};
};

static void grid_dims(const les_params& p, GridDims& g)
{
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
    g.n = (long long)p.nx * p.ny * p.nz;
    g.sx = (long long)p.ny * p.nz;
    g.sy = p.nz;
    g.dx = p.lx / (double)p.nx;
    g.dy = p.ly / (double)p.ny;
    g.dz = p.lz / (double)p.nz;
}

static DuctConsts make_duct_consts(const les_params& p)
{
    DuctConsts k;
    memset(&k, 0, sizeof(k));
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
    k.reinit_interval = p.reinit_interval; k.reinit_iters = p.reinit_iters;
    k.proj_iters = p.proj_iters;
    k.stats_start = p.stats_start_step; k.stats_per_cell = p.stats_per_cell;
    k.L = p.turbulent_length_scale;
    k.dx4 = pow(p.lx / (double)p.nx, 4.0);
    k.ly = p.ly; k.lz = p.lz;
    k.seed = (unsigned long long)p.seed;
    GridDims g;
    grid_dims(p, g);
    const double dd[3] = {g.dx, g.dy, g.dz};
    for (int d = 0; d < 3; ++d) {
        k.rc[DUCT_RC_ID + d] = 1.0 / dd[d];
        k.rc[DUCT_RC_HID + d] = 0.5 / dd[d];
        k.rc[DUCT_RC_ID2 + d] = 1.0 / (dd[d] * dd[d]);
    }
    k.rc[DUCT_RC_IRUT] = 1.0 / k.rho_u_Tu;
    k.rc[DUCT_RC_IDEN] = 1.0 / (6.0 + k.wc);
    k.rc[DUCT_RC_ISC] = 0.5 / k.sc_t;
    k.rc[DUCT_RC_IPR] = 0.5 / k.pr_t;
    k.rc[DUCT_RC_INVD] = 1.0 / (2.0 / (g.dx * g.dx) + 2.0 / (g.dy * g.dy) + 2.0 / (g.dz * g.dz));
    k.rc[DUCT_RC_ITDT] = 1.0 / k.table_dt;
    k.rc[DUCT_RC_IPRAN] = 1.0 / k.prandtl;
    k.rc[DUCT_RC_IW2] = 1.0 / (2.0 * k.accel_width * k.accel_width);
    k.rc[DUCT_RC_DF] = k.nu_unburnt / (k.s_l + LES_SMALL);
    k.rc[DUCT_RC_ITSL] = (k.tau * k.s_l > LES_SMALL) ? 1.0 / (k.tau * k.s_l) : 0.0;
    k.rc[DUCT_RC_ILY] = 1.0 / k.ly;
    k.rc[DUCT_RC_ILZ] = 1.0 / k.lz;
    for (int r = 0; r < DUCT_NRC; ++r) k.rcf[r] = (float)k.rc[r];
    return k;
}

static void* dalloc_bytes(les_gpu* c, size_t bytes)
{
    void* p = nullptr;
    if (hipMalloc(&p, bytes + 256) != hipSuccess) return nullptr;
    c->allocs.push_back(p);
    (void)hipMemset(p, 0, bytes + 256);
    return p;
}

// Allocates the pool of one context.  cfg limits come from the batch config
// (a single-run context is a pool of one slot sized to its grid).
static int pool_create(les_gpu* c, const les_batch_config& cfg)
{
    DuctState* d = new (std::nothrow) DuctState();
    if (!d) return LES_ERR_HIP;
    c->d = d;
    if (hipStreamCreateWithFlags(&d->aux, hipStreamNonBlocking) != hipSuccess) return LES_ERR_HIP;
    for (int a = 0; a < 2; ++a)
        if (hipEventCreateWithFlags(&d->evf[a], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&d->evj[a], hipEventDisableTiming) != hipSuccess)
            return LES_ERR_HIP;
    d->precision = cfg.precision;
    d->esize = (cfg.precision == LES_PREC_FP32) ? 4 : 8;
    d->max_slots = cfg.max_slots;
    d->pool_cells = cfg.pool_cells;
    d->nx_max = cfg.max_nx; d->ny_max = cfg.max_ny; d->nz_max = cfg.max_nz;
    d->nvort_max = cfg.max_vortices > 0 ? cfg.max_vortices : 1;
    d->point_cap = cfg.point_cap > 0 ? cfg.point_cap : 1024;
    d->probe_cap = cfg.probe_cap > 0 ? cfg.probe_cap : 0;
    d->stats_cells = cfg.stats_per_cell;
    d->nzb = (d->nz_max + 255) / 256;
    d->tk = cfg.tile_k;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        d->cus = prop.multiProcessorCount;
    d->xcd_map = cfg.xcd_map;
    // tiles of all resident runs, for any tile width: per run at most
    // n/256 + nx*ny/2 + nx*nz/32 + nx (ceil bounds of each tile shape)
    {
        long long cap = d->pool_cells / 256 +
                        (long long)d->max_slots * d->nx_max * (d->ny_max / 2 + d->nz_max / 32 + 2) + 64;
        if (cap > (1LL << 30)) return LES_ERR_PARAM;
        d->tile_cap = (int)cap;
    }
    const size_t nb = (size_t)d->pool_cells * d->esize;
    Pool& P = d->P;
    memset(&P, 0, sizeof(P));
    for (int f = 0; f < DNF; ++f)
        if (!(P.A[f] = dalloc_bytes(c, nb)) || !(P.B[f] = dalloc_bytes(c, nb)) || !(P.F[f] = dalloc_bytes(c, nb)))
            return LES_ERR_HIP;
    if (!(P.nut = dalloc_bytes(c, nb)) || !(P.st = dalloc_bytes(c, nb))) return LES_ERR_HIP;
    if (d->stats_cells)
        for (int a = 0; a < 10; ++a)
            if (!(P.stat[a] = (double*)dalloc_bytes(c, 8 * (size_t)d->pool_cells))) return LES_ERR_HIP;
    const int tn = cfg.table_n;
    std::vector<double> cp(tn), mu(tn);
    for (int t = 0; t < tn; ++t) {
        double T = cfg.table_t0 + t * cfg.table_dt;
        cp[t] = 1000.0 + 0.2 * (T - 300.0);
        mu[t] = 1.8e-5 * pow(T / 300.0, 0.7);
    }
    P.cp_tab = (double*)dalloc_bytes(c, 8 * (size_t)tn);
    P.mu_tab = (double*)dalloc_bytes(c, 8 * (size_t)tn);
    if (!P.cp_tab || !P.mu_tab) return LES_ERR_HIP;
    DHIPC(hipMemcpy(P.cp_tab, cp.data(), 8 * (size_t)tn, hipMemcpyHostToDevice));
    DHIPC(hipMemcpy(P.mu_tab, mu.data(), 8 * (size_t)tn, hipMemcpyHostToDevice));
    const size_t S = (size_t)d->max_slots;
    P.eddy = (double*)dalloc_bytes(c, 8 * 7 * S * d->nvort_max);
    P.inlet = (double*)dalloc_bytes(c, 8 * 3 * S * d->ny_max * d->nz_max);
    P.prof = (double*)dalloc_bytes(c, 8 * 10 * S * d->nx_max);
    P.points = (double*)dalloc_bytes(c, 8 * 4 * DUCT_NPTS * S * d->point_cap);
    P.rho_part = (double*)dalloc_bytes(c, 8 * (size_t)d->tile_cap);
    P.vpart = (double*)dalloc_bytes(c, 8 * (size_t)d->tile_cap);
    P.dpart = (double*)dalloc_bytes(c, 8 * 2 * 3 * (size_t)d->tile_cap);
    P.bcpart = (double*)dalloc_bytes(c, 8 * S * d->ny_max * d->nzb);
    P.spart = (double*)dalloc_bytes(c, 8 * 10 * S * d->nx_max * STAT_CHUNKS);
    P.ppart = (double*)dalloc_bytes(c, 8 * 2 * (size_t)d->tile_cap);
    d->d_desc = (SlotDesc*)dalloc_bytes(c, sizeof(SlotDesc) * S);
    d->d_dyn = (SlotDyn*)dalloc_bytes(c, sizeof(SlotDyn) * S);
    d->d_tile_slot = (int*)dalloc_bytes(c, 4 * (size_t)d->tile_cap);
    d->d_ntile = (int*)dalloc_bytes(c, 64);
    d->mon_part = (double*)dalloc_bytes(c, 8 * 6 * 1024);
    if (!P.eddy || !P.inlet || !P.prof || !P.points || !P.rho_part || !P.vpart || !P.dpart || !P.bcpart ||
        !P.spart || !P.ppart || !d->d_desc || !d->d_dyn || !d->d_tile_slot || !d->d_ntile || !d->mon_part)
        return LES_ERR_HIP;
    if (d->precision == LES_PREC_FP32 && !(d->stage64 = (double*)dalloc_bytes(c, 8 * (size_t)d->pool_cells)))
        return LES_ERR_HIP;
    P.desc = d->d_desc; P.dyn = d->d_dyn; P.tile_slot = d->d_tile_slot; P.ntile = d->d_ntile;
    P.tile_cap = d->tile_cap;
    P.ny_max = d->ny_max; P.nz_max = d->nz_max; P.nx_max = d->nx_max; P.nzb = d->nzb;
    P.point_cap = d->point_cap;
    P.probe_cap = d->probe_cap;
    P.sem_split = (d->nvort_max + SEM_CHUNK - 1) / SEM_CHUNK;
    if (P.sem_split < 1) P.sem_split = 1;
    if ((long long)S * P.sem_split > 65535) {
        snprintf(c->err, sizeof(c->err), "max_slots x eddy chunks exceeds the launch grid (65535)");
        return LES_ERR_PARAM;
    }
    if (!(P.sem_part = (double*)dalloc_bytes(c, 8 * S * P.sem_split * 3 * (size_t)d->ny_max * d->nz_max)))
        return LES_ERR_HIP;
    if (d->probe_cap > 0 &&
        !(P.probe = (double*)dalloc_bytes(c, 8 * S * d->probe_cap * LES_NPROBE_MAX * (size_t)PROBE_STRIDE(P))))
        return LES_ERR_HIP;
    d->desc.assign(S, SlotDesc());
    for (auto& x : d->desc) memset(&x, 0, sizeof(x));
    d->tile_slot.assign(d->tile_cap, -1);
    d->taken.assign(S, 0);
    d->host_step.assign(S, 0);
    d->point_base.assign(S, 0);
    d->probe_base.assign(S, 0);
    d->probe_rows.assign(S, std::vector<les_probe_row>());
    d->probe_x.assign(S, std::vector<double>(LES_NPROBE_MAX, 0.0));
    d->cells.reset(d->pool_cells);
    d->tiles.reset(d->tile_cap);
    std::vector<int> neg(d->tile_cap, -1);
    DHIPC(hipMemcpy(d->d_tile_slot, neg.data(), 4 * (size_t)d->tile_cap, hipMemcpyHostToDevice));
    return LES_OK;
}

// Uploads the descriptor table, tile list and tile high-water mark.
static int pool_sync_tables(les_gpu* c)
{
    DuctState* d = c->d;
    DHIPC(hipStreamSynchronize(c->stream));     // no step may be reading the tables
    int hwm = 0;
    for (int t = 0; t < d->tile_cap; ++t) if (d->tile_slot[t] >= 0) hwm = t + 1;
    d->ntile_hwm = hwm;
    const int grid = hwm > 0 ? (hwm + 7) / 8 * 8 : 8;
    if (grid != d->grid) { d->grid = grid; d->graph_ok[0] = d->graph_ok[1] = false; }
    DHIPC(hipMemcpyAsync(d->d_desc, d->desc.data(), sizeof(SlotDesc) * d->max_slots, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->d_tile_slot, d->tile_slot.data(), 4 * (size_t)d->tile_cap, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->d_ntile, &d->ntile_hwm, 4, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    int maxit = 0, maxpj = 0;
    for (auto& x : d->desc) {
        if (x.active && x.k.reinit_interval > 0 && x.k.reinit_iters > maxit) maxit = x.k.reinit_iters;
        if (x.active && x.k.proj_iters > maxpj) maxpj = x.k.proj_iters;
    }
    if (maxit != d->max_reinit_iters || maxpj != d->max_proj_iters) d->graph_ok[0] = d->graph_ok[1] = false;
    d->max_reinit_iters = maxit;
    d->max_proj_iters = maxpj;
    return LES_OK;
}

// Admits run p into a free slot (first-fit cells and tiles); the slot starts
// empty (state zero) until init or upload.
static int pool_admit(les_gpu* c, const les_params& p, int n_steps, int* slot_out)
{
    DuctState* d = c->d;
    if (p.mode != LES_MODE_DUCT || p.integrator != LES_INT_SSPRK3 ||
        (p.advection != LES_ADV_WENO5 && p.advection != LES_ADV_MUSCL_MINMOD))
        return LES_ERR_PARAM;
    if (p.precision != d->precision || p.nx > d->nx_max || p.ny > d->ny_max || p.nz > d->nz_max ||
        p.num_vortices > d->nvort_max || p.reinit_iters < 0 || p.reinit_iters > 64 || p.proj_iters < 0 ||
        p.proj_iters > 256)
        return LES_ERR_ADMISSION;
    // kernels address a run's field with 32-bit byte offsets (RunArr)
    if ((long long)p.nx * p.ny * p.nz * (long long)d->esize >= (1ll << 32)) return LES_ERR_PARAM;
    int s = -1;
    for (int a = 0; a < d->max_slots; ++a) if (!d->taken[a]) { s = a; break; }
    if (s < 0) return LES_ERR_ADMISSION;
    GridDims g;
    grid_dims(p, g);
    int tiles_j, tiles_k;
    long long ntiles;
    run_tiles(d->tk, p.nx, p.ny, p.nz, tiles_j, tiles_k, ntiles);
    long long c0 = 0, t0 = 0;
    if (!d->cells.alloc(g.n, c0)) return LES_ERR_ADMISSION;
    if (!d->tiles.alloc(ntiles, t0)) { d->cells.release(c0, g.n); return LES_ERR_ADMISSION; }
    SlotDesc& D = d->desc[s];
    memset(&D, 0, sizeof(D));
    D.g = g;
    D.off = c0;
    D.poff = (long long)s * 3 * d->ny_max * d->nz_max;
    D.xoff = (long long)s * 10 * d->nx_max;
    D.eoff = (long long)s * d->nvort_max;
    D.ptoff = (long long)s * 4 * DUCT_NPTS * d->point_cap;
    D.tile0 = (int)t0; D.ntiles = (int)ntiles; D.tiles_j = tiles_j; D.tiles_k = tiles_k;
    D.active = 0;               // set once the run has a state (init or upload)
    D.n_steps = n_steps;
    int t = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int e = 0; e < 4; ++e) {
                int ii = (a + 1) * p.nx / 5, jj = (2 * b + 1) * p.ny / 8, kk = (2 * e + 1) * p.nz / 8;
                D.point_idx[t++] = (int)(((long long)ii * p.ny + jj) * p.nz + kk);
            }
    D.npts = DUCT_NPTS;
    D.k = make_duct_consts(p);
    D.nprobe = p.num_probe_x;
    for (int n = 0; n < p.num_probe_x; ++n) {
        D.probe_i[n] = (int)(p.probe_x[n] / g.dx);
        d->probe_x[s][n] = p.probe_x[n];
    }
    d->probe_rows[s].clear();
    D.proff = (long long)s * d->probe_cap * LES_NPROBE_MAX * (2 + 6LL * d->ny_max);
    for (long long a = 0; a < ntiles; ++a) d->tile_slot[t0 + a] = s;
    d->host_step[s] = 0;
    d->point_base[s] = 0;
    d->probe_base[s] = 0;
    d->taken[s] = 1;
    SlotDyn y;
    memset(&y, 0, sizeof(y));
    DHIPC(hipMemcpyAsync(d->d_dyn + s, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    int e = pool_sync_tables(c);
    if (e != LES_OK) return e;
    *slot_out = s;
    return LES_OK;
}

static int pool_release(les_gpu* c, int s)
{
    DuctState* d = c->d;
    if (s < 0 || s >= d->max_slots || !d->taken[s]) return LES_ERR_PARAM;
    SlotDesc& D = d->desc[s];
    d->cells.release(D.off, D.g.n);
    d->tiles.release(D.tile0, D.ntiles);
    for (int a = 0; a < D.ntiles; ++a) d->tile_slot[D.tile0 + a] = -1;
    D.active = 0;
    d->taken[s] = 0;
    return pool_sync_tables(c);
}

// ---- launches ----
template <typename real, int TK>
static int stage_launch(les_gpu* c, int stage, real* const* Lout)
{
    DuctState* d = c->d;
    const int G = d->grid;
    const int qin_B = stage > 0, qout_B = stage < 2;
    hipLaunchKernelGGL((k_duct_filter<real, TK>), dim3(G), dim3(256), 0, c->stream, d->P, qin_B, d->xcd_map);
    hipLaunchKernelGGL(k_duct_rho_final, dim3(d->max_slots), dim3(1024), 0, c->stream, d->P);
    hipLaunchKernelGGL((k_duct_sgs<real, TK>), dim3(G), dim3(256), 0, c->stream, d->P, stage, d->xcd_map);
    DLAUNCH_CHECK();
    RhsArgs<real> a;
    a.qin_is_B = qin_B; a.qout_is_B = qout_B; a.stage = stage;
    a.ca = (stage == 1) ? 0.75 : (1.0 / 3.0);
    a.cb = (stage == 1) ? 0.25 : (2.0 / 3.0);
    a.track_vel = (stage == 2);
    a.xcd_map = d->xcd_map;
    for (int f = 0; f < DNF; ++f) a.Lout[f] = Lout ? Lout[f] : nullptr;
    if (c->timing) DHIPC(hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL((k_duct_rhs_scalar<real, TK>), dim3(G), dim3(256), 0, c->stream, d->P, a);
    hipLaunchKernelGGL((k_duct_rhs_momentum<real, TK>), dim3(G), dim3(256), 0, c->stream, d->P, a);
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

// One engine step of every active, unfinished slot; all launches on c->stream.
// Forks the side stream off the main stream (f) or joins it back (j); in the
// per-kernel timing mode everything stays on the main stream.
static int side_fork(les_gpu* c, int f)
{
    DuctState* d = c->d;
    if (c->timing) return LES_OK;
    DHIPC(hipEventRecord(d->evf[f], c->stream));
    DHIPC(hipStreamWaitEvent(d->aux, d->evf[f], 0));
    return LES_OK;
}

static int side_join(les_gpu* c, int f)
{
    DuctState* d = c->d;
    if (c->timing) return LES_OK;
    DHIPC(hipEventRecord(d->evj[f], d->aux));
    DHIPC(hipStreamWaitEvent(c->stream, d->evj[f], 0));
    return LES_OK;
}

template <typename real, int TK>
static int step_launches(les_gpu* c, bool reinit)
{
    DuctState* d = c->d;
    const int S = d->max_slots;
    hipStream_t side = c->timing ? c->stream : d->aux;
    hipLaunchKernelGGL(k_duct_dt, dim3(S), dim3(1024), 0, c->stream, d->P);
    // the inflow turbulence (eddies -> SEM plane) is read only by the boundary
    // update after the stages, so it runs beside them
    int e = side_fork(c, 0);
    if (e != LES_OK) return e;
    hipLaunchKernelGGL(k_duct_eddy_update, dim3((d->nvort_max + 255) / 256, S), dim3(256), 0, side, d->P);
    hipLaunchKernelGGL(k_duct_sem_plane, dim3((d->nz_max + 63) / 64 * d->P.sem_split, d->ny_max, S), dim3(64), 0,
                       side, d->P);
    hipLaunchKernelGGL(k_duct_sem_norm, dim3(3, S), dim3(1024), 0, side, d->P);
    DLAUNCH_CHECK();
    for (int s = 0; s < 3; ++s) {
        e = stage_launch<real, TK>(c, s, nullptr);
        if (e != LES_OK) return e;
    }
    if (d->max_proj_iters > 0) {
        hipLaunchKernelGGL((k_duct_proj_div<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, d->xcd_map);
        for (int it = 0; it < d->max_proj_iters; ++it)
            hipLaunchKernelGGL((k_duct_proj_sweep<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, it,
                               d->xcd_map);
        hipLaunchKernelGGL((k_duct_proj_correct<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, d->xcd_map);
        hipLaunchKernelGGL(k_duct_proj_final, dim3(S), dim3(1024), 0, c->stream, d->P);
        DLAUNCH_CHECK();
    }
    if ((e = side_join(c, 0)) != LES_OK) return e;
    hipLaunchKernelGGL((k_duct_bc<real>), dim3(d->nzb, d->ny_max, S), dim3(256), 0, c->stream, d->P);
    DLAUNCH_CHECK();
    if (reinit) {
        hipLaunchKernelGGL((k_duct_reinit<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, 0, 0, d->xcd_map);
        for (int it = 0; it < d->max_reinit_iters; ++it) {
            hipLaunchKernelGGL((k_duct_reinit<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, 1, it,
                               d->xcd_map);
            hipLaunchKernelGGL((k_duct_reinit<real, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, 2, it,
                               d->xcd_map);
        }
        DLAUNCH_CHECK();
    }
    // outputs of the new state: point samples and probe rows beside the statistics
    if ((e = side_fork(c, 1)) != LES_OK) return e;
    hipLaunchKernelGGL((k_duct_points<real>), dim3(S), dim3(64), 0, side, d->P);
    if (d->probe_cap > 0)
        hipLaunchKernelGGL((k_duct_probe<real>), dim3(LES_NPROBE_MAX, S), dim3(256), 0, side, d->P);
    {
        const long long plm = (long long)d->ny_max * d->nz_max;
        const int sch = (int)std::min<long long>(STAT_CHUNKS, std::max<long long>(1, (plm + STAT_CHUNK_CELLS - 1) / STAT_CHUNK_CELLS));
        hipLaunchKernelGGL((k_duct_stats<real>), dim3(sch, d->nx_max, S), dim3(256), 0, c->stream, d->P);
    }
    hipLaunchKernelGGL(k_duct_stats_final, dim3((d->nx_max * 10 + 255) / 256, S), dim3(256), 0, c->stream, d->P);
    if ((e = side_join(c, 1)) != LES_OK) return e;
    hipLaunchKernelGGL(k_duct_step_end, dim3((S + 255) / 256), dim3(256), 0, c->stream, d->P, S);
    DLAUNCH_CHECK();
    return LES_OK;
}

static int pool_step_launches(les_gpu* c, bool reinit)
{
    return tk_dispatch(c->d->tk, [&](auto K) {
        return c->d->precision == LES_PREC_FP32 ? step_launches<float, decltype(K)::value>(c, reinit)
                                                : step_launches<double, decltype(K)::value>(c, reinit);
    });
}

// Advances the pool by n engine steps: graph replay, one graph with and one
// without the reinitialisation launches (chosen from the host's step mirror).
static int pool_steps(les_gpu* c, int n)
{
    DuctState* d = c->d;
    for (int a = 0; a < n; ++a) {
        bool reinit = false;
        for (int s = 0; s < d->max_slots; ++s) {
            const SlotDesc& D = d->desc[s];
            if (!D.active || (D.n_steps > 0 && d->host_step[s] >= D.n_steps)) continue;
            if (D.k.reinit_interval > 0 && (d->host_step[s] + 1) % D.k.reinit_interval == 0) reinit = true;
        }
        const int gi = reinit ? 1 : 0;
        bool done = false;
        if (d->use_graph && !c->timing) {
            if (!d->graph_ok[gi]) {
                if (d->gexec[gi]) { (void)hipGraphExecDestroy(d->gexec[gi]); d->gexec[gi] = nullptr; }
                if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                    int e = pool_step_launches(c, reinit);
                    hipGraph_t gr = nullptr;
                    hipError_t ce = hipStreamEndCapture(c->stream, &gr);
                    if (e == LES_OK && ce == hipSuccess && gr &&
                        hipGraphInstantiate(&d->gexec[gi], gr, nullptr, nullptr, 0) == hipSuccess)
                        d->graph_ok[gi] = true;
                    if (gr) (void)hipGraphDestroy(gr);
                }
                (void)hipGetLastError();
                if (!d->graph_ok[gi]) d->use_graph = 0;
            }
            if (d->graph_ok[gi]) {
                DHIPC(hipGraphLaunch(d->gexec[gi], c->stream));
                done = true;
            }
        }
        if (!done) {
            int e = pool_step_launches(c, reinit);
            if (e != LES_OK) return e;
        }
        for (int s = 0; s < d->max_slots; ++s) {
            const SlotDesc& D = d->desc[s];
            if (!D.active || (D.n_steps > 0 && d->host_step[s] >= D.n_steps)) continue;
            d->host_step[s]++;
        }
    }
    return LES_OK;
}

// ---- state exchange of one slot (fields rho*u, rho*v, rho*w, phi, T as double) ----
static int slot_xfer(les_gpu* c, int s, void* arena, double* host, bool upload)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    const size_t n = (size_t)D.g.n;
    char* dev = (char*)arena + (size_t)D.off * d->esize;
    if (d->precision == LES_PREC_FP64) {
        if (upload) DHIPC(hipMemcpyAsync(dev, host, n * 8, hipMemcpyHostToDevice, c->stream));
        else DHIPC(hipMemcpyAsync(host, dev, n * 8, hipMemcpyDeviceToHost, c->stream));
        return LES_OK;
    }
    double* tmp = d->stage64;
    int nb = (int)((n + 255) / 256);
    if (upload) {
        DHIPC(hipMemcpyAsync(tmp, host, n * 8, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_to_real<float>, dim3(nb), dim3(256), 0, c->stream, (long long)n, tmp, (float*)dev);
    } else {
        hipLaunchKernelGGL(k_to_double<float>, dim3(nb), dim3(256), 0, c->stream, (long long)n, (const float*)dev, tmp);
        DHIPC(hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

// After new state in slot s: DT maxima of the state (|u|), no diffusive history.
static int slot_state_changed(les_gpu* c, int s)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    DHIPC(hipMemsetAsync(d->P.bcpart + (size_t)s * d->ny_max * d->nzb, 0, 8 * (size_t)d->ny_max * d->nzb, c->stream));
    tk_dispatch(d->tk, [&](auto K) {
        constexpr int TK = decltype(K)::value;
        if (d->precision == LES_PREC_FP32)
            hipLaunchKernelGGL((k_duct_velmax<float, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, s, d->xcd_map);
        else
            hipLaunchKernelGGL((k_duct_velmax<double, TK>), dim3(d->grid), dim3(256), 0, c->stream, d->P, s, d->xcd_map);
        return 0;
    });
    DLAUNCH_CHECK();
    DHIPC(hipStreamSynchronize(c->stream));
    (void)D;
    return LES_OK;
}

// New state of slot s (the step counter is kept, as the oracle's set_state
// keeps its step); the run becomes active if it was only admitted.
static int slot_upload(les_gpu* c, int s, const double* const* f, int nf)
{
    DuctState* d = c->d;
    if (nf < DNF) { snprintf(c->err, sizeof(c->err), "duct mode state has 5 fields"); return LES_ERR_PARAM; }
    if (!d->desc[s].active) {
        d->desc[s].active = 1;
        int e = pool_sync_tables(c);
        if (e != LES_OK) return e;
    }
    for (int a = 0; a < DNF; ++a) {
        int e = slot_xfer(c, s, d->P.A[a], (double*)f[a], true);
        if (e != LES_OK) return e;
    }
    return slot_state_changed(c, s);
}

// initial state of a slot: plug flow at U, planar flame at flame_x0 (host
// computed: runs once per run), Gaussian eddies, cleared statistics
static int slot_init(les_gpu* c, int s, const les_params& p)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    const size_t n = (size_t)D.g.n;
    std::vector<double> h(5 * n);
    const double U = p.inlet_velocity, Tu = p.inlet_temperature, Tb = p.burnt_temperature;
    for (int i = 0; i < p.nx; ++i) {
        double phi = i * D.g.dx - p.flame_x0;
        double cc = 0.5 * (1.0 + tanh(phi / (2.0 * D.g.dx)));
        double T = Tu + cc * (Tb - Tu);
        double rho = p.rho_unburnt * Tu / T;
        for (size_t q = (size_t)i * p.ny * p.nz; q < (size_t)(i + 1) * p.ny * p.nz; ++q) {
            h[q] = rho * U; h[n + q] = 0.0; h[2 * n + q] = 0.0; h[3 * n + q] = phi; h[4 * n + q] = T;
        }
    }
    SlotDyn y;
    memset(&y, 0, sizeof(y));
    DHIPC(hipMemcpyAsync(d->d_dyn + s, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    d->host_step[s] = 0;
    d->point_base[s] = 0;
    d->probe_base[s] = 0;
    const double* f[5] = {&h[0], &h[n], &h[2 * n], &h[3 * n], &h[4 * n]};
    int e = slot_upload(c, s, f, 5);
    if (e != LES_OK) return e;
    if (D.k.nvort > 0) {
        hipLaunchKernelGGL(k_duct_eddy_init, dim3((D.k.nvort + 255) / 256), dim3(256), 0, c->stream, d->P, s);
        DLAUNCH_CHECK();
    }
    DHIPC(hipMemsetAsync(d->P.inlet + D.poff, 0, 8 * 3 * (size_t)D.g.ny * D.g.nz, c->stream));
    DHIPC(hipMemsetAsync(d->P.prof + D.xoff, 0, 8 * 10 * (size_t)D.g.nx, c->stream));
    if (d->P.stat[0])
        for (int a = 0; a < 10; ++a) DHIPC(hipMemsetAsync(d->P.stat[a] + D.off, 0, 8 * n, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

template <typename real>
static int slot_monitor_t(les_gpu* c, int s, les_step_report* rep)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    int nb = (int)((D.g.n + 255) / 256);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL((k_duct_monitor<real>), dim3(nb), dim3(256), 0, c->stream, d->P, s, d->mon_part);
    DLAUNCH_CHECK();
    std::vector<double> h((size_t)nb * 6);
    SlotDyn y;
    DHIPC(hipMemcpyAsync(h.data(), d->mon_part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipMemcpyAsync(&y, d->d_dyn + s, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    double tm = 0, mnT = 1e9, mxT = -1e9, mnG = 1e9, mxG = -1e9, bad = 0;
    for (int b = 0; b < nb; ++b) {
        const double* q = &h[6 * (size_t)b];
        tm += q[0]; mnT = fmin(mnT, q[1]); mxT = fmax(mxT, q[2]); mnG = fmin(mnG, q[3]); mxG = fmax(mxG, q[4]);
        bad = fmax(bad, q[5]);
    }
    rep->step = y.step - 1;
    rep->dt = y.dt; rep->max_velocity = y.maxvel;
    rep->total_mass = tm; rep->min_T = mnT; rep->max_T = mxT; rep->min_G = mnG; rep->max_G = mxG;
    rep->min_Y_fuel = 0.0; rep->nonfinite = bad > 0.0;
    rep->proj_residual = y.proj_res; rep->proj_residual0 = y.proj_res0;
    return LES_OK;
}

static int slot_monitor(les_gpu* c, int s, les_step_report* rep)
{
    return c->d->precision == LES_PREC_FP32 ? slot_monitor_t<float>(c, s, rep) : slot_monitor_t<double>(c, s, rep);
}

// ---------------------------------------------------------------------------
// autotuner: times each (tile width, XCD mapping) candidate on a scratch pool
// filled with replicas of the probe run (as many as the pool's slots and cells
// hold, so the timed launches have the pool's occupancy) and keeps the fastest
// ---------------------------------------------------------------------------
static int pool_autotune(les_gpu* c, const les_params& probe)
{
    DuctState* d = c->d;
    const int tks[4] = {0, 32, 64, 128};
    double best = 1e30;
    int btk = d->tk, bx = d->xcd_map;
    for (int a = 0; a < 4; ++a)
        for (int x = 0; x < 2; ++x) {
            const long long pc = (long long)probe.nx * probe.ny * probe.nz;
            long long reps = d->pool_cells / pc;
            if (reps > d->max_slots) reps = d->max_slots;
            if (reps < 1) reps = 1;
            les_batch_config cfg;
            memset(&cfg, 0, sizeof(cfg));
            cfg.max_slots = (int)reps;
            cfg.pool_cells = reps * pc;
            cfg.max_nx = probe.nx; cfg.max_ny = probe.ny; cfg.max_nz = probe.nz;
            cfg.max_vortices = probe.num_vortices > 0 ? probe.num_vortices : 1;
            cfg.precision = d->precision;
            cfg.tile_k = tks[a];
            cfg.xcd_map = x;
            cfg.table_t0 = probe.table_t0; cfg.table_dt = probe.table_dt; cfg.table_n = probe.table_n;
            cfg.point_cap = 16;
            les_gpu* t = new (std::nothrow) les_gpu();
            if (!t) return LES_ERR_HIP;
            t->p = probe;
            t->stream = c->stream;
            int e = pool_create(t, cfg);
            for (long long r = 0; r < reps && e == LES_OK; ++r) {
                int slot = -1;
                e = pool_admit(t, probe, 0, &slot);
                if (e == LES_OK) e = slot_init(t, slot, probe);
            }
            if (e == LES_OK) e = pool_steps(t, 2);
            float ms = 0.f;
            if (e == LES_OK && hipEventRecord(c->ev[2], c->stream) == hipSuccess) {
                e = pool_steps(t, 3);
                (void)hipEventRecord(c->ev[3], c->stream);
                (void)hipEventSynchronize(c->ev[3]);
                (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
            }
            if (e == LES_OK && ms > 0 && ms < best) {
                best = ms;
                btk = tks[a];
                bx = x;
            }
            (void)hipStreamSynchronize(c->stream);
            t->stream = nullptr;
            les_gpu_destroy(t);
            if (e != LES_OK) {
                snprintf(c->err, sizeof(c->err), "autotune candidate tile_k=%d xcd_map=%d failed (%d)", tks[a], x, e);
                return e;
            }
        }
    d->tk = btk;
    d->xcd_map = bx;
    d->tuned_ms = best / 3.0;
    d->graph_ok[0] = d->graph_ok[1] = false;
    return LES_OK;
}

// ---------------------------------------------------------------------------
// downloads of one slot
// ---------------------------------------------------------------------------
static int slot_download_field(les_gpu* c, int s, const char* name, double* out)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    struct { const char* nm; void* p; } tab[] = {
        {"u_f", d->P.F[0]}, {"v_f", d->P.F[1]}, {"w_f", d->P.F[2]}, {"T_f", d->P.F[3]}, {"phi_f", d->P.F[4]},
        {"nu_t", d->P.nut}, {"ST", d->P.st},
        {"rhou", d->P.A[0]}, {"rhov", d->P.A[1]}, {"rhow", d->P.A[2]}, {"phi", d->P.A[3]}, {"T", d->P.A[4]}};
    for (auto& t : tab)
        if (strcmp(t.nm, name) == 0) {
            int e = slot_xfer(c, s, t.p, out, false);
            if (e != LES_OK) return e;
            DHIPC(hipStreamSynchronize(c->stream));
            return LES_OK;
        }
    const size_t PL = (size_t)D.g.ny * D.g.nz;
    struct { const char* nm; const double* p; size_t cnt; } dt[] = {
        {"inlet", d->P.inlet + D.poff, 3 * PL}, {"eddies", d->P.eddy + 7 * D.eoff, 7 * (size_t)D.k.nvort},
        {"profiles", d->P.prof + D.xoff, 10 * (size_t)D.g.nx}};
    for (auto& t : dt)
        if (strcmp(t.nm, name) == 0) {
            DHIPC(hipMemcpyAsync(out, t.p, t.cnt * 8, hipMemcpyDeviceToHost, c->stream));
            DHIPC(hipStreamSynchronize(c->stream));
            return LES_OK;
        }
    if (strcmp(name, "dyn") == 0) {          // dt, mean density, max |u|, time of the last step
        SlotDyn y;
        DHIPC(hipMemcpyAsync(&y, d->d_dyn + s, sizeof(y), hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        out[0] = y.dt; out[1] = y.rho_avg; out[2] = y.maxvel; out[3] = y.time;
        return LES_OK;
    }
    if (strncmp(name, "stat", 4) == 0 && d->P.stat[0]) {
        int a = atoi(name + 4);
        if (a < 0 || a > 9) return LES_ERR_PARAM;
        DHIPC(hipMemcpyAsync(out, d->P.stat[a] + D.off, 8 * (size_t)D.g.n, hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        return LES_OK;
    }
    snprintf(c->err, sizeof(c->err), "unknown duct field %s", name);
    return LES_ERR_PARAM;
}

static int slot_download(les_gpu* c, int s, double* const* f, int nf)
{
    for (int a = 0; a < nf && a < DNF; ++a) {
        int e = slot_xfer(c, s, c->d->P.A[a], f[a], false);
        if (e != LES_OK) return e;
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

// fixed-point series since the last fetch: out[step][point][u, v, w, G]; the
// ring keeps the newest point_cap steps (the sample of step t sits at t % cap)
static int slot_point_fetch(les_gpu* c, int s, double* out, int max_steps, int* n_steps)
{
    DuctState* d = c->d;
    const int total = d->host_step[s] - d->point_base[s];
    const int avail = total < d->point_cap ? total : d->point_cap;
    const int n = avail < max_steps ? avail : max_steps;
    const int first = d->host_step[s] - avail;
    const double* ring = d->P.points + d->desc[s].ptoff;
    for (int a = 0; a < n; ++a) {
        const int r = (first + a) % d->point_cap;
        const int np = d->desc[s].npts;
        DHIPC(hipMemcpyAsync(out + (size_t)a * np * 4, ring + (size_t)r * DUCT_NPTS * 4, (size_t)np * 4 * 8,
                             hipMemcpyDeviceToHost, c->stream));
    }
    DHIPC(hipStreamSynchronize(c->stream));
    d->point_base[s] = d->host_step[s];
    *n_steps = n;
    return LES_OK;
}

// probe rows recorded since the last drain (the newest probe_cap steps), in
// the row format of the reference's probe CSV (SAMPLING_RANGE 1.0 m, :280)
static int slot_probe_drain(les_gpu* c, int s, const double* probe_x, std::vector<les_probe_row>& out)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[s];
    if (d->probe_cap <= 0 || D.nprobe <= 0) { d->probe_base[s] = d->host_step[s]; return LES_OK; }
    const int total = d->host_step[s] - d->probe_base[s];
    const int avail = total < d->probe_cap ? total : d->probe_cap;
    const int first = d->host_step[s] - avail;
    const size_t stride = 2 + 6 * (size_t)d->ny_max;
    std::vector<double> h((size_t)LES_NPROBE_MAX * stride);
    for (int a = 0; a < avail; ++a) {
        const int step = first + a;
        const double* src = d->P.probe + D.proff + (size_t)(step % d->probe_cap) * LES_NPROBE_MAX * stride;
        DHIPC(hipMemcpyAsync(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
        DHIPC(hipStreamSynchronize(c->stream));
        for (int n = 0; n < D.nprobe; ++n) {
            const double* e = &h[(size_t)n * stride];
            if (e[0] != 1.0) continue;
            for (int j = 0; j < D.g.ny; ++j) {
                const double* r = e + 2 + 6 * j;
                double yp = j * D.g.dy, yr = yp - e[1];
                if (fabs(yr) > 1.0) continue;
                les_probe_row pr;
                pr.timestep = step; pr.station = n; pr.x_location = probe_x[n];
                pr.y_flame = e[1]; pr.y_physical = yp; pr.y_relative = yr;
                pr.u = r[0]; pr.v = r[1]; pr.w = r[2]; pr.G = r[3]; pr.nu_t = r[4]; pr.S_mag = r[5];
                out.push_back(pr);
            }
        }
    }
    d->probe_base[s] = d->host_step[s];
    return LES_OK;
}

// ---------------------------------------------------------------------------
// single-run duct context: a pool of one slot sized to its grid, 64 x 4 tiles
// ---------------------------------------------------------------------------
int duct_create(les_gpu* c)
{
    const les_params& p = c->p;
    if (p.integrator != LES_INT_SSPRK3 || (p.advection != LES_ADV_WENO5 && p.advection != LES_ADV_MUSCL_MINMOD))
        return LES_ERR_PARAM;
    les_batch_config cfg;
    les_batch_config_default(&cfg);
    cfg.max_slots = 1;
    cfg.pool_cells = (long long)p.nx * p.ny * p.nz;
    cfg.max_nx = p.nx; cfg.max_ny = p.ny; cfg.max_nz = p.nz;
    cfg.max_vortices = p.num_vortices > 0 ? p.num_vortices : 1;
    cfg.precision = p.precision;
    // the widest tile row that divides nz (no idle lanes), else flat tiles
    cfg.tile_k = (p.nz % 64 == 0) ? 64 : ((p.nz % 32 == 0) ? 32 : LES_TILE_FLAT);
    cfg.xcd_map = 1;
    cfg.point_cap = 4096;
    cfg.probe_cap = p.num_probe_x > 0 ? 64 : 0;
    cfg.stats_per_cell = (p.stats_start_step >= 0 && p.stats_per_cell) ? 1 : 0;
    cfg.table_t0 = p.table_t0; cfg.table_dt = p.table_dt; cfg.table_n = p.table_n;
    int e = pool_create(c, cfg);
    if (e != LES_OK) return e;
    int slot = -1;
    e = pool_admit(c, p, 0, &slot);
    if (e != LES_OK) {
        snprintf(c->err, sizeof(c->err), "duct run does not fit its own pool (status %d)", e);
        return e;
    }
    c->attr_h.assign((size_t)c->g.n, 3);
    return LES_OK;
}

void duct_destroy(les_gpu* c)
{
    if (!c->d) return;
    for (auto& g : c->d->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->d->aux) {
        (void)hipStreamSynchronize(c->d->aux);
        (void)hipStreamDestroy(c->d->aux);
    }
    for (int a = 0; a < 2; ++a) {
        if (c->d->evf[a]) (void)hipEventDestroy(c->d->evf[a]);
        if (c->d->evj[a]) (void)hipEventDestroy(c->d->evj[a]);
    }
    delete c->d;
    c->d = nullptr;
}

void duct_params_changed(les_gpu* c)
{
    if (!c->d) return;
    c->d->desc[0].k = make_duct_consts(c->p);
    (void)pool_sync_tables(c);
}

int duct_init_state(les_gpu* c)
{
    int e = slot_init(c, 0, c->p);
    if (e != LES_OK) return e;
    c->step = 0;
    c->stat_count = 0;
    c->probe.clear();
    return LES_OK;
}

int duct_upload_state(les_gpu* c, const double* const* f, int nf) { return slot_upload(c, 0, f, nf); }
int duct_download_state(les_gpu* c, double* const* f, int nf) { return slot_download(c, 0, f, nf); }
int duct_download_field(les_gpu* c, const char* name, double* out) { return slot_download_field(c, 0, name, out); }
int duct_point_fetch(les_gpu* c, double* out, int max_steps, int* n_steps)
{
    return slot_point_fetch(c, 0, out, max_steps, n_steps);
}

int duct_npoints(const les_gpu* c) { return c->d->desc[0].npts; }

int duct_set_points(les_gpu* c, const int* cells, int n)
{
    SlotDesc& D = c->d->desc[0];
    for (int a = 0; a < n; ++a) D.point_idx[a] = cells[a];
    D.npts = n;
    c->d->point_base[0] = c->d->host_step[0];
    return pool_sync_tables(c);
}

int duct_step(les_gpu* c, int n_steps, les_step_report* rep)
{
    DuctState* d = c->d;
    const int st0 = c->p.stats_start_step;
    for (int a = 0; a < n_steps; ++a)
        if (st0 >= 0 && d->host_step[0] + a >= st0) c->stat_count++;
    // probe rows are drained into the context's row list at least every
    // probe_cap steps (the ring holds that many); the step time excludes the
    // drains
    const int chunk = d->probe_cap > 0 ? d->probe_cap : (n_steps > 0 ? n_steps : 1);
    double ms_sum = 0.0;
    for (int done = 0; done < n_steps;) {
        const int n = n_steps - done < chunk ? n_steps - done : chunk;
        DHIPC(hipEventRecord(c->ev[2], c->stream));
        int e = pool_steps(c, n);
        if (e != LES_OK) return e;
        DHIPC(hipEventRecord(c->ev[3], c->stream));
        DHIPC(hipEventSynchronize(c->ev[3]));
        float ms = 0.f;
        DHIPC(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        ms_sum += ms;
        done += n;
        if (d->probe_cap > 0) {
            e = slot_probe_drain(c, 0, c->p.probe_x, c->probe);
            if (e != LES_OK) return e;
        }
    }
    c->step_ms = n_steps > 0 ? ms_sum / n_steps : 0.0;
    c->step += n_steps;
    int e = LES_OK;
    if (rep) {
        e = slot_monitor(c, 0, rep);
        if (e != LES_OK) return e;
        rep->step = c->step - 1;
        if (rep->nonfinite) {
            snprintf(c->err, sizeof(c->err), "non-finite value in the state at step %d", rep->step);
            return LES_ERR_NAN;
        }
    }
    return LES_OK;
}

// Testing hook: the RHS L(q) of the current state (a stage-0 evaluation whose
// combined output goes to the dead stage buffers B).
int duct_rhs(les_gpu* c, double* const* out, int n_out)
{
    if (n_out < DNF) return LES_ERR_PARAM;
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    const size_t n = (size_t)D.g.n;
    std::vector<void*> Lbuf(DNF, nullptr);
    for (int f = 0; f < DNF; ++f) {
        DHIPC(hipMalloc(&Lbuf[f], n * d->esize));
        DHIPC(hipMemsetAsync(Lbuf[f], 0, n * d->esize, c->stream));   // boundary planes read as 0
    }
    int e = tk_dispatch(d->tk, [&](auto K) {
        constexpr int TK = decltype(K)::value;
        if (d->precision == LES_PREC_FP32) return stage_launch<float, TK>(c, 0, (float* const*)Lbuf.data());
        return stage_launch<double, TK>(c, 0, (double* const*)Lbuf.data());
    });
    // slot 0 starts at cell 0 of the arenas, so slot-local buffers xfer as slot 0
    for (int f = 0; f < DNF && e == LES_OK; ++f) e = slot_xfer(c, 0, (char*)Lbuf[f] - (size_t)D.off * d->esize, out[f], false);
    DHIPC(hipStreamSynchronize(c->stream));
    for (int f = 0; f < DNF; ++f) (void)hipFree(Lbuf[f]);
    return e;
}

// ---------------------------------------------------------------------------
// batch ABI
// ---------------------------------------------------------------------------
struct les_batch {
    les_gpu* c = nullptr;
};

static int batch_config_check(const les_batch_config* cfg)
{
    if (!cfg || cfg->max_slots < 1 || cfg->max_slots > 8192 || cfg->pool_cells < 1 || cfg->max_nx < 6 ||
        cfg->max_ny < 1 || cfg->max_nz < 1 || cfg->max_vortices < 0 || !tk_valid(cfg->tile_k) ||
        (cfg->precision != LES_PREC_FP64 && cfg->precision != LES_PREC_FP32) || cfg->point_cap < 1 ||
        cfg->table_n < 2 || cfg->table_dt <= 0.0)
        return LES_ERR_PARAM;
    return LES_OK;
}

extern "C" int les_batch_pool_bytes(const les_batch_config* cfg, size_t* bytes)
{
    if (batch_config_check(cfg) != LES_OK || !bytes) return LES_ERR_PARAM;
    const size_t es = cfg->precision == LES_PREC_FP32 ? 4 : 8;
    const size_t n = (size_t)cfg->pool_cells, S = (size_t)cfg->max_slots;
    const size_t tiles = n / 256 + S * cfg->max_nx * (cfg->max_ny / 2 + cfg->max_nz / 32 + 2) + 64;
    size_t b = 17 * n * es;                                   // A, B, F (5 each), nu_t, S_T
    if (cfg->stats_per_cell) b += 80 * n;
    if (cfg->precision == LES_PREC_FP32) b += 8 * n;           // FP64 staging of one run
    b += tiles * (8 + 8 + 48 + 16 + 4);                       // rho, |u|, nu/alpha, residual partials, tile list
    const size_t nv = cfg->max_vortices > 0 ? cfg->max_vortices : 1;
    b += S * 8 * (7 * nv + 3 * (size_t)cfg->max_ny * cfg->max_nz + 10 * (size_t)cfg->max_nx +
                  4 * (size_t)DUCT_NPTS * cfg->point_cap + (size_t)cfg->max_ny * ((cfg->max_nz + 255) / 256) +
                  10 * (size_t)cfg->max_nx * STAT_CHUNKS);
    b += S * (sizeof(SlotDesc) + sizeof(SlotDyn)) + 16 * (size_t)cfg->table_n + 8 * 6 * 1024;
    b += S * 8 * 3 * (size_t)cfg->max_ny * cfg->max_nz * ((nv + SEM_CHUNK - 1) / SEM_CHUNK);   // SEM partials
    *bytes = b;
    return LES_OK;
}

extern "C" int les_batch_create(const les_batch_config* cfg, const les_params* probe, les_batch** out)
{
    if (!out) return LES_ERR_PARAM;
    *out = nullptr;
    if (batch_config_check(cfg) != LES_OK) return LES_ERR_PARAM;
    if (cfg->autotune && (!probe || les_params_validate(probe) != LES_OK || probe->mode != LES_MODE_DUCT))
        return LES_ERR_PARAM;
    les_batch* b = new (std::nothrow) les_batch();
    if (!b) return LES_ERR_HIP;
    les_gpu* c = new (std::nothrow) les_gpu();
    if (!c) { delete b; return LES_ERR_HIP; }
    b->c = c;
    if (probe) c->p = *probe;
    else les_params_default(&c->p, LES_PRESET_ENSEMBLE);
    c->p.precision = cfg->precision;
    les_mixture_synthetic(&c->mix);
    int rc = LES_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = LES_ERR_HIP;
    for (int a = 0; a < 4 && rc == LES_OK; ++a)
        if (hipEventCreate(&c->ev[a]) != hipSuccess) rc = LES_ERR_HIP;
    if (rc == LES_OK) rc = pool_create(c, *cfg);
    if (rc == LES_OK && cfg->autotune) rc = pool_autotune(c, *probe);
    if (rc != LES_OK) {
        fprintf(stderr, "les_batch_create: %s\n", c->err);
        les_gpu_destroy(c);
        delete b;
        return rc;
    }
    *out = b;
    return LES_OK;
}

extern "C" void les_batch_destroy(les_batch* b)
{
    if (!b) return;
    les_gpu_destroy(b->c);
    delete b;
}

extern "C" const char* les_batch_last_error(const les_batch* b) { return b && b->c ? b->c->err : "null context"; }

extern "C" int les_batch_tuning(const les_batch* b, int* tile_k, int* grid_blocks, int* xcd_map, double* tuned_ms)
{
    if (!b) return LES_ERR_PARAM;
    const DuctState* d = b->c->d;
    if (tile_k) *tile_k = d->tk;
    if (grid_blocks) *grid_blocks = d->grid;
    if (xcd_map) *xcd_map = d->xcd_map;
    if (tuned_ms) *tuned_ms = d->tuned_ms;
    return LES_OK;
}

static int batch_slot_ok(les_batch* b, int slot)
{
    if (!b || slot < 0 || slot >= b->c->d->max_slots || !b->c->d->taken[slot]) {
        if (b) snprintf(b->c->err, sizeof(b->c->err), "slot %d is not admitted", slot);
        return 0;
    }
    return 1;
}

extern "C" int les_batch_admit(les_batch* b, const les_params* p, int n_steps, int init, int* slot)
{
    if (!b || !p || !slot || n_steps < 0) return LES_ERR_PARAM;
    les_gpu* c = b->c;
    *slot = -1;
    if (les_params_validate(p) != LES_OK || p->mode != LES_MODE_DUCT) {
        snprintf(c->err, sizeof(c->err), "batched runs are duct-mode runs with valid parameters");
        return LES_ERR_PARAM;
    }
    const DuctState* d = c->d;
    if (p->table_t0 != c->p.table_t0 || p->table_dt != c->p.table_dt || p->table_n != c->p.table_n ||
        (p->stats_per_cell && !d->stats_cells)) {
        snprintf(c->err, sizeof(c->err), "run needs a property table or per-cell statistics the pool lacks");
        return LES_ERR_ADMISSION;
    }
    int e = pool_admit(c, *p, n_steps, slot);
    if (e == LES_ERR_ADMISSION) {
        snprintf(c->err, sizeof(c->err), "no free slot, cells or tiles for a %dx%dx%d run", p->nx, p->ny, p->nz);
        return e;
    }
    if (e != LES_OK) return e;
    if (init) {
        e = slot_init(c, *slot, *p);
        if (e != LES_OK) { (void)pool_release(c, *slot); *slot = -1; return e; }
    }
    return LES_OK;
}

extern "C" int les_batch_release(les_batch* b, int slot)
{
    if (!batch_slot_ok(b, slot)) return LES_ERR_PARAM;
    return pool_release(b->c, slot);
}

extern "C" int les_batch_step(les_batch* b, int n_steps, double* ms_per_step)
{
    if (!b || n_steps < 0) return LES_ERR_PARAM;
    les_gpu* c = b->c;
    DHIPC(hipEventRecord(c->ev[2], c->stream));
    int e = pool_steps(c, n_steps);
    if (e != LES_OK) return e;
    DHIPC(hipEventRecord(c->ev[3], c->stream));
    DHIPC(hipEventSynchronize(c->ev[3]));
    float ms = 0.f;
    DHIPC(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
    c->step_ms = n_steps > 0 ? ms / n_steps : 0.0;
    if (ms_per_step) *ms_per_step = c->step_ms;
    return LES_OK;
}

extern "C" int les_batch_slot_info(les_batch* b, int slot, les_slot_info* info)
{
    if (!b || !info || slot < 0 || slot >= b->c->d->max_slots) return LES_ERR_PARAM;
    les_gpu* c = b->c;
    DuctState* d = c->d;
    memset(info, 0, sizeof(*info));
    if (!d->taken[slot]) return LES_OK;
    const SlotDesc& D = d->desc[slot];
    SlotDyn y;
    DHIPC(hipMemcpyAsync(&y, d->d_dyn + slot, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    info->active = D.active;
    info->done = y.done;
    info->step = y.step;
    info->n_steps = D.n_steps;
    info->time = y.time;
    info->dt = y.dt;
    info->max_velocity = y.maxvel;
    info->nx = D.g.nx; info->ny = D.g.ny; info->nz = D.g.nz;
    info->cells = D.g.n;
    return LES_OK;
}

extern "C" int les_batch_poll(les_batch* b, int* slots, int max_slots, int* n_done)
{
    if (!b || !n_done || (max_slots > 0 && !slots)) return LES_ERR_PARAM;
    const DuctState* d = b->c->d;
    int n = 0;
    for (int s = 0; s < d->max_slots; ++s) {
        const SlotDesc& D = d->desc[s];
        if (d->taken[s] && D.active && D.n_steps > 0 && d->host_step[s] >= D.n_steps) {
            if (n < max_slots) slots[n] = s;
            ++n;
        }
    }
    *n_done = n;
    return LES_OK;
}

extern "C" int les_batch_upload(les_batch* b, int slot, const double* const* fields, int n_fields)
{
    if (!batch_slot_ok(b, slot) || !fields) return LES_ERR_PARAM;
    return slot_upload(b->c, slot, fields, n_fields);
}

extern "C" int les_batch_download(les_batch* b, int slot, double* const* fields, int n_fields)
{
    if (!batch_slot_ok(b, slot) || !fields) return LES_ERR_PARAM;
    return slot_download(b->c, slot, fields, n_fields);
}

extern "C" int les_batch_download_field(les_batch* b, int slot, const char* name, double* out)
{
    if (!batch_slot_ok(b, slot) || !name || !out) return LES_ERR_PARAM;
    return slot_download_field(b->c, slot, name, out);
}

extern "C" int les_batch_report(les_batch* b, int slot, les_step_report* rep)
{
    if (!batch_slot_ok(b, slot) || !rep) return LES_ERR_PARAM;
    return slot_monitor(b->c, slot, rep);
}

extern "C" int les_batch_probe_fetch(les_batch* b, int slot, les_probe_row* rows, int max_rows, int* n_rows)
{
    if (!batch_slot_ok(b, slot) || !rows || !n_rows || max_rows < 0) return LES_ERR_PARAM;
    les_gpu* c = b->c;
    DuctState* d = c->d;
    std::vector<les_probe_row>& v = d->probe_rows[slot];
    int e = slot_probe_drain(c, slot, d->probe_x[slot].data(), v);
    if (e != LES_OK) return e;
    const int n = (int)v.size() < max_rows ? (int)v.size() : max_rows;
    if (n > 0) memcpy(rows, v.data(), (size_t)n * sizeof(les_probe_row));
    v.erase(v.begin(), v.begin() + n);
    *n_rows = n;
    return LES_OK;
}

extern "C" int les_batch_point_fetch(les_batch* b, int slot, double* out, int max_steps, int* n_steps)
{
    if (!batch_slot_ok(b, slot) || !out || !n_steps || max_steps < 0) return LES_ERR_PARAM;
    return slot_point_fetch(b->c, slot, out, max_steps, n_steps);
}

// ---------------------------------------------------------------------------
// resume support of a single-run duct context (slot 0)
// ---------------------------------------------------------------------------
int duct_upload_field(les_gpu* c, const char* name, const double* data)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    if (strcmp(name, "profiles") == 0) {
        DHIPC(hipMemcpyAsync(d->P.prof + D.xoff, data, 8 * 10 * (size_t)D.g.nx, hipMemcpyHostToDevice, c->stream));
    } else if (strcmp(name, "eddies") == 0) {
        DHIPC(hipMemcpyAsync(d->P.eddy + 7 * D.eoff, data, 8 * 7 * (size_t)D.k.nvort, hipMemcpyHostToDevice, c->stream));
    } else if (strncmp(name, "stat", 4) == 0 && d->P.stat[0]) {
        int a = atoi(name + 4);
        if (a < 0 || a > 9) return LES_ERR_PARAM;
        DHIPC(hipMemcpyAsync(d->P.stat[a] + D.off, data, 8 * (size_t)D.g.n, hipMemcpyHostToDevice, c->stream));
    } else {
        snprintf(c->err, sizeof(c->err), "duct field %s cannot be uploaded", name);
        return LES_ERR_PARAM;
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}

struct AuxDuct {
    char magic[8];
    int mode, step, stat_count, nvort, nx, pad;
    double time, max_nu, max_alpha;
};

static const char DUCT_AUX_MAGIC[8] = {'L', 'E', 'S', 'A', 'U', 'X', '1', 0};

int duct_export_aux(les_gpu* c, void* buf, size_t* bytes)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    const size_t ne = 7 * (size_t)D.k.nvort, np = 10 * (size_t)D.g.nx;
    const size_t need = sizeof(AuxDuct) + (ne + np) * 8;
    if (!buf || *bytes < need) { *bytes = need; return buf ? LES_ERR_PARAM : LES_OK; }
    AuxDuct a;
    memset(&a, 0, sizeof(a));
    memcpy(a.magic, DUCT_AUX_MAGIC, 8);
    a.mode = LES_MODE_DUCT; a.step = c->step; a.stat_count = c->stat_count; a.nvort = D.k.nvort; a.nx = D.g.nx;
    SlotDyn y;
    std::vector<double> dp(6 * (size_t)D.ntiles);
    DHIPC(hipMemcpyAsync(&y, d->d_dyn, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    for (int stg = 0; stg < 3; ++stg)
        DHIPC(hipMemcpyAsync(dp.data() + 2 * (size_t)stg * D.ntiles, d->P.dpart + 2 * ((size_t)stg * d->tile_cap + D.tile0),
                             16 * (size_t)D.ntiles, hipMemcpyDeviceToHost, c->stream));
    char* o = (char*)buf + sizeof(a);
    DHIPC(hipMemcpyAsync(o, d->P.eddy + 7 * D.eoff, ne * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipMemcpyAsync(o + ne * 8, d->P.prof + D.xoff, np * 8, hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    // the diffusive DT history: maxima of nu and alpha over the last step's stages
    double mn = 0.0, ma = 0.0;
    for (size_t t = 0; t < 3 * (size_t)D.ntiles; ++t) { mn = fmax(mn, dp[2 * t]); ma = fmax(ma, dp[2 * t + 1]); }
    a.time = y.time; a.max_nu = mn; a.max_alpha = ma;
    memcpy(buf, &a, sizeof(a));
    *bytes = need;
    return LES_OK;
}

// after les_gpu_upload_state (which recomputes the velocity maxima of the new
// state and clears the diffusive history): restores step, run time, eddies,
// profiles and the nu/alpha maxima the next DT needs
int duct_import_aux(les_gpu* c, const void* buf, size_t bytes)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    AuxDuct a;
    if (bytes < sizeof(a)) { snprintf(c->err, sizeof(c->err), "aux blob too short"); return LES_ERR_PARAM; }
    memcpy(&a, buf, sizeof(a));
    const size_t ne = 7 * (size_t)D.k.nvort, np = 10 * (size_t)D.g.nx;
    if (memcmp(a.magic, DUCT_AUX_MAGIC, 8) != 0 || a.mode != LES_MODE_DUCT || a.nvort != D.k.nvort || a.nx != D.g.nx ||
        bytes != sizeof(a) + (ne + np) * 8 || a.step < 0) {
        snprintf(c->err, sizeof(c->err), "aux blob does not match this context");
        return LES_ERR_PARAM;
    }
    const char* in = (const char*)buf + sizeof(a);
    DHIPC(hipMemcpyAsync(d->P.eddy + 7 * D.eoff, in, ne * 8, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->P.prof + D.xoff, in + ne * 8, np * 8, hipMemcpyHostToDevice, c->stream));
    SlotDyn y;
    DHIPC(hipMemcpyAsync(&y, d->d_dyn, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    y.step = a.step;
    y.time = a.time;
    y.done = 0;
    const double h[2] = {a.max_nu, a.max_alpha};
    DHIPC(hipMemcpyAsync(d->d_dyn, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    DHIPC(hipMemcpyAsync(d->P.dpart + 2 * (size_t)D.tile0, h, 16, hipMemcpyHostToDevice, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    d->host_step[0] = a.step;
    d->point_base[0] = a.step;
    d->probe_base[0] = a.step;
    c->step = a.step;
    c->stat_count = a.stat_count;
    c->probe.clear();
    return LES_OK;
}

int duct_set_step(les_gpu* c, int step)
{
    DuctState* d = c->d;
    SlotDyn y;
    DHIPC(hipMemcpyAsync(&y, d->d_dyn, sizeof(y), hipMemcpyDeviceToHost, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    y.step = step;
    y.done = 0;
    DHIPC(hipMemcpyAsync(d->d_dyn, &y, sizeof(y), hipMemcpyHostToDevice, c->stream));
    DHIPC(hipStreamSynchronize(c->stream));
    d->host_step[0] = d->point_base[0] = d->probe_base[0] = step;
    c->step = step;
    return LES_OK;
}

int duct_init_sem(les_gpu* c)
{
    DuctState* d = c->d;
    const SlotDesc& D = d->desc[0];
    if (D.k.nvort > 0) {
        hipLaunchKernelGGL(k_duct_eddy_init, dim3((D.k.nvort + 255) / 256), dim3(256), 0, c->stream, d->P, 0);
        DLAUNCH_CHECK();
    }
    DHIPC(hipStreamSynchronize(c->stream));
    return LES_OK;
}
