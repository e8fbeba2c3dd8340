// les_device.hpp - shared device-side definitions of the gfx950 LES engine.
//
// Layout: every field is one contiguous [i][j][k] block (k fastest), the
// reference's Grid3D (les_combustion_solver.c:286).  One thread owns one cell;
// consecutive threads own consecutive k, so every stencil load of a wavefront
// is a 64-wide coalesced row.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/les_gpu.h"

#define LES_NSP LES_NSPECIES
#define LES_NRHS (5 + LES_NSPECIES)

// physical constants of the reference (les_combustion_solver.c:136-141)
#define LES_R_UNIV 8.314462
#define LES_PATM 101325.0
#define LES_P_CONST 101325.0
#define LES_G_GRAV (-9.81)
#define LES_SMALL 1e-10

// Grid geometry and scheme constants handed to every kernel by value.
struct GridDims {
    int nx, ny, nz;
    long long n;          // nx*ny*nz
    long long sx, sy;     // strides of i and j (sz = 1)
    double dx, dy, dz;
};

// Scalars the reference keeps in globals or #defines, resolved on the host.
struct RefConsts {
    double inletY[LES_NSP], ambY[LES_NSP];
    double st_o2, st_co2, st_h2o;
    double mw[LES_NSP];          // g/mol
    double inlet_velocity, inlet_temperature, ambient_temperature, burnt_temperature;
    double noise_amp;            // white-noise amplitude (m/s)
    double filter_wc;
    double cs_u, cs_b, cs_f;
    double s_l, nu_unburnt, tau, d3;
    double pr_t, sc_t;
    double accel_strength, accel_width, accel_center, weight_x;
    double tip_radius, max_force, ramp_distance, lz_half;
    double relaxation;
    int    transition_step;
};

__host__ __device__ __forceinline__ long long lidx(const GridDims& g, int i, int j, int k)
{
    return ((long long)i * g.ny + j) * g.nz + k;
}

// 3-D launch helper: blockDim (64, 4, 1) over (k, j, i).
struct Launch3 {
    dim3 grid, block;
};

static inline Launch3 launch3(int nx, int ny, int nz)
{
    Launch3 L;
    L.block = dim3(64, 4, 1);
    L.grid = dim3((nz + 63) / 64, (ny + 3) / 4, nx);
    return L;
}

// Positive doubles order like their bit patterns: an unsigned atomicMax on the
// bits is an exact, order-independent max reduction.
__device__ __forceinline__ void atomic_max_pos(unsigned long long* dst, double v)
{
    if (v > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(v));
}
