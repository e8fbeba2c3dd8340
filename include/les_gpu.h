/*
 * les_gpu.h - C ABI of the MI355X (gfx950) LES time-step engine.
 *
 * The reference solver (les_combustion_solver.c) has no library API: its hot
 * path is a set of void functions over process-global Grid3D arrays, called
 * only from main() (les_combustion_solver.c:4390-4603).  This header is the
 * drop-in boundary a host program calls instead.  Every entry point is plain C:
 * POD structs, pointers and sizes, int status codes, no exceptions, no exit().
 *
 * Field layout everywhere: row-major [i][j][k], k fastest, exactly the
 * reference's Grid3D (les_combustion_solver.c:286), one contiguous block of
 * nx*ny*nz values per field.
 *
 * Reference interfaces replaced (file:line in les_combustion_solver.c):
 *   les_gpu_step          <- the time loop body :4390-4603 (DT :4399-4411, SEM :4414,
 *                            RK4 :4420-4544, BCs :4555, sampling :4558-4574)
 *   les_gpu_rhs           <- calculate_RHS :2494-2963 (one RHS evaluation)
 *   les_gpu_upload_state  <- load_quick_checkpoint :3701-3737 (field order u,v,w,T,G,rho,Y0..Y4)
 *   les_gpu_download_state<- save_quick_checkpoint :3671-3698 (same order)
 *   les_gpu_init_state    <- initialize :1085-1133 + setting :1135-1285
 *   les_gpu_probe_fetch   <- record_flame_region_data :3403-3471 (device sampling, host CSV)
 *   les_gpu_download_field("stat0".."stat9")
 *                         <- update_time_averages :3167-3192 accumulators (FP64)
 *   les_gpu_set_params    <- the #define block :11-284 (constants become runtime values)
 *   les_mixture_load      <- load_chem_bin :324-356 / load_tran_bin :359-382
 */
#ifndef LES_GPU_H
#define LES_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference exits instead; see :1299, :3858) ---- */
#define LES_OK               0
#define LES_ERR_HIP         -1   /* a HIP runtime call failed */
#define LES_ERR_PARAM       -2   /* invalid argument or parameter */
#define LES_ERR_NAN         -3   /* a non-finite value appeared in the state */
#define LES_ERR_ADMISSION   -4   /* serving: request does not fit */
#define LES_ERR_IO          -5   /* file could not be read or written */
#define LES_ERR_STATE       -6   /* call not valid in the current state */

/* ---- scheme selectors ---- */
#define LES_MODE_REFERENCE   0   /* the reference's scheme, quirks included */
#define LES_MODE_DUCT        1   /* periodic duct, SSP-RK3, table properties, level-set G */

#define LES_INT_RK4          0   /* classic RK4 (:4416-4544) */
#define LES_INT_SSPRK3       1   /* Shu-Osher SSP-RK3 */

#define LES_ADV_TENO5        0   /* TENO5 for G (:2190-2298), MUSCL-superbee for T/Y (:2037-2159) */
#define LES_ADV_WENO5        1   /* Jiang-Shu WENO5 upwind reconstruction */
#define LES_ADV_MUSCL_MINMOD 2   /* MUSCL with the minmod limiter */

#define LES_PREC_FP64        0
#define LES_PREC_FP32        1

#define LES_NSPECIES         5   /* C3H8, O2, N2, CO2, H2O (:66, :129-133) */
#define LES_NFIELDS          11  /* u, v, w, T, G, rho, Y0..Y4 */
#define LES_NPROBE_MAX       16  /* flame-region x-stations */
#define LES_NPOINT_MAX       64  /* point time series (:198) */

/* Per-species data as read from chem.bin / tran.bin (:292-300, :305). */
typedef struct les_species {
    char   name[17];
    double mw;              /* g/mol */
    double t_mid;           /* K */
    double nasa_high[7];
    double nasa_low[7];
    double lambda_fit[4];   /* ln(lambda[erg/cm/s/K]) cubic in ln T */
    double eta_fit[4];      /* ln(eta[poise]) cubic in ln T */
} les_species;

typedef struct les_mixture {
    les_species sp[LES_NSPECIES];
    double diff_fit[LES_NSPECIES][LES_NSPECIES][4];   /* ln(D_kj[cm^2/s]) */
} les_mixture;

/*
 * Run parameters.  Names follow the reference's #defines (:46-284 and the
 * in-function constants :2648, :2740, :2834, :4101); les_params_default()
 * fills the values of a preset.
 */
typedef struct les_params {
    /* grid and domain (:46-51) */
    int    nx, ny, nz;
    double lx, ly, lz;

    /* scheme selection */
    int    mode;                 /* LES_MODE_* */
    int    integrator;           /* LES_INT_* */
    int    advection;            /* LES_ADV_* */
    int    precision;            /* LES_PREC_* (bulk arithmetic; statistics stay FP64) */

    /* inflow / boundary (:53-63) */
    double inlet_radius;
    double inlet_velocity;
    double inlet_temperature;
    double ambient_temperature;
    double turbulence_intensity;
    int    num_vortices;
    double turbulent_length_scale;
    double inlet_noise_amp;      /* white-noise amplitude / U (:2987) */

    /* combustion (:66-70, :4101) */
    double s_l;
    double burnt_temperature;
    int    initial_flame_front_i;
    double equivalence_ratio;

    /* LES / turbulence (:75-78, :151-157) */
    double pr_t, sc_t, fractal_d3;
    double filter_weight_center;
    double nu_unburnt_approx;
    double rho_unburnt_approx, rho_burnt_approx;   /* SSR tau (:2701-2703) */

    /* time step (:82-83) */
    double courant, courant_diffusion;

    /* flame-acceleration model and the optimiser's constants (:234-262, :2648) */
    double flame_accel_strength, flame_accel_width, flame_accel_center;
    double cs_unburned, cs_burnt, cs_flame_boost;
    double tip_protection_radius, max_flame_force, flame_ramp_distance;
    double weight_x;
    double relaxation_coeff;     /* nudging rate (:2834) */
    int    scheme_transition_step;  /* WENO5 |grad G| from this step (:2740) */

    /* duct mode (mode LES_MODE_DUCT) */
    double rho_unburnt;          /* kg/m^3 at T_u; rho = rho_unburnt*T_u/T */
    double prandtl;              /* lambda = mu*cp/Pr */
    double flame_x0;             /* initial planar flame position (m) */
    int    reinit_interval;      /* level-set reinitialisation every n steps (0 = off) */
    int    reinit_iters;         /* pseudo-time iterations per reinitialisation */
    int    proj_iters;           /* projection relaxation sweeps per step (0 = off) */
    double table_t0, table_dt;   /* cp/mu table start and spacing (K) */
    int    table_n;

    /* sampling / statistics */
    int    num_probe_x;          /* flame-region x-stations */
    double probe_x[LES_NPROBE_MAX];
    int    stats_start_step;     /* FP64 statistics from this step on (<0 = off) */
    int    stats_per_cell;       /* 1: also the per-cell accumulators of :3167-3192 */

    uint64_t seed;               /* SEM / noise seed */

    /* duct mode: > 0 gives every step this DT_sim instead of the CFL rule of
     * :4399-4411 (runs that must share a time-step sequence, e.g. an FP32 run
     * checked against its FP64 twin); 0: the rule */
    double fixed_dt;
} les_params;

/* Presets for les_params_default(). */
#define LES_PRESET_REFERENCE   0   /* reference #define values, 60x20x40, mode R */
#define LES_PRESET_DUCT64      1   /* 64x32x32 duct, WENO5, SSP-RK3, FP64 */
#define LES_PRESET_DUCT128     2   /* 128^3 duct */
#define LES_PRESET_ENSEMBLE    3   /* 64x48x48 duct member, FP32 */
#define LES_PRESET_LARGE       4   /* 256x192x192 duct, FP32 */

/* Report of one les_gpu_step() call (the last step it ran). */
typedef struct les_step_report {
    int    step;                 /* index of the last step run */
    double dt;                   /* DT_sim of that step (:4411) */
    double max_velocity;         /* (:4399) */
    double total_mass;           /* monitor reductions (:3885-3897) */
    double min_T, max_T, min_G, max_G, min_Y_fuel;
    int    nonfinite;            /* 1 if a non-finite value was found */
    double proj_residual;        /* duct mode: FP64 RMS residual of the projection's
                                    Poisson relaxation after its sweeps */
    double proj_residual0;       /* the same before the sweeps (RMS of div(rho u)) */
} les_step_report;

/* One flame-region probe row (the CSV line of :3453). */
typedef struct les_probe_row {
    int    timestep;
    int    station;
    double x_location, y_flame, y_physical, y_relative;
    double u, v, w, G, nu_t, S_mag;
} les_probe_row;

typedef struct les_gpu les_gpu;   /* opaque context */

/* ---- parameters and property data ---- */
int les_params_default(les_params* p, int preset);
int les_params_validate(const les_params* p);
/* The engine's synthetic 5-species property set (chem.bin/tran.bin stand-in). */
int les_mixture_synthetic(les_mixture* m);
/* Read the reference's Fortran-record property files (:311-382). */
int les_mixture_load(les_mixture* m, const char* chem_bin, const char* tran_bin);
/* Write property files in the same format (the reader above reads them back). */
int les_mixture_save(const les_mixture* m, const char* chem_bin, const char* tran_bin);

/* ---- context ---- */
int  les_gpu_create(const les_params* p, const les_mixture* mix, les_gpu** out);
void les_gpu_destroy(les_gpu* ctx);
const char* les_gpu_last_error(const les_gpu* ctx);

/* Initial state on the device (initialize + setting + initialize_vortices). */
int les_gpu_init_state(les_gpu* ctx);
/* Host <-> device state in checkpoint order u,v,w,T,G,rho,Y0..Y4 (n_fields <= 11). */
int les_gpu_upload_state(les_gpu* ctx, const double* const* fields, int n_fields);
int les_gpu_download_state(les_gpu* ctx, double* const* fields, int n_fields);
/* Named diagnostic fields: "nu_t", "S_mag", "nu", "alpha", "attribute" (as double). */
int les_gpu_download_field(les_gpu* ctx, const char* name, double* out);

/* Advance n_steps time steps; rep (may be NULL) describes the last one. */
int les_gpu_step(les_gpu* ctx, int n_steps, les_step_report* rep);
/* Testing hook: one RHS evaluation on the current state.  out holds
 * 5 + LES_NSPECIES fields in the order drhou, drhov, drhow, dT, dG, dY0..dY4. */
int les_gpu_rhs(les_gpu* ctx, int t_loop, double* const* out, int n_out);

/* Largest device staging les_gpu_rhs_noalloc uses: 32 MiB, i.e. duct grids up
 * to 838 860 cells in FP64 or 1 677 721 cells in FP32 (five RHS fields). */
#define LES_RHS_NOALLOC_MAX_BYTES (32u << 20)
/* les_gpu_rhs without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory between steps (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  les_gpu_rhs stages the duct RHS in buffers it allocates and
 * frees on every call; this variant stages it in a static device array of
 * LES_RHS_NOALLOC_MAX_BYTES that each GPU holds until the process ends, and
 * returns LES_ERR_PARAM (with a message in les_gpu_last_error) for a grid
 * whose staging would exceed it.  Results are identical to les_gpu_rhs.  All
 * calls on one GPU share that array: calls that overlap on one GPU (contexts
 * used from several host threads at once) race.  The reference scheme
 * allocates nothing per call already and takes the les_gpu_rhs path. */
int les_gpu_rhs_noalloc(les_gpu* ctx, int t_loop, double* const* out, int n_out);

/* Probe rows recorded since the last fetch (up to max_rows); *n_rows = count. */
int les_gpu_probe_fetch(les_gpu* ctx, les_probe_row* rows, int max_rows, int* n_rows);
/* Fixed-point time series recorded since the last fetch (the newest 4096
 * steps): out[step][point][u, v, w, G], oldest first (point sampling
 * :4557-4569; flush_point_data :3824-3850 writes them out). */
int les_gpu_point_fetch(les_gpu* ctx, double* out, int max_steps, int* n_steps, int* n_points);
/* Choose the point-series cells (n <= 64 triples i, j, k); mode R starts
 * with the reference's setup_measurement_points (:3780-3822) list, duct mode
 * with 4 x 4 x 4 evenly spaced points.  Replaces setup_measurement_points. */
int les_gpu_set_points(les_gpu* ctx, const int* ijk, int n);
int les_gpu_get_points(les_gpu* ctx, int* ijk, int max_n, int* n);
/* Resume support (load_quick_checkpoint :3701-3737, load_statistics_backup
 * :3759-3777, with the state those files leave out).  upload_field writes
 * "stat0".."stat9" (FP64 accumulators), mode R "nu"/"alpha" (the diffusive
 * history the next DT reads), duct "profiles"/"eddies".  export_aux writes
 * an opaque versioned blob of the engine state that is not a field (step,
 * SEM state: mode R's glibc rand() stream and eddies, duct eddies, run time,
 * profiles and DT history); *bytes is the capacity in, the size out (buf NULL
 * asks for the size).  import_aux restores it after les_gpu_upload_state;
 * with the fields, the run then continues exactly as if never interrupted. */
int les_gpu_upload_field(les_gpu* ctx, const char* name, const double* data);
int les_gpu_set_stats_count(les_gpu* ctx, int count);
int les_gpu_export_aux(les_gpu* ctx, void* buf, size_t* bytes);
int les_gpu_import_aux(les_gpu* ctx, const void* buf, size_t bytes);
/* For a restart without aux state (VTK): the step index, and fresh SEM
 * eddies (initialize_vortices :861-877, as the reference's restart does). */
int les_gpu_set_step(les_gpu* ctx, int step);
int les_gpu_init_sem(les_gpu* ctx);
/* Reference scheme: give the run the attribute map main() sets before its
 * restart detection (:4123-4136: the j = 0 plane is a wall, not the symmetry
 * plane of setting :1171-1172), which the reference's restart paths keep
 * because they skip setting() (binary :4184-4190, VTK :4269-4280).  A resume
 * from files the reference wrote (no aux blob) calls it to continue as the
 * reference does; les_gpu_init_state restores setting()'s map, and the aux
 * blob records which map a run uses.  LES_ERR_PARAM in duct mode. */
int les_gpu_use_resume_attributes(les_gpu* ctx);
/* Step recorded in a les_gpu_export_aux blob, to match it to the step in
 * checkpoint.bin (load_quick_checkpoint :3701-3737) before importing it. */
int les_aux_step(const void* buf, size_t bytes, int* step);
/* Change the optimiser-tunable constants between steps. */
int les_gpu_set_params(les_gpu* ctx, const les_params* p);
/* Host-visible SEM state (mode R keeps the eddies and glibc rand() stream on the host). */
int les_gpu_get_step(const les_gpu* ctx);
/* Average kernel time (ms) of the RHS kernel over the last les_gpu_step() call,
 * measured with HIP events on the engine's stream; *launches = launches timed. */
int les_gpu_kernel_timing(const les_gpu* ctx, double* rhs_ms, int* launches, double* step_ms);
/* Turn per-launch HIP-event timing of the RHS kernel on (1) or off (0). */
int les_gpu_set_timing(les_gpu* ctx, int on);
/* Launch shape a duct-mode run uses (no counterpart in the reference: the
 * reference has no device launches): tile_k (LES_TILE_FLAT or the tile row
 * width) and xcd_map (how tiles are dealt to the GPU's 8 XCDs: 0 round-robin,
 * 1 one contiguous range per XCD, 2 and 3 chunks of 16 and 64 consecutive
 * tiles round-robin, 4 slabs: each XCD the same eighth of every x-plane).  A
 * run times the mappings at creation and keeps the fastest; the environment
 * variable LES_XCD_MAP=0..4 (and LES_TILE_K for the tile width), set before
 * les_gpu_create, fixes the mapping instead (results do not depend on it).
 * LES_ERR_PARAM in mode R, whose kernels launch one thread per cell. */
int les_gpu_launch_config(const les_gpu* ctx, int* tile_k, int* xcd_map);
/* Number of steps accumulated in the FP64 statistics (:3191 time_avg_steps). */
int les_gpu_stats_count(const les_gpu* ctx);
/* Achievable HBM bandwidth: a plain device copy of `bytes` (read + write),
 * timed with HIP events (the calibration point of the roofline figures). */
int les_copy_bandwidth(size_t bytes, int reps, double* ms_per_copy, double* gbps);
/* sizeof(les_params), sizeof(les_mixture), sizeof(les_step_report),
 * sizeof(les_probe_row), sizeof(les_batch_config), sizeof(les_slot_info):
 * lets bindings check their struct mirrors. */
int les_abi_struct_sizes(size_t out[6]);

/* ------------------------------------------------------------------------
 * Batched runs (duct mode).  The reference runs one simulation per process
 * (main :4016-4632), one optimiser trial at a time: the optimiser's
 * interactive main (bayesian_optimizer.py:719-844) scores a finished run and
 * prints the next constants as #defines for the next build.  A batch context holds
 * a POOL of run slots sharing field arenas in device memory; every step
 * launch advances all resident, unfinished runs at once, so an ensemble of
 * trials, or a stream of requests admitted and retired while others run
 * (continuous batching), costs one launch sequence per step.  Each run keeps
 * its own grid, parameters, time step and step counter, and its results are
 * bitwise the same as when it runs alone in a pool with the same launch
 * configuration.
 * ------------------------------------------------------------------------ */
#define LES_TILE_FLAT        0   /* tiles of 256 consecutive cells of one x-plane */

typedef struct les_batch_config {
    int       max_slots;       /* runs resident at once */
    long long pool_cells;      /* cells of the shared arenas (sum over resident runs) */
    int       max_nx, max_ny, max_nz, max_vortices;   /* per-run limits */
    int       precision;       /* LES_PREC_*: one storage type per pool */
    int       tile_k;          /* tile width in k: LES_TILE_FLAT, 16, 32, 64 or 128 */
    int       xcd_map;         /* tiles to XCDs: 0 round-robin, 1 one contiguous range per XCD,
                                  2 / 3 chunks of 16 / 64 tiles round-robin, 4 slabs (each XCD the
                                  same eighth of every x-plane of slot 0's layout); + 8: the stencil
                                  launches alternate the order in which each XCD walks its tiles */
    int       autotune;        /* 1: time tile-width and XCD-mapping candidates at create
                                  on the probe grid and keep the fastest */
    int       point_cap;       /* steps kept in each run's point-series ring */
    int       probe_cap;       /* steps kept in each run's flame-region probe ring (0: off) */
    int       stats_per_cell;  /* 1: per-cell FP64 accumulators (10 per cell) */
    double    table_t0, table_dt;   /* cp/mu tables, shared by the pool */
    int       table_n;
} les_batch_config;

typedef struct les_slot_info {
    int    active;             /* 1 while admitted */
    int    done;               /* 1 once step == n_steps */
    int    step, n_steps;
    double time, dt, max_velocity;
    int    nx, ny, nz;
    long long cells;
} les_slot_info;

typedef struct les_batch les_batch;   /* opaque pool context */

int les_batch_config_default(les_batch_config* cfg);
/* Device bytes a pool of this configuration allocates (admission budgeting:
 * size pool_cells from the free HBM before creating the pool).  With
 * autotune = 1, creation also holds one scratch pool of at most the same
 * arena size per candidate while it times it; a candidate whose scratch pool
 * does not fit is skipped, and with none timed the configured tile width and
 * mapping stay. */
int les_batch_pool_bytes(const les_batch_config* cfg, size_t* bytes);
/* probe (may be NULL) is the run shape the autotuner times. */
int  les_batch_create(const les_batch_config* cfg, const les_params* probe, les_batch** out);
void les_batch_destroy(les_batch* b);
const char* les_batch_last_error(const les_batch* b);
/* Launch configuration in use (after autotuning): tile_k, blocks per stencil
 * launch (one per tile of the resident runs), XCD mapping, and the tuned step
 * time on the probe grid (ms, 0 if not tuned). */
int les_batch_tuning(const les_batch* b, int* tile_k, int* grid_blocks, int* xcd_map, double* tuned_ms);
/* Groups of runs the stage launches advance in turn (no counterpart in the
 * reference, which steps one run): a pool whose fields span more than
 * 12288 x 256 FP32 values (half that in FP64) is split at run boundaries into
 * up to four contiguous tile ranges, and each step runs the three RK stages
 * of one range before the next's, so each group's fields are a smaller share
 * of the caches; results are bitwise those of one group (LES_GROUPS=1 keeps
 * one).  Writes the group count and, for the first cap groups, each one's
 * first tile and tile count. */
int les_batch_groups(const les_batch* b, int cap, int* ngroups, int* tile_base, int* tile_count);

/* Admit a duct-mode run of n_steps steps (0: unbounded) into a free slot.
 * LES_ERR_ADMISSION when no slot is free, its cells or tiles do not fit in
 * the free arena ranges, or it exceeds the pool's per-run limits or
 * precision.  init = 1 sets the initial state (else upload one). */
int les_batch_admit(les_batch* b, const les_params* p, int n_steps, int init, int* slot);
int les_batch_release(les_batch* b, int slot);
/* Advance every resident, unfinished run by n_steps steps; *ms_per_step
 * (may be NULL) is the device time per step (HIP events).  With probe_cap > 0
 * the steps run in chunks of probe_cap and each run's probe ring is drained
 * into its row list between chunks, so no probe row is lost however many
 * steps one call runs. */
int les_batch_step(les_batch* b, int n_steps, double* ms_per_step);
int les_batch_slot_info(les_batch* b, int slot, les_slot_info* info);
/* Slots whose runs have finished (done and still resident). */
int les_batch_poll(les_batch* b, int* slots, int max_slots, int* n_done);
/* State of one run: rho*u, rho*v, rho*w, phi, T as double. */
int les_batch_upload(les_batch* b, int slot, const double* const* fields, int n_fields);
int les_batch_download(les_batch* b, int slot, double* const* fields, int n_fields);
/* "profiles" (10 per x-plane), "inlet", "eddies", "stat0".."stat9", "nu_t", "ST". */
int les_batch_download_field(les_batch* b, int slot, const char* name, double* out);
int les_batch_report(les_batch* b, int slot, les_step_report* rep);
int les_batch_point_fetch(les_batch* b, int slot, double* out, int max_steps, int* n_steps);
/* Flame-region probe rows of one run recorded since the last fetch (the newest
 * probe_cap steps; record_flame_region_data :3403-3471 rows). */
int les_batch_probe_fetch(les_batch* b, int slot, les_probe_row* rows, int max_rows, int* n_rows);
/* x-plane copy between two resident runs of one pool: `planes` x-planes of the
 * five state fields from planes [src_plane, src_plane + planes) of src_slot to
 * [dst_plane, dst_plane + planes) of dst_slot (equal ny and nz, ranges inside
 * each run, no overlap within one run), repeated `reps` (>= 1) times on the
 * pool's stream; *ms_per_copy (optional) receives the mean duration of one
 * copy.  The destination's dt maxima are refreshed afterwards, as after an
 * upload.  No counterpart in the reference, whose domain is one process's
 * global arrays: this is the local part of an x-slab halo exchange for a run
 * split across GPUs (an x-face of a run's state is one contiguous range per
 * field, so a ghost-plane fill copies five contiguous ranges, one launch
 * here, the ranges an RCCL send/receive would move straight from the arena). */
int les_batch_xface_copy(les_batch* b, int src_slot, int src_plane, int dst_slot, int dst_plane, int planes,
                         int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* LES_GPU_H */
