"""ctypes mirror of include/les_gpu.h (POD structs and constants).

The C header is the boundary; this module only restates its structs so that
Python callers (tests, bench, the served loop) can fill them.  Field order and
types must match the header exactly; ``tests/test_abi.py`` checks the sizes
against the compiled library.
"""
import ctypes as C

LES_OK = 0
LES_ERR_HIP = -1
LES_ERR_PARAM = -2
LES_ERR_NAN = -3
LES_ERR_ADMISSION = -4
LES_ERR_IO = -5
LES_ERR_STATE = -6

LES_MODE_REFERENCE = 0
LES_MODE_DUCT = 1
LES_INT_RK4 = 0
LES_INT_SSPRK3 = 1
LES_ADV_TENO5 = 0
LES_ADV_WENO5 = 1
LES_ADV_MUSCL_MINMOD = 2
LES_PREC_FP64 = 0
LES_PREC_FP32 = 1

LES_NSPECIES = 5
LES_NFIELDS = 11
LES_NPROBE_MAX = 16
LES_NPOINT_MAX = 64

LES_PRESET_REFERENCE = 0
LES_PRESET_DUCT64 = 1
LES_PRESET_DUCT128 = 2
LES_PRESET_ENSEMBLE = 3
LES_PRESET_LARGE = 4

FIELD_NAMES = ["u", "v", "w", "T", "G", "rho", "Y0", "Y1", "Y2", "Y3", "Y4"]
RHS_NAMES = ["drhou", "drhov", "drhow", "dT", "dG", "dY0", "dY1", "dY2", "dY3", "dY4"]


class Species(C.Structure):
    _fields_ = [
        ("name", C.c_char * 17),
        ("mw", C.c_double),
        ("t_mid", C.c_double),
        ("nasa_high", C.c_double * 7),
        ("nasa_low", C.c_double * 7),
        ("lambda_fit", C.c_double * 4),
        ("eta_fit", C.c_double * 4),
    ]


class Mixture(C.Structure):
    _fields_ = [
        ("sp", Species * LES_NSPECIES),
        ("diff_fit", ((C.c_double * 4) * LES_NSPECIES) * LES_NSPECIES),
    ]


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("mode", C.c_int), ("integrator", C.c_int), ("advection", C.c_int), ("precision", C.c_int),
        ("inlet_radius", C.c_double), ("inlet_velocity", C.c_double),
        ("inlet_temperature", C.c_double), ("ambient_temperature", C.c_double),
        ("turbulence_intensity", C.c_double),
        ("num_vortices", C.c_int),
        ("turbulent_length_scale", C.c_double),
        ("inlet_noise_amp", C.c_double),
        ("s_l", C.c_double), ("burnt_temperature", C.c_double),
        ("initial_flame_front_i", C.c_int),
        ("equivalence_ratio", C.c_double),
        ("pr_t", C.c_double), ("sc_t", C.c_double), ("fractal_d3", C.c_double),
        ("filter_weight_center", C.c_double),
        ("nu_unburnt_approx", C.c_double),
        ("rho_unburnt_approx", C.c_double), ("rho_burnt_approx", C.c_double),
        ("courant", C.c_double), ("courant_diffusion", C.c_double),
        ("flame_accel_strength", C.c_double), ("flame_accel_width", C.c_double),
        ("flame_accel_center", C.c_double),
        ("cs_unburned", C.c_double), ("cs_burnt", C.c_double), ("cs_flame_boost", C.c_double),
        ("tip_protection_radius", C.c_double), ("max_flame_force", C.c_double),
        ("flame_ramp_distance", C.c_double),
        ("weight_x", C.c_double),
        ("relaxation_coeff", C.c_double),
        ("scheme_transition_step", C.c_int),
        ("rho_unburnt", C.c_double),
        ("prandtl", C.c_double),
        ("flame_x0", C.c_double),
        ("reinit_interval", C.c_int),
        ("reinit_iters", C.c_int),
        ("proj_iters", C.c_int),
        ("table_t0", C.c_double), ("table_dt", C.c_double),
        ("table_n", C.c_int),
        ("num_probe_x", C.c_int),
        ("probe_x", C.c_double * LES_NPROBE_MAX),
        ("stats_start_step", C.c_int),
        ("stats_per_cell", C.c_int),
        ("seed", C.c_uint64),
        ("fixed_dt", C.c_double),
    ]


class StepReport(C.Structure):
    _fields_ = [
        ("step", C.c_int),
        ("dt", C.c_double),
        ("max_velocity", C.c_double),
        ("total_mass", C.c_double),
        ("min_T", C.c_double), ("max_T", C.c_double),
        ("min_G", C.c_double), ("max_G", C.c_double),
        ("min_Y_fuel", C.c_double),
        ("nonfinite", C.c_int),
        ("proj_residual", C.c_double),
        ("proj_residual0", C.c_double),
    ]


class ProbeRow(C.Structure):
    _fields_ = [
        ("timestep", C.c_int),
        ("station", C.c_int),
        ("x_location", C.c_double), ("y_flame", C.c_double),
        ("y_physical", C.c_double), ("y_relative", C.c_double),
        ("u", C.c_double), ("v", C.c_double), ("w", C.c_double), ("G", C.c_double),
        ("nu_t", C.c_double), ("S_mag", C.c_double),
    ]


LES_TILE_FLAT = 0


class BatchConfig(C.Structure):
    _fields_ = [
        ("max_slots", C.c_int),
        ("pool_cells", C.c_longlong),
        ("max_nx", C.c_int), ("max_ny", C.c_int), ("max_nz", C.c_int), ("max_vortices", C.c_int),
        ("precision", C.c_int),
        ("tile_k", C.c_int),
        ("xcd_map", C.c_int),
        ("autotune", C.c_int),
        ("point_cap", C.c_int),
        ("probe_cap", C.c_int),
        ("stats_per_cell", C.c_int),
        ("table_t0", C.c_double), ("table_dt", C.c_double),
        ("table_n", C.c_int),
    ]


class SlotInfo(C.Structure):
    _fields_ = [
        ("active", C.c_int), ("done", C.c_int),
        ("step", C.c_int), ("n_steps", C.c_int),
        ("time", C.c_double), ("dt", C.c_double), ("max_velocity", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("cells", C.c_longlong),
    ]


def declare_common(lib):
    """Declare the parameter/mixture entry points shared by both libraries."""
    lib.les_params_default.argtypes = [C.POINTER(Params), C.c_int]
    lib.les_params_default.restype = C.c_int
    lib.les_params_validate.argtypes = [C.POINTER(Params)]
    lib.les_params_validate.restype = C.c_int
    lib.les_mixture_synthetic.argtypes = [C.POINTER(Mixture)]
    lib.les_mixture_synthetic.restype = C.c_int
    lib.les_mixture_load.argtypes = [C.POINTER(Mixture), C.c_char_p, C.c_char_p]
    lib.les_mixture_load.restype = C.c_int
    lib.les_mixture_save.argtypes = [C.POINTER(Mixture), C.c_char_p, C.c_char_p]
    lib.les_mixture_save.restype = C.c_int


def probe_row_csv(r):
    """Format one probe row exactly as the reference's fprintf (les_combustion_solver.c:3453)."""
    return "%d,%.3f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6e,%.6e\n" % (
        r.timestep, r.x_location, r.y_flame, r.y_physical, r.y_relative,
        r.u, r.v, r.w, r.G, r.nu_t, r.S_mag)
