/*
 * les_params.c - run-parameter presets and property-file I/O (host C).
 *
 * The reference keeps every run parameter in compile-time #defines
 * (les_combustion_solver.c:11-284) and reads species data from the Fortran
 * record files chem.bin / tran.bin (load_chem_bin :324-356, load_tran_bin
 * :359-382).  Here the same values are runtime fields of les_params, and the
 * file readers/writers follow the reference's byte layout so that property
 * files stay interchangeable.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/les_gpu.h"

/* Reference values of the #define block (les_combustion_solver.c:46-283). */
static void set_reference_values(les_params* p)
{
    memset(p, 0, sizeof(*p));
    p->nx = 60; p->ny = 20; p->nz = 40;                 /* :46-48 */
    p->lx = 0.12; p->ly = 0.04; p->lz = 0.08;           /* :49-51 */
    p->mode = LES_MODE_REFERENCE;
    p->integrator = LES_INT_RK4;                        /* :4416-4544 */
    p->advection = LES_ADV_TENO5;                       /* :2738, :2844, :2918 */
    p->precision = LES_PREC_FP64;
    p->inlet_radius = 0.013;                            /* :54 */
    p->inlet_velocity = 4.0;                            /* :56 */
    p->inlet_temperature = 300.0;                       /* :57 */
    p->ambient_temperature = 300.0;                     /* :58 */
    p->turbulence_intensity = 0.20;                     /* :59 */
    p->num_vortices = 2000;                             /* :62 */
    p->turbulent_length_scale = 0.2 * 0.013;            /* :63 */
    p->inlet_noise_amp = 0.10;                          /* :2987 */
    p->s_l = 0.4;                                       /* :68 */
    p->burnt_temperature = 2200.0;                      /* :70 */
    p->initial_flame_front_i = 100;                     /* :69 */
    p->equivalence_ratio = 1.1;                         /* :4101 */
    p->pr_t = 0.5; p->sc_t = 0.8; p->fractal_d3 = 2.35; /* :75-78 */
    p->filter_weight_center = 50.0;                     /* :157 */
    p->nu_unburnt_approx = 1.5e-5;                      /* :154 */
    p->rho_unburnt_approx = 1.2; p->rho_burnt_approx = 0.2;  /* :2701-2702 */
    p->courant = 0.1; p->courant_diffusion = 0.1;       /* :82-83 */
    p->flame_accel_strength = 2.0361;                   /* :257 */
    p->flame_accel_width = 0.1126;                      /* :258 */
    p->flame_accel_center = 0.6697;                     /* :259 */
    p->cs_unburned = 0.1374;                            /* :260 */
    p->cs_burnt = 0.4223;                               /* :261 */
    p->cs_flame_boost = 0.2371;                         /* :262 */
    p->tip_protection_radius = 0.0;                     /* :234 */
    p->max_flame_force = 500.0;                         /* :240 */
    p->flame_ramp_distance = 0.03;                      /* :245 */
    p->weight_x = 0.05;                                 /* :2648 */
    p->relaxation_coeff = 5000.0;                       /* :2834 */
    p->scheme_transition_step = 500;                    /* :2740 */
    p->rho_unburnt = 1.2;
    p->prandtl = 0.7;
    p->flame_x0 = 0.016;
    p->reinit_interval = 0;
    p->reinit_iters = 0;
    p->proj_iters = 0;
    p->table_t0 = 300.0; p->table_dt = 1.0; p->table_n = 2001;
    p->num_probe_x = 7;                                 /* :270-274 */
    {
        const double xs[7] = {0.010, 0.015, 0.020, 0.025, 0.030, 0.035, 0.040};
        for (int n = 0; n < 7; ++n) p->probe_x[n] = xs[n];
    }
    p->stats_start_step = -1;                           /* dead in the reference (:4382) */
    p->stats_per_cell = 1;
    p->seed = 0;
}

/* Periodic duct of the duct-mode presets: dx = 1 mm, U = 5 m/s, 200 Gaussian
 * eddies of 4-cell length scale, 10 % intensity, planar flame at 16 mm,
 * S_L = 0.3 m/s, T_u = 300 K, T_b = 2100 K (density ratio 7), Cs = 0.12,
 * WENO5, SSP-RK3 with CFL 0.3. */
static void set_duct_values(les_params* p, int nx, int ny, int nz)
{
    set_reference_values(p);
    p->mode = LES_MODE_DUCT;
    p->integrator = LES_INT_SSPRK3;
    p->advection = LES_ADV_WENO5;
    p->nx = nx; p->ny = ny; p->nz = nz;
    p->lx = nx * 1e-3; p->ly = ny * 1e-3; p->lz = nz * 1e-3;
    p->inlet_velocity = 5.0;
    p->turbulence_intensity = 0.10;
    p->num_vortices = 200;
    p->turbulent_length_scale = 4e-3;
    p->inlet_noise_amp = 0.0;
    p->s_l = 0.3;
    p->burnt_temperature = 2100.0;
    p->cs_unburned = 0.12; p->cs_burnt = 0.12; p->cs_flame_boost = 0.0;
    p->courant = 0.3;
    p->courant_diffusion = 0.1;
    p->rho_unburnt = 1.2;
    p->prandtl = 0.7;
    p->flame_x0 = 0.016;
    p->reinit_interval = 10;
    p->reinit_iters = 4;
    p->proj_iters = 0;
    p->num_probe_x = 7;
    {
        const double xs[7] = {0.010, 0.015, 0.020, 0.025, 0.030, 0.035, 0.040};
        for (int n = 0; n < 7; ++n) p->probe_x[n] = xs[n];
    }
    p->stats_start_step = 0;
    p->stats_per_cell = 0;
    p->seed = 0;
}

int les_params_default(les_params* p, int preset)
{
    if (!p) return LES_ERR_PARAM;
    switch (preset) {
    case LES_PRESET_REFERENCE: set_reference_values(p); return LES_OK;
    case LES_PRESET_DUCT64:    set_duct_values(p, 64, 32, 32); return LES_OK;
    case LES_PRESET_DUCT128:   set_duct_values(p, 128, 128, 128); return LES_OK;
    case LES_PRESET_ENSEMBLE:
        set_duct_values(p, 64, 48, 48);
        p->precision = LES_PREC_FP32;
        return LES_OK;
    case LES_PRESET_LARGE:
        set_duct_values(p, 256, 192, 192);
        p->precision = LES_PREC_FP32;
        p->turbulence_intensity = 0.15;
        p->num_vortices = 2000;
        p->proj_iters = 8;           /* projection relaxation with FP64 residuals */
        p->num_probe_x = 16;
        for (int n = 0; n < 16; ++n) p->probe_x[n] = 0.010 + 0.015 * n;
        return LES_OK;
    default: return LES_ERR_PARAM;
    }
}

int les_params_validate(const les_params* p)
{
    if (!p) return LES_ERR_PARAM;
    if (p->nx < 8 || p->ny < 8 || p->nz < 8) return LES_ERR_PARAM;
    if ((double)p->nx * p->ny * p->nz > 4.0e9) return LES_ERR_PARAM;
    if (!(p->lx > 0 && p->ly > 0 && p->lz > 0)) return LES_ERR_PARAM;
    if (p->mode != LES_MODE_REFERENCE && p->mode != LES_MODE_DUCT) return LES_ERR_PARAM;
    if (p->integrator != LES_INT_RK4 && p->integrator != LES_INT_SSPRK3) return LES_ERR_PARAM;
    if (p->advection < LES_ADV_TENO5 || p->advection > LES_ADV_MUSCL_MINMOD) return LES_ERR_PARAM;
    if (p->precision != LES_PREC_FP64 && p->precision != LES_PREC_FP32) return LES_ERR_PARAM;
    if (p->num_vortices < 0 || p->num_vortices > 65536) return LES_ERR_PARAM;
    if (p->num_probe_x < 0 || p->num_probe_x > LES_NPROBE_MAX) return LES_ERR_PARAM;
    if (!(p->courant > 0) || !(p->courant_diffusion > 0)) return LES_ERR_PARAM;
    if (!(p->fixed_dt >= 0) || p->fixed_dt > 1.0) return LES_ERR_PARAM;
    if (p->mode == LES_MODE_REFERENCE && p->fixed_dt != 0) return LES_ERR_PARAM;
    if (p->mode == LES_MODE_DUCT) {
        if (p->table_n < 2 || !(p->table_dt > 0)) return LES_ERR_PARAM;
        if (!(p->rho_unburnt > 0) || !(p->prandtl > 0)) return LES_ERR_PARAM;
        if (p->proj_iters < 0 || p->proj_iters > 256) return LES_ERR_PARAM;
        if (p->reinit_iters < 0 || p->reinit_iters > 64 || p->reinit_interval < 0) return LES_ERR_PARAM;
    }
    if (p->mode == LES_MODE_REFERENCE && p->precision != LES_PREC_FP64) return LES_ERR_PARAM;
    return LES_OK;
}

/* ------------------------------------------------------------------------ */
/* Synthetic 5-species property set.                                          */
/* NASA-7 coefficients are the public GRI-Mech 3.0 thermodynamic polynomials; */
/* transport fits are power laws written as cubic fits in ln T so that every  */
/* term of the reference's evaluator (:385-391) is exercised.                 */
/* ------------------------------------------------------------------------ */
static void fit_power_law(double out[4], double value_300, double exponent,
                          double a2, double a3)
{
    double L = log(300.0);
    out[1] = exponent;
    out[2] = a2;
    out[3] = a3;
    out[0] = log(value_300) - (exponent * L + a2 * L * L + a3 * L * L * L);
}

int les_mixture_synthetic(les_mixture* m)
{
    if (!m) return LES_ERR_PARAM;
    memset(m, 0, sizeof(*m));
    static const char* names[LES_NSPECIES] = {"C3H8", "O2", "N2", "CO2", "H2O"};
    static const double mw[LES_NSPECIES] = {44.09652, 31.9988, 28.0134, 44.0095, 18.01528};
    static const double lo[LES_NSPECIES][7] = {
        {0.93355381, 0.026424579, 6.1059727e-6, -2.1977499e-8, 9.5149253e-12, -13958.52, 19.201691},
        {3.78245636, -2.99673416e-3, 9.84730201e-6, -9.68129509e-9, 3.24372837e-12, -1063.94356, 3.65767573},
        {3.298677, 1.4082404e-3, -3.963222e-6, 5.641515e-9, -2.444854e-12, -1020.8999, 3.950372},
        {2.35677352, 8.98459677e-3, -7.12356269e-6, 2.45919022e-9, -1.43699548e-13, -48371.9697, 9.90105222},
        {4.19864056, -2.0364341e-3, 6.52040211e-6, -5.48797062e-9, 1.77197817e-12, -30293.7267, -0.849032208}};
    static const double hi[LES_NSPECIES][7] = {
        {7.5341368, 0.018872239, -6.2718491e-6, 9.1475649e-10, -4.7838069e-14, -16467.516, -17.892349},
        {3.28253784, 1.48308754e-3, -7.57966669e-7, 2.09470555e-10, -2.16717794e-14, -1088.45772, 5.45323129},
        {2.92664, 1.4879768e-3, -5.68476e-7, 1.0097038e-10, -6.753351e-15, -922.7977, 5.980528},
        {3.85746029, 4.41437026e-3, -2.21481404e-6, 5.23490188e-10, -4.72084164e-14, -48759.166, 2.27163806},
        {3.03399249, 2.17691804e-3, -1.64072518e-7, -9.7041987e-11, 1.68200992e-14, -30004.2971, 4.9667701}};
    static const double eta300[LES_NSPECIES] = {8.2e-5, 2.07e-4, 1.78e-4, 1.50e-4, 1.00e-4};   /* poise */
    static const double lam300[LES_NSPECIES] = {1800.0, 2660.0, 2600.0, 1660.0, 1860.0};       /* erg/cm/s/K */
    static const double d300[LES_NSPECIES] = {0.10, 0.20, 0.21, 0.16, 0.25};                   /* cm^2/s scale */
    for (int s = 0; s < LES_NSPECIES; ++s) {
        snprintf(m->sp[s].name, sizeof(m->sp[s].name), "%s", names[s]);
        m->sp[s].mw = mw[s];
        m->sp[s].t_mid = 1000.0;
        for (int c = 0; c < 7; ++c) { m->sp[s].nasa_low[c] = lo[s][c]; m->sp[s].nasa_high[c] = hi[s][c]; }
        fit_power_law(m->sp[s].eta_fit, eta300[s], 0.70, 2.0e-3, -1.0e-4);
        fit_power_law(m->sp[s].lambda_fit, lam300[s], 0.80, 1.5e-3, -8.0e-5);
    }
    for (int a = 0; a < LES_NSPECIES; ++a)
        for (int b = 0; b < LES_NSPECIES; ++b) {
            /* symmetric binary diffusivity ~ mean of the two species scales */
            double d = 0.5 * (d300[a] + d300[b]);
            fit_power_law(m->diff_fit[a][b], d, 1.75, 1.0e-3, -5.0e-5);
        }
    return LES_OK;
}

/* ---- Fortran unformatted sequential records: int32 len, payload, int32 len ---- */
static int write_record(FILE* fp, const void* data, int len)
{
    if (fwrite(&len, 4, 1, fp) != 1) return 0;
    if (len > 0 && fwrite(data, 1, (size_t)len, fp) != (size_t)len) return 0;
    if (fwrite(&len, 4, 1, fp) != 1) return 0;
    return 1;
}

static unsigned char* read_record(FILE* fp, int* len_out)
{
    int len = 0, end = 0;
    if (fread(&len, 4, 1, fp) != 1 || len < 0 || len > (1 << 26)) return NULL;
    unsigned char* buf = (unsigned char*)malloc(len > 0 ? (size_t)len : 1);
    if (!buf) return NULL;
    if (len > 0 && fread(buf, 1, (size_t)len, fp) != (size_t)len) { free(buf); return NULL; }
    if (fread(&end, 4, 1, fp) != 1) { free(buf); return NULL; }
    *len_out = len;
    return buf;
}

/*
 * chem.bin layout consumed by load_chem_bin (:324-356): record 1 (ignored),
 * record 2 with n_elements at byte 12 and n_species at byte 16, record 3
 * (ignored), record 4 with an 8-byte header followed by one block per species:
 * name[16], composition int32[n_el], 2 int32, mw double, int32, 3 doubles
 * (the second of which is followed by t_mid), nasa_high[7], nasa_low[7].
 */
#define CHEM_N_ELEMENTS 4

int les_mixture_save(const les_mixture* m, const char* chem_bin, const char* tran_bin)
{
    if (!m || !chem_bin || !tran_bin) return LES_ERR_PARAM;
    FILE* fp = fopen(chem_bin, "wb");
    if (!fp) return LES_ERR_IO;
    int ok = 1;
    {
        char hdr[16] = "LES-SYNTHETIC";
        ok &= write_record(fp, hdr, 16);
        unsigned char r2[20];
        memset(r2, 0, sizeof(r2));
        int nel = CHEM_N_ELEMENTS, nsp = LES_NSPECIES;
        memcpy(r2 + 12, &nel, 4);
        memcpy(r2 + 16, &nsp, 4);
        ok &= write_record(fp, r2, 20);
        ok &= write_record(fp, hdr, 16);
        const size_t per = 16 + 4 * CHEM_N_ELEMENTS + 8 + 8 + 4 + 24 + 2 * 56;
        size_t len = 8 + per * LES_NSPECIES;
        unsigned char* r4 = (unsigned char*)calloc(1, len);
        if (!r4) { fclose(fp); return LES_ERR_IO; }
        for (int s = 0; s < LES_NSPECIES; ++s) {
            unsigned char* b = r4 + 8 + per * s;
            memset(b, ' ', 16);
            memcpy(b, m->sp[s].name, strnlen(m->sp[s].name, 16));
            size_t mw_off = 16 + 4 * CHEM_N_ELEMENTS + 8;
            memcpy(b + mw_off, &m->sp[s].mw, 8);
            size_t tmid_off = mw_off + 8 + 4 + 16;
            memcpy(b + tmid_off, &m->sp[s].t_mid, 8);
            memcpy(b + tmid_off + 8, m->sp[s].nasa_high, 56);
            memcpy(b + tmid_off + 8 + 56, m->sp[s].nasa_low, 56);
        }
        ok &= write_record(fp, r4, (int)len);
        free(r4);
    }
    if (fclose(fp) != 0) ok = 0;
    if (!ok) return LES_ERR_IO;

    /* tran.bin (:359-382): 4 ignored records, then lambda fits [K][4],
     * eta fits [K][4], binary diffusion fits [K][K][4]. */
    fp = fopen(tran_bin, "wb");
    if (!fp) return LES_ERR_IO;
    {
        char hdr[16] = "LES-SYNTHETIC";
        for (int r = 0; r < 4; ++r) ok &= write_record(fp, hdr, 16);
        double lam[LES_NSPECIES * 4], eta[LES_NSPECIES * 4];
        for (int s = 0; s < LES_NSPECIES; ++s)
            for (int c = 0; c < 4; ++c) {
                lam[s * 4 + c] = m->sp[s].lambda_fit[c];
                eta[s * 4 + c] = m->sp[s].eta_fit[c];
            }
        ok &= write_record(fp, lam, (int)sizeof(lam));
        ok &= write_record(fp, eta, (int)sizeof(eta));
        ok &= write_record(fp, m->diff_fit, (int)sizeof(m->diff_fit));
    }
    if (fclose(fp) != 0) ok = 0;
    return ok ? LES_OK : LES_ERR_IO;
}

int les_mixture_load(les_mixture* m, const char* chem_bin, const char* tran_bin)
{
    if (!m || !chem_bin || !tran_bin) return LES_ERR_PARAM;
    memset(m, 0, sizeof(*m));
    FILE* fp = fopen(chem_bin, "rb");
    if (!fp) return LES_ERR_IO;
    int len = 0, n_el = 0, n_sp = 0;
    unsigned char* b = read_record(fp, &len);
    if (!b) { fclose(fp); return LES_ERR_IO; }
    free(b);
    b = read_record(fp, &len);
    if (!b || len < 20) { free(b); fclose(fp); return LES_ERR_IO; }
    memcpy(&n_el, b + 12, 4);
    memcpy(&n_sp, b + 16, 4);
    free(b);
    if (n_sp != LES_NSPECIES || n_el < 0 || n_el > 64) { fclose(fp); return LES_ERR_IO; }
    b = read_record(fp, &len);
    if (!b) { fclose(fp); return LES_ERR_IO; }
    free(b);
    b = read_record(fp, &len);
    if (!b) { fclose(fp); return LES_ERR_IO; }
    const size_t per = 16 + 4 * (size_t)n_el + 8 + 8 + 4 + 24 + 2 * 56;
    if ((size_t)len < 8 + per * (size_t)n_sp) { free(b); fclose(fp); return LES_ERR_IO; }
    for (int s = 0; s < n_sp; ++s) {
        const unsigned char* q = b + 8 + per * s;
        memcpy(m->sp[s].name, q, 16);
        m->sp[s].name[16] = '\0';
        for (int c = 15; c >= 0 && m->sp[s].name[c] == ' '; --c) m->sp[s].name[c] = '\0';
        size_t mw_off = 16 + 4 * (size_t)n_el + 8;
        memcpy(&m->sp[s].mw, q + mw_off, 8);
        size_t tmid_off = mw_off + 8 + 4 + 16;
        memcpy(&m->sp[s].t_mid, q + tmid_off, 8);
        memcpy(m->sp[s].nasa_high, q + tmid_off + 8, 56);
        memcpy(m->sp[s].nasa_low, q + tmid_off + 8 + 56, 56);
    }
    free(b);
    fclose(fp);

    fp = fopen(tran_bin, "rb");
    if (!fp) return LES_ERR_IO;
    for (int r = 0; r < 4; ++r) {
        b = read_record(fp, &len);
        if (!b) { fclose(fp); return LES_ERR_IO; }
        free(b);
    }
    b = read_record(fp, &len);
    if (!b || len < (int)(LES_NSPECIES * 4 * 8)) { free(b); fclose(fp); return LES_ERR_IO; }
    for (int s = 0; s < LES_NSPECIES; ++s) memcpy(m->sp[s].lambda_fit, b + s * 32, 32);
    free(b);
    b = read_record(fp, &len);
    if (!b || len < (int)(LES_NSPECIES * 4 * 8)) { free(b); fclose(fp); return LES_ERR_IO; }
    for (int s = 0; s < LES_NSPECIES; ++s) memcpy(m->sp[s].eta_fit, b + s * 32, 32);
    free(b);
    b = read_record(fp, &len);
    if (!b || len < (int)sizeof(m->diff_fit)) { free(b); fclose(fp); return LES_ERR_IO; }
    memcpy(m->diff_fit, b, sizeof(m->diff_fit));
    free(b);
    fclose(fp);
    return LES_OK;
}

int les_abi_struct_sizes(size_t out[6])
{
    if (!out) return LES_ERR_PARAM;
    out[0] = sizeof(les_params);
    out[1] = sizeof(les_mixture);
    out[2] = sizeof(les_step_report);
    out[3] = sizeof(les_probe_row);
    out[4] = sizeof(les_batch_config);
    out[5] = sizeof(les_slot_info);
    return LES_OK;
}

int les_batch_config_default(les_batch_config* cfg)
{
    if (!cfg) return LES_ERR_PARAM;
    memset(cfg, 0, sizeof(*cfg));
    cfg->max_slots = 64;
    cfg->pool_cells = 64LL * 64 * 48 * 48;
    cfg->max_nx = 128;
    cfg->max_ny = 128;
    cfg->max_nz = 128;
    cfg->max_vortices = 2000;
    cfg->precision = LES_PREC_FP32;
    cfg->tile_k = 64;
    cfg->xcd_map = 1;
    cfg->autotune = 0;
    cfg->point_cap = 256;
    cfg->probe_cap = 0;
    cfg->stats_per_cell = 0;
    cfg->table_t0 = 300.0;
    cfg->table_dt = 1.0;
    cfg->table_n = 2001;
    return LES_OK;
}
