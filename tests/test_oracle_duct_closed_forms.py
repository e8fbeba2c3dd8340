"""Closed-form pins of the oracle's duct RHS assembly, filter, SGS closure,
DT rule, boundary planes, SEM kernel shape, progress map and sampling.

Every expected value below is derived here from the scheme's definition
(filter weights, central differences, linear WENO5 weights, Smagorinsky
closure, SSR flame-speed model, force model, DT rule), evaluated on states
whose filtered fields, gradients and reconstructions have closed forms; none
reads an intermediate back from the oracle except where a test says so.  The
mistakes each test is built to catch are named in its docstring.

Reference anchors (les_combustion_solver.c): conservative face fluxes
:683-727, momentum assembly :2552-2691 (gravity G_GRAVITY = -9.81 :139,
force clip MAX_FLAME_FORCE_LIMIT :240), box filter :1307-1470 (centre weight
50 :157), nu_t :1569-1633, SSR model :1913-1965 (constants :151-154), DT rule
:4399-4411, boundary conditions :2967-3161, SEM :943-1075, statistics
:3167-3219, probe rows :3403-3471.
"""
import math

import numpy as np
import pytest

DUCT = 1            # LES_PRESET_DUCT64: 64 x 32 x 32, dx = 1 mm
G_GRAVITY = -9.81   # les_combustion_solver.c:139
SMALL = 1e-10       # les_combustion_solver.c:141


def make(oracle, **over):
    base = dict(reinit_interval=0, num_vortices=0)
    base.update(over)
    p = oracle.params(DUCT, **base)
    o = oracle.Oracle(p)
    o.init_state()
    return o, p


def coords(p):
    dx, dy, dz = p.lx / p.nx, p.ly / p.ny, p.lz / p.nz
    i, j, k = np.meshgrid(np.arange(p.nx), np.arange(p.ny), np.arange(p.nz), indexing="ij")
    return dx, dy, dz, i * dx, j * dy, k * dz


def rho_of(p, T):
    return p.rho_unburnt * p.inlet_temperature / T


def set_primitives(o, p, u, v, w, phi, T):
    shape = (p.nx, p.ny, p.nz)
    u, v, w, phi, T = (np.broadcast_to(np.asarray(a, dtype=np.float64), shape) for a in (u, v, w, phi, T))
    r = rho_of(p, T)
    o.set_state([r * u, r * v, r * w, phi.copy(), T.copy()])


def mu_table(p, T):
    """cp/mu table definition: nodes mu(T_n) = 1.8e-5 (T_n/300)^0.7 at 1 K
    steps, linear interpolation between them."""
    T = np.asarray(T, dtype=np.float64)
    s = np.clip((T - p.table_t0) / p.table_dt, 0.0, p.table_n - 1)
    i0 = np.minimum(s.astype(np.int64), p.table_n - 2)
    f = s - i0
    n0 = p.table_t0 + i0 * p.table_dt
    m0 = 1.8e-5 * (n0 / 300.0) ** 0.7
    m1 = 1.8e-5 * ((n0 + p.table_dt) / 300.0) ** 0.7
    return m0 + f * (m1 - m0)


def progress(phi, dx):
    """c(phi) = 1 / (1 + exp(-phi / dx)), the logistic equal to (1 + tanh(phi / 2dx)) / 2."""
    return 1.0 / (1.0 + np.exp(-np.asarray(phi) / dx))


def filter_gain(kd, wc=50.0):
    """Gain of the 7-point box filter (centre weight wc, six neighbours 1) on
    a mode sin(k y) that varies along one periodic axis only."""
    return (wc + 4.0 + 2.0 * math.cos(kd)) / (wc + 6.0)


def weno5_linear_diff(kd):
    """Complex factor H with f_{j+1/2} - f_{j-1/2} = H e^{i k y_j} for the
    upwind-biased linear WENO5 face value (2, -13, 47, 27, -3) / 60 of data
    e^{i k y}: the limit of WENO5 when every smoothness indicator is far below
    eps (linear weights 0.1, 0.6, 0.3 on the three third-order candidates)."""
    e = np.exp(1j * kd * np.arange(-2, 3))
    face = np.dot(np.array([2.0, -13.0, 47.0, 27.0, -3.0]) / 60.0, e)
    return face * (1.0 - np.exp(-1j * kd))


def close(a, b, rtol, atol=0.0):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max()
    assert err <= rtol * scale + atol, (err, scale, err / scale)


# --------------------------------------------------------------------------
# box filter (:1307-1470)
# --------------------------------------------------------------------------

def test_filter_constant_linear_and_brute_force(oracle):
    """A constant field and a field linear in x are fixed points of the
    filter; on a random state every filtered primitive equals the 7-point
    weighted mean (centre 50, neighbours 1, wrap in y/z, copy on the x end
    planes) evaluated here.  Catches a wrong weight, a missing wrap, a
    neighbour taken from the wrong axis, or filtering rho*u instead of u."""
    o, p = make(oracle)
    dx, dy, dz, x, y, z = coords(p)
    set_primitives(o, p, 3.0, -1.0, 0.5, 0.25, 1500.0)
    o.rhs(0)
    for nm, want in [("u_f", 3.0), ("v_f", -1.0), ("w_f", 0.5), ("phi_f", 0.25), ("T_f", 1500.0)]:
        close(o.field(nm), np.full(x.shape, want), 4e-16 * 8)
    set_primitives(o, p, 5.0 + 100.0 * x, -2.0 * x, 0.0, x - 0.02, 900.0 + 2e4 * x)
    o.rhs(0)
    for nm, want in [("u_f", 5.0 + 100.0 * x), ("v_f", -2.0 * x), ("phi_f", x - 0.02), ("T_f", 900.0 + 2e4 * x)]:
        close(o.field(nm), want, 1e-14, 1e-17)
    rng = np.random.default_rng(5)
    u, v, w = rng.uniform(-3, 3, (3,) + x.shape)
    phi = 1e-3 * rng.standard_normal(x.shape)
    T = rng.uniform(300, 2100, x.shape)
    set_primitives(o, p, u, v, w, phi, T)
    o.rhs(0)
    wc = 50.0
    for nm, f in [("u_f", u), ("v_f", v), ("w_f", w), ("phi_f", phi), ("T_f", T)]:
        s = wc * f
        for ax in (1, 2):
            s = s + np.roll(f, 1, axis=ax) + np.roll(f, -1, axis=ax)
        s = s + np.roll(f, 1, axis=0) + np.roll(f, -1, axis=0)
        want = s / (wc + 6.0)
        want[0], want[-1] = f[0], f[-1]
        close(o.field(nm), want, 1e-14)


# --------------------------------------------------------------------------
# nu_t and S_T (:1569-1633, :1913-1965)
# --------------------------------------------------------------------------

def test_nu_t_under_normal_strain_and_cs_of_progress(oracle):
    """u = U + a x, v = b x: filtered velocities are exact (linear), so
    S11 = a, S12 = b/2, |S| = sqrt(2 SijSij) = sqrt(2 a^2 + b^2) and
    nu_t = (Cs dx)^2 |S| with Cs(G) = Cu + (Cb - Cu) G + 4 Cf G (1 - G); the
    level set sits at phi = 0 (G = 1/2) so all three constants count.
    Catches a missing factor 2 in |S|, a wrong S12, or a wrong Cs(G)."""
    cu, cb, cf = 0.10, 0.20, 0.05
    o, p = make(oracle, cs_unburned=cu, cs_burnt=cb, cs_flame_boost=cf)
    dx, dy, dz, x, y, z = coords(p)
    a, b = 120.0, -70.0
    set_primitives(o, p, 5.0 + a * x, b * x, 0.0, 0.0, 2100.0)
    o.rhs(0)
    S = math.sqrt(2 * a * a + b * b)
    g = 0.5
    cs = cu + (cb - cu) * g + 4.0 * cf * g * (1.0 - g)
    sl = np.s_[1:-1]
    close(o.field("S_mag")[sl], np.full(x[sl].shape, S), 1e-12)
    close(o.field("nu_t")[sl], np.full(x[sl].shape, (cs * dx) ** 2 * S), 1e-12)
    assert not o.field("nu_t")[0].any() and not o.field("nu_t")[-1].any()


def test_nu_t_under_shear_sine(oracle):
    """u = a sin(k y): the filter scales the mode by (54 + 2 cos k dy)/56, the
    central difference gives du/dy = a' cos(k y) sin(k dy)/dy, and with no
    other gradient |S| = |du/dy|.  Catches a filter applied on the wrong
    axis, a one-sided difference, or |S| built from S12 alone."""
    o, p = make(oracle)
    dx, dy, dz, x, y, z = coords(p)
    kk = 2 * np.pi / p.ly
    a = 2.0
    set_primitives(o, p, a * np.sin(kk * y), 0.0, 0.0, x + 1.0, 2100.0)
    o.rhs(0)
    ap = a * filter_gain(kk * dy)
    S = np.abs(ap * np.cos(kk * y) * math.sin(kk * dy) / dy)
    sl = np.s_[2:-2]
    close(o.field("S_mag")[sl], S[sl], 1e-12, 1e-12)
    close(o.field("nu_t")[sl], (p.cs_burnt * dx) ** 2 * S[sl], 1e-12, 1e-20)


def ssr_flame_speed(p, dx, nu, ss, div):
    """SSR fractal flame speed restated from les_combustion_solver.c:1913-1965
    with its constants (A_eta 0.2340, B_eta 0.2187, C_alpha 10 :151-153), the
    duct's expansion ratio tau = T_b/T_u - 1 (rho proportional to 1/T)."""
    SL, D3 = p.s_l, p.fractal_d3
    tau = p.burnt_temperature / p.inlet_temperature - 1.0
    X = (dx ** 4 / (nu * nu)) * (ss - div * div / 3.0)
    if X < SMALL:
        d_eta = 1.0
    else:
        px = X ** 0.75
        d_eta = 0.2340 * px * (math.log(0.2187 * px) if 0.2187 * px > 1.0 else 0.0)
    eta = dx / (d_eta + SMALL)
    dF = p.nu_unburnt_approx / (SL + SMALL)
    asf = 8.0 * math.exp(10.0 * dF / (8.0 * eta + SMALL))
    wr = (asf * eta / dx) ** (2.0 - D3)
    te = dx * max(0.0, div) / (tau * SL) if tau * SL > SMALL else 0.0
    return max(SL, SL * max(0.0, wr + te)), X, wr, te


@pytest.mark.parametrize("a", [0.0, 300.0, 5600.0], ids=["X0", "no_log", "log_branch"])
def test_ssr_flame_speed_at_strain(oracle, a):
    """S_T of u = U + a x at T = 2100 K (SijSij = a^2, div = a): the three
    regimes of the model -- X = 0, 0 < X below the log threshold, and the
    log branch (a = 5600 1/s gives X ~ 113, where the wrinkling and
    dilatation terms both matter and S_T > S_L).  Catches a wrong constant,
    exponent or branch, or the dilatation term's sign."""
    o, p = make(oracle)
    dx, dy, dz, x, y, z = coords(p)
    set_primitives(o, p, 5.0 + a * x, 0.0, 0.0, x + 1.0, 2100.0)
    o.rhs(0)
    T = 2100.0
    nu = float(mu_table(p, T)) / rho_of(p, T)
    want, X, wr, te = ssr_flame_speed(p, dx, nu, a * a, a)
    if a == 5600.0:
        assert X > 50 and wr > 0.3 and te > 1.0 and want > 3 * p.s_l
    got = o.field("ST")[1:-1]
    close(got, np.full(got.shape, want), 1e-12)


# --------------------------------------------------------------------------
# momentum RHS (:2552-2691 with the conservative flux :683-727)
# --------------------------------------------------------------------------

def test_momentum_viscous_term(oracle):
    """u = a sin(k y), v = w = 0, uniform T, G = 1: convection vanishes
    exactly (no transverse mass flux, u uniform along x), so
    d(rho u)/dt = (mu + rho nu_t) * a' sin(k y) (2 cos(k dy) - 2)/dy^2 and the
    other components are zero.  Catches a wrong sign or scale of diffusion,
    mu_eff without rho*nu_t, or a Laplacian on the unfiltered field."""
    o, p = make(oracle, cs_unburned=0.1, cs_burnt=0.2, cs_flame_boost=0.3)
    dx, dy, dz, x, y, z = coords(p)
    kk = 2 * np.pi / p.ly
    a, T = 2.0, 2100.0
    set_primitives(o, p, a * np.sin(kk * y), 0.0, 0.0, x + 1.0, T)
    out = o.rhs(0)
    ap = a * filter_gain(kk * dy)
    rho = rho_of(p, T)
    S = np.abs(ap * np.cos(kk * y) * math.sin(kk * dy) / dy)
    mu_eff = float(mu_table(p, T)) + rho * (0.2 * dx) ** 2 * S
    lap = ap * np.sin(kk * y) * (2 * math.cos(kk * dy) - 2) / dy ** 2
    sl = np.s_[4:-4]   # x-lines of 7 points stay off the unfiltered end planes
    close(out[0][sl], (mu_eff * lap)[sl], 1e-10)
    # the rho*nu_t part is ~1 % of mu_eff here: it is resolved at this tolerance
    assert np.abs(rho * (0.2 * dx) ** 2 * S * lap).max() > 1e-3 * np.abs(mu_eff * lap).max()
    close(out[1][sl], np.zeros(x[sl].shape), 0.0, 1e-12)
    close(out[2][sl], np.zeros(x[sl].shape), 0.0, 1e-12)


def test_momentum_x_flux_linear_velocity(oracle):
    """u = U + a x at uniform T: the faces are exact (WENO5 and first-order
    upwind reproduce a linear field at the cell, central averages at the
    face), so d(rho u)/dt = -rho [u_{i+1/2} f_{i+1/2} - u_{i-1/2} f_{i-1/2}]/dx
    with f = U + a x_{i+-1/2} on planes 3..nx-4 and f = u_i, u_{i-1} (upwind
    cell values) in the first-order band (planes 1, 2, nx-3, nx-2).  Catches a
    wrong face average, a missing band, or a flux without rho."""
    o, p = make(oracle)
    dx, dy, dz, x, y, z = coords(p)
    U, a, T = 5.0, 100.0, 2100.0
    set_primitives(o, p, U + a * x, 0.0, 0.0, x + 1.0, T)
    out = o.rhs(0)
    rho = rho_of(p, T)
    xi = np.arange(p.nx) * dx
    uc = U + a * xi
    want = np.zeros(p.nx)
    for i in range(1, p.nx - 1):
        up, um = 0.5 * (uc[i + 1] + uc[i]), 0.5 * (uc[i] + uc[i - 1])
        if i < 3 or i > p.nx - 4:
            fp, fm = uc[i], uc[i - 1]
        else:
            fp, fm = U + a * (xi[i] + 0.5 * dx), U + a * (xi[i] - 0.5 * dx)
        want[i] = -rho * (up * fp - um * fm) / dx
    got = out[0]
    for i in range(1, p.nx - 1):
        close(got[i], np.full(got[i].shape, want[i]), 1e-11)
    # interior planes: -rho d(u^2)/dx = -2 rho a u exactly
    assert want[10] == pytest.approx(-2 * rho * a * uc[10], rel=1e-12)
    # the level set: -u dphi/dx + S_T |grad phi| with slope 1 (planes 1 and nx-2
    # also carry an SGS flux: nu_t is 0 on the end planes)
    close(out[3][2:-2], (-(U + a * x) + o.field("ST"))[2:-2], 1e-11)


def test_momentum_transverse_convection_linear_weno_limit(oracle):
    """u = a sin(k y) with a = 1e-6 (every WENO5 smoothness indicator far below
    eps, so the nonlinear weights equal the linear ones) carried by v = V:
    d(rho u)/dt = -rho V Im(H e^{iky}) a'/dy + mu_eff a' sin(ky)(2cos(k dy)-2)/dy^2,
    H the linear WENO5 face-difference factor, upwinded for V > 0 and, mirrored,
    for V < 0.  Catches a wrong upwind side, a wrong linear weight or
    candidate, or a missing 1/dy."""
    for V in (3.0, -3.0):
        o, p = make(oracle)
        dx, dy, dz, x, y, z = coords(p)
        kk = 2 * np.pi / p.ly
        a, T = 1e-6, 2100.0
        set_primitives(o, p, a * np.sin(kk * y), V, 0.0, x + 1.0, T)
        out = o.rhs(0)
        ap = a * filter_gain(kk * dy)
        rho = rho_of(p, T)
        H = weno5_linear_diff(kk * dy) if V > 0 else -np.conj(weno5_linear_diff(kk * dy))
        conv = rho * V * ap * np.imag(H * np.exp(1j * kk * y)) / dy
        S = np.abs(ap * np.cos(kk * y) * math.sin(kk * dy) / dy)
        mu_eff = float(mu_table(p, T)) + rho * (p.cs_burnt * dx) ** 2 * S
        lap = ap * np.sin(kk * y) * (2 * math.cos(kk * dy) - 2) / dy ** 2
        sl = np.s_[4:-4]   # x-lines of 7 points stay off the unfiltered end planes
        close(out[0][sl], (-conv + mu_eff * lap)[sl], 1e-7)
        close(out[1][sl], np.zeros(x[sl].shape), 0.0, 1e-13)


def test_buoyancy_and_face_density(oracle):
    """T = T0 + b sin(k y), v = V, u = w = 0: the x-momentum is the buoyancy
    (rho~ - rho_bar) g alone, rho_bar the mean of rho(T~) over every cell
    (unfiltered on the x end planes), and the y-momentum is the mass-flux
    divergence of uniform v through face-averaged densities,
    -V^2 (rho~_{j+1} - rho~_{j-1}) / (2 dy).  Catches a wrong gravity sign or
    value, a rho_bar over the interior only, or a face density taken from
    one side."""
    o, p = make(oracle)
    dx, dy, dz, x, y, z = coords(p)
    kk = 2 * np.pi / p.ly
    T0, b, V = 1200.0, 400.0, 2.0
    T = T0 + b * np.sin(kk * y)
    set_primitives(o, p, 0.0, V, 0.0, x + 1.0, T)
    out = o.rhs(0)
    Tf = T0 + b * filter_gain(kk * dy) * np.sin(kk * y)
    Tf[0], Tf[-1] = T[0], T[-1]
    rf = rho_of(p, Tf)
    rbar = rf.mean()
    sl = np.s_[1:-1]
    close(out[0][sl], ((rf - rbar) * G_GRAVITY)[sl], 1e-12)
    want1 = -V * V * (np.roll(rf, -1, axis=1) - np.roll(rf, 1, axis=1)) / (2 * dy)
    close(out[1][sl], want1[sl], 1e-12)
    close(out[2][sl], np.zeros(x[sl].shape), 0.0, 1e-12)


@pytest.mark.parametrize("strength", [2.0361, 1e5], ids=["model", "clipped"])
def test_flame_force(oracle, strength):
    """phi = x - x0 + A sin(k y), zero velocity, uniform T: the only momentum
    source is the flame-acceleration force, F = A_f rho W(G) grad G with the
    x part weighted by 0.05, the window W = exp(-(G - c)^2 / 2w^2), central
    differences of G = c(phi~), only inside 0.01 < G < 0.99, each component
    clipped to +-500 (strength 1e5 saturates it).  Catches a wrong window,
    band, x weight, gradient, or clip."""
    o, p = make(oracle, flame_accel_strength=strength)
    dx, dy, dz, x, y, z = coords(p)
    kk = 2 * np.pi / p.ly
    x0, A, T = 0.0325, 0.5e-3, 2100.0
    phi = x - x0 + A * np.sin(kk * y)
    set_primitives(o, p, 0.0, 0.0, 0.0, phi, T)
    out = o.rhs(0)
    pf = x - x0 + A * filter_gain(kk * dy) * np.sin(kk * y)
    pf[0], pf[-1] = phi[0], phi[-1]
    g = progress(pf, dx)
    gx = (np.roll(g, -1, axis=0) - np.roll(g, 1, axis=0)) / (2 * dx)
    gy = (np.roll(g, -1, axis=1) - np.roll(g, 1, axis=1)) / (2 * dy)
    win = np.exp(-((g - p.flame_accel_center) ** 2) / (2 * p.flame_accel_width ** 2))
    band = (g > 0.01) & (g < 0.99)
    common = strength * rho_of(p, T) * win
    Fm = p.max_flame_force
    fx = np.where(band, np.clip(common * gx * p.weight_x, -Fm, Fm), 0.0)
    fy = np.where(band, np.clip(common * gy, -Fm, Fm), 0.0)
    sl = np.s_[1:-1]
    assert band[sl].sum() > 100
    if strength > 1e3:
        assert np.abs(fx[sl]).max() == Fm and np.abs(fy[sl]).max() == Fm
    close(out[0][sl], fx[sl], 1e-10, 1e-10)
    close(out[1][sl], fy[sl], 1e-10, 1e-10)
    close(out[2][sl], np.zeros(x[sl].shape), 0.0, 1e-10)


# --------------------------------------------------------------------------
# scalar RHS: T (:2909-2962) and the level set (:2706-2825)
# --------------------------------------------------------------------------

def test_temperature_equation_terms(oracle):
    """T = T* + b sin(k y) with b = 1e-5 K (linear WENO5 limit), v = V, and a
    shear u = a sin(k y) that makes nu_t vary along y:
    dT/dt = -V Im(H e^{iky}) b'/dy + alpha(T~) lap T~
            + [g_{j+1/2}(T~_{j+1} - T~_j) - g_{j-1/2}(T~_j - T~_{j-1})]/dy^2
            + k_r (T* - T~),
    alpha = mu / (rho Pr) (lambda = mu cp / Pr), g = (nu_t + nu_t')/(2 Pr_t)
    face-averaged, T* = T_b at G = 1.  Catches a wrong SGS face average or
    Pr_t, alpha with cp left in, or a wrong nudging target or rate."""
    o, p = make(oracle, cs_unburned=0.1, cs_burnt=0.3, cs_flame_boost=0.2)
    dx, dy, dz, x, y, z = coords(p)
    kk = 2 * np.pi / p.ly
    b, V, a = 1e-5, 3.0, 5.0
    Ts = p.burnt_temperature
    T = Ts + b * np.sin(kk * y)
    set_primitives(o, p, a * np.sin(kk * y), V, 0.0, x + 1.0, T)
    out = o.rhs(0)
    gain = filter_gain(kk * dy)
    Tf = Ts + b * gain * np.sin(kk * y)
    rf = rho_of(p, Tf)
    alpha = mu_table(p, Tf) / (rf * p.prandtl)
    conv = V * b * gain * np.imag(weno5_linear_diff(kk * dy) * np.exp(1j * kk * y)) / dy
    lap = b * gain * np.sin(kk * y) * (2 * math.cos(kk * dy) - 2) / dy ** 2
    S = np.abs(a * gain * np.cos(kk * y) * math.sin(kk * dy) / dy)
    nut = (0.3 * dx) ** 2 * S
    gp = (np.roll(nut, -1, axis=1) + nut) / (2 * p.pr_t)
    gm = (nut + np.roll(nut, 1, axis=1)) / (2 * p.pr_t)
    dT = np.roll(Tf, -1, axis=1) - Tf, Tf - np.roll(Tf, 1, axis=1)
    sgs = (gp * dT[0] - gm * dT[1]) / dy ** 2
    nudge = p.relaxation_coeff * (Ts - Tf)
    want = -conv + alpha * lap + sgs + nudge
    sl = np.s_[4:-4]   # x-lines of 7 points stay off the unfiltered end planes
    # the SGS term is about 4e-4 of the total: resolved at this tolerance
    assert np.abs(sgs[sl]).max() > 1e-4 * np.abs(want[sl]).max()
    close(out[4][sl], want[sl], 1e-6)


def test_level_set_sgs_and_front_speed(oracle):
    """phi = x - x0 + A sin(k y) far outside the flame band (G = 1: no
    artificial diffusion), a shear u = a sin(k z) (nu_t varies along z only)
    and v = w = 0:
    dphi/dt = -u~ + S_T |grad phi~| + (nu_t / Sc_t) d2phi~/dy2 (nu_t uniform along y),
    with |grad phi~|^2 = 1 + (A' cos(ky) sin(k dy)/dy)^2 from central
    differences; the SSR speed S_T is read from the oracle (pinned above).
    Catches Sc_t swapped for Pr_t, a gradient magnitude without the
    transverse part, or a wrong sign of the propagation term."""
    o, p = make(oracle)
    dx, dy, dz, x, y, z = coords(p)
    ky, kz = 2 * np.pi / p.ly, 2 * np.pi / p.lz
    x0, A, a = -1.0, 2e-4, 3.0
    phi = x - x0 + A * np.sin(ky * y)
    set_primitives(o, p, a * np.sin(kz * z), 0.0, 0.0, phi, p.burnt_temperature)
    out = o.rhs(0)
    gy = filter_gain(ky * dy)
    Ap = A * gy
    grad_y = Ap * np.cos(ky * y) * math.sin(ky * dy) / dy
    gmag = np.sqrt(1.0 + grad_y ** 2)
    S = np.abs(a * filter_gain(kz * dz) * np.cos(kz * z) * math.sin(kz * dz) / dz)
    nut = (p.cs_burnt * dx) ** 2 * S
    lap_y = Ap * np.sin(ky * y) * (2 * math.cos(ky * dy) - 2) / dy ** 2
    uf = a * filter_gain(kz * dz) * np.sin(kz * z)          # carries phi along x, slope 1
    want = -uf + o.field("ST") * gmag + nut / p.sc_t * lap_y
    sl = np.s_[4:-4]   # x-lines of 7 points stay off the unfiltered end planes
    close(out[3][sl], want[sl], 1e-9)


def test_nudging_target_inside_the_flame(oracle):
    """Uniform T and phi = 0 (G = 1/2 everywhere, no gradients): the only T
    source is the relaxation to the G-weighted target,
    dT/dt = k_r (T_u + G (T_b - T_u) - T).  Catches a target fixed at T_b or
    weighted by 1 - G."""
    o, p = make(oracle)
    T = 1000.0
    set_primitives(o, p, 0.0, 0.0, 0.0, 0.0, T)
    out = o.rhs(0)
    Ts = p.inlet_temperature + 0.5 * (p.burnt_temperature - p.inlet_temperature)
    got = out[4][1:-1]
    close(got, np.full(got.shape, p.relaxation_coeff * (Ts - T)), 1e-13)


# --------------------------------------------------------------------------
# SSP-RK3 stage clip (:4440-4443 clip structure)
# --------------------------------------------------------------------------

def test_ssprk3_stage_clip(oracle):
    """Uniform T = 3500 K relaxing to T* = 2100 K: each Shu-Osher stage is
    clipped to [200, 3000] before the next one uses it, so one step gives
    y1 = clip(T0 + dt r(T0)), y2 = clip(3/4 T0 + 1/4 (y1 + dt r(y1))),
    T1 = clip(1/3 T0 + 2/3 (y2 + dt r(y2))) with r(T) = k (T* - T).
    Catches a clip applied only at the end, or not at all."""
    o, p = make(oracle)
    dx, dy, dz, x, y, z = coords(p)
    T0 = 3500.0
    set_primitives(o, p, 5.0, 0.0, 0.0, x + 1.0, T0)
    dt = o.step(1).dt
    k, Ts = p.relaxation_coeff, p.burnt_temperature

    def clip(t):
        return min(3000.0, max(200.0, t))

    y1 = clip(T0 + dt * k * (Ts - T0))
    y2 = clip(0.75 * T0 + 0.25 * (y1 + dt * k * (Ts - y1)))
    want = clip(T0 / 3.0 + 2.0 / 3.0 * (y2 + dt * k * (Ts - y2)))
    unclipped = Ts + (T0 - Ts) * (1 - k * dt + (k * dt) ** 2 / 2 - (k * dt) ** 3 / 6)
    assert abs(want - unclipped) > 10.0
    got = o.get_state(5)[4][12:-12]
    close(got, np.full(got.shape, want), 1e-12)


# --------------------------------------------------------------------------
# DT rule (:4399-4411)
# --------------------------------------------------------------------------

def test_dt_rule_convective_diffusive_and_fallback(oracle):
    """Plug flow at U and uniform T: the first step takes C dx / (U + 1e-10)
    (no diffusive history yet); the second takes the minimum of that and
    C_d dx^2 / max nu, C_d dx^2 / max alpha, with nu = mu/rho and
    alpha = nu / Pr at T.  With C = 50 the diffusive limit of alpha binds.
    A state at rest falls back to U_in.  Catches nu in place of alpha, a
    missing Courant number, or dx instead of dx^2."""
    T, U = 2100.0, 5.0
    nu = float(mu_table(oracle.params(DUCT), T)) / (1.2 * 300.0 / T)
    for C in (0.3, 50.0):
        o, p = make(oracle, courant=C)
        dx, dy, dz, x, y, z = coords(p)
        set_primitives(o, p, U, 0.0, 0.0, x + 1.0, T)
        dt1 = o.step(1).dt
        assert dt1 == pytest.approx(C * dx / (U + SMALL), rel=1e-14)
        dt2 = o.step(1).dt
        dtc = C * dx / (U + SMALL)
        dtn = p.courant_diffusion * dx * dx / (nu + SMALL)
        dta = p.courant_diffusion * dx * dx / (nu / p.prandtl + SMALL)
        assert dt2 == pytest.approx(min(dtc, dtn, dta), rel=1e-12)
        if C > 1:
            assert dt2 == pytest.approx(dta, rel=1e-12) and dta < dtn
    o, p = make(oracle)
    dx = p.lx / p.nx
    set_primitives(o, p, 0.0, 0.0, 0.0, coords(p)[3] + 1.0, T)
    assert o.step(1).dt == pytest.approx(p.courant * dx / (p.inlet_velocity + SMALL), rel=1e-14)


# --------------------------------------------------------------------------
# inflow / outflow planes (:2967-3161 roles)
# --------------------------------------------------------------------------

def test_inflow_and_outflow_planes(oracle):
    """After a step: the inflow plane carries rho_u (U + u'), rho_u v',
    rho_u w' from the SEM plane, T = T_u and phi continued with slope 1 from
    plane 1 (phi_0 = phi_1 - dx); the outflow plane repeats plane nx-2 in
    every field.  Catches a fluctuation applied without rho_u, an inflow
    phi copied instead of extrapolated, or an outflow copy from the wrong
    plane."""
    o, p = make(oracle, num_vortices=200)
    o.step(1)
    q = o.get_state(5)
    P = p.ny * p.nz
    pl = o.raw_field("inlet", 3 * P).reshape(3, p.ny, p.nz)
    assert np.abs(pl).max() > 0.1
    ru = rho_of(p, p.inlet_temperature)
    assert np.array_equal(q[0][0], ru * (p.inlet_velocity + pl[0]))
    assert np.array_equal(q[1][0], ru * pl[1])
    assert np.array_equal(q[2][0], ru * pl[2])
    assert np.array_equal(q[3][0], q[3][1] - p.lx / p.nx)
    assert np.all(q[4][0] == p.inlet_temperature)
    for f in range(5):
        assert np.array_equal(q[f][-1], q[f][-2]), f


# --------------------------------------------------------------------------
# SEM (:943-1075 structure, Gaussian eddies)
# --------------------------------------------------------------------------

def test_sem_single_gaussian_eddy_shape(oracle):
    """One eddy at (x_e, y_e, z_e): before normalisation the inlet plane is
    s * exp(-pi d^2 / L^2) with d the minimum-image distance in y and z
    (periodic) and x_e in x; the normalisation then removes the plane mean and
    scales each component to rms U I, so component c is
    sign(s_c) U I (f - <f>) / rms(f - <f>).  The eddy sits half a cell from
    y = 0 and one cell below z = l_z so the wrap matters.  Catches the cone
    shape, a missing pi, or a plane that does not wrap."""
    o, p = make(oracle, num_vortices=1)
    L = p.turbulent_length_scale
    dy, dz = p.ly / p.ny, p.lz / p.nz
    xe, ye, ze = 0.3 * L, 0.5 * dy, p.lz - dz
    e = np.array([[xe, ye, ze, 0.8, -0.4, 0.2, 0.0]])
    L_ = oracle.lib()
    import ctypes as C
    assert L_.ora_duct_set_eddies(o.h, e.ctypes.data_as(C.POINTER(C.c_double)), 1) == 0
    assert L_.ora_duct_sem(o.h, 0.0) == 0
    yy = np.arange(p.ny)[:, None] * dy
    zz = np.arange(p.nz)[None, :] * dz
    ddy = yy - ye
    ddy -= p.ly * np.rint(ddy / p.ly)
    ddz = zz - ze
    ddz -= p.lz * np.rint(ddz / p.lz)
    f = np.exp(-np.pi * (xe ** 2 + ddy ** 2 + ddz ** 2) / L ** 2)
    fh = f - f.mean()
    fh = fh / np.sqrt((fh ** 2).mean()) * p.inlet_velocity * p.turbulence_intensity
    pl = o.raw_field("inlet", 3 * p.ny * p.nz).reshape(3, p.ny, p.nz)
    for c, s in enumerate((0.8, -0.4, 0.2)):
        close(pl[c], np.sign(s) * fh, 1e-12)


def test_sem_eddy_advection_and_recycling(oracle):
    """Eddies advance by U dt in x; one that passes x = +L re-enters at
    x - 2L with generation + 1 and fresh position and strengths in range."""
    o, p = make(oracle, num_vortices=2)
    L, U, dt = p.turbulent_length_scale, p.inlet_velocity, 2e-4
    e = np.array([[0.1 * L, 0.01, 0.01, 0.5, 0.5, 0.5, 3.0],
                  [L - 0.25 * U * dt, 0.02, 0.02, -0.5, 0.5, -0.5, 7.0]])
    import ctypes as C
    L_ = oracle.lib()
    assert L_.ora_duct_set_eddies(o.h, e.ctypes.data_as(C.POINTER(C.c_double)), 2) == 0
    assert L_.ora_duct_sem(o.h, dt) == 0
    got = o.raw_field("eddies", 14).reshape(2, 7)
    assert got[0, 0] == pytest.approx(0.1 * L + U * dt, rel=1e-14)
    assert np.array_equal(got[0, 1:], e[0, 1:])
    assert got[1, 0] == pytest.approx(L - 0.25 * U * dt + U * dt - 2 * L, rel=1e-12)
    assert got[1, 6] == 8.0
    assert 0 <= got[1, 1] < p.ly and 0 <= got[1, 2] < p.lz
    assert np.all(np.abs(got[1, 3:6]) <= 1.0) and not np.array_equal(got[1, 1:6], e[1, 1:6])


# --------------------------------------------------------------------------
# progress map width
# --------------------------------------------------------------------------

def test_progress_map_width(oracle):
    """c(phi) = (1 + tanh(phi / 2 dx)) / 2 is the logistic 1/(1 + e^{-phi/dx}):
    c(+-dx) = 1/(1 + e^{-+1}), c(3 dx) = 1/(1 + e^{-3}).  Catches a front
    width of dx/2 or 2 dx."""
    L = oracle.lib()
    dx = 1e-3
    for m in (-3.0, -1.0, 1.0, 3.0):
        assert L.ora_duct_progress(m * dx, dx) == pytest.approx(1.0 / (1.0 + math.exp(-m)), rel=2e-15)


# --------------------------------------------------------------------------
# statistics (:3167-3219) and probe rows (:3403-3471)
# --------------------------------------------------------------------------

def primitives(p, q):
    r = rho_of(p, q[4])
    return q[0] / r, q[1] / r, q[2] / r


def test_plane_profiles_and_per_cell_statistics(oracle):
    """After two steps the y-z plane sums of u, v, w, G, uu, vv, ww, uv, uw, vw
    at every x equal the sums over the two post-step states, and the per-cell
    FP64 accumulators equal the per-cell sums of the same samples.  Catches a
    product of the wrong components, a missed sample, or samples taken before
    the boundary update."""
    o, p = make(oracle, num_vortices=50, stats_per_cell=1)
    samples = []
    for s in range(2):
        o.step(1)
        samples.append(o.get_state(5))
    dx = p.lx / p.nx
    acc = np.zeros((p.nx, 10))
    cell = [np.zeros((p.nx, p.ny, p.nz)) for _ in range(10)]
    for q in samples:
        u, v, w = primitives(p, q)
        g = progress(q[3], dx)
        vals = [u, v, w, g, u * u, v * v, w * w, u * v, u * w, v * w]
        for a_, f in enumerate(vals):
            acc[:, a_] += f.sum(axis=(1, 2))
            cell[a_] += f
    prof = o.raw_field("profiles", 10 * p.nx).reshape(p.nx, 10)
    for a_ in range(10):
        close(prof[:, a_], acc[:, a_], 1e-12, 1e-15)
    st, cnt = o.stats()
    assert cnt == 2
    for a_ in range(10):
        close(st[a_], cell[a_], 1e-13, 1e-15)


def test_probe_rows_brute_force(oracle):
    """A flame front along y (phi = phi(y), periodic): at each station's
    column (i = int(x/dx), k = nz/2) the first G = 1/2 crossing scanning j
    upward is interpolated linearly, and every j within 1 m of it yields a
    row (x, y_flame, y, y - y_flame, u, v, w, G, nu_t, |S|) from the post-step
    state and the last stage's nu_t and |S|."""
    o, p = make(oracle, num_vortices=50)
    dx, dy, dz, x, y, z = coords(p)
    ky, kz = 2 * np.pi / p.ly, 2 * np.pi / p.lz
    # the front and the velocity vary along z too, so the column k = nz/2 matters
    phi = -4e-3 * np.cos(ky * y) + 1e-3 + 1e-3 * np.sin(kz * z)
    set_primitives(o, p, 5.0 + 0.5 * np.cos(kz * z), 0.3 * np.sin(ky * y), 0.2 * np.sin(kz * z), phi, 1200.0)
    o.step(1)
    q = o.get_state(5)
    u, v, w = primitives(p, q)
    nut, smag = o.field("nu_t"), o.field("S_mag")
    kp = p.nz // 2
    want = []
    for n in range(p.num_probe_x):
        xp = p.probe_x[n]
        ip = int(xp / dx)
        if ip <= 0 or ip >= p.nx - 1:
            continue
        col = progress(q[3][ip, :, kp], dx)
        yf = -1.0
        for j in range(p.ny - 1):
            g1, g2 = col[j], col[j + 1]
            if (g1 - 0.5) * (g2 - 0.5) <= 0.0:
                yf = j * dy if abs(g2 - g1) < 1e-9 else j * dy + (0.5 - g1) / (g2 - g1) * dy
                break
        if not yf > 0.0:
            continue
        for j in range(p.ny):
            yp = j * dy
            if abs(yp - yf) > 1.0:
                continue
            want.append((0, n, xp, yf, yp, yp - yf, u[ip, j, kp], v[ip, j, kp], w[ip, j, kp], col[j],
                         nut[ip, j, kp], smag[ip, j, kp]))
    rows = o.probe_rows()
    assert len(want) == p.num_probe_x * p.ny and len(rows) == len(want)
    for r, wv in zip(rows, want):
        got = (r.timestep, r.station, r.x_location, r.y_flame, r.y_physical, r.y_relative, r.u, r.v, r.w, r.G,
               r.nu_t, r.S_mag)
        assert got[:2] == wv[:2]
        np.testing.assert_allclose(got[2:], wv[2:], rtol=1e-12, atol=1e-15)
