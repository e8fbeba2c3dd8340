"""Closed-form pins of the oracle's duct scheme (no reference counterpart).

Each test is chosen so that a plausible mistake fails it: a wrong linear
weight or candidate stencil (WENO5 exactness, order of accuracy), a wrong
limiter branch (minmod table, TVD bound), a wrong Runge-Kutta coefficient or
nudging rate (exact amplification factor of the stiff relaxation), a wrong
wrap (translation invariance), a wrong sign in the level-set equation (planar
front speed), a wrong normalisation of the inflow (mean and rms).
"""
import ctypes as C

import numpy as np
import pytest

DUCT = 1   # LES_PRESET_DUCT64


def weno(L, v):
    return L.ora_weno5_face(*[float(x) for x in v])


def test_weno5_linear_exact(oracle):
    L = oracle.lib()
    for a, b in [(1.0, 0.3), (-2.0, 5.0), (0.0, -1e-3)]:
        v = [a + b * n for n in range(5)]          # point values on x = n
        assert weno(L, v) == pytest.approx(a + b * 2.5, rel=1e-13, abs=1e-15)


def test_weno5_quartic_cell_averages_exact_in_linear_limit(oracle):
    """Small amplitude (IS << eps): nonlinear weights equal the linear ones,
    and the linear 5-point reconstruction of cell averages is exact for
    polynomials of degree <= 4 at the face."""
    L = oracle.lib()
    rng = np.random.default_rng(2)
    for _ in range(50):
        c = 1e-9 * rng.standard_normal(5)          # IS ~ 1e-17 << eps = 1e-6
        P = np.polynomial.Polynomial(c)
        Q = P.integ()
        avg = [Q(n + 1) - Q(n) for n in range(5)]  # averages over [n, n+1]
        assert weno(L, avg) == pytest.approx(P(3.0), rel=1e-8, abs=1e-22)


def test_weno5_fifth_order(oracle):
    L = oracle.lib()
    errs = []
    for n in (32, 64, 128):
        h = 1.0 / n
        e = 0.0
        for i in range(n):
            x = [(i + m - 2) * h for m in range(5)]
            avg = [(np.cos(2 * np.pi * xx) - np.cos(2 * np.pi * (xx + h))) / (2 * np.pi * h) for xx in x]
            face = (i + 1) * h
            e = max(e, abs(weno(L, avg) - np.sin(2 * np.pi * face)))
        errs.append(e)
    order = np.log2(errs[0] / errs[1]), np.log2(errs[1] / errs[2])
    assert min(order) > 4.5, order


def test_minmod_and_muscl_tvd(oracle):
    L = oracle.lib()
    cases = [(1.0, 2.0, 1.0), (2.0, 1.0, 1.0), (-1.0, -3.0, -1.0), (1.0, -1.0, 0.0), (0.0, 5.0, 0.0)]
    for a, b, want in cases:
        assert L.ora_minmod(a, b) == want
    rng = np.random.default_rng(4)
    for _ in range(500):
        v1, v2, v3 = rng.standard_normal(3)
        f = L.ora_muscl_face(v1, v2, v3)
        assert min(v2, v3) - 1e-15 <= f <= max(v2, v3) + 1e-15   # no new extrema


def test_table_lookup(oracle):
    L = oracle.lib()
    T = np.arange(300.0, 2301.0)
    mu = 1.8e-5 * (T / 300.0) ** 0.7
    tab = (C.c_double * len(T))(*mu)
    for t in (300.0, 301.0, 1000.0, 2300.0):
        assert L.ora_duct_table(tab, 300.0, 1.0, len(T), t) == pytest.approx(1.8e-5 * (t / 300.0) ** 0.7, rel=1e-14)
    # midpoints: linear interpolation error <= dT^2/8 max|mu''|
    bound = 1.0 / 8.0 * 1.8e-5 * 0.7 * 0.3 / 300.0 ** 2 * 1.05
    for t in np.arange(300.5, 2300.0, 97.0):
        err = abs(L.ora_duct_table(tab, 300.0, 1.0, len(T), t) - 1.8e-5 * (t / 300.0) ** 0.7)
        assert err <= bound
    # clamped beyond the ends
    assert L.ora_duct_table(tab, 300.0, 1.0, len(T), 100.0) == mu[0]
    assert L.ora_duct_table(tab, 300.0, 1.0, len(T), 5000.0) == mu[-1]


def test_progress_map(oracle):
    L = oracle.lib()
    assert L.ora_duct_progress(0.0, 1e-3) == 0.5
    for phi in (1e-4, 5e-4, 3e-3):
        assert L.ora_duct_progress(phi, 1e-3) + L.ora_duct_progress(-phi, 1e-3) == pytest.approx(1.0, abs=1e-15)


def uniform_duct(oracle, T=2100.0, x0=-1.0, **over):
    """Plug flow at U with uniform T and a planar level set far from the domain
    (c(phi) = 1 everywhere: outside every flame band)."""
    p = oracle.params(DUCT, reinit_interval=0, num_vortices=0, **over)
    o = oracle.Oracle(p)
    o.init_state()
    nx, ny, nz = o.shape
    x = np.arange(nx)[:, None, None] * (p.lx / nx) * np.ones((nx, ny, nz))
    rho = 1.2 * 300.0 / T
    st = [rho * 5.0 * np.ones(o.shape), np.zeros(o.shape), np.zeros(o.shape), x - x0, T * np.ones(o.shape)]
    o.set_state(st)
    return o, p, st


def test_ssprk3_nudging_amplification(oracle):
    """dT/dt = k (T* - T) on a uniform state: one SSP-RK3 step multiplies the
    deviation by exactly 1 + z + z^2/2 + z^3/6 (z = -k dt); the planar level
    set moves at -U + S_T with S_T from the SSR model at zero strain."""
    T0 = 2000.0
    o, p, st = uniform_duct(oracle, T=T0)
    rep = o.step(1)
    dt = rep.dt
    z = -p.relaxation_coeff * dt
    Tstar = 2100.0                         # c(phi) = 1
    want = Tstar + (T0 - Tstar) * (1 + z + z * z / 2 + z ** 3 / 6)
    T = o.get_state(5)[4]
    inner = T[12:-12]          # 3 stages x (filter + WENO) reach <= 12 cells from the x planes
    assert np.allclose(inner, want, rtol=0, atol=1e-9 * Tstar)


def test_planar_level_set_speed(oracle):
    """At T = T* (no nudging, constant density so u stays U) the planar level
    set phi = x + 1 moves at -U + S_T, S_T from the SSR model at zero strain:
    X = 0 -> Delta/eta = 1, eta = dx/(1 + 1e-10)."""
    o, p, st = uniform_duct(oracle, T=2100.0)
    dt = o.step(1).dt
    dx = p.lx / p.nx
    eta = dx / (1.0 + 1e-10)
    dF = p.nu_unburnt_approx / (p.s_l + 1e-10)
    asf = 8.0 * np.exp(10.0 * dF / (8.0 * eta + 1e-10))
    ST = max(p.s_l, p.s_l * max(0.0, (asf * eta / dx) ** (2.0 - p.fractal_d3)))
    phi = o.get_state(5)[3]
    x = np.arange(p.nx)[:, None, None] * dx
    dphi = (phi - (x + 1.0))[12:-12]
    assert np.allclose(dphi, dt * (-p.inlet_velocity + ST), rtol=1e-9, atol=1e-15)


def test_translation_invariance_y(oracle):
    """Periodic y: the RHS of a state shifted by m rows is the shifted RHS."""
    p = oracle.params(DUCT, num_vortices=0)
    o = oracle.Oracle(p)
    o.init_state()
    rng = np.random.default_rng(9)
    st = [a.copy() for a in o.get_state(5)]
    rho = 1.2 * 300.0 / st[4]
    for f in range(3):
        st[f] = st[f] + rho * rng.uniform(-1, 1, st[f].shape)
    st[3] = st[3] + 1e-4 * rng.standard_normal(st[3].shape)
    o.set_state(st)
    r0 = [a.copy() for a in o.rhs(0)[:5]]
    o.set_state([np.roll(a, 5, axis=1) for a in st])
    r1 = o.rhs(0)[:5]
    for a, b in zip(r0, r1):
        assert np.allclose(np.roll(a, 5, axis=1), b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_sem_normalisation(oracle):
    p = oracle.params(DUCT)
    o = oracle.Oracle(p)
    o.init_state()
    oracle.lib().ora_duct_sem(o.h, 1e-5)
    P = p.ny * p.nz
    pl = o.raw_field("inlet", 3 * P).reshape(3, P)
    for a in range(3):
        assert abs(pl[a].mean()) < 1e-14
        assert np.sqrt((pl[a] ** 2).mean()) == pytest.approx(p.inlet_velocity * p.turbulence_intensity, rel=1e-12)


def test_reinit_fixed_point_and_convergence(oracle):
    p = oracle.params(DUCT, num_vortices=0)
    o = oracle.Oracle(p)
    o.init_state()
    st = o.get_state(5)
    dx = p.lx / p.nx
    phi_plane = st[3].copy()                       # x - x0: |grad phi| = 1
    oracle.lib().ora_duct_reinit(o.h, 5)
    assert np.allclose(o.get_state(5)[3][1:-1], phi_plane[1:-1], rtol=0, atol=1e-15)
    st[3] = 2.0 * phi_plane                        # |grad phi| = 2
    o.set_state(st)
    oracle.lib().ora_duct_reinit(o.h, 60)
    phi = o.get_state(5)[3]
    g = np.gradient(phi[:, 3, 3], dx)
    i0 = int(round(p.flame_x0 / dx))
    assert np.all(np.abs(g[i0 - 4: i0 + 5] - 1.0) < 0.05)   # slope relaxes to 1 near the front
    # the zero crossing stays in the same cell
    assert np.sign(phi[i0 - 1, 3, 3]) != np.sign(phi[i0 + 1, 3, 3])


def test_fixed_dt(oracle):
    """A prescribed DT_sim replaces the CFL rule: one step takes exactly that
    dt, and the stiff nudging follows its closed form with it."""
    T0, fdt = 2000.0, 2.0e-5
    o, p, st = uniform_duct(oracle, T=T0, fixed_dt=fdt)
    rep = o.step(1)
    assert rep.dt == fdt
    z = -p.relaxation_coeff * fdt
    want = 2100.0 + (T0 - 2100.0) * (1 + z + z * z / 2 + z ** 3 / 6)
    assert np.allclose(o.get_state(5)[4][12:-12], want, rtol=0, atol=1e-9 * 2100.0)
    assert o.step(3).dt == fdt


def test_fixed_dt_validation(lesgpu):
    """fixed_dt must be finite and non-negative, and the reference scheme keeps
    its own DT rule."""
    L = lesgpu.lib()
    ok = lesgpu.params(1, fixed_dt=1e-5)
    assert L.les_params_validate(C.byref(ok)) == 0
    for bad in (lesgpu.params(1, fixed_dt=-1e-5), lesgpu.params(1, fixed_dt=float("nan")),
                lesgpu.params(0, fixed_dt=1e-5)):
        assert L.les_params_validate(C.byref(bad)) != 0
