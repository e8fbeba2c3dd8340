"""Closed-form pins of the oracle's projection relaxation (duct mode; no
reference counterpart).

With b = L psi for a discrete eigenvector psi of the compact Laplacian under
the relaxation's ghost rules (Neumann mirror at the inflow plane, zero at the
outflow plane, periodic y/z), the Richardson sweeps from phi = 0 give
phi_n = (1 - g^n) psi with g = 1 + lambda/d, so the residual RMS after n sweeps
is exactly |g|^n times the initial one, and the mass-flux correction is
-(1 - g^n) grad psi.  A wrong diagonal d, a wrong ghost rule, a wrong
Laplacian term, a wrong sweep count or a wrong correction sign fails it.
"""
import ctypes as C

import numpy as np
import pytest

DUCT = 1


def eigen_state(oracle, mj=2, lk=1, p_x=1, nx=20, ny=16, nz=12):
    p = oracle.params(DUCT, nx=nx, ny=ny, nz=nz, lx=nx * 1e-3, ly=ny * 1e-3, lz=nz * 1.5e-3,
                      num_vortices=0, reinit_interval=0)
    o = oracle.Oracle(p)
    o.init_state()
    dx, dy, dz = p.lx / nx, p.ly / ny, p.lz / nz
    N = nx - 2
    th = (2 * p_x + 1) * np.pi / (2 * N + 1)
    ay, az = 2 * np.pi * mj / ny, 2 * np.pi * lk / nz
    i = np.arange(nx)[:, None, None]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[None, None, :]
    X = np.cos(th * (i - 0.5))
    lam = -(2 - 2 * np.cos(th)) / dx ** 2 - (2 - 2 * np.cos(ay)) / dy ** 2 - (2 - 2 * np.cos(az)) / dz ** 2
    # m_z = A X cos(ay j) sin(az k): its central z-difference is lam * psi with
    # psi = X cos(ay j) cos(az k)
    A = lam * dz / np.sin(az)
    st = o.get_state(5)
    mz = A * X * np.cos(ay * j) * np.sin(az * k) * np.ones((nx, ny, nz))
    st[0] = np.zeros((nx, ny, nz))
    st[1] = np.zeros((nx, ny, nz))
    st[2] = mz
    o.set_state(st)
    psi = X * np.cos(ay * j) * np.cos(az * k) * np.ones((nx, ny, nz))
    d = 2 / dx ** 2 + 2 / dy ** 2 + 2 / dz ** 2
    return o, p, psi, lam, d, (dx, dy, dz)


def project(oracle, o, n):
    r0, r1 = C.c_double(), C.c_double()
    assert oracle.lib().ora_duct_project(o.h, n, C.byref(r0), C.byref(r1)) == 0
    return r0.value, r1.value


@pytest.mark.parametrize("n", [1, 2, 5, 12])
def test_residual_decay_is_geometric(oracle, n):
    o, p, psi, lam, d, _ = eigen_state(oracle)
    r0, r1 = project(oracle, o, n)
    g = 1 + lam / d
    assert 0 < abs(g) < 1
    # initial residual: RMS of b = lam psi over the interior planes
    assert r0 == pytest.approx(abs(lam) * np.sqrt((psi[1:-1] ** 2).mean()), rel=1e-9)
    assert r1 == pytest.approx(r0 * abs(g) ** n, rel=1e-8)
    o.close()


def test_correction_is_minus_grad_phi(oracle):
    n = 3
    o, p, psi, lam, d, (dx, dy, dz) = eigen_state(oracle)
    before = [a.copy() for a in o.get_state(3)]
    project(oracle, o, n)
    after = o.get_state(3)
    cn = 1 - (1 + lam / d) ** n                      # phi_n = cn * psi
    ext = np.concatenate([psi[1:2], psi[1:-1], np.zeros_like(psi[:1])], axis=0)   # ghosts at 0 and nx-1
    gx = (ext[2:] - ext[:-2]) / (2 * dx)
    gy = (np.roll(psi, -1, 1) - np.roll(psi, 1, 1)) / (2 * dy)
    gz = (np.roll(psi, -1, 2) - np.roll(psi, 1, 2)) / (2 * dz)
    for f, g in enumerate([gx, gy[1:-1], gz[1:-1]]):
        want = before[f][1:-1] - cn * g
        scale = max(np.abs(before[2]).max(), 1e-30)
        assert np.abs(after[f][1:-1] - want).max() <= 1e-12 * scale, f
    for f in range(3):                              # boundary planes untouched
        assert np.array_equal(after[f][0], before[f][0]) and np.array_equal(after[f][-1], before[f][-1])
    o.close()


def test_divergence_free_flux_is_unchanged(oracle):
    """m = (0, 0, f(i, j)) has zero central divergence: nothing moves."""
    p = oracle.params(DUCT, num_vortices=0, reinit_interval=0)
    o = oracle.Oracle(p)
    o.init_state()
    st = o.get_state(5)
    rng = np.random.default_rng(3)
    prof = rng.standard_normal((p.nx, p.ny, 1))
    st[0] = np.zeros(o.shape)
    st[1] = np.zeros(o.shape)
    st[2] = prof * np.ones(o.shape)
    o.set_state(st)
    r0, r1 = project(oracle, o, 6)
    assert r0 == 0.0 and r1 == 0.0
    assert np.array_equal(o.get_state(5)[2], st[2])
    o.close()
