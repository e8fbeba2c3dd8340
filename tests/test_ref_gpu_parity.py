"""Parity of the HIP engine's reference scheme (FP64) against the CPU oracle.

Tolerance: 1e-10 relative to each field's own scale (the bound set for the
plain FP64 path); differences may come only from reduction order (the
domain-mean density) and device-libm ulps in exp/log/pow.  Initial states are
compared bit-exactly.
"""
import numpy as np
import pytest

from common import FIELDS, rel_err, synthetic_reference_state

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def pair(lesgpu, oracle):
    p = lesgpu.params()
    mix = lesgpu.mixture()
    g = lesgpu.LesGpu(p, mix)
    o = oracle.Oracle(oracle.params(), oracle.mixture())
    yield g, o
    g.close()
    o.close()


def test_init_state_bitwise(pair):
    g, o = pair
    print("case: init_state 60x20x40", flush=True)
    g.init_state()
    o.init_state()
    for name, a, b in zip(FIELDS, g.download_state(), o.get_state()):
        assert np.array_equal(a, b), name
    assert np.array_equal(g.field("attribute"), o.field("attribute"))


def test_rhs_synthetic_state(pair):
    g, o = pair
    print("case: rhs on synthetic state, t_loop=0 and 600", flush=True)
    g.init_state()
    o.init_state()
    st = synthetic_reference_state(o, seed=3)
    g.upload_state(st)
    o.set_state(st)
    for t_loop in (0, 600):
        rg = g.rhs(t_loop)
        ro = o.rhs(t_loop)
        names = ["drhou", "drhov", "drhow", "dT", "dG", "dY0", "dY1", "dY2", "dY3", "dY4"]
        for nm, a, b in zip(names, rg, ro):
            e = rel_err(a, b)
            assert e <= TOL, (t_loop, nm, e)
        for nm in ["nu_t", "S_mag", "nu", "alpha", "mu_mix", "ST", "u_f", "G_f", "rho_f", "D0"]:
            e = rel_err(g.field(nm), o.field(nm))
            assert e <= TOL, (t_loop, nm, e)


@pytest.mark.parametrize("grid", [(60, 20, 40), (36, 40, 72)], ids=["grouped_launches", "one_launch"])
def test_rhs_both_launch_forms(lesgpu, oracle, grid):
    """The property and RHS kernels run as output groups (one slab of blocks
    per group) on grids of fewer than 2048 line waves, 60x20x40 having 1200,
    and as one slab otherwise, 36x40x72 having 2880; every output of both
    forms against the oracle at 1e-10, on a state that takes the TENO, MUSCL
    and upwind branches."""
    nx, ny, nz = grid
    print("case: reference-scheme rhs %dx%dx%d vs oracle" % grid, flush=True)
    g = lesgpu.LesGpu(lesgpu.params(0, nx=nx, ny=ny, nz=nz), lesgpu.mixture())
    o = oracle.Oracle(oracle.params(0, nx=nx, ny=ny, nz=nz), oracle.mixture())
    try:
        g.init_state()
        o.init_state()
        st = synthetic_reference_state(o, seed=5)
        g.upload_state(st)
        o.set_state(st)
        names = ["drhou", "drhov", "drhow", "dT", "dG", "dY0", "dY1", "dY2", "dY3", "dY4"]
        for t_loop in (0, 600):
            for nm, a, b in zip(names, g.rhs(t_loop), o.rhs(t_loop)):
                assert np.abs(b).max() > 0, nm
                e = rel_err(a, b)
                assert e <= TOL, (t_loop, nm, e)
            for nm in ["nu", "alpha", "mu_mix", "cp_mix", "lambda_mix", "u_f", "w_f", "T_f", "G_f", "rho_f",
                       "D0", "D2", "D4", "gradG"]:
                e = rel_err(g.field(nm), o.field(nm))
                assert e <= TOL, (t_loop, nm, e)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("start", ["init", "synthetic"])
def test_steps_match_oracle(pair, start):
    g, o = pair
    nsteps = 4
    print("case: %d RK4 steps from %s state" % (nsteps, start), flush=True)
    g.init_state()
    o.init_state()
    if start == "synthetic":
        st = synthetic_reference_state(o, seed=7)
        g.upload_state(st)
        o.set_state(st)
    for s in range(nsteps):
        rg = g.step(1)
        ro = o.step(1)
        assert abs(rg.dt - ro.dt) <= 1e-12 * ro.dt, (s, rg.dt, ro.dt)
        for name, a, b in zip(FIELDS, g.download_state(), o.get_state()):
            e = rel_err(a, b)
            assert e <= TOL, (s, name, e)
    for nm in ["nu_t", "S_mag"]:
        assert rel_err(g.field(nm), o.field(nm)) <= TOL, nm
    # probe rows: same stations, same rows, values to the same tolerance
    pg, po = g.probe_rows(), o.probe_rows()
    assert len(pg) == len(po)
    for a, b in zip(pg, po):
        assert (a.timestep, a.station) == (b.timestep, b.station)
        for f in ["y_flame", "u", "v", "w", "G", "nu_t", "S_mag"]:
            assert abs(getattr(a, f) - getattr(b, f)) <= TOL * max(1.0, abs(getattr(b, f))), f


def test_probe_rows_of_a_multi_step_call(pair):
    """One les_gpu_step call of five steps: each step's probe rows are copied
    to pinned memory and turned into rows at the next step's wait, the last
    step's at the end of the call; the rows of all its steps equal the
    oracle's, in order."""
    g, o = pair
    print("case: probe rows of one 5-step call", flush=True)
    g.init_state()
    o.init_state()
    st = synthetic_reference_state(o, seed=11)      # a flame band at every station from the start
    g.upload_state(st)
    o.set_state(st)
    g.probe_rows()
    o.probe_rows()
    g.step(5)
    for _ in range(5):
        o.step(1)
    pg, po = g.probe_rows(), o.probe_rows()
    assert len(pg) == len(po) and len(po) > 0
    assert sorted({a.timestep for a in pg}) == sorted({b.timestep for b in po})
    assert len({a.timestep for a in pg}) >= 2            # rows of several steps of the one call
    for a, b in zip(pg, po):
        assert (a.timestep, a.station) == (b.timestep, b.station)
        for f in ["y_flame", "u", "v", "w", "G", "nu_t", "S_mag"]:
            assert abs(getattr(a, f) - getattr(b, f)) <= TOL * max(1.0, abs(getattr(b, f))), f


def test_step_report_monitor(pair):
    g, o = pair
    print("case: monitor reductions", flush=True)
    g.init_state()
    o.init_state()
    rg = g.step(2)
    ro = o.step(2)
    assert rg.nonfinite == 0
    assert abs(rg.total_mass - ro.total_mass) <= 1e-12 * ro.total_mass
    for f in ["min_T", "max_T", "min_G", "max_G", "min_Y_fuel"]:
        assert getattr(rg, f) == getattr(ro, f), f


@pytest.mark.parametrize("grid", [(60, 20, 40), (24, 96, 128)], ids=["reference_grid", "plane96x128"])
def test_inflow_plane_bitwise(lesgpu, oracle, grid):
    """The SEM inflow plane (generate_inlet_turbulence_SEM :943-1075) after
    three steps equals the oracle's bit for bit: the device sums each cell's
    in-radius eddies in eddy order and the normalisation sums the inlet cells
    in plane order, as the reference's serial loops do."""
    nx, ny, nz = grid
    print("case: reference-scheme inflow plane %dx%dx%d vs oracle, bitwise" % grid, flush=True)
    g = lesgpu.LesGpu(lesgpu.params(0, nx=nx, ny=ny, nz=nz))
    o = oracle.Oracle(oracle.params(0, nx=nx, ny=ny, nz=nz))
    try:
        g.init_state()
        o.init_state()
        g.step(3)
        o.step(3)
        P = ny * nz
        pg = g.field("inlet", 3 * P).reshape(3, ny, nz)
        po = o.inlet()
        for c in range(3):
            assert np.abs(po[c]).max() > 0, c
            assert np.array_equal(pg[c], po[c]), (c, float(np.abs(pg[c] - po[c]).max()))
    finally:
        g.close()
        o.close()
