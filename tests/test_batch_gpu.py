"""Batched duct runs (les_batch_*): a pool of runs advanced by one launch
sequence per step.

* a run in a pool gives bitwise the same state as the same run alone in a
  pool of the same configuration (tile width, launch mapping), whatever else
  is resident — other grids, parameters, step counts, runs admitted and
  retired mid-way (continuous batching).  The tile width groups the
  per-tile partial sums of the domain-mean density, so runs under different
  tile widths agree to rounding, not bitwise;
* every tile shape and launch configuration the autotuner may pick stays
  within the FP64 tolerance (1e-10 of each field's scale) of the oracle;
* admission refuses runs that do not fit and reuses freed arena ranges.
"""
import numpy as np
import pytest

from common import rel_err

pytestmark = pytest.mark.gpu
DUCT = 1        # LES_PRESET_DUCT64
TOL = 1e-10


def run_params(lesgpu, precision=0, **over):
    base = dict(precision=precision, reinit_interval=4, reinit_iters=2)
    base.update(over)
    return lesgpu.params(DUCT, **base)


def single_run(lesgpu, p, steps):
    """The run alone in a pool of the test pools' configuration."""
    with pool(lesgpu, p.precision, slots=1, cells=p.nx * p.ny * p.nz) as b:
        s = b.admit(p, n_steps=steps)
        b.step(steps)
        return b.download(s), b.field(s, "profiles", 10 * p.nx)


def pool(lesgpu, precision=0, slots=4, cells=None, **over):
    kw = dict(max_slots=slots, pool_cells=cells or 4 * 64 * 48 * 48, max_nx=96, max_ny=64, max_nz=64,
              max_vortices=400, precision=precision, tile_k=64, point_cap=64)
    kw.update(over)
    cfg = lesgpu.batch_config(**kw)
    return lesgpu.Batch(cfg)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_member_in_mix_equals_alone(lesgpu, precision):
    print("case: batch member vs single run, precision=%d" % precision, flush=True)
    pa = run_params(lesgpu, precision, seed=7)
    alone, prof_alone = single_run(lesgpu, pa, 9)
    with pool(lesgpu, precision) as b:
        others = [
            run_params(lesgpu, precision, nx=48, ny=40, nz=24, seed=3, inlet_velocity=4.0),
            run_params(lesgpu, precision, nx=96, ny=48, nz=48, seed=11, advection=2, reinit_interval=3,
                       reinit_iters=3),   # odd sweep count beside the even ones
        ]
        s1 = b.admit(others[0], n_steps=4)          # finishes first, then sits done
        sa = b.admit(pa, n_steps=9)
        b.step(3)
        s2 = b.admit(others[1], n_steps=20)         # admitted mid-way
        b.step(2)
        assert s1 in b.poll()
        b.release(s1)
        s3 = b.admit(run_params(lesgpu, precision, nx=40, ny=32, nz=32, seed=5), n_steps=5)
        b.step(4)
        assert b.info(sa).step == 9 and b.info(sa).done == 1
        st = b.download(sa)
        for nm, x, y in zip(["rhou", "rhov", "rhow", "phi", "T"], st, alone):
            assert np.array_equal(x, y), nm
        assert np.array_equal(b.field(sa, "profiles", 10 * pa.nx), prof_alone)
        assert b.info(s2).step == 6 and b.info(s3).step == 4
        for s in (s2, s3):
            rep = b.report(s)
            assert not rep.nonfinite


def test_runs_admitted_later_match_alone(lesgpu):
    print("case: run admitted into a reused slot vs alone", flush=True)
    pc = run_params(lesgpu, 0, nx=40, ny=32, nz=32, seed=21)
    alone, _ = single_run(lesgpu, pc, 6)
    with pool(lesgpu, 0, slots=2, cells=64 * 32 * 32 + 40 * 32 * 32) as b:
        s0 = b.admit(run_params(lesgpu, 0, seed=1), n_steps=2)
        b.step(2)
        b.release(s0)
        sc = b.admit(pc, n_steps=6)
        b.step(8)                                   # two steps past its end: it must not move
        assert b.poll() == [sc]
        for x, y in zip(b.download(sc), alone):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("tile_k,xcd", [(0, 1), (32, 1), (128, 0), (64, 0), (16, 2), (64, 4), (32, 12)],
                         ids=["flat_xcd", "tk32_xcd", "tk128", "tk64", "tk16_chunks", "tk64_slabs",
                              "tk32_slabs_alternating"])
def test_launch_shapes_vs_oracle(lesgpu, oracle, tile_k, xcd):
    print("case: batch tile_k=%d xcd=%d vs oracle" % (tile_k, xcd), flush=True)
    over = dict(reinit_interval=3, reinit_iters=2)
    p = lesgpu.params(DUCT, **over)
    o = oracle.Oracle(oracle.params(DUCT, **over))
    with pool(lesgpu, 0, slots=2, tile_k=tile_k, xcd_map=xcd) as b:
        s = b.admit(p, n_steps=0)
        # a second, differently shaped run shares every launch
        b.admit(lesgpu.params(DUCT, nx=48, ny=24, nz=40, seed=9), n_steps=0)
        o.init_state()
        b.step(6)
        o.step(6)
        for nm, x, y in zip(["rhou", "rhov", "rhow", "phi", "T"], b.download(s), o.get_state(5)):
            assert rel_err(x, y) <= TOL, (nm, rel_err(x, y))
    o.close()


def test_projection_runs_in_mix_equal_alone(lesgpu):
    print("case: runs with different projection sweep counts in one pool vs alone", flush=True)
    pa = run_params(lesgpu, 0, proj_iters=3, seed=2)
    pb = run_params(lesgpu, 0, nx=48, ny=40, nz=24, proj_iters=6, seed=4)
    pc = run_params(lesgpu, 0, nx=40, ny=32, nz=32, seed=6)           # no projection
    alone = [single_run(lesgpu, p, 5)[0] for p in (pa, pb, pc)]
    with pool(lesgpu, 0, slots=3) as b:
        slots = [b.admit(p, n_steps=5) for p in (pa, pb, pc)]
        b.step(5)
        for s, want in zip(slots, alone):
            for x, y in zip(b.download(s), want):
                assert np.array_equal(x, y)
        assert b.report(slots[0]).proj_residual > 0 and b.report(slots[2]).proj_residual == 0


def test_admission_limits(lesgpu):
    print("case: admission refusal and reuse", flush=True)
    with pool(lesgpu, 1, slots=3, cells=2 * 64 * 32 * 32) as b:
        p = run_params(lesgpu, 1)
        s0 = b.admit(p, 5)
        s1 = b.admit(p, 5)
        assert s0 is not None and s1 is not None
        assert b.admit(p, 5) is None                             # arena full
        assert b.admit(run_params(lesgpu, 1, nx=128), 5) is None  # beyond max_nx
        assert b.admit(run_params(lesgpu, 0), 5) is None          # wrong precision
        b.release(s0)
        s2 = b.admit(run_params(lesgpu, 1, nx=32), 5)            # fits the freed range
        assert s2 is not None
        b.step(1)
        assert b.info(s2).step == 1


def test_batch_point_series_and_autotune(lesgpu):
    print("case: batch point series + autotuner", flush=True)
    probe = run_params(lesgpu, 1, nx=64, ny=48, nz=48)
    cfg = lesgpu.batch_config(max_slots=2, pool_cells=2 * 64 * 48 * 48, max_nx=64, max_ny=48, max_nz=48,
                              max_vortices=400, precision=1, autotune=1, point_cap=8)
    with lesgpu.Batch(cfg, probe) as b:
        t = b.tuning()
        print("tuned:", t, flush=True)
        assert t["tile_k"] in (0, 16, 32, 64, 128) and t["grid_blocks"] > 0 and t["tuned_ms"] > 0
        assert (t["xcd_map"] & 7) in range(5) and t["xcd_map"] < 16    # mapping 0..4, +8 alternated sweeps
        s = b.admit(probe, 12)
        b.step(12)
        ser = b.point_series(s, 64)
        assert ser.shape == (8, 64, 4)                           # ring keeps the newest 8 steps
        st = b.download(s)
        rho = 1.2 * 300.0 / st[4]
        u = (st[0] / rho).reshape(-1)
        nx, ny, nz = 64, 48, 48
        idx = [((a + 1) * nx // 5 * ny + (2 * c + 1) * ny // 8) * nz + (2 * e + 1) * nz // 8
               for a in range(4) for c in range(4) for e in range(4)]
        assert np.allclose(ser[-1, :, 0], u[idx], rtol=1e-6, atol=0)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_xface_copy_fills_ghost_planes(lesgpu, precision):
    """les_batch_xface_copy moves x-planes of the state between two runs (two
    x-slabs' halo exchange on one GPU): the copied planes equal the source's
    bitwise, every other plane is untouched, and bad ranges are refused."""
    print("case: x-face copy between two runs, precision=%d" % precision, flush=True)
    with pool(lesgpu, precision) as b:
        sa = b.admit(run_params(lesgpu, precision, nx=40, ny=24, nz=32, seed=5), n_steps=0)
        sb = b.admit(run_params(lesgpu, precision, nx=48, ny=24, nz=32, seed=6), n_steps=0)
        b.step(3)
        qa, qb = b.download(sa), b.download(sb)
        ms = b.xface_copy(sa, 4, sb, 44, 4, reps=3)          # a's first interior planes -> b's high ghosts
        b.xface_copy(sa, 32, sb, 0, 4)                        # a's last interior planes -> b's low ghosts
        qb2, qa2 = b.download(sb), b.download(sa)
        assert ms > 0
        for f in range(5):
            assert np.array_equal(qa2[f], qa[f])
            assert np.array_equal(qb2[f][44:48], qa[f][4:8])
            assert np.array_equal(qb2[f][0:4], qa[f][32:36])
            assert np.array_equal(qb2[f][4:44], qb[f][4:44])
        b.step(1)                                             # the runs go on from the new state
        assert not b.report(sb).nonfinite
        for bad in [(sa, 38, sb, 0, 4), (sa, 0, sb, 46, 4), (sa, 0, sa, 2, 4), (sa, -1, sb, 0, 2), (sa, 0, sb, 0, 0)]:
            with pytest.raises(lesgpu.LesError):
                b.xface_copy(*bad)
        sc = b.admit(run_params(lesgpu, precision, nx=40, ny=24, nz=24, seed=7), n_steps=0)
        with pytest.raises(lesgpu.LesError):
            b.xface_copy(sa, 4, sc, 0, 4)                     # ny, nz differ


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_xcd_mapping_does_not_change_results(lesgpu, precision):
    """The block-to-tile mapping only changes which XCD runs a tile: the state
    and profiles are bitwise the same under all five mappings (round-robin,
    contiguous ranges, chunks of 16 and of 64 tiles, slabs) and with the
    stencil launches' sweep direction alternated (+8) (single runs tune the mapping
    per run and pools the sweep, so results must not depend on either)."""
    print("case: XCD mapping neutrality, precision=%d" % precision, flush=True)
    p = run_params(lesgpu, precision, nx=64, ny=32, nz=64, seed=9)
    out = []
    for xm in (0, 1, 2, 3, 4, 8, 10, 12):
        with pool(lesgpu, precision, slots=1, cells=p.nx * p.ny * p.nz, xcd_map=xm) as b:
            s = b.admit(p, n_steps=6)
            b.step(6)
            out.append((b.download(s), b.field(s, "profiles", 10 * p.nx)))
    for o in out[1:]:
        for f in range(5):
            assert np.array_equal(out[0][0][f], o[0][f]), f
        assert np.array_equal(out[0][1], o[1])


@pytest.mark.parametrize("precision,ngroups", [(0, 4), (1, 2)], ids=["fp64", "fp32"])
def test_run_groups_do_not_change_results(lesgpu, precision, ngroups, monkeypatch):
    """A pool large enough is stepped in groups of runs (contiguous tile
    ranges split at run boundaries, each group's three RK stages in turn):
    20 runs of 96x64x64 (7.9M cells; 4 groups in FP64, 2 in FP32) give
    bitwise the states and profiles of the same pool stepped as one group
    (LES_GROUPS=1), over plain steps and a reinitialisation step."""
    print("case: run groups vs one group, precision=%d" % precision, flush=True)
    runs = [run_params(lesgpu, precision, nx=96, ny=64, nz=64, seed=20 + i, advection=1 + (i % 3 == 2),
                       inlet_velocity=5.0 + 0.25 * (i % 4)) for i in range(20)]
    cells = 20 * 96 * 64 * 64

    def stepped(one_group):
        if one_group:
            monkeypatch.setenv("LES_GROUPS", "1")
        else:
            monkeypatch.delenv("LES_GROUPS", raising=False)
        with pool(lesgpu, precision, slots=20, cells=cells) as b:
            slots = [b.admit(p, n_steps=12) for p in runs]
            assert None not in slots
            b.step(5)
            grp = b.groups()
            return grp, [(b.download(s), b.field(s, "profiles", 10 * 96)) for s in slots]

    grp, split = stepped(False)
    one, whole = stepped(True)
    assert len(one) == 1 and len(grp) == ngroups, (one, grp)
    assert grp[0][0] == 0 and sum(c for _, c in grp) == one[0][1]
    for (b0, c0), (b1, _) in zip(grp, grp[1:]):
        assert b0 + c0 == b1 and c0 > 0
    for (fa, pa), (fb, pb) in zip(split, whole):
        for f in range(5):
            assert np.array_equal(fa[f], fb[f]), f
        assert np.array_equal(pa, pb)


def test_groups_with_continuous_batching(lesgpu):
    """Runs released and admitted while the pool steps in groups of runs (the
    resident tile ranges change, the captured steps are rebuilt): a run
    admitted mid-way into freed ranges, and a run resident throughout, equal
    the same runs alone, bitwise."""
    print("case: groups with releases and admissions, FP64", flush=True)
    runs = [run_params(lesgpu, 0, nx=96, ny=64, nz=64, seed=40 + i) for i in range(20)]
    late = run_params(lesgpu, 0, nx=64, ny=48, nz=48, seed=77, advection=2)
    kept_alone, _ = single_run(lesgpu, runs[7], 7)
    late_alone, _ = single_run(lesgpu, late, 4)
    with pool(lesgpu, 0, slots=20, cells=20 * 96 * 64 * 64) as b:
        slots = [b.admit(p, n_steps=7) for p in runs]
        assert None not in slots
        b.step(3)
        before = b.groups()
        for s in (slots[2], slots[3], slots[12]):
            b.release(s)
        sl = b.admit(late, n_steps=4)
        assert sl is not None
        b.step(4)
        after = b.groups()
        assert len(before) == 4 and len(after) >= 2, (before, after)
        for x, y in zip(b.download(slots[7]), kept_alone):
            assert np.array_equal(x, y)
        for x, y in zip(b.download(sl), late_alone):
            assert np.array_equal(x, y)


def wrinkled_front_state(p, amp=3e-3):
    """Duct state with the planar front wrinkled along y (the probe lines at
    k = nz/2 cross G = 0.5 at every station)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    dx = p.lx / nx
    x = np.arange(nx)[:, None, None] * dx
    y = np.arange(ny)[None, :, None] * (p.ly / ny)
    phi = (x - p.flame_x0 - amp * np.sin(2 * np.pi * y / p.ly)) * np.ones((nx, ny, nz))
    c = 0.5 * (1 + np.tanh(phi / (2 * dx)))
    T = p.inlet_temperature + c * (p.burnt_temperature - p.inlet_temperature)
    rho = p.rho_unburnt * p.inlet_temperature / T
    return [rho * p.inlet_velocity, np.zeros_like(T), np.zeros_like(T), phi, T]


def test_probe_rows_survive_one_long_step_call(lesgpu):
    """A pool whose probe rings hold 4 steps, advanced 10 steps in one
    les_batch_step call, returns the same rows as one advanced a step at a
    time and fetched after each step: every step's rows, none lost."""
    print("case: batch probe rings (probe_cap 4) over 10 steps in one call", flush=True)
    p = run_params(lesgpu, 0, reinit_interval=0)
    st = wrinkled_front_state(p)
    got = {}
    for mode in ("one_call", "per_step"):
        with pool(lesgpu, 0, slots=2, probe_cap=4) as b:
            b.admit(run_params(lesgpu, 0, nx=48, ny=40, nz=24, seed=3), n_steps=0)   # a second resident run
            s = b.admit(p, n_steps=0)
            b.upload(s, st)
            rows = []
            if mode == "one_call":
                b.step(10)
                rows = b.probe_rows(s)
            else:
                for _ in range(10):
                    b.step(1)
                    rows += b.probe_rows(s)
            got[mode] = [(r.timestep, r.station, r.y_flame, r.y_physical, r.u, r.v, r.w, r.G, r.nu_t, r.S_mag)
                         for r in rows]
    steps = sorted({r[0] for r in got["per_step"]})
    assert len(steps) == 10, steps
    assert got["one_call"] == got["per_step"]
