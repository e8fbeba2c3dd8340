"""Write tests/golden/reference_golden.npz from the reference solver.

Runs the reference (les_combustion_solver.c compiled from its own source by
``make -C oracle ref`` into oracle/_ref/les_ref_golden, serial build) under the
harness oracle/ref_harness.c and stores small vectors:

* a fresh run (srand seed 0, synthetic chem.bin/tran.bin written by the
  engine's les_mixture_save) for 3 steps: the 11 state fields plus nu_t and
  S_mag at 1000 sampled cells after each step, full-field sums and sums of
  squares, the first 64 SEM eddies and the three inlet planes after step 0;
* one calculate_RHS call (t_loop = 600, so the WENO5 |grad G| branch is live)
  on the synthetic state of tests/common.synthetic_reference_state(seed=3):
  the 10 RHS fields and 8 intermediates at the sampled cells plus sums;
* the probe CSV rows record_flame_region_data writes for that state.

With the argument ``probe-csv`` it writes tests/golden/reference_probe_csv.npz
instead: the text of the seven probe_data_flame_region_loc{n}.csv files and of
simulation_monitor.csv that the reference writes over a fresh 10-step run of
seed 0 (rows at steps 3-9: the initial state is burnt everywhere, :69, so the
G = 0.5 crossing first appears after the inflow has entered).

Usage:  python tests/golden/make_golden.py [probe-csv]    (needs oracle/_ref built)
"""
import glob
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pyoracle as po  # noqa: E402
from common import FIELDS, synthetic_reference_state  # noqa: E402

OUT = os.path.join(HERE, "reference_golden.npz")
SHAPE = (60, 20, 40)
N = 60 * 20 * 40
RHS_NAMES = ["rhs_rhou", "rhs_rhov", "rhs_rhow", "rhs_T", "rhs_G",
             "rhs_Y0", "rhs_Y1", "rhs_Y2", "rhs_Y3", "rhs_Y4"]
MID_NAMES = ["rhs_nu_t", "rhs_S_mag", "rhs_ST", "rhs_mu_mix", "rhs_alpha", "rhs_gradG", "rhs_D0", "rhs_cp_mix"]


def sample_indices():
    rng = np.random.default_rng(1234)
    base = rng.choice(N, 800, replace=False)
    # plus every attribute class: the inlet plane, walls, symmetry plane, outflow
    i, j, k = np.unravel_index(np.arange(N), SHAPE)
    extra = np.concatenate([
        np.flatnonzero(i == 0)[::8], np.flatnonzero(i == 59)[::16],
        np.flatnonzero((j == 0) & (i > 0) & (i < 59))[::24], np.flatnonzero((j == 19) & (i > 0))[::24],
        np.flatnonzero((k == 0) & (i > 0))[::24], np.flatnonzero((i == 1) | (i == 2))[::16]])
    return np.unique(np.concatenate([base, extra]))[:1000]


def main():
    if not os.path.exists(po.REF_GOLDEN):
        sys.exit("reference build missing: make -C oracle ref")
    idx = sample_indices()
    out = {"idx": idx, "seed": np.array([0]), "steps": np.array([3])}
    with tempfile.TemporaryDirectory() as wd:
        run = os.path.join(wd, "run")
        inp = os.path.join(wd, "rhs_in")
        os.makedirs(inp)
        o = po.Oracle()
        o.init_state()
        st = synthetic_reference_state(o, seed=3)
        for nm, a in zip(FIELDS, st):
            np.ascontiguousarray(a, dtype=np.float64).tofile(os.path.join(inp, nm + ".bin"))
        out["rhs_input_sums"] = np.array([a.sum() for a in st])
        po.run_reference(run, steps=3, seed=0, rhs_in=inp, rhs_t=600)
        for s in range(3):
            for nm in FIELDS + ["nu_t", "S_mag", "nu", "alpha"]:
                a = po.load_dump(run, nm, s).reshape(-1)
                out["s%d_%s" % (s, nm)] = a[idx]
                out["s%d_%s_sum" % (s, nm)] = np.array([a.sum(), (a * a).sum()])
        out["s0_vortices"] = po.load_dump(run, "vortices", 0, shape=(2000, 7))[:64]
        out["s0_vortices_sum"] = np.array([po.load_dump(run, "vortices", 0, shape=None).sum()])
        for nm in ["u_inlet", "v_inlet", "w_inlet"]:
            out["s0_" + nm] = po.load_dump(run, nm, 0, shape=(20, 40))
        for nm in RHS_NAMES + MID_NAMES:
            a = po.load_dump(run, nm).reshape(-1)
            out[nm] = a[idx]
            out[nm + "_sum"] = np.array([a.sum(), (a * a).sum()])
        rows = []
        for f in sorted(glob.glob(os.path.join(run, "*probe_data_flame_region_loc*.csv"))):
            n = int(f.rsplit("loc", 1)[1].split(".")[0])
            for line in open(f).read().splitlines()[1:]:
                if line.startswith("999999,"):
                    rows.append("%d|%s" % (n, line))
        out["probe_rows"] = np.array(rows)
    np.savez_compressed(OUT, **out)
    print("wrote %s (%d bytes)" % (OUT, os.path.getsize(OUT)))


def probe_csv_fixture(steps=10):
    out_path = os.path.join(HERE, "reference_probe_csv.npz")
    with tempfile.TemporaryDirectory() as wd:
        po.run_reference(wd, steps=steps, seed=0, dump=[])
        out = {"steps": np.array([steps]), "seed": np.array([0])}
        for f in sorted(glob.glob(os.path.join(wd, "*probe_data_flame_region_loc*.csv"))):
            n = int(f.rsplit("loc", 1)[1].split(".")[0])
            out["loc%d" % n] = np.array(open(f).read())
        mon = glob.glob(os.path.join(wd, "*simulation_monitor.csv"))
        out["monitor"] = np.array(open(mon[0]).read())
    np.savez_compressed(out_path, **out)
    print("wrote %s (%d bytes)" % (out_path, os.path.getsize(out_path)))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "probe-csv":
        probe_csv_fixture()
    else:
        main()
