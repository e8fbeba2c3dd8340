"""Served and ensemble LES runs on one GPU over the les_batch_* C ABI.

The reference's optimiser drives one simulation at a time: its interactive
main (bayesian_optimizer.py:719-844) scores one finished run and prints the
next six constants as #defines (:703-716, bounds :574-581), which a human
pastes into the solver, recompiles and runs.  Here a request is a
parameter set plus a step count; requests are admitted into a pool of
device-resident runs as slots and arena cells free up (admission by HBM
footprint: a run needs its cells and tiles in the pool's arenas), every pool
step advances all resident runs with one launch sequence, finished runs are
read back (FP64 plane profiles, state) and their slots released while the
others keep running (continuous batching).

Workload samplers follow the ensemble and served-stream definitions: SGS
constant ~U[0.08, 0.20], momentum-forcing amplitude ~U[0, 2], inflow
intensity ~U[0.05, 0.15] per run, drawn from one recorded seed.
"""
import time

import numpy as np

from . import abi
from . import lesgpu

SERVED_GRIDS = [(48, 48, 48), (64, 48, 48), (96, 64, 64), (128, 64, 64)]


def member_params(rng, nx, ny, nz, precision=abi.LES_PREC_FP32, seed=0):
    """One run's parameters: the duct physics at dx = 1 mm plus the per-run
    draws (SGS constant, forcing amplitude, inflow intensity)."""
    p = lesgpu.params(abi.LES_PRESET_DUCT64, nx=nx, ny=ny, nz=nz, lx=nx * 1e-3, ly=ny * 1e-3, lz=nz * 1e-3,
                      precision=precision)
    cs = float(rng.uniform(0.08, 0.20))
    p.cs_unburned = cs
    p.cs_burnt = cs
    p.flame_accel_strength = float(rng.uniform(0.0, 2.0))
    p.turbulence_intensity = float(rng.uniform(0.05, 0.15))
    p.seed = seed
    return p


def ensemble_members(n=64, grid=(64, 48, 48), seed=0, precision=abi.LES_PREC_FP32):
    rng = np.random.default_rng(seed)
    return [member_params(rng, *grid, precision=precision, seed=m) for m in range(n)]


def served_requests(n=96, mean_gap_s=2.0, seed=0, precision=abi.LES_PREC_FP32):
    """(arrival time s, params, n_steps) of the served stream, from one seed."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(mean_gap_s, n))
    out = []
    for r in range(n):
        nx, ny, nz = SERVED_GRIDS[int(rng.integers(0, len(SERVED_GRIDS)))]
        p = member_params(rng, nx, ny, nz, precision=precision, seed=1000 + r)
        steps = int(rng.integers(500, 2001))
        out.append((float(t[r]), p, steps))
    return out


def reynolds_profiles(prof, nx, ny, nz, steps):
    """Plane profiles (10 FP64 sums per x-plane, each over `steps` steps x the
    plane's ny*nz cells) -> means and resolved Reynolds stresses per x-plane,
    the calculate_resolved_reynolds_stresses (les_combustion_solver.c:3195-3219)
    formulas <u_i u_j> - <u_i><u_j> applied to plane-time averages:
    columns u, v, w, c, uu, vv, ww, uv, uw, vw."""
    s = np.asarray(prof, dtype=np.float64).reshape(nx, 10) / (float(steps) * ny * nz)
    m = s[:, :4]
    r = np.empty((nx, 6))
    r[:, 0] = s[:, 4] - m[:, 0] * m[:, 0]
    r[:, 1] = s[:, 5] - m[:, 1] * m[:, 1]
    r[:, 2] = s[:, 6] - m[:, 2] * m[:, 2]
    r[:, 3] = s[:, 7] - m[:, 0] * m[:, 1]
    r[:, 4] = s[:, 8] - m[:, 0] * m[:, 2]
    r[:, 5] = s[:, 9] - m[:, 1] * m[:, 2]
    return np.concatenate([m, r], axis=1)


PROFILE_COLUMNS = ["u", "v", "w", "c", "uu", "vv", "ww", "uv", "uw", "vw"]


def profile_errors(a, b, U, intensity):
    """Errors of profiles a against reference profiles b (rows: x-planes,
    columns PROFILE_COLUMNS).  "rel": max |a - b| / max |b| per column.
    "scaled": the same against the column's physical scale where the column is
    near zero: U for the means, 1 for c, max(peak, 1 % of (U I)^2) for the
    stresses (where the inflow turbulence has not arrived, <v>, <u'v'> ... are
    far below the rounding of the means they are computed from)."""
    out = {"rel": {}, "scaled": {}}
    floor = 1e-2 * (U * intensity) ** 2
    for k, nm in enumerate(PROFILE_COLUMNS):
        d = float(np.abs(a[:, k] - b[:, k]).max())
        peak = float(np.abs(b[:, k]).max())
        out["rel"][nm] = d / max(peak, 1e-300)
        scale = U if k < 3 else (1.0 if k == 3 else max(peak, floor))
        out["scaled"][nm] = d / scale
    out["max_scaled"] = max(out["scaled"].values())
    return out


def pool_config(max_slots, pool_cells, max_grid, precision=abi.LES_PREC_FP32, autotune=True, point_cap=64):
    return lesgpu.batch_config(max_slots=max_slots, pool_cells=pool_cells, max_nx=max_grid[0],
                               max_ny=max_grid[1], max_nz=max_grid[2], max_vortices=400, precision=precision,
                               autotune=1 if autotune else 0, point_cap=point_cap)


def pool_for_budget(budget_bytes, max_slots, max_grid, precision=abi.LES_PREC_FP32, autotune=True, point_cap=64):
    """Largest pool (in arena cells) whose device allocation fits budget_bytes:
    the HBM footprint admission works against."""
    lo, hi = 1, 1 << 34
    while lo < hi:
        mid = (lo + hi + 1) // 2
        cfg = pool_config(max_slots, mid, max_grid, precision, autotune, point_cap)
        if lesgpu.pool_bytes(cfg) <= budget_bytes:
            lo = mid
        else:
            hi = mid - 1
    return pool_config(max_slots, lo, max_grid, precision, autotune, point_cap)


class Server:
    """Continuous batching of duct-mode runs in one pool.

    submit() queues a request; pump() admits queued requests in arrival order
    while they fit, advances the pool by `chunk` steps and collects finished
    runs.  A finished run's result holds its FP64 plane profiles, its final
    monitor report and its timing."""

    def __init__(self, batch, chunk=10):
        self.b = batch
        self.chunk = chunk
        self.queue = []             # (request id, params, n_steps, arrival wall time)
        self.resident = {}          # slot -> (request id, params, n_steps, arrival, admit time)
        self.results = {}
        self.steps_run = 0
        self.gpu_ms = 0.0

    def submit(self, rid, p, n_steps, arrival=None):
        self.queue.append((rid, p, n_steps, time.perf_counter() if arrival is None else arrival))

    def _admit(self):
        while self.queue:
            rid, p, n, arr = self.queue[0]
            s = self.b.admit(p, n_steps=n, init=True)
            if s is None:
                if not self.resident:
                    raise RuntimeError("request %r does not fit an empty pool" % (rid,))
                return                      # head of line waits for memory (FIFO)
            self.queue.pop(0)
            self.resident[s] = (rid, p, n, arr, time.perf_counter())

    def _collect(self):
        for s in self.b.poll():
            rid, p, n, arr, adm = self.resident.pop(s)
            prof = self.b.field(s, "profiles", 10 * p.nx)
            rep = self.b.report(s)
            self.b.release(s)
            done = time.perf_counter()
            self.results[rid] = {"profiles": prof, "steps": n, "grid": (p.nx, p.ny, p.nz),
                                 "arrival": arr, "admitted": adm, "done": done, "latency_s": done - arr,
                                 "nonfinite": bool(rep.nonfinite), "max_T": rep.max_T}

    def pump(self):
        self._admit()
        if self.resident:
            self.gpu_ms += self.b.step(self.chunk) * self.chunk
            self.steps_run += self.chunk
        self._collect()
        return bool(self.resident or self.queue)


def run_stream(server, requests, time_scale=1.0, progress=None, progress_every_s=15.0):
    """Feed (arrival, params, steps) requests at their arrival times (scaled by
    time_scale) and pump until every request finished; returns wall seconds."""
    t0 = time.perf_counter()
    nxt = 0
    last = t0
    while nxt < len(requests) or server.resident or server.queue:
        now = time.perf_counter() - t0
        while nxt < len(requests) and requests[nxt][0] * time_scale <= now:
            arr, p, n = requests[nxt]
            server.submit(nxt, p, n, arrival=t0 + arr * time_scale)
            nxt += 1
        if server.resident or server.queue:
            server.pump()
        elif nxt < len(requests):
            time.sleep(min(0.005, max(0.0, requests[nxt][0] * time_scale - now)))
        if progress and time.perf_counter() - last > progress_every_s:
            last = time.perf_counter()
            progress("served %d/%d, resident %d, queued %d, %.0f s"
                     % (len(server.results), len(requests), len(server.resident), len(server.queue),
                        last - t0))
    return time.perf_counter() - t0
