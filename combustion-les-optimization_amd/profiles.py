"""Reynolds-stress profiles the optimiser scores, from probe rows in FP64.

bayesian_optimizer.py:71-117 (load_and_compute_reynolds_stress) reads
probe_data_flame_region_loc{n}.csv, bins the rows by G into num_bins equal
bins on [0, 1) (a row belongs to bin b when edge_b <= G < edge_{b+1}), and
for every bin with at least 10 rows takes <u'v'> = <uv> - <u><v>; bins with
fewer rows are NaN.  The same statistic is computed here from les_probe_row
records (full FP64, not the CSV's 6 decimals), from a CSV written by the
solver program, or from raw arrays.
"""
import numpy as np

MIN_SAMPLES = 10     # bayesian_optimizer.py:106 (bins with fewer rows are skipped)


def binned_uv(G, u, v, num_bins=20):
    """G-binned <u'v'>: (bin centres, values with NaN for sparse bins, counts)."""
    G = np.asarray(G, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    edges = np.linspace(0.0, 1.0, num_bins + 1)
    centres = 0.5 * (edges[:-1] + edges[1:])
    vals = np.full(num_bins, np.nan)
    counts = np.zeros(num_bins, dtype=np.int64)
    for b in range(num_bins):
        m = (G >= edges[b]) & (G < edges[b + 1])
        counts[b] = int(m.sum())
        if counts[b] < MIN_SAMPLES:
            continue
        ub, vb = u[m], v[m]
        vals[b] = (ub * vb).mean() - ub.mean() * vb.mean()
    return centres, vals, counts


def rows_to_arrays(rows, station=None):
    """les_probe_row records (optionally of one station) -> dict of arrays."""
    sel = [r for r in rows if station is None or r.station == station]
    names = ["timestep", "station", "x_location", "y_flame", "y_physical", "y_relative",
             "u", "v", "w", "G", "nu_t", "S_mag"]
    return {n: np.array([getattr(r, n) for r in sel]) for n in names}


def binned_uv_from_rows(rows, station=0, num_bins=20):
    a = rows_to_arrays(rows, station)
    return binned_uv(a["G"], a["u"], a["v"], num_bins)


def binned_uv_from_csv(path, num_bins=20):
    """The same statistic from a probe CSV (the optimiser's input file)."""
    data = np.genfromtxt(path, delimiter=",", names=True)
    data = np.atleast_1d(data)
    return binned_uv(data["G"], data["u"], data["v"], num_bins)
