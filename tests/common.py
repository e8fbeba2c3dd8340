"""Shared helpers of the test-suite: synthetic states and comparison metrics."""
import numpy as np

FIELDS = ["u", "v", "w", "T", "G", "rho", "Y0", "Y1", "Y2", "Y3", "Y4"]


def rel_err(a, b, floor=1e-300):
    """max |a-b| / max |b| over a field (the field's own scale)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), floor)
    return float(np.abs(a - b).max() / scale)


def eos_rho(T, Y, mw):
    """Reference EOS (les_combustion_solver.c:1893-1905), vectorised."""
    s = np.zeros_like(T)
    for k in range(5):
        s = s + Y[k] / (mw[k] / 1000.0)
    wtm = 1.0 / s
    rmix = 8.314462 / wtm
    return 101325.0 / (rmix * T)


def synthetic_reference_state(o, seed=1, front_sharp=True):
    """A non-trivial state on the reference grid for per-operator parity.

    Starts from the oracle's initial state ``o`` (attribute map, composition)
    and overlays: a curved flame band in G with a sharp edge (TENO cut-off
    branches), smooth random velocity with both signs (upwind branches),
    T and Y consistent with G plus noise, and rho from the EOS.
    """
    rng = np.random.default_rng(seed)
    st = o.get_state()
    nx, ny, nz = st[0].shape
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x = i / nx
    y = j / ny
    z = k / nz
    attr = o.field("attribute")
    front = 0.35 + 0.15 * np.sin(2 * np.pi * y) * np.cos(2 * np.pi * z)
    G = 0.5 * (1 + np.tanh((x - front) / 0.04))
    if front_sharp:
        G = np.where(x > front + 0.25, 1.0, G)
    G = np.clip(G + 0.01 * rng.standard_normal(G.shape), 0.0, 1.0)
    u = 4.0 + 1.5 * np.sin(2 * np.pi * (x + 0.3 * y)) + 0.4 * rng.standard_normal(G.shape)
    v = 0.8 * np.sin(2 * np.pi * (z + x)) + 0.3 * rng.standard_normal(G.shape)
    w = 0.8 * np.cos(2 * np.pi * (y - x)) + 0.3 * rng.standard_normal(G.shape)
    T = 300.0 + 1900.0 * G + 20.0 * rng.standard_normal(G.shape)
    inletY = st[6:11]
    # composition: interpolate between the inlet state (G=0) and burnt state
    Yin = np.array([inletY[s][0, 1, 20] for s in range(5)])
    Yin = np.array([0.0551, 0.2041, 0.7408, 0.0, 0.0]) if Yin.sum() == 0 else Yin
    Yb = Yin.copy()
    mw = np.array([44.09652, 31.9988, 28.0134, 44.0095, 18.01528])
    cons = Yb[0]
    Yb = np.array([0.0, Yb[1] - cons * 5 * mw[1] / mw[0], Yb[2], cons * 3 * mw[3] / mw[0], cons * 4 * mw[4] / mw[0]])
    Y = [Yin[s] * (1 - G) + Yb[s] * G for s in range(5)]
    Y = [np.clip(y_ + 0.002 * rng.standard_normal(G.shape), 0, 1) for y_ in Y]
    tot = sum(Y)
    Y = [y_ / tot for y_ in Y]
    rho = eos_rho(T, Y, mw)
    # keep the boundary cells of the initial state (walls, inlet, outflow)
    out = []
    new = [u, v, w, T, G, rho] + Y
    for f0, f1 in zip(st, new):
        out.append(np.where(attr == 3, f1, f0).astype(np.float64))
    return out
