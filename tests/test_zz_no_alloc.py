"""les_gpu_rhs_noalloc against les_gpu_rhs through the C ABI.

The variant stages the duct RHS in the library's static device scratch instead
of per-call device buffers.  Both run the same kernels, so the bar is bitwise
equality: on two calls in a row, the second on a smaller grid (the scratch
still holds the first call's values, so a missing clear of the planes the RHS
does not write would show), in FP64 and FP32.  A grid whose staging exceeds the
scratch is refused with LES_ERR_PARAM and a message; the reference scheme takes
the les_gpu_rhs path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LES_ERR_PARAM = -2
DUCT = ["rhou", "rhov", "rhow", "phi", "T"]


def perturbed_context(lesgpu, nx, ny, nz, precision, seed):
    p = lesgpu.params(1, nx=nx, ny=ny, nz=nz, lx=nx * 1e-3, ly=ny * 1e-3, lz=nz * 1e-3, precision=precision)
    g = lesgpu.LesGpu(p)
    g.init_state()
    st = g.download_state(5)
    rng = np.random.default_rng(seed)
    rho = 1.2 * 300.0 / st[4]
    for f in range(3):
        st[f] = st[f] + rho * rng.uniform(-1.5, 1.5, st[f].shape)   # both upwind signs
    st[3] = st[3] + 2e-4 * rng.standard_normal(st[3].shape)
    g.upload_state(st)
    return g


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32"])
def test_rhs_noalloc_equals_rhs_two_calls(lesgpu, precision):
    grids = [(64, 32, 32), (40, 24, 16)]          # second call on a smaller grid
    for call, (nx, ny, nz) in enumerate(grids):
        print("case: rhs_noalloc precision=%d call %d grid %dx%dx%d" % (precision, call, nx, ny, nz), flush=True)
        g = perturbed_context(lesgpu, nx, ny, nz, precision, seed=20 + call)
        try:
            a = g.rhs_noalloc(fill=np.nan)[:5]
            b = g.rhs()[:5]
            for nm, x, y in zip(DUCT, a, b):
                assert np.all(np.isfinite(x)), (call, nm)
                assert np.array_equal(x, y), (call, nm, float(np.abs(x - y).max()))
                # the planes the RHS does not write come back as zero
                assert not x[0].any() and not x[-1].any(), (call, nm)
            assert any(np.abs(x[1:-1]).max() > 0 for x in a)
        finally:
            g.close()


def test_rhs_noalloc_refuses_oversize_grid(lesgpu):
    print("case: rhs_noalloc 128^3 FP64 (80 MiB of staging) refused", flush=True)
    g = lesgpu.LesGpu(lesgpu.params(2))
    try:
        g.init_state()
        out = [np.zeros(g.n) for _ in range(10)]
        rc = lesgpu.lib().les_gpu_rhs_noalloc(g.h, 0, lesgpu._ptrs(out), 10)
        assert rc == LES_ERR_PARAM
        assert b"static scratch" in lesgpu.lib().les_gpu_last_error(g.h)
        # the context stays usable
        b = g.rhs()[:5]
        assert all(np.all(np.isfinite(x)) for x in b)
    finally:
        g.close()


def test_rhs_noalloc_reference_scheme(lesgpu):
    print("case: rhs_noalloc reference scheme 60x20x40", flush=True)
    g = lesgpu.LesGpu(lesgpu.params(0))
    try:
        g.init_state()
        g.step(1)
        a = g.rhs_noalloc(t_loop=600)
        b = g.rhs(t_loop=600)
        for f, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x, y), f
    finally:
        g.close()
