"""dim2d_time_dev_noalloc: the allocation-free time derivative (device arrays
in and out, working memory in the library's device scratch array) against
dim2d_time_dev of a context, which allocates its own buffers.  Collected
last: it runs alone and after every other GPU test."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R, P, W, O = 0, 1, 2, 3


def _hip():
    """The HIP runtime the library itself runs on (torch bundles a runtime of
    its own, which cannot open the GPU once this one has)."""
    import dim2d_finite_volume_amd as d
    d.capi.load()
    paths = [ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln]
    path = next((q for q in paths if "/torch/" not in q), paths[0])
    H = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    H.hipFree.argtypes = [vp]
    H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    H.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
    return H


def _noalloc(p, nodes):
    """dim2d_time_dev_noalloc on device copies of `nodes` (device arrays
    allocated by the test); returns the derivative and the status."""
    import dim2d_finite_volume_amd as d
    lib = d.capi.load()
    H = _hip()
    src = np.ascontiguousarray(nodes, dtype=np.float64)
    nbytes = src.nbytes
    dev_in, dev_out = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dev_in), nbytes) == 0
    assert H.hipMalloc(ctypes.byref(dev_out), nbytes) == 0
    try:
        assert H.hipMemcpy(dev_in, src.ctypes.data_as(ctypes.c_void_p), nbytes, 1) == 0
        # 0xFF bytes: NaN everywhere, so cells the call does not write show
        assert H.hipMemset(dev_out, 0xFF, nbytes) == 0
        st = lib.dim2d_time_dev_noalloc(ctypes.byref(p.to_c()), dev_in, dev_out, None)
        out = np.empty_like(src)
        assert H.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev_out, nbytes, 2) == 0
    finally:
        H.hipFree(dev_in)
        H.hipFree(dev_out)
    return out, st


@pytest.mark.parametrize("case,shapes", [
    ("periodic_x_wall_y", [(300, 210), (70, 41)]),
    ("open", [(260, 133), (40, 52)]),
    ("reflection", [(1020, 1020), (75, 64)]),
])
def test_time_dev_noalloc_matches_time_dev(orc, case, shapes):
    """Two calls in a row on one GPU, the second on a smaller grid (the
    scratch array still holds the first call's data): each equals the
    context's dim2d_time_dev bitwise (same kernels, same inputs), which the
    GPU parity tests hold to the oracle at 1e-11; the largest grid the scratch
    array takes (1020 x 1020) is among them."""
    import dim2d_finite_volume_amd as d
    from test_gpu_parity import BC_CASES, setup_case
    for nx, ny in shapes:
        _, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], nx=nx, ny=ny)
        got, st = _noalloc(p, nodes)
        assert st == 0, d.capi.load().dim2d_last_error()
        f = d.Field(p, nodes)
        want = f.compute_time_dev()
        f.close()
        print(case, (nx, ny), "max |noalloc - time_dev|", float(np.abs(got - want).max()))
        assert np.array_equal(got, want), (case, nx, ny)


def test_time_dev_noalloc_rejects_grids_beyond_its_scratch():
    """A grid whose two interleaved fields exceed the 64 MiB scratch array is
    rejected with DIM2D_EINVAL and a message naming the limit; the output is
    left untouched (still the NaN fill)."""
    import dim2d_finite_volume_amd as d
    p = d.ModelParams(nx=1030, ny=1030, bc=(R, R, R, R))
    nodes = np.ones((4, p.ny, p.nx))
    got, st = _noalloc(p, nodes)
    assert st == d.capi.EINVAL
    msg = d.capi.load().dim2d_last_error().decode()
    assert "scratch" in msg and "67108864" in msg, msg
    assert np.isnan(got).all()
