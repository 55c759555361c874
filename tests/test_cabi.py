"""The C ABI library (no GPU needed): it loads, exports every symbol that
include/dim2d_hip.h declares, and rejects invalid parameters before touching
the device."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dim2d_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dim2d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import dim2d_finite_volume_amd as d
    L = d.capi.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(d.capi.EXPORTS)
    nm = subprocess.run(["nm", "-D", "--defined-only", d.capi.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dim2d_\w+)", nm))
    assert set(names) <= exported
    # nothing from the oracle in the product library
    assert "orc_" not in nm


def test_version_and_error_text():
    import dim2d_finite_volume_amd as d
    L = d.capi.load()
    assert L.dim2d_version().startswith(b"dim2d-mi355x")
    assert isinstance(L.dim2d_last_error(), bytes)


@pytest.mark.parametrize("field,value,msg", [
    ("bc_size", 3, b"bc_size"),
    ("nx", 6, b"too small"),
    ("bc_w", 7, b"unknown boundary"),
    ("precision", 5, b"precision"),
    ("re", -1.0, b"positive"),
    ("integrator", 7, b"unknown integrator"),
])
def test_invalid_params_rejected(field, value, msg):
    import dim2d_finite_volume_amd as d
    L = d.capi.load()
    p = d.ModelParams(nx=40, ny=40).to_c()
    setattr(p, field, value)
    ctx = ctypes.c_void_p()
    st = L.dim2d_ctx_create(ctypes.byref(p), ctypes.byref(ctx))
    assert st == d.capi.EINVAL and not ctx.value
    assert msg in L.dim2d_last_error()


def test_unsupported_boundary_combinations():
    import dim2d_finite_volume_amd as d
    L = d.capi.load()
    ctx = ctypes.c_void_p()
    # periodic must pair W/E
    p = d.ModelParams(nx=40, ny=40, bc=(1, 0, 0, 0)).to_c()
    assert L.dim2d_ctx_create(ctypes.byref(p), ctypes.byref(ctx)) == d.capi.EINVAL
    # walls on both axes (the reference's library_config.py rejects wall_xy)
    p = d.ModelParams(nx=40, ny=40, bc=(2, 0, 2, 0)).to_c()
    assert L.dim2d_ctx_create(ctypes.byref(p), ctypes.byref(ctx)) == d.capi.EINVAL
    assert b"wall_xy" in L.dim2d_last_error()


def test_low_storage_rk_limits():
    """The low-storage RK3 runs in single contexts only, with the Mattsson
    operators."""
    import dim2d_finite_volume_amd as d
    L = d.capi.load()
    ctx = ctypes.c_void_p()
    p = d.ModelParams(nx=40, ny=40, operators=d.OPS_COCKBURN_GAU,
                      integrator=d.INT_LSRK3).to_c()
    assert L.dim2d_ctx_create(ctypes.byref(p), ctypes.byref(ctx)) == d.capi.EINVAL
    assert b"RK3-TVD only" in L.dim2d_last_error()
    ps = (d.capi.Params * 2)(d.ModelParams(nx=40, ny=40).to_c(),
                             d.ModelParams(nx=40, ny=40, integrator=d.INT_LSRK3).to_c())
    b = ctypes.c_void_p()
    assert L.dim2d_batch_create(ps, 2, ctypes.byref(b)) == d.capi.EINVAL
    assert b"RK3-TVD" in L.dim2d_last_error()


def test_null_arguments():
    import dim2d_finite_volume_amd as d
    L = d.capi.load()
    assert L.dim2d_ctx_create(None, None) == d.capi.EINVAL
    assert L.dim2d_upload(None, None) == d.capi.EINVAL
    assert L.dim2d_rk3_step(None, 0.0, 1e-3) == d.capi.EINVAL
    assert L.dim2d_batch_create(None, 0, None) == d.capi.EINVAL


def test_python_binding_raises():
    import dim2d_finite_volume_amd as d
    with pytest.raises(d.Dim2dError) as e:
        d.Field(d.ModelParams(nx=40, ny=40, bc=(2, 2, 2, 2)))
    assert e.value.code == d.capi.EINVAL


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_cockburn_gau_limits():
    """The Cockburn-Gau operators run in RK3-TVD single contexts (FP64 or
    mixed precision) with reflection or periodic edges; anything else is
    rejected before the device."""
    import dim2d_finite_volume_amd as d
    L = d.capi.load()
    ctx = ctypes.c_void_p()
    cases = [
        (dict(operators=7), b"unknown operator set"),
        (dict(operators=1, integrator=1), b"with RK3-TVD only"),
        (dict(operators=1, bc=(1, 1, 2, 2)), b"reflection and periodic"),
        (dict(operators=1, bc=(3, 3, 3, 3)), b"reflection and periodic"),
    ]
    for kw, msg in cases:
        p = d.ModelParams(nx=40, ny=40, **kw).to_c()
        assert L.dim2d_ctx_create(ctypes.byref(p), ctypes.byref(ctx)) == d.capi.EINVAL, kw
        assert msg in L.dim2d_last_error(), (kw, L.dim2d_last_error())
    n = 1
    p = (d.capi.Params * 1)(d.ModelParams(nx=40, ny=40, operators=1).to_c())
    b = ctypes.c_void_p()
    assert L.dim2d_batch_create(p, n, ctypes.byref(b)) == d.capi.EINVAL
    assert b"Mattsson" in L.dim2d_last_error()
