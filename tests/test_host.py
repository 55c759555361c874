"""Host-side logic (no GPU): coordinates, initial conditions, the time step
rule, and the Fortran binding's build."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_coordinates_match_reference_rule(orc):
    """field_abstract_class.f:286-315: dx = (x_max-x_min)/(nx-1-2 bc_size),
    x_map(bc_size+1) = x_min, x_map(nx-bc_size) = x_max."""
    import dim2d_finite_volume_amd as d
    p = d.ModelParams(nx=37, ny=29, x_min=-0.3, x_max=0.7, y_min=0.0, y_max=0.5)
    xm, ym = p.coordinates()
    oxm, oym = orc.coordinates(37, 29, -0.3, 0.7, 0.0, 0.5)
    assert np.array_equal(xm, oxm) and np.array_equal(ym, oym)
    assert math.isclose(xm[2], -0.3) and math.isclose(xm[-3], 0.7)
    dx, dy = p.spacing()
    assert dx == xm[1] - xm[0] and dy == ym[1] - ym[0]


def test_initial_conditions_are_seeded():
    from dim2d_finite_volume_amd import initial_conditions as ic
    p1, a = ic.drop64(n=64, seed=7)
    p2, b = ic.drop64(n=64, seed=7)
    _, c = ic.drop64(n=64, seed=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    rho_l, rho_v = ic.saturated_densities(0.95)
    # tanh profile of width 2 sqrt(We) dx around a drop of radius 0.15
    assert abs(a[0].max() - rho_l) < 5e-3 and abs(a[0].min() - rho_v) < 1e-5
    # energy of a fluid at rest at T0: E = rho (8/3 cv T0 - 3 rho) + |grad rho|^2/(2 We)
    far = a[:, 2, 2]
    assert math.isclose(far[3], far[0] * (8 / 3 * 2.5 * 0.95 - 3 * far[0]), rel_tol=1e-6)


def test_two_bubbles_and_batch_configs():
    from dim2d_finite_volume_amd import initial_conditions as ic
    p, nodes = ic.two_bubbles(n=128)
    assert nodes.shape == (4, 132, 132) and p.bc == (1, 1, 2, 2)
    assert np.allclose(nodes[1], 0.1 * nodes[0]) and np.all(nodes[2] == 0)
    sims = ic.batch_configs(nsims=40, seed=3)
    sizes = [s[0].nx - 4 for s in sims]
    assert min(sizes) >= 64 and max(sizes) <= 512
    wes = [s[0].we for s in sims]
    assert 5 <= min(wes) and max(wes) <= 20
    for p, n in sims[:5]:
        assert ic.cfl_dt(p, n, 0.1) > 0


@pytest.mark.skipif(shutil.which("/opt/rocm/lib/llvm/bin/flang") is None, reason="no flang")
def test_fortran_binding_builds():
    fdir = os.path.join(ROOT, "dim2d_finite_volume_amd", "fortran")
    r = subprocess.run(["make", "-s", "-C", fdir], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(os.path.join(fdir, "build", "sim_dim2d_hip"))
