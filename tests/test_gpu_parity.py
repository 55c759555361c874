"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances (written per test): FP64 single evaluations 1e-11 and 200-step
trajectories 1e-10 in max relative error per variable (max |gpu - oracle| /
max |oracle| over the interior); mixed precision 1e-6 relative density error.
"""
import math

import numpy as np
import pytest

from conftest import max_rel_err

pytestmark = pytest.mark.gpu

R, P, W, O = 0, 1, 2, 3
BC_CASES = {
    "reflection": (R, R, R, R),
    "periodic": (P, P, P, P),
    "periodic_x_wall_y": (P, P, W, W),
    "wall_s_reflection": (R, R, W, R),
    "wall_x_reflection_y": (W, W, R, R),
    "wall_x_periodic_y": (W, W, P, P),
    "open": (O, O, O, O),
}
ALL = (slice(None), slice(None))
INT = (slice(2, -2), slice(2, -2))


def ghost_errors(got, ref, bc):
    """Max relative error over the ghost cells the reference defines: every
    ghost cell of reflection and periodic edges away from wall corners; for a
    wall edge, the density and momenta of the inner ghost line and the density
    of the outer one (the reference leaves the ghost energy and the outer
    momenta unset, wall_xy_equilibrium_module.f:948-1437)."""
    errs = []
    wx = bc[0] == W or bc[1] == W
    wy = bc[2] == W or bc[3] == W
    if not wx and not wy:
        errs += max_rel_err(got, ref, region=(slice(0, 2), slice(None)))
        errs += max_rel_err(got, ref, region=(slice(None), slice(0, 2)))
        return errs
    lines = []  # (region of the inner line, region of the outer line)
    if bc[2] == W:
        lines.append(((slice(1, 2), slice(2, -2)), (slice(0, 1), slice(2, -2))))
    if bc[3] == W:
        lines.append(((slice(-2, -1), slice(2, -2)), (slice(-1, None), slice(2, -2))))
    if bc[0] == W:
        lines.append(((slice(2, -2), slice(1, 2)), (slice(2, -2), slice(0, 1))))
    if bc[1] == W:
        lines.append(((slice(2, -2), slice(-2, -1)), (slice(2, -2), slice(-1, None))))
    for inner, outer in lines:
        errs += max_rel_err(got, ref, region=inner, ks=range(3))
        errs += max_rel_err(got, ref, region=outer, ks=[0])
    # the reflection or periodic ghost lines of the other edges, over the
    # interior span
    for k, region in enumerate(((slice(2, -2), slice(0, 2)), (slice(2, -2), slice(-2, None)),
                                (slice(0, 2), slice(2, -2)), (slice(-2, None), slice(2, -2)))):
        if bc[k] != W:
            errs += max_rel_err(got, ref, region=region)
    return errs


def _lib():
    import dim2d_finite_volume_amd as d
    return d


def perturbed_state(params, seed=3):
    """Drop with mean flow plus smooth momentum/energy perturbations, so every
    flux term is active."""
    from dim2d_finite_volume_amd import initial_conditions as ic
    dx, _ = params.spacing()
    nodes = ic.circles_field(params, 0.95, [(0.05, -0.03)], 0.2, "liquid",
                             2.0 * math.sqrt(params.we) * dx, flow=(0.1, -0.05),
                             noise=1e-4, seed=seed)
    x_map, y_map = params.coordinates()
    X, Y = np.meshgrid(x_map, y_map)
    nodes[1] += 0.02 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    nodes[2] += 0.03 * np.cos(2 * np.pi * X + 0.3) * np.sin(2 * np.pi * Y)
    nodes[3] += 0.01 * np.sin(2 * np.pi * (X + Y))
    return nodes


def oracle_model(orc, p):
    return orc.make_model(p.nx, p.ny, re=p.re, we=p.we, pr=p.pr, viscous_r=p.viscous_r,
                          cv_r=p.cv_r, gravity=p.gravity, bc=p.bc,
                          wall_contact_angle=p.wall_contact_angle,
                          wall_heat_flux=p.wall_heat_flux, integrator=p.integrator,
                          operators=p.operators)


def setup_case(orc, bc, nx=74, ny=41, precision=0, integrator=0, operators=0, **physics):
    d = _lib()
    p = d.ModelParams(nx=nx, ny=ny, bc=bc, precision=precision, integrator=integrator,
                      operators=operators, **physics)
    m = oracle_model(orc, p)
    x_map, y_map = p.coordinates()
    nodes = perturbed_state(p)
    orc.apply_bc(m, nodes, x_map, y_map)
    return d, p, m, x_map, y_map, nodes


@pytest.mark.parametrize("case", list(BC_CASES))
def test_time_dev_fp64(orc, case):
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case])
    ref = orc.time_dev(m, nodes, xm, ym)
    f = d.Field(p, nodes)
    f.set_tiles(64, 16)
    got = f.compute_time_dev()
    err = max_rel_err(got, ref)
    print(case, "time_dev rel err", err)
    assert max(err) < 1e-11, err
    if case == "open":
        # the ghost frame has the characteristic time derivative
        assert max(max_rel_err(got, ref, region=ALL)) < 1e-11
    else:
        # ghost cells carry debug_real like the reference
        assert np.all(got[:, :2, :] == ref[:, :2, :])


@pytest.mark.parametrize("case", ["reflection", "periodic_x_wall_y", "wall_x_periodic_y", "open"])
def test_time_dev_mixed_context_is_fp64(orc, case):
    """dim2d_time_dev (compute_time_dev) of a mixed-precision context: the
    host-facing derivative is evaluated in FP64, so it meets the FP64 bar
    (1e-11 relative, the frame too with open edges) and the steady-state
    check of such a context reports the FP64 maxima."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], precision=1)
    ref = orc.time_dev(m, nodes, xm, ym)
    f = d.Field(p, nodes)
    got = f.compute_time_dev()
    region = ALL if case == "open" else INT
    err = max_rel_err(got, ref, region=region)
    print(case, "mixed-context time_dev rel err", err)
    assert max(err) < 1e-11, err
    want = np.array([np.abs(ref[k][region]).max() for k in range(4)])
    steady, mx = f.check_steady_state(1e-12)
    assert not steady
    assert np.allclose(mx, want, rtol=1e-11, atol=0), (mx, want)
    f.close()


@pytest.mark.parametrize("case", ["periodic_x_wall_y", "wall_s_reflection", "wall_x_reflection_y",
                                  "wall_x_periodic_y"])
@pytest.mark.parametrize("prec", [0, 1])
def test_wall_closed_forms(orc, case, prec):
    """Closed forms of the wall on the GPU itself (the oracle's wall terms are
    pinned by the same ones): a uniform fluid at rest between 90-degree walls
    keeps a zero time derivative over 5 steps, and with a wall heat flux q the
    interior energy grows at (number of walls) x q x (wall length) while the
    mass stays constant (FP64 to rounding; mixed to its FP32 increments)."""
    d = _lib()
    bc = BC_CASES[case]
    n = 66
    p = d.ModelParams(nx=n, ny=n, bc=bc, precision=prec)
    rho, T = 1.2, 0.95
    nodes = np.zeros((4, n, n))
    nodes[0] = rho
    nodes[3] = rho * 8.0 * p.cv_r / 3.0 * T - 3.0 * rho ** 2
    f = d.Field(p, nodes)
    got = f.compute_time_dev()
    dx, dy = p.spacing()
    P = 8 * rho * T / (3 - rho) - 3 * rho ** 2
    assert np.abs(got[:, 2:-2, 2:-2]).max() < 1e-13 * P / dx
    f.run(5, dt=1e-4)
    after = f.download()
    # FP64 keeps the state to rounding; mixed precision to the FP32 rounding of
    # its fluxes, ~1e-7 P/dx per unit time
    tol = 1e-13 * P if prec == 0 else 1e-6 * P / dx * 5 * 1e-4
    assert np.abs(after[:, 2:-2, 2:-2] - nodes[:, 2:-2, 2:-2]).max() < tol
    f.close()
    import dataclasses
    q = 0.01
    walls = sum(1 for k in bc if k == W)
    L = (n - 4) * (dx if bc[2] == W else dy)
    pq = dataclasses.replace(p, wall_heat_flux=q)
    _, _, m, xm, ym, state = setup_case(orc, bc, nx=n, ny=n, precision=prec, wall_heat_flux=q)
    g = d.Field(pq, state)
    s0 = g.diagnostics()
    dt, steps = 2e-5, 10
    g.run(steps, dt=dt)
    s1 = g.diagnostics()
    dE = (s1[3] - s0[3]) * dx * dy
    want = walls * q * L * steps * dt
    tol = 1e-7 if prec == 0 else 1e-4
    assert abs(dE - want) < tol * want, (dE, want)
    assert abs(s1[0] - s0[0]) < (1e-13 if prec == 0 else 1e-9) * s0[0]
    g.close()


@pytest.mark.parametrize("case", list(BC_CASES))
@pytest.mark.parametrize("tiles", [(64, 16), (64, 7), (128, 9)])
def test_steps_fp64(orc, case, tiles):
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case])
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    f = d.Field(p, nodes.copy())
    f.set_tiles(*tiles)
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(10, dt)) == 0
    for _ in range(10):
        f.integrate(dt)
    got = f.download()
    err = max_rel_err(got, ref)
    print(case, tiles, "10-step rel err", err)
    assert max(err) < 1e-10, err
    # ghost cells the reference defines
    if case != "open":
        e = ghost_errors(got, ref, BC_CASES[case])
        assert max(e) < 1e-10, e


def test_drop64_200_steps_fp64(orc):
    """The 64^2 drop for 200 RK3 steps: FP64 parity 1e-10."""
    from dim2d_finite_volume_amd import initial_conditions as ic
    d = _lib()
    p, nodes = ic.drop64()
    m = oracle_model(orc, p)
    xm, ym = p.coordinates()
    orc.apply_bc(m, nodes, xm, ym)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(200, dt)) == 0
    f = d.Field(p, nodes)
    f.run(200, dt=dt)
    got = f.download()
    err = max_rel_err(got, ref)
    print("drop64 200 steps rel err", err)
    assert max(err) < 1e-10, err
    assert max(max_rel_err(got, ref, region=(slice(None), slice(None)))) < 1e-10


def test_drop64_mixed_precision(orc):
    """Mixed FP32/FP64 path against the FP64 oracle after 200 steps: density
    error <= 1e-6 relative; mass conserved to 1e-10 relative (the stage
    increments are stored in FP32, so the telescoping of the face fluxes holds
    to FP32 rounding of dt*L per cell)."""
    from dim2d_finite_volume_amd import initial_conditions as ic
    d = _lib()
    p, nodes = ic.drop64(precision=1)
    m = oracle_model(orc, p)
    xm, ym = p.coordinates()
    orc.apply_bc(m, nodes, xm, ym)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    orc.run(m, ref, xm, ym, np.full(200, dt))
    f = d.Field(p, nodes.copy())
    mass0 = f.diagnostics()[0]
    f.run(200, dt=dt)
    got = f.download()
    err = max_rel_err(got, ref)
    print("mixed drop64 200 steps rel err", err)
    assert err[0] < 1e-6 and err[3] < 1e-6, err
    assert abs(f.diagnostics()[0] - mass0) / mass0 < 1e-10


def test_cfl_integrate(orc):
    """Device-side CFL time step fused into stage 3 against the oracle's rule."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["periodic_x_wall_y"])
    f = d.Field(p, nodes.copy())
    ref = nodes.copy()
    dts = []
    for _ in range(12):
        dt = orc.cfl_dt(m, ref, *p.spacing(), 0.1)[0]
        dts.append(dt)
        assert orc.run(m, ref, xm, ym, np.array([dt])) == 0
    f.run(12, cfl=0.1)
    got = f.download()
    t, last_dt = f.device_time()
    assert abs(t - sum(dts)) < 1e-12 * sum(dts)
    assert abs(last_dt - dts[-1]) < 1e-12 * dts[-1]
    err = max_rel_err(got, ref)
    print("cfl run rel err", err)
    assert max(err) < 1e-10, err


@pytest.mark.parametrize("case", ["periodic_x_wall_y", "open"])
def test_cfl_integrate_mixed(orc, case):
    """Mixed precision with the device CFL time step (its sound speed in FP32
    after an FP64 temperature) against the FP64 oracle stepping with the FP64
    rule: elapsed time and last dt within 1e-6 relative; density and energy
    within 1e-6 of the oracle."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], precision=1)
    f = d.Field(p, nodes.copy())
    ref = nodes.copy()
    dts = []
    for _ in range(12):
        dt = orc.cfl_dt(m, ref, *p.spacing(), 0.1)[0]
        dts.append(dt)
        assert orc.run(m, ref, xm, ym, np.array([dt])) == 0
    f.run(12, cfl=0.1)
    got = f.download()
    t, last_dt = f.device_time()
    print(case, "mixed cfl: t", t / sum(dts) - 1, "last dt", last_dt / dts[-1] - 1)
    assert abs(t - sum(dts)) < 1e-6 * sum(dts)
    assert abs(last_dt - dts[-1]) < 1e-6 * dts[-1]
    err = max_rel_err(got, ref)
    print(case, "mixed cfl run rel err", err)
    assert err[0] < 1e-6 and err[3] < 1e-6, err


@pytest.mark.parametrize("case", ["reflection", "periodic_x_wall_y", "wall_s_reflection", "open"])
def test_lsrk3_steps_fp64(orc, case):
    """Low-storage RK3 (2N-storage, Williamson 1980): 20 fixed-dt steps against
    the oracle's restatement of the same scheme, FP64 parity 1e-10, with its
    stages 1 and 2 fused (the default; open edges keep the separate stages)
    and apart; fused and apart agree to 1e-12 (walls: 1e-11, the ghost root
    of u2 starting from u0 instead of u1 converges to within the root's
    tolerance of the same value)."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], integrator=1)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(20, dt)) == 0
    outs = []
    for fuse in ((True, False) if case != "open" else (False,)):
        f = d.Field(p, nodes.copy())
        f.set_stage_fusion(fuse)
        f.run(20, dt=dt)
        got = f.download()
        f.close()
        err = max_rel_err(got, ref)
        print(case, "fused" if fuse else "apart", "lsrk3 20-step rel err", err)
        assert max(err) < 1e-10, err
        outs.append(got)
    if len(outs) == 2:
        tol = 1e-11 if W in BC_CASES[case] else 1e-12
        assert max(max_rel_err(outs[0], outs[1])) < tol


@pytest.mark.parametrize("case", ["reflection", "periodic_x_wall_y", "wall_s_reflection", "open"])
def test_lsrk3_steps_mixed(orc, case):
    """Low-storage RK3 in mixed precision (FP32 fluxes and register q, FP64
    stage states and state updates): 20 fixed-dt steps against the oracle's
    FP64 restatement of the scheme, density and energy to 1e-6 (frame cells
    included for open edges); stages 1 and 2 fused (the default; open edges
    keep the separate stages) and apart agree to 1e-7; the FP64 GPU run of
    the same scheme to 1e-6."""
    import dataclasses
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], precision=1, integrator=1)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(20, dt)) == 0
    region = ALL if case == "open" else INT
    outs = []
    for fuse in ((True, False) if case != "open" else (False,)):
        f = d.Field(p, nodes.copy())
        f.set_stage_fusion(fuse)
        f.run(20, dt=dt)
        got = f.download()
        f.close()
        err = max_rel_err(got, ref, region=region)
        print(case, "fused" if fuse else "apart", "mixed lsrk3 20-step rel err", err)
        assert np.isfinite(got).all()
        assert err[0] < 1e-6 and err[3] < 1e-6, err
        outs.append(got)
    if len(outs) == 2:
        e = max_rel_err(outs[0], outs[1])
        print(case, "mixed lsrk3 fused vs apart", e)
        assert max(e) < 1e-7, e
    f = d.Field(dataclasses.replace(p, precision=0), nodes.copy())
    f.run(20, dt=dt)
    g64 = f.download()
    f.close()
    e = max_rel_err(outs[0], g64, region=region)
    print(case, "mixed vs FP64 lsrk3", e)
    assert e[0] < 1e-6 and e[3] < 1e-6, e


@pytest.mark.parametrize("fuse,prec", [(True, 0), (False, 0), (True, 1), (False, 1)])
def test_lsrk3_cfl_integrate(orc, fuse, prec):
    """Low-storage RK3 with the device CFL time step (graph replays), stages 1
    and 2 fused and apart, FP64 (parity 1e-10) and mixed (density and energy
    1e-6, time 1e-6)."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["periodic"], integrator=1, precision=prec)
    f = d.Field(p, nodes.copy())
    f.set_stage_fusion(fuse)
    ref = nodes.copy()
    dts = []
    for _ in range(10):
        dt = orc.cfl_dt(m, ref, *p.spacing(), 0.1)[0]
        dts.append(dt)
        assert orc.run(m, ref, xm, ym, np.array([dt])) == 0
    f.run(10, cfl=0.1)
    got = f.download()
    t, _ = f.device_time()
    err = max_rel_err(got, ref)
    print("lsrk3 cfl run rel err", prec, err)
    if prec == 0:
        assert abs(t - sum(dts)) < 1e-12 * sum(dts)
        assert max(err) < 1e-10, err
    else:
        assert abs(t - sum(dts)) < 1e-6 * sum(dts)
        assert err[0] < 1e-6 and err[3] < 1e-6, err


def test_open_boundaries_whole_field(orc):
    """Hedstrom open edges (hedstrom_xy): the ghost frame is integrated with
    its characteristic time derivative; 30 RK3 steps of a drop in a mean flow
    against the oracle over every cell, frame included, FP64 parity 1e-10;
    then the device CFL time step over 10 more steps."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["open"], nx=70, ny=53)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(30, dt)) == 0
    f = d.Field(p, nodes.copy())
    f.run(30, dt=dt)
    got = f.download()
    err = max_rel_err(got, ref, region=ALL)
    print("open 30-step rel err (all cells)", err)
    assert max(err) < 1e-10, err
    dts = []
    for _ in range(10):
        dtc = orc.cfl_dt(m, ref, *p.spacing(), 0.1)[0]
        dts.append(dtc)
        assert orc.run(m, ref, xm, ym, np.array([dtc])) == 0
    f.run(10, cfl=0.1)
    got = f.download()
    err = max_rel_err(got, ref, region=ALL)
    print("open CFL rel err (all cells)", err)
    assert max(err) < 1e-10, err


def test_open_boundaries_mixed(orc):
    """Mixed precision with open edges: density and energy within 1e-6 of the
    FP64 oracle over every cell after 50 steps."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["open"], nx=70, ny=53, precision=1)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(50, dt)) == 0
    f = d.Field(p, nodes.copy())
    f.run(50, dt=dt)
    err = max_rel_err(f.download(), ref, region=ALL)
    print("open mixed 50-step rel err", err)
    assert err[0] < 1e-6 and err[3] < 1e-6, err


@pytest.mark.parametrize("shape", [(7, 7), (200, 9), (9, 150), (132, 68)])
def test_open_frame_segments(orc, shape):
    """Open edges on the smallest grid and on grids whose frame lines hold
    one, several, partial and exactly filled segments of 64 cells (the frame
    kernel's workgroups): time derivative over every cell (1e-11) and 5 RK3
    steps (1e-10) against the oracle.  Liquid at T0 = 0.95 with the
    perturbations of perturbed_state, so the sound speed is real everywhere.
    On the 7x7 grid (dx = 1/3) the derivatives are ~6e-4, differences of
    fluxes ~1e3 times larger, so the time derivative's tolerance there is
    1e-11 x 1e2."""
    d = _lib()
    from dim2d_finite_volume_amd import initial_conditions as ic
    p = d.ModelParams(nx=shape[0], ny=shape[1], bc=BC_CASES["open"])
    m = oracle_model(orc, p)
    dx, _ = p.spacing()
    nodes = ic.circles_field(p, 0.95, [(9.0, 9.0)], 0.1, "vapor", 2.0 * math.sqrt(p.we) * dx,
                             flow=(0.1, -0.05), noise=1e-4, seed=3)
    xm, ym = p.coordinates()
    X, Y = np.meshgrid(xm, ym)
    nodes[1] += 0.02 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    nodes[2] += 0.03 * np.cos(2 * np.pi * X + 0.3) * np.sin(2 * np.pi * Y)
    nodes[3] += 0.01 * np.sin(2 * np.pi * (X + Y))
    orc.apply_bc(m, nodes, xm, ym)
    ref = orc.time_dev(m, nodes, xm, ym)
    f = d.Field(p, nodes.copy())
    err = max_rel_err(f.compute_time_dev(), ref, region=ALL)
    print(shape, "open time_dev rel err (all cells)", err)
    assert max(err) < (1e-9 if shape == (7, 7) else 1e-11), err
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(5, dt)) == 0
    f.run(5, dt=dt)
    err = max_rel_err(f.download(), ref, region=ALL)
    print(shape, "open 5-step rel err (all cells)", err)
    assert max(err) < 1e-10, err
    f.close()


@pytest.mark.parametrize("case,prec,integ,ops,physics", [
    ("reflection", 0, 0, 0, dict(gravity=0.5)),
    ("periodic_x_wall_y", 0, 0, 0, dict(gravity=0.5)),
    ("open", 0, 0, 0, dict(gravity=0.5)),
    ("periodic_x_wall_y", 1, 0, 0, dict(gravity=0.5)),
    ("periodic_x_wall_y", 0, 1, 0, dict(gravity=0.5)),
    ("periodic", 0, 0, 1, dict(gravity=0.5)),
    ("periodic_x_wall_y", 0, 0, 0, dict(wall_contact_angle=60.0)),
    ("wall_s_reflection", 0, 0, 0, dict(wall_contact_angle=120.0, wall_heat_flux=0.01)),
    ("periodic_x_wall_y", 1, 0, 0, dict(wall_contact_angle=60.0, wall_heat_flux=0.01)),
    ("wall_s_reflection", 0, 1, 0, dict(wall_contact_angle=45.0, wall_heat_flux=-0.01,
                                        gravity=0.5)),
    ("wall_x_reflection_y", 0, 0, 0, dict(wall_contact_angle=120.0, wall_heat_flux=0.01,
                                          gravity=0.5)),
    ("wall_x_periodic_y", 1, 0, 0, dict(wall_contact_angle=70.0, wall_heat_flux=-0.01)),
    ("wall_x_periodic_y", 0, 1, 0, dict(wall_contact_angle=110.0)),
])
def test_body_force_and_wall_physics(orc, case, prec, integ, ops, physics):
    """Gravity (compute_body_forces, pmodel_eq_class.f:1840-1918), contact
    angles other than 90 degrees (the wetting term of the wall root function,
    wall_xy_equilibrium_module.f:2238-2293) and a wall heat flux
    (get_wall_heat_flux, :1552-1569, in the wall temperature and the wall
    energy flux), across boundary sets, precisions, both integrators and the
    Cockburn-Gau operators: the time derivative (FP64: 1e-11) and 10 steps
    (FP64 1e-10 on every variable; mixed 1e-6 on density and energy) against
    the oracle, ghost rows of the walls included (density and momenta of the
    inner one)."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], precision=prec, integrator=integ,
                                        operators=ops, **physics)
    if prec == 0:
        ref = orc.time_dev(m, nodes, xm, ym)
        got = d.Field(p, nodes).compute_time_dev()
        region = ALL if case == "open" else (slice(2, -2), slice(2, -2))
        err = max_rel_err(got, ref, region=region)
        print(case, physics, "time_dev rel err", err)
        assert max(err) < 1e-11, err
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(10, dt)) == 0
    f = d.Field(p, nodes.copy())
    f.run(10, dt=dt)
    got = f.download()
    err = max_rel_err(got, ref, region=ALL if case == "open" else (slice(2, -2), slice(2, -2)))
    print(case, prec, integ, ops, physics, "10-step rel err", err)
    if prec == 0:
        assert max(err) < 1e-10, err
    else:
        assert err[0] < 1e-6 and err[3] < 1e-6, err
    if W in BC_CASES[case]:
        tol = 1e-10 if prec == 0 else 1e-6
        e = ghost_errors(got, ref, BC_CASES[case])
        assert max(e) < tol, e
    f.close()


def test_cfl_and_diagnostics(orc):
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["reflection"])
    f = d.Field(p, nodes)
    ref = orc.cfl_dt(m, nodes, *p.spacing(), 0.3)
    got = f.cfl_dt(0.3)
    assert np.allclose(got, ref, rtol=1e-13, atol=0), (got, ref)
    diag = f.diagnostics()
    sums = orc.sums(m, nodes)
    assert np.allclose(diag[:4], sums, rtol=1e-12, atol=1e-12 * np.abs(sums).max())
    it = nodes[0, 2:-2, 2:-2]
    assert diag[4] == it.min() and diag[5] == it.max() and diag[6] == 0


@pytest.mark.parametrize("fuse", [False, True, "pipe"])
def test_batch_matches_single(orc, fuse):
    """Persistent batch kernels under a HIP graph give the single-simulation
    result, with stages 1 and 2 fused (against single runs that fuse them),
    apart, and pipelined (two groups half a step apart, a stage pair beside a
    stage 3 in each launch; the steps come in calls of 1, 4 and 1 so that the
    first-pair-alone, graph replay, leftover step and odd-count buffer swap
    all run)."""
    d = _lib()
    from dim2d_finite_volume_amd import initial_conditions as ic
    sims = ic.batch_configs(nsims=5, seed=11, nmin=20, nmax=90)
    dts = [ic.cfl_dt(p, n, 0.1) for p, n in sims]
    b = d.Batch([p for p, _ in sims])
    # every edge here allows fusion; FP64 batches are pipelined by default
    assert b.tiles()["stage_fusion"] and b.tiles()["pipelined"]
    b.set_tiles(64, 16)
    b.set_stage_fusion(bool(fuse), pipeline=fuse == "pipe")
    for i, (p, n) in enumerate(sims):
        b.upload(i, n)
    # the autotuner runs real steps on candidate tilings, then restores the state
    rows, wg, ms = b.autotune(dts, 4)
    assert rows in (8, 12, 16, 24, 32) and wg > 0 and ms > 0
    t = b.tiles()
    assert t["stage_fusion"] == bool(fuse) and t["rows"] == rows
    assert t["pipelined"] == (fuse == "pipe")
    if fuse:
        assert t["pair_rows"] in (16, 24, 32, 48, 60, 96, 128) and t["pair_workgroups"] > 0
    if fuse == "pipe":
        for k in (1, 4, 1):
            b.integrate(dts, k)
    else:
        b.integrate(dts, 6)
    for i, (p, n) in enumerate(sims):
        f = d.Field(p, n)
        f.set_tiles(64, 16)
        f.set_stage_fusion(bool(fuse))
        f.run(6, dt=dts[i])
        single = f.download()
        got = b.download(i)
        assert np.array_equal(got[:, 2:-2, 2:-2], single[:, 2:-2, 2:-2]), i


def test_pipelined_batch_edge_cases(orc):
    """A pipelined batch of one simulation (its second group is empty: each
    step is the stage pair, then stage 3, from the two mixed launches) equals
    the single fused run bitwise, over calls of 1, 2 and 3 steps; fusion modes
    other than 0, 1 and 2 are rejected."""
    import ctypes
    d = _lib()
    _, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["periodic_x_wall_y"], nx=61, ny=47)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    b = d.Batch([p])
    assert not b.tiles()["pipelined"]  # one simulation: not pipelined by default
    b.set_stage_fusion(True, pipeline=True)
    assert b.tiles()["pipelined"]
    b.upload(0, nodes)
    for k in (1, 2, 3):
        b.integrate([dt], k)
    f = d.Field(p, nodes.copy())
    f.run(6, dt=dt)
    single = f.download()
    f.close()
    got = b.download(0)
    assert np.array_equal(got[:, 2:-2, 2:-2], single[:, 2:-2, 2:-2])
    assert np.array_equal(got[:3, 1, 2:-2], single[:3, 1, 2:-2])
    L = d.capi.load()
    for bad in (3, -1):
        assert L.dim2d_batch_set_stage_fusion(b._b, bad, 0, 0) == d.capi.EINVAL
    assert b.tiles()["pipelined"]  # a rejected call changes nothing
    b.close()


@pytest.mark.parametrize("prec,mode", [(0, True), (1, True), (0, "pipe")])
def test_batch_pair_tiles_of_one_to_three_rows(orc, prec, mode):
    """Batch stage-pair tiles of 1, 2 and 3 rows (the single-context API takes
    4 and up): the second tile of a 1-row tiling starts at row 3 and gathers
    the inner wall ghost row, which its own consumer then computes; every
    tiling equals single fused runs bitwise, wall ghost rows included."""
    d = _lib()
    specs = [
        dict(bc=BC_CASES["wall_s_reflection"], nx=66, ny=30, wall_contact_angle=120.0,
             wall_heat_flux=0.01),
        dict(bc=BC_CASES["periodic_x_wall_y"], nx=48, ny=25),
        dict(bc=BC_CASES["reflection"], nx=40, ny=21),
    ]
    sims = []
    for sp in specs:
        sp = dict(sp)
        bc, nx, ny = sp.pop("bc"), sp.pop("nx"), sp.pop("ny")
        _, p, m, xm, ym, nodes = setup_case(orc, bc, nx=nx, ny=ny, precision=prec, **sp)
        sims.append((p, nodes, orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]))
    singles = []
    for p, n, dt in sims:
        f = d.Field(p, n.copy())
        f.run(5, dt=dt)
        singles.append(f.download())
        f.close()
    for rows in (1, 2, 3):
        b = d.Batch([s[0] for s in sims])
        b.set_stage_fusion(True, rows=rows, pipeline=mode == "pipe")
        assert b.tiles()["pair_rows"] == rows
        for i, s in enumerate(sims):
            b.upload(i, s[1])
        b.integrate([s[2] for s in sims], 5)
        for i, (p, n, dt) in enumerate(sims):
            got = b.download(i)
            assert np.isfinite(got[:, 2:-2, 2:-2]).all(), (rows, i)
            assert np.array_equal(got[:, 2:-2, 2:-2], singles[i][:, 2:-2, 2:-2]), (rows, i)
            if p.bc[2] == W:
                assert np.array_equal(got[:3, 1, 2:-2], singles[i][:3, 1, 2:-2]), (rows, i)
        b.close()


@pytest.mark.parametrize("prec,fuse", [(0, False), (1, False), (0, True), (1, True),
                                       (0, "pipe"), (1, "pipe")])
def test_batch_mixed_edges_matches_single(orc, prec, fuse):
    """A batch whose simulations have different edges (reflection, periodic,
    periodic x with walls y, a south wall with a 120-degree contact angle and
    heat flux, gravity, walls on W/E) and sizes: 7 steps of the persistent
    batch kernels and the batched wall kernel equal single-simulation runs
    bitwise on every interior cell and on the inner wall ghost row (density
    and momenta), and the oracle to 1e-10 (mixed: 1e-6).  With stages 1 and 2
    fused (the W/E-wall simulation left out: the stage pair takes S/N walls
    only, and a batch holding it runs the stages apart) against single runs
    that fuse them.  (A 60-degree angle on the 66x70 drop goes non-finite
    within 2 steps in the oracle too, so that case is not used.)"""
    d = _lib()
    specs = [
        dict(bc=BC_CASES["reflection"], nx=40, ny=33),
        dict(bc=BC_CASES["periodic"], nx=70, ny=45),
        dict(bc=BC_CASES["periodic_x_wall_y"], nx=131, ny=52),
        dict(bc=BC_CASES["wall_s_reflection"], nx=66, ny=70, wall_contact_angle=120.0,
             wall_heat_flux=0.01),
        dict(bc=BC_CASES["periodic_x_wall_y"], nx=48, ny=90, gravity=0.5),
        dict(bc=BC_CASES["wall_x_reflection_y"], nx=57, ny=61, wall_contact_angle=70.0),
    ]
    if fuse:
        specs = specs[:-1]
    sims = []
    for sp in specs:
        sp = dict(sp)
        bc, nx, ny = sp.pop("bc"), sp.pop("nx"), sp.pop("ny")
        _, p, m, xm, ym, nodes = setup_case(orc, bc, nx=nx, ny=ny, precision=prec, **sp)
        sims.append((p, nodes, orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0], m, xm, ym))
    b = d.Batch([s[0] for s in sims])
    assert b.tiles()["stage_fusion"] == bool(fuse)
    assert b.tiles()["pipelined"] == (bool(fuse) and prec == 0)
    if fuse:
        b.set_stage_fusion(True, pipeline=fuse == "pipe")
    for i, s in enumerate(sims):
        b.upload(i, s[1])
    dts = [s[2] for s in sims]
    b.integrate(dts, 7)
    for i, (p, n, dt, m, xm, ym) in enumerate(sims):
        f = d.Field(p, n.copy())
        f.set_stage_fusion(bool(fuse))
        f.run(7, dt=dt)
        single = f.download()
        f.close()
        got = b.download(i)
        assert np.isfinite(single[:, 2:-2, 2:-2]).all(), i
        assert np.array_equal(got[:, 2:-2, 2:-2], single[:, 2:-2, 2:-2]), i
        if p.bc[2] == W:
            assert np.array_equal(got[:3, 1, 2:-2], single[:3, 1, 2:-2]), i
        if p.bc[0] == W:
            assert np.array_equal(got[:3, 2:-2, 1], single[:3, 2:-2, 1]), i
        # the batch against the oracle directly, not only through single runs
        ref = n.copy()
        assert orc.run(m, ref, xm, ym, np.full(7, dt)) == 0
        err = max_rel_err(got, ref)
        print("batch sim", i, p.bc, "vs oracle", err)
        if prec == 0:
            assert max(err) < 1e-10, (i, err)
            if W in p.bc:
                assert max(ghost_errors(got, ref, p.bc)) < 1e-10, i
        else:
            assert err[0] < 1e-6 and err[3] < 1e-6, (i, err)


@pytest.mark.parametrize("case,prec", [("periodic_x_wall_y", 0), ("reflection", 1), ("open", 0),
                                       ("open", 1)])
def test_graph_replay_matches_eager(orc, case, prec):
    """dim2d_integrate replays steps from captured HIP graphs of 8 and of 2
    steps (one per step parity); profiling makes it launch every step
    eagerly.  Several replays with host work in between, odd step counts,
    counts that take 8-step graphs and a dt change give bitwise the eager
    results, fixed-dt and CFL."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], precision=prec)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    g = d.Field(p, nodes.copy())
    e = d.Field(p, nodes.copy())
    e.profile(True, kinds=["stage1"])
    for n, kw in ((4, {"dt": dt}), (5, {"dt": dt}), (2, {"dt": 0.5 * dt}), (7, {"cfl": 0.1}),
                  (3, {"cfl": 0.1}), (19, {"cfl": 0.1}), (10, {"dt": dt}), (8, {"dt": dt})):
        g.run(n, **kw)
        a = g.download()
        e.run(n, **kw)
        b = e.download()
        assert np.array_equal(a, b), (n, kw)
    assert np.array_equal(g.device_time(), e.device_time())


@pytest.mark.parametrize("precision", [0, 1])
def test_tile_invariance(orc, precision):
    """Results are bitwise independent of the tile height and of one or two
    column strips per workgroup (131 columns: 3 strips, so the second
    workgroup of a pair holds one strip and one wave that exits at once)."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["periodic"], nx=131, ny=70,
                                        precision=precision)
    outs = []
    for tiles in [(64, 64), (64, 5), (64, 33), (128, 16), (128, 5), (128, 70)]:
        f = d.Field(p, nodes.copy())
        f.set_tiles(*tiles)
        f.run(3, dt=1e-6)
        outs.append(f.download())
        f.close()
    for k, o in enumerate(outs[1:]):
        assert np.array_equal(outs[0], o), k


FUSED_CASES = ["reflection", "periodic", "periodic_x_wall_y", "wall_s_reflection", "open"]


@pytest.mark.parametrize("case", FUSED_CASES)
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("shape", [(131, 70), (76, 45), (12, 12)])
def test_stage_fusion_matches_separate_stages(orc, case, prec, shape):
    """The fused stages 1 and 2 (one launch, stage-1 state on chip, its ghost
    cells from periodic images, reflection mirrors and in-kernel wall roots)
    against the separate stage kernels with the wall kernel between them:
    equal results after 6 steps up to rounding (the two launch shapes let the
    compiler contract products into FMAs differently: FP64 to 1e-12 of each
    variable's scale, mixed to 1e-7), for tile heights from 4 rows to one tile for the whole
    grid, on grids whose last strip is partial (131 and 76 columns) and on
    the smallest grid the fused launch takes (8 interior points), and with
    last tiles of one row (41 interior rows in 4- and 8-row tiles), whose
    neighbour tile's stencil reaches the north ghost rows; the fused run
    against the oracle too (FP64 1e-10, mixed 1e-6 on density and energy).
    Open edges: the ghost frame (state) included, its stage 1 before the
    fused launch and its stage 2 after it."""
    nx, ny = shape
    kw = dict(wall_contact_angle=70.0, wall_heat_flux=0.01) if W in BC_CASES[case] else {}
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], nx=nx, ny=ny, precision=prec, **kw)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    f = d.Field(p, nodes.copy())
    f.set_stage_fusion(False)
    f.run(6, dt=dt)
    sep = f.download()
    f.close()
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(6, dt)) == 0
    tol = 1e-12 if prec == 0 else 1e-7
    first = None
    for rows in (4, 7, 8, 32, ny):
        f = d.Field(p, nodes.copy())
        f.set_stage_fusion(True, rows)
        f.run(6, dt=dt)
        got = f.download()
        f.close()
        # every output row runs the same instruction stream whatever the tiling
        if first is None:
            first = got
        assert np.array_equal(got[:, 2:-2, 2:-2], first[:, 2:-2, 2:-2]), rows
        err = max_rel_err(got, sep, region=ALL if W not in BC_CASES[case] else INT)
        print(case, prec, shape, rows, "fused vs separate", err)
        assert max(err) < tol, (rows, err)
        if W in BC_CASES[case]:
            assert max(ghost_errors(got, sep, BC_CASES[case])) < tol, rows
        e = max_rel_err(got, ref, region=ALL if case == "open" else INT)
        if prec == 0:
            assert max(e) < 1e-10, (rows, e)
        else:
            assert e[0] < 1e-6 and e[3] < 1e-6, (rows, e)


@pytest.mark.parametrize("case", ["periodic_x_wall_y", "open"])
def test_stage_fusion_with_cfl_and_graphs(orc, case):
    """The fused launch under graph replay with the device CFL time step
    (the stage pair takes dt from the reductions and clears the next slot)
    equals the separate launches up to rounding: elapsed time, last dt and
    state to 1e-12 (open edges: the frame kernels of stages 1 and 2 on the
    context's stream around it, inside the captured graphs)."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], nx=200, ny=150)
    out = []
    for fuse in (False, True):
        f = d.Field(p, nodes.copy())
        f.set_stage_fusion(fuse)
        f.run(9, cfl=0.1)
        out.append((f.download(), f.device_time()))
        f.close()
    (a, ta), (b, tb) = out
    assert abs(ta[0] - tb[0]) <= 1e-12 * ta[0] and abs(ta[1] - tb[1]) <= 1e-12 * ta[1]
    assert max(max_rel_err(b, a)) < 1e-12


def test_stage_fusion_rejects_unsupported_edges():
    """Walls on W/E, the low-storage RK3 with open edges, the Cockburn-Gau
    operators and tiny grids keep the separate stage kernels: turning fusion
    on there is an error."""
    d = _lib()
    for kw in (dict(bc=BC_CASES["wall_x_reflection_y"]),
               dict(bc=BC_CASES["open"], integrator=1),
               dict(bc=BC_CASES["periodic"], operators=1),
               dict(bc=BC_CASES["reflection"], nx=11, ny=40)):
        kw = dict(kw)
        p = d.ModelParams(nx=kw.pop("nx", 40), ny=kw.pop("ny", 36), **kw)
        f = d.Field(p, np.ones((4, p.ny, p.nx)))
        with pytest.raises(RuntimeError):
            f.set_stage_fusion(True)
        f.set_stage_fusion(False)
        f.close()
    # a batch holding one such simulation runs the stages apart
    ps = [d.ModelParams(nx=40, ny=36, bc=BC_CASES["reflection"]),
          d.ModelParams(nx=40, ny=36, bc=BC_CASES["wall_x_reflection_y"])]
    b = d.Batch(ps)
    assert not b.tiles()["stage_fusion"]
    with pytest.raises(RuntimeError):
        b.set_stage_fusion(True)
    b.close()


def test_set_tiles_rejects_other_widths():
    """block_x is 64 or 128 (one or two strips per workgroup); 0 keeps it."""
    from dim2d_finite_volume_amd import initial_conditions as ic
    d = _lib()
    p, nodes = ic.drop64()
    f = d.Field(p, nodes)
    for bad in (32, 96, 192, 256):
        with pytest.raises(RuntimeError):
            f.set_tiles(bad, 0)
    f.set_tiles(128, 0)
    f.set_tiles(0, 8)
    f.close()


def test_smallest_grid(orc):
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES["reflection"], nx=7, ny=7)
    ref = orc.time_dev(m, nodes, xm, ym)
    got = d.Field(p, nodes).compute_time_dev()
    assert max(max_rel_err(got, ref)) < 1e-11


@pytest.mark.parametrize("case", ["reflection", "periodic_x_wall_y", "open"])
def test_check_steady_state_max_matches_oracle(orc, case):
    """field%check_steady_state: the max |time_dev| per variable over the cells
    with a derivative (interior; the frame too with open edges) against the
    oracle's time derivative, 1e-11 relative; limits on either side of it."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case])
    ref = orc.time_dev(m, nodes, xm, ym)
    region = ALL if case == "open" else (slice(2, -2), slice(2, -2))
    want = np.array([np.abs(ref[k][region]).max() for k in range(4)])
    f = d.Field(p, nodes)
    steady, got = f.check_steady_state(1e-12)
    assert not steady
    assert np.allclose(got, want, rtol=1e-11, atol=0), (got, want)
    assert f.check_steady_state(1.01 * want.max())[0]
    assert not f.check_steady_state(0.99 * want.max())[0]
    # the state is unchanged by the check
    assert np.array_equal(f.download()[:, 2:-2, 2:-2], nodes[:, 2:-2, 2:-2])
    f.close()


def test_check_steady_state_uniform_and_nonfinite():
    """A uniform fluid at rest has an exactly zero time derivative; a NaN in the
    state is never steady (the maximum reads +inf)."""
    d = _lib()
    p = d.ModelParams(nx=40, ny=36, bc=(R, R, R, R))
    rho, T = 1.2, 0.97
    E = rho * 8.0 * p.cv_r / 3.0 * T - 3.0 * rho ** 2
    nodes = np.zeros((4, p.ny, p.nx))
    nodes[0], nodes[3] = rho, E
    f = d.Field(p, nodes)
    steady, got = f.check_steady_state(0.0)
    assert steady and np.array_equal(got, np.zeros(4))
    nodes[3, 10, 12] = np.nan
    f.upload(nodes)
    steady, got = f.check_steady_state(1e30)
    assert not steady and np.isinf(got).any()
    f.close()


@pytest.mark.parametrize("case", ["periodic_x_wall_y", "open"])
def test_restart_from_output_file_continues_exactly(orc, tmp_path, case):
    """field%write_data then a restart from that file (-r) continues the run
    bit for bit: the file holds the state, ghost cells included, and the time.
    A restart that takes its parameters from the file's header (wall angle
    and heat flux through the description of a wall initial condition) does
    the same."""
    from dim2d_finite_volume_amd import io_netcdf
    phys = {"wall_contact_angle": 70.0, "wall_heat_flux": 0.01} if case != "open" else {}
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], **phys)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    f = d.Field(p, nodes)
    f.run(5, dt=dt)
    path = str(tmp_path / "data1.nc")
    ic = io_netcdf.InitialCondition("bubble_next_to_wall" if phys else "bubble_transported")
    f.write_data(path, dt=dt, ic=ic)
    g = d.Field.restart(p, path)
    h = d.Field.restart(None, path)
    assert h.params == p
    assert g.time == f.time == h.time
    f.run(7, dt=dt)
    g.run(7, dt=dt)
    h.run(7, dt=dt)
    assert np.array_equal(f.download(), g.download())
    assert np.array_equal(f.download(), h.download())
    assert np.array_equal(f.device_time(), g.device_time())
    for x in (f, g, h):
        x.close()


# ---- Cockburn and Gau operators ----

CG_CASES = ["reflection", "periodic"]


@pytest.mark.parametrize("case", CG_CASES)
@pytest.mark.parametrize("shape", [(74, 41), (37, 52)])
def test_cockburn_gau_time_dev(orc, case, shape):
    """The Cockburn-Gau time derivative (sd_choice cockburnandgau) against the
    oracle with the same operator set, 1e-11, on grids that are and are not
    multiples of the 16x16 tile."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], nx=shape[0], ny=shape[1],
                                        operators=1)
    ref = orc.time_dev(m, nodes, xm, ym)
    f = d.Field(p, nodes)
    got = f.compute_time_dev()
    err = max_rel_err(got, ref)
    print(case, shape, "CG time_dev rel err", err)
    assert max(err) < 1e-11, err
    f.close()


@pytest.mark.parametrize("case", CG_CASES)
def test_cockburn_gau_steps(orc, case):
    """30 RK3-TVD steps with the Cockburn-Gau operators, fixed dt, then 10 steps
    with the device CFL rule, against the oracle: 1e-10."""
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], operators=1)
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(30, dt)) == 0
    f = d.Field(p, nodes.copy())
    f.run(30, dt=dt)
    err = max_rel_err(f.download(), ref)
    print(case, "CG 30-step rel err", err)
    assert max(err) < 1e-10, err
    for _ in range(10):
        dtc = orc.cfl_dt(m, ref, *p.spacing(), 0.1)[0]
        assert orc.run(m, ref, xm, ym, np.array([dtc])) == 0
    f.run(10, cfl=0.1)
    err = max_rel_err(f.download(), ref)
    print(case, "CG CFL rel err", err)
    assert max(err) < 1e-10, err
    f.close()


@pytest.mark.parametrize("case", CG_CASES)
@pytest.mark.parametrize("shape", [(74, 41), (37, 52)])
def test_cockburn_gau_steps_mixed(orc, case, shape):
    """The Cockburn-Gau operators in mixed precision (FP32 fluxes and stage
    increments, FP64 densities and state): 30 fixed-dt steps, then 10 with
    the device CFL rule, against the FP64 oracle with the same operators,
    density and energy to 1e-6, and against the FP64 GPU run of the same
    operators to 1e-6; a mixed context's time derivative stays FP64 (1e-11
    against the oracle)."""
    import dataclasses
    d, p, m, xm, ym, nodes = setup_case(orc, BC_CASES[case], nx=shape[0], ny=shape[1],
                                        operators=1, precision=1)
    got = d.Field(p, nodes.copy()).compute_time_dev()
    assert max(max_rel_err(got, orc.time_dev(m, nodes, xm, ym))) < 1e-11
    dt = orc.cfl_dt(m, nodes, *p.spacing(), 0.1)[0]
    ref = nodes.copy()
    assert orc.run(m, ref, xm, ym, np.full(30, dt)) == 0
    f = d.Field(p, nodes.copy())
    f.run(30, dt=dt)
    got = f.download()
    err = max_rel_err(got, ref)
    print(case, shape, "mixed CG 30-step rel err", err)
    assert np.isfinite(got).all()
    assert err[0] < 1e-6 and err[3] < 1e-6, err
    g64 = d.Field(dataclasses.replace(p, precision=0), nodes.copy())
    g64.run(30, dt=dt)
    e = max_rel_err(got, g64.download())
    g64.close()
    print(case, shape, "mixed vs FP64 CG", e)
    assert e[0] < 1e-6 and e[3] < 1e-6, e
    for _ in range(10):
        dtc = orc.cfl_dt(m, ref, *p.spacing(), 0.1)[0]
        assert orc.run(m, ref, xm, ym, np.array([dtc])) == 0
    f.run(10, cfl=0.1)
    err = max_rel_err(f.download(), ref)
    print(case, shape, "mixed CG CFL rel err", err)
    assert err[0] < 1e-6 and err[3] < 1e-6, err
    f.close()
