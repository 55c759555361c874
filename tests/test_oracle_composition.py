"""Pins of the oracle's step composition: the order of the ghost fills within a
step and across the three RK3-TVD stages (td_integrator_class.f:103-161,
bc_operators_gen_class.f:238-381), checked by equivalences that the
mathematics fixes rather than by restating the composition:

- reflection on a box is the periodic problem on the box mirrored across its
  edges (the normal momentum negated in each mirror image), so an
  all-reflection run equals a quadrant of an all-periodic run on the doubled
  domain, ghost cells and corners included;
- walls in y with reflection in x (the reference's wall_S_reflection choice,
  and walls on S and N) equal walls in y with periodic x on the x-mirrored domain:
  this pins the periodic-x / wall-y composition, which the reference has no
  dispatcher for, against one it has;
- walls on W/E are the transposes of walls on S/N (momenta swapped);
- periodic edges commute with a shift of the field;
- RK3-TVD is exact for a state whose derivative is polynomial in t.

A test then compiles copies of the oracle with one planted composition
mistake each (-DORC_PLANT=n, dim2d_oracle.c) and requires every copy to fail
at least one of these checks.  The time-step rule, which has no counterpart
in the reference, is pinned against the thermodynamic sound speed of the van
der Waals fluid.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import dim2d_finite_volume_amd as d
from dim2d_finite_volume_amd import initial_conditions as ic

REFL, PER, WALL = 0, 1, 2
N, NY, STEPS = 16, 12, 8


def _field(orc, bc, n=N, ny=NY, flow=(0.1, -0.07), seed=3, angle=90.0, q=0.0):
    """A drop off-centre with flow and noise (no symmetry to hide a slip), with
    its ghost cells filled for edges `bc` = (W, E, S, N)."""
    p = d.ModelParams(nx=n + 4, ny=ny + 4, bc=bc)
    dx, _ = p.spacing()
    nodes = ic.circles_field(p, 0.95, [(0.1, -0.15)], 0.2, "liquid", 2 * math.sqrt(p.we) * dx,
                             flow=flow, noise=1e-3, seed=seed)
    m = orc.make_model(n + 4, ny + 4, bc=bc, wall_contact_angle=angle, wall_heat_flux=q)
    xm, ym = orc.coordinates(n + 4, ny + 4, -0.5, 0.5, -0.5, 0.5)
    orc.apply_bc(m, nodes, xm, ym)
    return m, xm, ym, nodes


def _maps(xm, ym, nx, ny):
    dx, dy = xm[1] - xm[0], ym[1] - ym[0]
    return xm[0] + dx * np.arange(nx), ym[0] + dy * np.arange(ny)


def _mirror_x(f):
    g = f[:, :, ::-1].copy()
    g[1] = -g[1]
    return np.concatenate([f, g], axis=2)


def _mirror_y(f):
    g = f[:, ::-1, :].copy()
    g[2] = -g[2]
    return np.concatenate([f, g], axis=1)


def _transpose(f):
    """(rho, qx, qy, E)(x, y) -> (rho, qy, qx, E)(y, x)"""
    return np.ascontiguousarray(f[[0, 2, 1, 3]].transpose(0, 2, 1))


def _rel(a, b, ks=range(4)):
    return max(float(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max()) for k in ks)


def _dt(orc, m, nodes, xm, ym):
    return orc.cfl_dt(m, nodes, xm[1] - xm[0], ym[1] - ym[0], 0.1)[0]


# ---- the checks: each returns (error, tolerance) ----

def check_reflection_is_mirrored_periodic(orc):
    m, xm, ym, a = _field(orc, (REFL,) * 4)
    dt = _dt(orc, m, a, xm, ym)
    big = np.zeros((4, 2 * NY + 4, 2 * N + 4))
    big[:, 2:-2, 2:-2] = _mirror_y(_mirror_x(a[:, 2:-2, 2:-2]))
    mb = orc.make_model(2 * N + 4, 2 * NY + 4, bc=(PER,) * 4)
    xb, yb = _maps(xm, ym, 2 * N + 4, 2 * NY + 4)
    orc.apply_bc(mb, big, xb, yb)
    if orc.run(m, a, xm, ym, np.full(STEPS, dt)) or orc.run(mb, big, xb, yb, np.full(STEPS, dt)):
        return math.inf, 0.0
    # the whole reflection array, ghosts and corners included, is the window of
    # the periodic one that starts two cells before its first interior cell
    # (the periodic ghost columns 0, 1 hold the E-mirror's last columns, which
    # are the mirror of the reflection field's first two interior columns)
    return _rel(a, big[:, :NY + 4, :N + 4]), 1e-12


def _wall_pair(orc, angle, q, north):
    bc = (REFL, REFL, WALL, north)
    m, xm, ym, a = _field(orc, bc, flow=(0.1, 0.0), seed=4, angle=angle, q=q)
    dt = _dt(orc, m, a, xm, ym)
    big = np.zeros((4, NY + 4, 2 * N + 4))
    big[:, :, 2:-2] = _mirror_x(a[:, :, 2:-2])
    mb = orc.make_model(2 * N + 4, NY + 4, bc=(PER, PER, WALL, north), wall_contact_angle=angle,
                        wall_heat_flux=q)
    xb, _ = _maps(xm, ym, 2 * N + 4, NY + 4)
    orc.apply_bc(mb, big, xb, ym)
    if orc.run(m, a, xm, ym, np.full(STEPS, dt)) or orc.run(mb, big, xb, ym, np.full(STEPS, dt)):
        return None
    return a, big


def check_wall_y_reflection_x_is_mirrored_periodic_x(orc):
    """A wall on S with reflection on W, E and N (the reference's
    wall_S_reflection_choice, library_config.py:102-115), and walls on S and N:
    interior rows, and the ghost rows' density and momenta (the wall fill
    never writes the ghost energy), at 70 degrees with heat flux.  The wall
    roots converge to xacc = 1e-12, so the match is to ~1e-13."""
    err = 0.0
    for north in (REFL, WALL):
        r = _wall_pair(orc, 70.0, 0.01, north)
        if r is None:
            return math.inf, 0.0
        a, big = r
        err = max(err, _rel(a[:, 2:-2, 2:-2], big[:, 2:-2, 2:N + 2]))
        err = max(err, _rel(a[:3, :, 2:-2], big[:3, :, 2:N + 2], ks=range(3)))
        # the corner ghost cells: reflection of the wall ghosts vs periodic
        # images (a wall's outer ghost row holds its density and debug momenta,
        # which a reflection negates and a periodic copy does not: density only)
        err = max(err, _rel(a[:3, 1:-1, :2], big[:3, 1:-1, :2], ks=range(3)))
        err = max(err, _rel(a[:1, :, :2], big[:1, :, :2], ks=range(1)))
    return err, 1e-11


def check_walls_we_are_transposed_walls_sn(orc):
    err = 0.0
    for xbc in (REFL, PER):
        m, xm, ym, a = _field(orc, (xbc, xbc, WALL, WALL), n=NY, ny=N, seed=5, angle=60.0, q=-0.01)
        dt = _dt(orc, m, a, xm, ym)
        t = _transpose(a)
        mt = orc.make_model(N + 4, NY + 4, bc=(WALL, WALL, xbc, xbc), wall_contact_angle=60.0,
                            wall_heat_flux=-0.01)
        xt, yt = ym.copy(), xm.copy()
        orc.apply_bc(mt, t, xt, yt)
        if orc.run(m, a, xm, ym, np.full(STEPS, dt)) or orc.run(mt, t, xt, yt, np.full(STEPS, dt)):
            return math.inf, 0.0
        at = _transpose(a)
        err = max(err, _rel(t[:, 2:-2, 2:-2], at[:, 2:-2, 2:-2]))
        err = max(err, _rel(t[:3, 2:-2, :], at[:3, 2:-2, :], ks=range(3)))
    return err, 1e-11


def _shift(f, sx, sy):
    """The field shifted by sx columns (and sy rows) across its periodic edges,
    ghost cells included, so that the shifted field is filled as the original
    was (a wall's root guess reads the ghost cells of u^n)."""
    g = f.copy()
    g[:, :, 2:-2] = np.roll(f[:, :, 2:-2], sx, axis=2)
    g[:, :, :2], g[:, :, -2:] = g[:, :, -4:-2], g[:, :, 2:4]
    if sy:
        g[:, 2:-2, :] = np.roll(g[:, 2:-2, :], sy, axis=1)
        g[:, :2, :], g[:, -2:, :] = g[:, -4:-2, :], g[:, 2:4, :]
    return g


def _shift_check(orc, bc, sx, sy):
    m, xm, ym, a = _field(orc, bc, seed=6)
    dt = _dt(orc, m, a, xm, ym)
    b = _shift(a, sx, sy)
    if orc.run(m, a, xm, ym, np.full(STEPS, dt)) or orc.run(m, b, xm, ym, np.full(STEPS, dt)):
        return math.inf
    return _rel(_shift(a, sx, sy)[:, 2:-2, 2:-2], b[:, 2:-2, 2:-2])


def check_periodic_shift_invariance(orc):
    """Every cell evaluates the same expressions on the same neighbourhood
    values after a periodic shift: bitwise."""
    return _shift_check(orc, (PER,) * 4, 5, 3), 0.0


def check_periodic_x_wall_y_shift_invariance(orc):
    return _shift_check(orc, (PER, PER, WALL, WALL), 5, 0), 0.0


def check_rk3_exact_for_gravity(orc):
    """rho v linear and E quadratic in t under gravity: RK3-TVD is exact."""
    g, rho, E0, dt, n = 2.0, 1.1, 5.0, 0.01, 7
    m = orc.make_model(14, 14, bc=(PER,) * 4, gravity=g)
    xm, ym = orc.coordinates(14, 14, -0.5, 0.5, -0.5, 0.5)
    nodes = np.empty((4, 14, 14))
    nodes[0], nodes[1], nodes[2], nodes[3] = rho, 0.0, 0.0, E0
    if orc.run(m, nodes, xm, ym, np.full(n, dt)):
        return math.inf, 0.0
    t = n * dt
    err = max(float(np.abs(nodes[2, 2:-2, 2:-2] / (-rho * g * t) - 1).max()),
              float(np.abs(nodes[3, 2:-2, 2:-2] / (E0 + 0.5 * rho * g * g * t * t) - 1).max()))
    return err, 1e-13


def check_reflection_conservation(orc):
    m, xm, ym, a = _field(orc, (REFL,) * 4)
    s0 = orc.sums(m, a)
    if orc.run(m, a, xm, ym, np.full(STEPS, _dt(orc, m, a, xm, ym))):
        return math.inf, 0.0
    s1 = orc.sums(m, a)
    return max(abs(s1[0] - s0[0]) / s0[0], abs(s1[3] - s0[3]) / abs(s0[3])), 1e-13


CHECKS = [check_reflection_is_mirrored_periodic, check_wall_y_reflection_x_is_mirrored_periodic_x,
          check_walls_we_are_transposed_walls_sn, check_periodic_shift_invariance,
          check_periodic_x_wall_y_shift_invariance, check_rk3_exact_for_gravity,
          check_reflection_conservation]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__[6:] for c in CHECKS])
def test_composition(orc, check):
    err, tol = check(orc)
    assert err <= tol, (check.__name__, err, tol)


PLANTS = {
    1: "no ghost fill after stage 1",
    2: "stage 2 blends the stage-1 state instead of u^n",
    3: "stage-3 weight 1/4 instead of 1/3",
    4: "periodic x / walls y: wall rows before the x-periodic fill",
    5: "periodic x / walls y: ghost rows not x-periodic",
    6: "reflection: corners before the W/E edges",
    7: "stages 2 and 3 differentiate u^n",
    8: "walls W/E, periodic y: ghost columns not y-periodic",
    9: "periodic x and y: the y fill over the interior columns only",
}


@pytest.fixture(scope="module")
def planted_libs(orc, tmp_path_factory):
    """The oracle compiled once per planted mistake (in parallel)."""
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(here, "oracle", "dim2d_oracle.c")
    out = tmp_path_factory.mktemp("planted")
    procs, libs = [], {}
    for n in PLANTS:
        lib = str(out / f"liborc_plant{n}.so")
        libs[n] = lib
        procs.append(subprocess.Popen(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-std=c11",
                                       "-shared", f"-DORC_PLANT={n}", "-o", lib, src, "-lm"]))
    for pr in procs:
        assert pr.wait() == 0
    return libs


@pytest.mark.parametrize("plant", sorted(PLANTS), ids=[PLANTS[k] for k in sorted(PLANTS)])
def test_planted_composition_mistake_fails_a_check(orc, planted_libs, plant):
    failed = []
    with orc.using(planted_libs[plant]):
        for check in CHECKS:
            with np.errstate(all="ignore"):
                err, tol = check(orc)
            if not err <= tol:
                failed.append(check.__name__)
    assert failed, f"planted mistake {plant} ({PLANTS[plant]}) passes every composition check"


# ---- the time-step rule (no counterpart in the reference) ----

def test_cfl_rule_against_thermodynamic_sound_speed(orc):
    """dt = cfl * min(h / max(|u| + c), Re min(rho) h^2, sqrt(We / max(rho)) h^2)
    on uniform states, with c^2 = (dP/drho)_T + T (dP/dT)_rho^2 / (rho^2 de/dT),
    the isentropic sound speed of any fluid, from central differences of the
    reduced van der Waals P(rho, T) = 8 rho T/(3 - rho) - 3 rho^2 and the
    specific internal energy e(rho, T) = 8 cv T/3 - 3 rho.  States on both
    sides of the spinodal-free region, with each limit of the rule binding."""
    cv = 2.5

    def P(r, T):
        return 8 * r * T / (3 - r) - 3 * r * r

    def e(r, T):
        return 8 * cv * T / 3 - 3 * r

    bound = set()
    for (rho, T, u, v, re, we, L) in [(0.3, 0.95, 0.4, -0.2, 5.0, 10.0, 1.0),
                                      (1.7, 1.1, -0.1, 0.05, 5.0, 10.0, 1.0),
                                      (0.8, 1.4, 0.7, 0.3, 1e4, 1e4, 40.0),
                                      (0.5, 2.0, 0.0, 0.0, 1e6, 0.5, 10.0),
                                      (1.2, 0.99, 0.2, 0.2, 200.0, 1e5, 8.0)]:
        h = 1e-4 * max(rho, T)
        dPdr = (P(rho + h, T) - P(rho - h, T)) / (2 * h)
        dPdT = (P(rho, T + h) - P(rho, T - h)) / (2 * h)
        dedT = (e(rho, T + h) - e(rho, T - h)) / (2 * h)
        c = math.sqrt(dPdr + T * dPdT ** 2 / (rho ** 2 * dedT))
        n = 10
        m = orc.make_model(n + 4, n + 4, re=re, we=we, cv_r=cv)
        xm, ym = orc.coordinates(n + 4, n + 4, -L / 2, L / 2, -L / 2, L / 2)
        dx = xm[1] - xm[0]
        nodes = np.empty((4, n + 4, n + 4))
        nodes[0], nodes[1], nodes[2] = rho, rho * u, rho * v
        nodes[3] = rho * (8 * cv * T / 3 - 3 * rho) + 0.5 * rho * (u * u + v * v)
        limits = [dx / (math.hypot(u, v) + c), re * rho * dx * dx, math.sqrt(we / rho) * dx * dx]
        bound.add(int(np.argmin(limits)))
        got = orc.cfl_dt(m, nodes, dx, dx, 0.3)[0]
        assert math.isclose(got, 0.3 * min(limits), rel_tol=1e-8), (rho, T, got, limits)
    assert bound == {0, 1, 2}
