"""Pins of the oracle's wall boundary (no GPU): the wall face flux
(orc_wall_flux, compute_wall_flux_x/_y of wall_xy_equilibrium_module.f:249-506)
and the wall ghost-cell fill (apply_bc_on_nodes of bc_operators_wall_xy_class.f
:232-330 with wall_xy_equilibrium_module.f:948-1437).

Two kinds of checks:
- the symmetry relations of the reference's own test
  (test_bc_operators/test_bc_operators_wall_xy.f): the W face flux of the
  x-mirrored field is the negated, prefactored E flux (:66-125), the S flux of
  the transposed field is the W flux with the momenta swapped (:127-149), the
  N flux of the y-mirrored field the negated, prefactored S flux (:151-166);
  the same mirror and transpose relations for the ghost cells (:246-420);
- closed forms a sign, factor or index slip in the wall terms fails: a fluid
  at rest between 90-degree walls has no time derivative; a wall heat flux q
  changes the total energy at rate q per unit wall length; a no-slip wall
  decelerates a uniform stream at the viscous rate 2 U / (Re dy) per wall;
  the mass and energy budgets hold over whole RK3 steps.
"""
import numpy as np
import pytest

R, P, W = 0, 1, 2
PF_X = np.array([1.0, -1.0, 1.0, 1.0])  # reflect_x_nodes prefactors (:535)
PF_Y = np.array([1.0, 1.0, -1.0, 1.0])  # reflect_y_nodes prefactors (:566)


def reflect_x(a):
    return np.ascontiguousarray(a[:, :, ::-1] * PF_X[:, None, None])


def reflect_y(a):
    return np.ascontiguousarray(a[:, ::-1, :] * PF_Y[:, None, None])


def transpose(a):
    """transpose_nodes (:582-598): x <-> y, momentum_x <-> momentum_y."""
    t = np.transpose(a, (0, 2, 1))
    return np.ascontiguousarray(t[[0, 2, 1, 3]])


def random_state(n, seed, rho=(0.3, 1.6), T=(0.9, 1.0), u=0.15, cv=2.5):
    """A random field (ghosts included) with positive temperatures: E from
    rho, the velocities and T (T_cl = (E - rho |u|^2/2 + 3 rho^2)/rho = 8 cv T / 3)."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(*rho, size=(n, n))
    ux = rng.uniform(-u, u, size=(n, n))
    uy = rng.uniform(-u, u, size=(n, n))
    t = rng.uniform(*T, size=(n, n))
    E = r * (8.0 * cv / 3.0 * t) - 3.0 * r * r + 0.5 * r * (ux * ux + uy * uy)
    return np.ascontiguousarray(np.stack([r, r * ux, r * uy, E]))


def wall_flux(orc, m, nodes, dir_y, i, j, dx, dy, right):
    import ctypes
    out = np.zeros(4)
    orc.lib().orc_wall_flux(ctypes.byref(m), orc._ptr(nodes), dir_y, i, j, dx, dy, right,
                            orc._ptr(out))
    return out


@pytest.mark.parametrize("q", [0.0, 0.013])
def test_wall_flux_mirror_and_transpose_relations(orc, q):
    """test_bc_operators_wall_xy.f:66-166 on a 10x10 field with dx = 0.5,
    dy = 0.8 (a dx/dy swap fails it): E, W (x-mirror), S (transpose), N
    (y-mirror of the transpose), faces of the interior rows/columns 3..8."""
    n, dx, dy = 10, 0.5, 0.8
    m = orc.make_model(n, n, wall_heat_flux=q)
    nodes = random_state(n, seed=11)
    E = np.array([wall_flux(orc, m, nodes, 0, n - 1, j, dx, dy, 1) for j in range(3, n - 1)])
    mx = reflect_x(nodes)
    Wf = np.array([wall_flux(orc, m, mx, 0, 3, j, dx, dy, 0) for j in range(3, n - 1)])
    scale = np.abs(E).max(axis=0) + 1e-300
    assert np.all(np.abs(Wf + PF_X * E) <= 1e-13 * scale + 1e-300), (Wf, E)
    # the transposed problem: y <- x with the spacings swapped (:127-136)
    t = transpose(mx)
    S = np.array([wall_flux(orc, m, t, 1, i, 3, dy, dx, 0) for i in range(3, n - 1)])
    assert np.all(np.abs(S - Wf[:, [0, 2, 1, 3]]) <= 1e-13 * scale[[0, 2, 1, 3]] + 1e-300)
    ty = reflect_y(t)
    N = np.array([wall_flux(orc, m, ty, 1, i, n - 1, dy, dx, 1) for i in range(3, n - 1)])
    assert np.all(np.abs(N + PF_Y * S) <= 1e-13 * scale[[0, 2, 1, 3]] + 1e-300)
    # no mass crosses a wall; the energy flux is the heat flux, into the domain
    assert np.all(E[:, 0] == 0.0) and np.all(Wf[:, 0] == 0.0)
    assert np.all(Wf[:, 3] == q) and np.all(E[:, 3] == -q)
    # the normal momentum carries the wall pressure: nonzero on this field
    assert np.all(np.abs(E[:, 1]) > 0.0)


def _fill(orc, bc, nodes, dx, dy, angle=90.0, q=0.0):
    n = nodes.shape[1]
    m = orc.make_model(n, n, bc=bc, wall_contact_angle=angle, wall_heat_flux=q)
    xm = dx * np.arange(n, dtype=float)
    ym = dy * np.arange(n, dtype=float)
    out = nodes.copy()
    orc.apply_bc(m, out, xm, ym, nodes_tmp=nodes.copy())
    return out


@pytest.mark.parametrize("angle,q", [(90.0, 0.0), (60.0, 0.0), (120.0, 0.01)])
def test_wall_ghost_mirror_and_transpose_relations(orc, angle, q):
    """test_bc_operators_wall_xy.f:246-420 with both walls of an axis at
    once: walls W and E with reflecting S and N edges (config X), walls S and
    N with reflecting W and E edges (config Y).  The x-mirror of the field
    mirrors the W/E ghosts, the transpose (spacings swapped) gives the S/N
    ghosts of config Y with the momenta swapped, and its y-mirror mirrors them.
    Compared: the inner ghost cell's density and momenta, the outer ghost
    cell's density (the reference writes no other ghost values)."""
    n, dx, dy = 10, 0.5, 0.8
    base = random_state(n, seed=3, rho=(0.5, 1.5), T=(0.85, 0.95))
    # reflecting edges first, so that the fill order of the two configs does
    # not matter for the transverse neighbours of the first and last rows
    nodes = _fill(orc, (R, R, R, R), base, dx, dy)
    X, Y = (W, W, R, R), (R, R, W, W)
    gx = _fill(orc, X, nodes, dx, dy, angle, q)
    assert np.isfinite(gx[:3, 2:-2, [0, 1, n - 2, n - 1]]).all()
    rows = slice(2, n - 2)

    def close(a, b, what):
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        assert err < 1e-12, (what, err)

    # E ghosts of the field against W ghosts of its mirror
    gxm = _fill(orc, X, reflect_x(nodes), dx, dy, angle, q)
    ref = reflect_x(gx)
    for k in range(3):
        close(gxm[k, rows, 1], ref[k, rows, 1], f"W inner k={k}")
        close(gxm[k, rows, n - 2], ref[k, rows, n - 2], f"E inner k={k}")
    close(gxm[0, rows, 0], ref[0, rows, 0], "W outer")
    close(gxm[0, rows, n - 1], ref[0, rows, n - 1], "E outer")
    # the transposed field with walls on S and N
    gy = _fill(orc, Y, transpose(nodes), dy, dx, angle, q)
    ref = transpose(gx)
    for k in range(3):
        close(gy[k, 1, rows], ref[k, 1, rows], f"S inner k={k}")
        close(gy[k, n - 2, rows], ref[k, n - 2, rows], f"N inner k={k}")
    close(gy[0, 0, rows], ref[0, 0, rows], "S outer")
    # its y-mirror
    gym = _fill(orc, Y, reflect_y(transpose(nodes)), dy, dx, angle, q)
    ref = reflect_y(gy)
    for k in range(3):
        close(gym[k, 1, rows], ref[k, 1, rows], f"S mirror k={k}")
    # at 90 degrees with no heat flux the ghost density equals the first
    # interior density (the equilibrium condition reduces to a zero normal
    # gradient), and the outer ghost density follows the extrapolation
    if angle == 90.0 and q == 0.0:
        assert np.abs(gx[0, rows, 1] - gx[0, rows, 2]).max() < 1e-12
        assert np.abs(gx[0, rows, 0] - gx[0, rows, 3]).max() < 1e-11
    else:
        assert np.abs(gx[0, rows, 1] - gx[0, rows, 2]).max() > 1e-6
    # the outer ghost density (wall_dmddx_high_order, :1439-1465) lies on the
    # parabola through the ghost, first and far interior densities, i.e. the
    # four densities along the wall normal have a zero third difference
    for j in range(2, n - 2):
        for cols in ((0, 1, 2, 3), (n - 1, n - 2, n - 3, n - 4)):
            pts = gx[0, j, list(cols[1:])]
            fit = np.polyfit([0.0, 1.0, 2.0], pts, 2)
            assert abs(np.polyval(fit, -1.0) - gx[0, j, cols[0]]) < 1e-12


def _coords(n, lo=-0.5, hi=0.5):
    from oracle import oracle as orc_mod
    return orc_mod.coordinates(n, n, lo, hi, lo, hi)


def _uniform(n, rho, T, u=0.0, v=0.0, cv=2.5):
    E = rho * (8.0 * cv / 3.0 * T) - 3.0 * rho * rho + 0.5 * rho * (u * u + v * v)
    a = np.empty((4, n, n))
    a[0], a[1], a[2], a[3] = rho, rho * u, rho * v, E
    return a


WALL_SETS = {
    "periodic_x_wall_y": (P, P, W, W),
    "wall_s_reflection": (R, R, W, R),
    "wall_x_reflection_y": (W, W, R, R),
    "wall_x_periodic_y": (W, W, P, P),
    "wall_e_reflection": (R, W, R, R),
}


@pytest.mark.parametrize("case", list(WALL_SETS))
def test_rest_state_between_walls_has_no_time_derivative(orc, case):
    """A uniform fluid at rest, 90-degree walls, no heat flux: every interior
    cell's time derivative vanishes to rounding.  At the wall face the normal
    momentum flux is the pressure P(T, rho) of the wall formula (8 rho T /
    (3 - rho) - 3 rho^2 with T = 3/(8 cv) T_eff), which must equal the
    interior classical pressure; a wrong factor there leaves a residual of
    order P/dy in the first interior row."""
    n = 20
    m = orc.make_model(n, n, bc=WALL_SETS[case])
    xm, ym = _coords(n)
    nodes = _uniform(n, 1.2, 0.95)
    orc.apply_bc(m, nodes, xm, ym)
    td = orc.time_dev(m, nodes, xm, ym)
    P = 8 * 1.2 * 0.95 / (3 - 1.2) - 3 * 1.2 ** 2
    dx = xm[1] - xm[0]
    assert np.abs(td[:, 2:-2, 2:-2]).max() < 1e-13 * P / dx


@pytest.mark.parametrize("case,walls", [("wall_s_reflection", 1), ("periodic_x_wall_y", 2),
                                        ("wall_e_reflection", 1), ("wall_x_periodic_y", 2)])
def test_wall_heat_flux_energy_budget(orc, case, walls):
    """With a wall heat flux q and no gravity, reflecting and periodic faces
    carry no energy and every wall injects q per unit length, whatever the
    state: d/dt sum(E) dx dy = walls * q * L exactly up to rounding; the mass
    is conserved.  A flipped sign of the S/N or W/E heat flux, or a flux
    scaled by 1/2 or by dx, fails it."""
    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic
    q = 0.01
    n = 24
    bc = WALL_SETS[case]
    p = d.ModelParams(nx=n + 4, ny=n + 4, bc=bc, wall_heat_flux=q)
    dx, dy = p.spacing()
    nodes = ic.circles_field(p, 0.95, [(0.1, -0.05)], 0.2, "liquid", 2 * np.sqrt(p.we) * dx,
                             flow=(0.05, 0.02), noise=1e-4, seed=2)
    m = orc.make_model(p.nx, p.ny, bc=bc, wall_heat_flux=q)
    xm, ym = p.coordinates()
    orc.apply_bc(m, nodes, xm, ym)
    td = orc.time_dev(m, nodes, xm, ym)[:, 2:-2, 2:-2]
    L = n * (dx if bc[2] == W else dy)
    rate = td[3].sum() * dx * dy
    mag = np.abs(td[3]).sum() * dx * dy
    assert abs(rate - walls * q * L) < 1e-13 * mag, (rate, walls * q * L)
    assert abs(td[0].sum() * dx * dy) < 1e-13 * np.abs(td[0]).sum() * dx * dy


def test_no_slip_wall_drag(orc):
    """A uniform stream u = U between no-slip walls (periodic x, walls S and
    N, 90 degrees, no heat flux): the ghost velocity is -U, so each wall face
    carries the viscous stress -(1/Re) 2U/dy against the stream and the total
    x-momentum changes at d/dt sum(rho u) dx dy = -4 U L / (Re dy); the normal
    momentum and the energy budgets vanish (the wall does no work in the
    reference's wall flux)."""
    n, U = 24, 0.3
    m = orc.make_model(n, n, bc=(P, P, W, W))
    xm, ym = _coords(n)
    dx, dy = xm[1] - xm[0], ym[1] - ym[0]
    nodes = _uniform(n, 1.3, 0.93, u=U)
    orc.apply_bc(m, nodes, xm, ym)
    # the no-slip ghost: tangential velocity reversed, normal velocity kept
    assert np.allclose(nodes[1, 1, 2:-2] / nodes[0, 1, 2:-2], -U, rtol=1e-14)
    td = orc.time_dev(m, nodes, xm, ym)[:, 2:-2, 2:-2]
    L = (n - 4) * dx
    rate = td[1].sum() * dx * dy
    exact = -4.0 * U * L / (m.re * dy)
    assert abs(rate - exact) < 1e-12 * abs(exact), (rate, exact)
    mag = np.abs(td).sum(axis=(1, 2)) * dx * dy
    assert abs(td[2].sum() * dx * dy) < 1e-13 * mag[2]
    assert abs(td[3].sum() * dx * dy) < 1e-13 * max(mag[3], 1.0)
    # the drag acts in the rows next to the walls only
    assert np.abs(td[1, 2:-2, :]).max() < 1e-12 * abs(td[1, 0, :]).max()


def test_budgets_over_rk3_steps_periodic_x_wall_y(orc):
    """Over several RK3 steps of periodic x + walls y with heat flux q: the
    mass stays constant and sum(E) dx dy grows by 2 q L per unit time (every
    stage's energy rate is 2 q L, and the Shu-Osher weights integrate a
    constant rate exactly)."""
    import dim2d_finite_volume_amd as d
    from dim2d_finite_volume_amd import initial_conditions as ic
    q, n = 0.01, 24
    p = d.ModelParams(nx=n + 4, ny=n + 4, bc=(P, P, W, W), wall_heat_flux=q,
                      wall_contact_angle=70.0)
    dx, dy = p.spacing()
    nodes = ic.circles_field(p, 0.95, [(0.0, -0.2)], 0.2, "liquid", 2 * np.sqrt(p.we) * dx,
                             flow=(0.1, 0.0), noise=1e-4, seed=9)
    m = orc.make_model(p.nx, p.ny, bc=p.bc, wall_heat_flux=q, wall_contact_angle=70.0)
    xm, ym = p.coordinates()
    orc.apply_bc(m, nodes, xm, ym)
    s0 = orc.sums(m, nodes)
    dt = orc.cfl_dt(m, nodes, dx, dy, 0.1)[0]
    steps = 12
    assert orc.run(m, nodes, xm, ym, np.full(steps, dt)) == 0
    s1 = orc.sums(m, nodes)
    L = n * dx
    dE = (s1[3] - s0[3]) * dx * dy
    assert abs(dE - 2 * q * L * steps * dt) < 1e-9 * abs(2 * q * L * steps * dt), (dE,)
    assert abs(s1[0] - s0[0]) < 1e-13 * s0[0]
