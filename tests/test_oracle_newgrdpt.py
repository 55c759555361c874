"""Pins of the oracle's new grid points of the buffer layers (no GPU): the
characteristic extrapolation of bf_layer_newgrdpt/bf_newgrdpt_class.f
(compute_newgrdpt_x/_y/_xy, the corner and edge procedures of
compute_newgrdpt_local, the interpolation and Newton-Cotes helpers) and the
one-layer growth of a rectangular grid built on it (orc_extend_newgrdpt).

Three kinds of checks:
- the known answers of the reference's own test,
  test_files/test_stymphalianbirds/test_bf_newgrdpt_prim.f: the helpers
  (:156-275 and the 2D cases at its end), compute_newgrdpt_x and its mirrored
  left case (:276-506), compute_newgrdpt_y and its mirror (:509-720),
  compute_newgrdpt_xy and its mirror (:722-925), for both eigen-quantity
  strategies, with the DIM far field that test requires (T0 = 0.95 with the
  low-temperature saturation fit, flow 0.1 in x, cv_r = 2.5);
- closed forms of the scheme: a state equal to the far field is reproduced;
  a density profile linear in x and y advected by a uniform subsonic flow
  leaves through an outflow edge exactly (the new point is the profile at
  (x - u dt, y - v dt)), and enters through an inflow edge as the far-field
  density less half a step of transverse advection (the trapezoid of
  compute_NewtonCotes_integration with no transverse term in the far field);
- every new point of a rectangle grown at all four edges at once follows
  those closed forms, whatever the flow direction, so the configuration ->
  procedure table and every procedure's indices are exercised; points that
  a wrong stencil would read are NaN.
"""
import math

import numpy as np
import pytest

CV = 2.5


def far_field_of_test():
    """get_far_field of bubble_transported at T0 = 0.95, u = 0.1, with the
    low-temperature liquid density 1 + 2.08 sqrt(1 - T) (dim2d_state_eq_module.f)."""
    rho = 1.0 + 2.08 * math.sqrt(0.05)
    u = 0.1
    return np.array([rho, rho * u, 0.0,
                     0.5 * rho * u * u + rho * (8.0 / 3.0 * CV * 0.95 - 3.0 * rho)])


def cons_from_prim(rho, u, v, P):
    """Conservative variables of the reduced van der Waals fluid:
    T = (P + 3 rho^2)(3 - rho) / (8 rho), E = rho (u^2+v^2)/2 + rho (8/3 cv T - 3 rho)."""
    T = (P + 3.0 * rho ** 2) * (3.0 - rho) / (8.0 * rho)
    return np.array([rho, rho * u, rho * v,
                     0.5 * rho * (u * u + v * v) + rho * (8.0 / 3.0 * CV * T - 3.0 * rho)])


def _arr(lit, shape):
    return np.array(lit, dtype=float).reshape(shape)


# bf_nodes0 and bf_nodes1 of test_sym_compute_newgrdpt_x / _xy (:322-365), in
# the memory order of nodes(i,j,k): shape (k, j, i)
NODES0 = _arr([1.48, 1.30, 1.35, 1.26, 1.45, 1.4, 1.46, 1.27, 1.47,
               0.128, 0.127, 0.142, 1.138, 0.148, 0.132, 0.146, 0.143, 0.145,
               0.0050, 0.020, 0.060, 0.0025, 0.001, 0.015, 0.0100, 0.002, 0.050,
               4.88, 4.870, 4.855, 4.85, 4.865, 4.845, 4.89, 4.870, 4.860], (4, 3, 3))
NODES1 = _arr([1.50, 1.455, 1.48, 0, 1.20, 1.350, 1.25, 0, 1.49, 1.250, 1.40, 0, 0, 0, 0, 0,
               0.128, 0.450, 0.135, 0, 0.148, 0.150, 0.122, 0, 0.142, 1.152, 0.236, 0,
               0, 0, 0, 0,
               0.006, 0.0600, 0.020, 0, 0.000, 0.0028, 0.035, 0, 0.020, 0.0030, 0.040, 0,
               0, 0, 0, 0,
               4.876, 4.825, 4.862, 0, 4.890, 4.871, 4.892, 0, 4.865, 4.757, 4.895, 0,
               0, 0, 0, 0], (4, 4, 4))

# expected new points: {strategy: values}, :381-392, :476-490, :741-752, :869-880
WANT_X = {"bc": [1.22383078395524, 0.39531842478603, -0.21050217290879, 4.19684181018688],
          "lin": [1.21167555521982, 0.35901827468671, -0.20475732388332, 4.20002914561351]}
WANT_X_SYM = {"bc": [1.279162121036500, 0.285275187690879, -0.220019311093461,
                     4.48274912783283],
              "lin": [1.63776226955242, 0.318825439971472, -8.980834765273495e-002,
                      2.97904706359555]}
WANT_XY = {"bc": [1.55742091189301, 0.56224898725725, -0.14504765609975, 4.63647450016455],
           "lin": [1.42742921947925, 0.74488540650052, -0.03274712450770, 4.27121569817512]}
TOL = 1e-10  # is_real_validated, check_data_module.f:37

STRATEGIES = [("bc", False), ("lin", True)]


@pytest.fixture(scope="module")
def m(orc):
    return orc.make_model(3, 3, cv_r=CV)


def test_far_field_is_the_tests(orc):
    ff = far_field_of_test()
    want = [1.46510213931996, 0.146510214, 0.0, 2.84673289046992]  # :82-86
    assert np.all(np.abs(ff - want) < TOL)


def test_interpolation_and_newton_cotes_known_answers(orc):
    c = orc.interp_coeff_1d([1.0, 2.0], _arr([1, 2, .5, -.5, .25, -.75, .25, -.75], (4, 2)).T)
    assert np.allclose(c, _arr([1, 0, -1, 1.5, -1, 1.25, -1, 1.25], (4, 2)).T, atol=TOL, rtol=0)
    v = orc.interpolate_1d(1.0, _arr([1, 0, -1, 1.5, -1, 1.25, -1, 1.25], (4, 2)).T)
    assert np.allclose(v, [1.0, 0.5, 0.25, 0.25], atol=TOL, rtol=0)
    nc = orc.newton_cotes([1.0, 1.25, 0.5, 0.5], [2.0, -0.25, 1.5, 1.5], 0.5)
    assert np.allclose(nc, [0.75, 0.25, 0.5, 0.5], atol=TOL, rtol=0)
    c2 = orc.interp_coeff_2d([1.0, 1.0, 2.0], [1.0, 2.0, 2.0],
                             _arr([1, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1], (4, 3)).T)
    assert np.allclose(c2, _arr([0, 0, 1, -1, 0, 2, 0, 1, -1, 0, 1, -1], (4, 3)).T, atol=TOL,
                       rtol=0)
    v2 = orc.interpolate_2d(0.25, 0.5, _arr([0, 0, 1, -1, 0, 1, 0, -1, 1, 0, -1, 1], (4, 3)).T)
    assert np.allclose(v2, [1.0, 0.75, 0.5, 0.5], atol=TOL, rtol=0)


def _x_grids(orc):
    g0 = orc.bf_grid(NODES0, [0.5, 1.5, 2.5], [0.0, 0.25, 0.5], (0, 0))
    g1 = orc.bf_grid(NODES1, [0.5, 1.5, 2.5, 3.5], [0.0, 0.25, 0.5, 0.75], (0, 0))
    return g0, g1


def _x_sym_grids(orc):
    """The data of :412-459: the x data mirrored (x -> -x, momentum_x
    negated).  The far field keeps its flow in +x, so the mirrored case is a
    second known answer rather than a mirror image of the first."""
    pf = np.array([1, -1, 1, 1])[:, None, None]
    n0 = np.ascontiguousarray(NODES0[:, :, ::-1] * pf)
    n1 = np.zeros((4, 4, 4))
    n1[:, :3, 1:] = NODES1[:, :3, :3][:, :, ::-1] * pf
    g0 = orc.bf_grid(n0, [-2.5, -1.5, -0.5], [0.0, 0.25, 0.5], (1, 0))
    g1 = orc.bf_grid(n1, [-3.5, -2.5, -1.5, -0.5], [0.0, 0.25, 0.5, 0.75], (0, 0))
    return g0, g1


def transpose_grid_data(n):
    """x <-> y with the momenta swapped (the data of test_sym_compute_newgrdpt_y)."""
    return np.ascontiguousarray(np.transpose(n, (0, 2, 1))[[0, 2, 1, 3]])


@pytest.mark.parametrize("name,lin", STRATEGIES)
def test_newgrdpt_x_known_answers(orc, m, name, lin):
    ff = far_field_of_test()
    g0, g1 = _x_grids(orc)
    got = orc.newgrdpt_x(m, ff, 0.25, g0, g1, 4, 2, 1, orc.GRAD_I, eigen_lin=lin)
    assert np.all(np.abs(got - WANT_X[name]) < TOL), got
    g0, g1 = _x_sym_grids(orc)
    sym = orc.newgrdpt_x(m, ff, 0.25, g0, g1, 1, 2, 0, orc.GRAD_I, eigen_lin=lin)
    sym[1] = -sym[1]
    assert np.all(np.abs(sym - WANT_X_SYM[name]) < TOL), sym


@pytest.mark.parametrize("name,lin", STRATEGIES)
def test_newgrdpt_y_known_answers(orc, m, name, lin):
    """test_sym_compute_newgrdpt_y (:509-720): the x data transposed; the
    expected values are the x ones with the momenta swapped.  The reference's
    main program does not run this case; its values hold with the far field
    transposed too (flow 0.1 in y), which is how it is checked here."""
    ff = far_field_of_test()[[0, 2, 1, 3]]
    g0 = orc.bf_grid(transpose_grid_data(NODES0), [0.0, 0.25, 0.5], [0.5, 1.5, 2.5], (0, 0))
    g1 = orc.bf_grid(transpose_grid_data(NODES1), [0.0, 0.25, 0.5, 0.75],
                     [0.5, 1.5, 2.5, 3.5], (0, 0))
    got = orc.newgrdpt_y(m, ff, 0.25, g0, g1, 2, 4, 1, orc.GRAD_I, eigen_lin=lin)
    want = np.array(WANT_X[name])[[0, 2, 1, 3]]
    assert np.all(np.abs(got - want) < TOL), got
    s0, s1 = _x_sym_grids(orc)
    g0 = orc.bf_grid(transpose_grid_data(s0.nodes), [0.0, 0.25, 0.5], [-2.5, -1.5, -0.5], (0, 1))
    g1 = orc.bf_grid(transpose_grid_data(s1.nodes), [0.0, 0.25, 0.5, 0.75],
                     [-3.5, -2.5, -1.5, -0.5], (0, 0))
    sym = orc.newgrdpt_y(m, ff, 0.25, g0, g1, 2, 1, 0, orc.GRAD_I, eigen_lin=lin)
    sym[2] = -sym[2]
    want = np.array(WANT_X_SYM[name])[[0, 2, 1, 3]]
    assert np.all(np.abs(sym - want) < TOL), sym


@pytest.mark.parametrize("name,lin", STRATEGIES)
def test_newgrdpt_xy_known_answers(orc, m, name, lin):
    ff = far_field_of_test()
    I, L0, R0 = orc.GRAD_I, orc.GRAD_L0, orc.GRAD_R0
    g0, g1 = _x_grids(orc)
    got = orc.newgrdpt_xy(m, ff, 0.25, g0, g1, 4, 4, orc.N2_DIRECTION, 1,
                          [R0, I, I, R0, R0, R0], [3, 3], [(3, 2), (2, 3), (3, 3)],
                          eigen_lin=lin)
    assert np.all(np.abs(got - WANT_XY[name]) < TOL), got
    # the y-mirrored data (:768-868) through the SE-corner form gives the same
    # point with momentum_y negated
    n0 = NODES0[:, ::-1, :] * np.array([1, 1, -1, 1])[:, None, None]
    n1 = np.zeros((4, 4, 4))
    n1[:, 1:, :] = NODES1[:, :3, :][:, ::-1, :] * np.array([1, 1, -1, 1])[:, None, None]
    g0 = orc.bf_grid(np.ascontiguousarray(n0), [0.5, 1.5, 2.5], [-0.5, -0.25, 0.0], (0, 1))
    g1 = orc.bf_grid(n1, [0.5, 1.5, 2.5, 3.5], [-0.75, -0.5, -0.25, 0.0], (0, 0))
    sym = orc.newgrdpt_xy(m, ff, 0.25, g0, g1, 4, 1, orc.N1_DIRECTION, 1,
                          [I, L0, R0, L0, R0, I], [3, 2], [(2, 2), (3, 2), (3, 3)],
                          eigen_lin=lin)
    sym[2] = -sym[2]
    assert np.all(np.abs(got - sym) < TOL), (got, sym)


# ---------------------------------------------------------------- closed forms
RHO0, P0 = 1.45, 8.0 * 1.45 * 0.95 / (3.0 - 1.45) - 3.0 * 1.45 ** 2  # liquid at T = 0.95
A_X, B_Y = 0.05, -0.03  # density slopes
DT = 0.02
RHO_FF = 1.30


def advected_state(x_map, y_map, u, v, shift):
    """rho = RHO0 + A_X (x - u shift) + B_Y (y - v shift), uniform u, v, P0;
    returned as nodes (4, ny, nx)."""
    X, Y = np.meshgrid(x_map, y_map)
    rho = RHO0 + A_X * (X - u * shift) + B_Y * (Y - v * shift)
    return cons_from_prim(rho, u, v, P0)


def expected_new_point(kind, x, y, u, v):
    """The closed form of a new point at (x, y), time t, for the procedure
    kind ('E','W','N','S','NE','NW','SE','SW')."""
    r = 1.0 / math.sqrt(2.0)
    un1, un2 = r * (u - v), r * (u + v)
    inflow = {"E": u < 0, "W": u > 0, "N": v < 0, "S": v > 0,
              "NE": un2 < 0, "SW": un2 > 0, "SE": un1 < 0, "NW": un1 > 0}[kind]
    if not inflow:
        rho = RHO0 + A_X * (x - u * DT) + B_Y * (y - v * DT)
    elif kind in ("E", "W"):
        rho = RHO_FF - 0.5 * DT * v * B_Y
    elif kind in ("N", "S"):
        rho = RHO_FF - 0.5 * DT * u * A_X
    elif kind in ("NE", "SW"):   # n2 direction: transverse u_n1 d/dn1
        rho = RHO_FF - 0.5 * DT * un1 * r * (A_X - B_Y)
    else:                        # n1 direction: transverse u_n2 d/dn2
        rho = RHO_FF - 0.5 * DT * un2 * r * (A_X + B_Y)
    return cons_from_prim(rho, u, v, P0)


# the last four put an eigenvalue at exactly 0 on some edge or corner: u = 0
# (E, W), v = 0 (N, S), u_n1 = 0 (SE, NW), u_n2 = 0 (NE, SW); there the
# characteristic comes from inside on both sides (lambda >= 0 right,
# lambda <= 0 left)
FLOWS = [(0.1, 0.05), (-0.1, 0.05), (0.05, -0.1), (-0.07, -0.12),
         (0.0, 0.06), (0.07, 0.0), (0.1, 0.1), (0.08, -0.08)]


@pytest.mark.parametrize("u,v", FLOWS)
def test_far_field_state_is_reproduced(orc, u, v):
    """A uniform state equal to the far field stays itself at every new point
    of a rectangle grown at all four edges."""
    nx, ny = 9, 8
    m = orc.make_model(nx, ny, cv_r=CV)
    x_map, y_map = orc.coordinates(nx, ny, -0.3, 0.2, -0.1, 0.3)
    st = cons_from_prim(np.full((ny, nx), RHO0), u, v, P0)
    ff = cons_from_prim(RHO0, u, v, P0)
    out = orc.extend_newgrdpt(m, ff, DT, st, st, x_map, y_map, 1, 1, 1, 1)
    assert np.isfinite(out).all()
    want = cons_from_prim(np.full(out.shape[1:], RHO0), u, v, P0)
    assert np.abs(out - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("u,v", FLOWS)
def test_grown_rectangle_follows_the_closed_forms(orc, u, v):
    """All four edges grown by one layer: outflow edges carry the advected
    profile out exactly, inflow edges bring the far-field density less half a
    step of transverse advection; the velocity and pressure stay uniform."""
    nx, ny = 10, 9
    m = orc.make_model(nx, ny, cv_r=CV)
    x_map, y_map = orc.coordinates(nx, ny, -0.35, 0.25, -0.2, 0.34)
    dx, dy = x_map[1] - x_map[0], y_map[1] - y_map[0]
    n0 = advected_state(x_map, y_map, u, v, 0.0)
    n1 = advected_state(x_map, y_map, u, v, DT)
    ff = cons_from_prim(RHO_FF, u, v, P0)
    out = orc.extend_newgrdpt(m, ff, DT, n0, n1, x_map, y_map, 1, 1, 1, 1)
    assert np.array_equal(out[:, 1:-1, 1:-1], n1)
    xs = np.concatenate([[x_map[0] - dx], x_map, [x_map[-1] + dx]])
    ys = np.concatenate([[y_map[0] - dy], y_map, [y_map[-1] + dy]])
    NX, NY = nx + 2, ny + 2
    for j in range(NY):
        for i in range(NX):
            if 0 < i < NX - 1 and 0 < j < NY - 1:
                continue
            kind = ("S" if j == 0 else "N" if j == NY - 1 else "") + \
                   ("W" if i == 0 else "E" if i == NX - 1 else "")
            want = expected_new_point(kind, xs[i], ys[j], u, v)
            got = out[:, j, i]
            assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), (kind, i, j, got, want)


def _poisoned_patch(orc, u, v, dom):
    """A 9x9 grid whose points outside the old domain dom = (i0, i1, j0, j1)
    (1-based, inclusive) are NaN: a procedure that reads them fails."""
    x_map, y_map = orc.coordinates(9, 9, -0.2, 0.3, -0.15, 0.35)
    n0 = advected_state(x_map, y_map, u, v, 0.0)
    n1 = advected_state(x_map, y_map, u, v, DT)
    i0, i1, j0, j1 = dom
    mask = np.zeros((9, 9), bool)
    mask[j0 - 1:j1, i0 - 1:i1] = True
    n0[:, ~mask] = np.nan
    n1[:, ~mask] = np.nan
    return x_map, y_map, orc.bf_grid(n0, x_map, y_map), orc.bf_grid(n1, x_map, y_map), mask


# procedure, flow, old domain, new point, expected configuration id
PROC_CASES = [
    ("E", "I", (0.1, 0.05), (2, 7, 2, 8), (8, 5), 41),
    ("E", "L0", (0.1, 0.05), (2, 7, 2, 8), (8, 2), 40),
    ("E", "R0", (0.1, 0.05), (2, 7, 2, 8), (8, 8), 9),
    ("W", "I", (-0.1, 0.05), (3, 8, 2, 8), (2, 5), 148),
    ("W", "L0", (-0.1, 0.05), (3, 8, 2, 8), (2, 2), 144),
    ("W", "R0", (-0.1, 0.05), (3, 8, 2, 8), (2, 8), 20),
    ("N", "I", (0.05, 0.1), (2, 8, 2, 7), (5, 8), 7),
    ("N", "L0", (0.05, 0.1), (2, 8, 2, 7), (2, 8), 6),
    ("N", "R0", (0.05, 0.1), (2, 8, 2, 7), (8, 8), 3),
    ("S", "I", (0.05, -0.1), (2, 8, 3, 8), (5, 2), 224),
    ("S", "L0", (0.05, -0.1), (2, 8, 3, 8), (2, 2), 192),
    ("S", "R0", (0.05, -0.1), (2, 8, 3, 8), (8, 2), 96),
    ("NE", "", (0.1, 0.05), (2, 7, 2, 7), (8, 8), 1),
    ("NW", "", (-0.1, 0.05), (3, 8, 2, 7), (2, 8), 4),
    ("SE", "", (0.1, 0.05), (2, 7, 3, 8), (8, 2), 32),
    ("SW", "", (-0.1, -0.05), (3, 8, 3, 8), (2, 2), 128),
]


@pytest.mark.parametrize("kind,grad,flow,dom,pt,config", PROC_CASES)
def test_each_procedure_reads_only_old_points(orc, kind, grad, flow, dom, pt, config):
    """Each edge and corner procedure, chosen from the configuration of old
    points around the new point, gives the outflow closed form and reads no
    point outside the old domain (those are NaN)."""
    u, v = flow
    x_map, y_map, g0, g1, mask = _poisoned_patch(orc, u, v, dom)
    bits, k = 0, 0
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            if di == 0 and dj == 0:
                continue
            if mask[pt[1] - 1 + dj, pt[0] - 1 + di]:
                bits |= 1 << k
            k += 1
    assert bits == config
    names = {"N": orc.NG_N_EDGE, "S": orc.NG_S_EDGE, "E": orc.NG_E_EDGE, "W": orc.NG_W_EDGE,
             "NE": orc.NG_NE_CORNER, "NW": orc.NG_NW_CORNER, "SE": orc.NG_SE_CORNER,
             "SW": orc.NG_SW_CORNER}
    grads = {"I": orc.GRAD_I, "L0": orc.GRAD_L0, "R0": orc.GRAD_R0, "": orc.GRAD_I}
    assert orc.newgrdpt_config_procedure(config) == (names[kind], grads[grad])
    m = orc.make_model(9, 9, cv_r=CV)
    ff = cons_from_prim(RHO_FF, u, v, P0)
    got = orc.newgrdpt_local(m, ff, DT, g0, g1, pt[0], pt[1], names[kind], grads[grad])
    want = expected_new_point(kind, x_map[pt[0] - 1], y_map[pt[1] - 1], u, v)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), (got, want)


def _mirror_x(n):
    return np.ascontiguousarray(n[:, :, ::-1] * np.array([1, -1, 1, 1])[:, None, None])


def _mirror_y(n):
    return np.ascontiguousarray(n[:, ::-1, :] * np.array([1, 1, -1, 1])[:, None, None])


@pytest.mark.parametrize("name,lin", STRATEGIES)
def test_procedures_reproduce_the_known_answers(orc, m, name, lin):
    """compute_newgrdpt_local's procedures on the data of the known answers:
    E edge (i1 = 4) and W edge (the mirrored case) as compute_newgrdpt_x, N
    and S edges on the transposed data, the NE corner on the xy data and the
    SE corner on its y-mirror (the argument sets of :722-925 are exactly
    those procedures'); NW and SW are the x-mirrors of NE and SE, with the
    far field mirrored as well."""
    ff = far_field_of_test()
    I = orc.GRAD_I
    g0, g1 = _x_grids(orc)
    got = orc.newgrdpt_local(m, ff, 0.25, g0, g1, 4, 2, orc.NG_E_EDGE, I, eigen_lin=lin)
    assert np.all(np.abs(got - WANT_X[name]) < TOL)
    ne = orc.newgrdpt_local(m, ff, 0.25, g0, g1, 4, 4, orc.NG_NE_CORNER, eigen_lin=lin)
    assert np.all(np.abs(ne - WANT_XY[name]) < TOL)
    s0, s1 = _x_sym_grids(orc)
    got = orc.newgrdpt_local(m, ff, 0.25, s0, s1, 1, 2, orc.NG_W_EDGE, I, eigen_lin=lin)
    assert np.all(np.abs(got * [1, -1, 1, 1] - WANT_X_SYM[name]) < TOL)
    ffT = ff[[0, 2, 1, 3]]
    g0 = orc.bf_grid(transpose_grid_data(NODES0), [0.0, 0.25, 0.5], [0.5, 1.5, 2.5], (0, 0))
    g1 = orc.bf_grid(transpose_grid_data(NODES1), [0.0, 0.25, 0.5, 0.75],
                     [0.5, 1.5, 2.5, 3.5], (0, 0))
    got = orc.newgrdpt_local(m, ffT, 0.25, g0, g1, 2, 4, orc.NG_N_EDGE, I, eigen_lin=lin)
    assert np.all(np.abs(got - np.array(WANT_X[name])[[0, 2, 1, 3]]) < TOL)
    g0 = orc.bf_grid(transpose_grid_data(s0.nodes), [0.0, 0.25, 0.5], [-2.5, -1.5, -0.5], (0, 1))
    g1 = orc.bf_grid(transpose_grid_data(s1.nodes), [0.0, 0.25, 0.5, 0.75],
                     [-3.5, -2.5, -1.5, -0.5], (0, 0))
    got = orc.newgrdpt_local(m, ffT, 0.25, g0, g1, 2, 1, orc.NG_S_EDGE, I, eigen_lin=lin)
    assert np.all(np.abs(got * [1, 1, -1, 1] - np.array(WANT_X_SYM[name])[[0, 2, 1, 3]]) < TOL)
    # SE corner: the xy data mirrored in y (:768-868)
    n1 = np.zeros((4, 4, 4))
    n1[:, 1:, :] = _mirror_y(NODES1[:, :3, :])
    h0 = orc.bf_grid(_mirror_y(NODES0), [0.5, 1.5, 2.5], [-0.5, -0.25, 0.0], (0, 1))
    h1 = orc.bf_grid(n1, [0.5, 1.5, 2.5, 3.5], [-0.75, -0.5, -0.25, 0.0], (0, 0))
    se = orc.newgrdpt_local(m, ff, 0.25, h0, h1, 4, 1, orc.NG_SE_CORNER, eigen_lin=lin)
    assert np.all(np.abs(se * [1, 1, -1, 1] - WANT_XY[name]) < TOL)
    # NW and SW: x-mirrors of NE and SE, far field mirrored (u -> -u)
    ffm = ff * [1, -1, 1, 1]
    g0, g1 = _x_grids(orc)
    k0 = orc.bf_grid(_mirror_x(g0.nodes), [-2.5, -1.5, -0.5], [0.0, 0.25, 0.5], (1, 0))
    k1 = orc.bf_grid(_mirror_x(g1.nodes), [-3.5, -2.5, -1.5, -0.5], [0.0, 0.25, 0.5, 0.75],
                     (0, 0))
    nw = orc.newgrdpt_local(m, ffm, 0.25, k0, k1, 1, 4, orc.NG_NW_CORNER, eigen_lin=lin)
    assert np.abs(nw * [1, -1, 1, 1] - ne).max() < 1e-12 * np.abs(ne).max()
    k0 = orc.bf_grid(_mirror_x(h0.nodes), [-2.5, -1.5, -0.5], [-0.5, -0.25, 0.0], (1, 1))
    k1 = orc.bf_grid(_mirror_x(h1.nodes), [-3.5, -2.5, -1.5, -0.5], [-0.75, -0.5, -0.25, 0.0],
                     (0, 0))
    sw = orc.newgrdpt_local(m, ffm, 0.25, k0, k1, 1, 1, orc.NG_SW_CORNER, eigen_lin=lin)
    assert np.abs(sw * [1, -1, 1, 1] - se).max() < 1e-12 * np.abs(se).max()


@pytest.mark.parametrize("u,v", FLOWS[:4])
def test_gravity_accelerates_every_new_point(orc, u, v):
    """With gravity g (the primitive body force -g on v), a uniform state
    equal to the far field gives new points whose velocity has gained -g dt
    in y, density and pressure unchanged, on every edge and corner."""
    nx, ny = 9, 8
    g = 0.7
    m = orc.make_model(nx, ny, cv_r=CV, gravity=g)
    x_map, y_map = orc.coordinates(nx, ny, -0.3, 0.2, -0.1, 0.3)
    st = cons_from_prim(np.full((ny, nx), RHO0), u, v, P0)
    ff = cons_from_prim(RHO0, u, v, P0)
    out = orc.extend_newgrdpt(m, ff, DT, st, st, x_map, y_map, 1, 1, 1, 1)
    want = cons_from_prim(RHO0, u, v - g * DT, P0)
    ring = np.ones(out.shape[1:], bool)
    ring[1:-1, 1:-1] = False
    got = out[:, ring]
    assert np.abs(got - want[:, None]).max() < 1e-12 * np.abs(want).max()


def test_unexpected_configuration_is_rejected(orc):
    assert orc.newgrdpt_config_procedure(0) is None
    assert orc.newgrdpt_config_procedure(255) is None
