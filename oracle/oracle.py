"""ctypes binding of the CPU oracle (oracle/dim2d_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker.  Arrays follow the reference's Fortran
layout nodes(nx, ny, ne); in numpy they are held as C arrays of shape
(4, ny, nx) (variable, j, i), which is the same memory.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libdim2d_oracle.so")

BC_REFLECTION, BC_PERIODIC, BC_WALL, BC_OPEN = 0, 1, 2, 3
INT_RK3TVD, INT_LSRK3 = 0, 1
OPS_MATTSSON, OPS_COCKBURN_GAU = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)


class Model(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("re", ctypes.c_double), ("we", ctypes.c_double), ("pr", ctypes.c_double),
        ("viscous_r", ctypes.c_double), ("cv_r", ctypes.c_double),
        ("gravity", ctypes.c_double), ("gravity_ac", ctypes.c_int),
        ("bc_w", ctypes.c_int), ("bc_e", ctypes.c_int), ("bc_s", ctypes.c_int),
        ("bc_n", ctypes.c_int),
        ("wall_contact_angle", ctypes.c_double), ("wall_heat_flux", ctypes.c_double),
        ("integrator", ctypes.c_int), ("operators", ctypes.c_int),
    ]


class BfGrid(ctypes.Structure):
    """orc_bf_grid: a grid handed to the new-point procedures."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("align_x", ctypes.c_int), ("align_y", ctypes.c_int),
        ("x_map", _dp), ("y_map", _dp), ("nodes", _dp),
    ]


GRAD_I, GRAD_L0, GRAD_R0 = 0, 1, 2
N1_DIRECTION, N2_DIRECTION = 1, 2
(NG_N_EDGE, NG_S_EDGE, NG_E_EDGE, NG_W_EDGE,
 NG_NE_CORNER, NG_NW_CORNER, NG_SE_CORNER, NG_SW_CORNER) = range(8)


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(path: str):
    """Route this module's calls to another build of the oracle at `path` (the
    copies with a planted mistake that the composition tests compile)."""
    global _lib
    saved = lib()
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved


def _load(path: str):
    L = ctypes.CDLL(path)
    M = ctypes.POINTER(Model)
    L.orc_flux_term.restype = ctypes.c_double
    L.orc_flux_term.argtypes = [M, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double]
    L.orc_mattsson_density_stencil.restype = ctypes.c_double
    L.orc_mattsson_density_stencil.argtypes = L.orc_flux_term.argtypes
    L.orc_cg_density_stencil.restype = ctypes.c_double
    L.orc_cg_density_stencil.argtypes = L.orc_flux_term.argtypes
    L.orc_compute_time_dev.argtypes = [M, ctypes.c_double, _dp, _dp, _dp, _dp]
    L.orc_apply_bc_on_nodes.argtypes = [M, ctypes.c_double, _dp, _dp, _dp, _dp]
    L.orc_rk3_step.argtypes = [M, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
    L.orc_lsrk3_step.argtypes = L.orc_rk3_step.argtypes
    L.orc_classical_pressure.restype = ctypes.c_double
    L.orc_classical_pressure.argtypes = [M, _dp, ctypes.c_int, ctypes.c_int]
    L.orc_openbc_timedev.argtypes = [M, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_void_p, _dp]
    L.orc_lsrk3_amplification.restype = ctypes.c_double
    L.orc_lsrk3_amplification.argtypes = [ctypes.c_double]
    L.orc_run.restype = ctypes.c_int
    L.orc_run.argtypes = [M, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.orc_coordinates.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [_dp, _dp]
    L.orc_cfl_dt.argtypes = [M, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
    L.orc_sums.argtypes = [M, _dp, _dp]
    L.orc_openbc_undermined_point.restype = ctypes.c_int
    L.orc_openbc_undermined_point.argtypes = [M, _dp]
    L.orc_openbc_undermined_at.restype = ctypes.c_int
    L.orc_openbc_undermined_at.argtypes = [M, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double]
    L.orc_openbc_undermined.argtypes = [M, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    for name in ("orc_mass_density_liquid", "orc_mass_density_vapor"):
        getattr(L, name).restype = ctypes.c_double
        getattr(L, name).argtypes = [ctypes.c_double]
    L.orc_surface_tension.restype = ctypes.c_double
    L.orc_surface_tension.argtypes = [M, ctypes.c_double]
    L.orc_speed_of_sound_local.restype = ctypes.c_double
    L.orc_speed_of_sound_local.argtypes = [M, _dp]
    L.orc_dwall_internal_energy_dmd.restype = ctypes.c_double
    L.orc_dwall_internal_energy_dmd.argtypes = [M] + [ctypes.c_double] * 3
    L.orc_wall_temperature.restype = ctypes.c_double
    L.orc_wall_temperature.argtypes = [M] + [ctypes.c_double] * 5
    L.orc_wall_root_f.restype = ctypes.c_double
    L.orc_wall_root_f.argtypes = [M, ctypes.c_double, _dp, _dp] + [ctypes.c_double] * 5
    L.orc_wall_md_ghost.restype = ctypes.c_double
    L.orc_wall_md_ghost.argtypes = [M, ctypes.c_double, ctypes.c_double, _dp, _dp] + \
        [ctypes.c_double] * 5
    L.orc_wall_flux.argtypes = [M, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
    ROOTFN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double)
    L.ROOTFN = ROOTFN
    L.orc_root_brackets.restype = ctypes.c_int
    L.orc_root_brackets.argtypes = [ROOTFN, ctypes.c_void_p, ctypes.c_double,
                                    ctypes.c_double, _dp]
    L.orc_ridders.restype = ctypes.c_double
    L.orc_ridders.argtypes = [ROOTFN, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    G = ctypes.POINTER(BfGrid)
    _ip = ctypes.POINTER(ctypes.c_int)
    L.orc_interp_coeff_1d.argtypes = [_dp, _dp, _dp]
    L.orc_interpolate_1d.argtypes = [ctypes.c_double, _dp, _dp]
    L.orc_interp_coeff_2d.argtypes = [_dp, _dp, _dp, _dp]
    L.orc_interpolate_2d.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.orc_newton_cotes.argtypes = [_dp, _dp, ctypes.c_double, _dp]
    ng = [M, _dp, ctypes.c_int, ctypes.c_double, G, G, ctypes.c_int, ctypes.c_int]
    L.orc_newgrdpt_x.argtypes = ng + [ctypes.c_int, ctypes.c_int, _dp]
    L.orc_newgrdpt_y.argtypes = ng + [ctypes.c_int, ctypes.c_int, _dp]
    L.orc_newgrdpt_xy.argtypes = ng + [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp]
    L.orc_newgrdpt_local.argtypes = ng + [ctypes.c_int, ctypes.c_int, _dp]
    L.orc_newgrdpt_config_procedure.restype = ctypes.c_int
    L.orc_newgrdpt_config_procedure.argtypes = [ctypes.c_int, _ip, _ip]
    L.orc_extend_newgrdpt.restype = ctypes.c_int
    L.orc_extend_newgrdpt.argtypes = [M, _dp, ctypes.c_double, _dp, _dp, _dp, _dp, _ip, _dp]
    return L


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def make_model(nx, ny, re=5.0, we=10.0, pr=20.0, viscous_r=-1.5, cv_r=2.5, gravity=0.0,
               bc=(BC_REFLECTION,) * 4, wall_contact_angle=90.0, wall_heat_flux=0.0,
               integrator=INT_RK3TVD, operators=OPS_MATTSSON) -> Model:
    """bc = (W, E, S, N).  Default physics is the reference's test set
    (dim2d_parameters.f:182-187)."""
    return Model(nx, ny, re, we, pr, viscous_r, cv_r, gravity, int(gravity != 0.0),
                 bc[0], bc[1], bc[2], bc[3], wall_contact_angle, wall_heat_flux, integrator,
                 operators)


def flux_term(m: Model, nodes: np.ndarray, which: int, i: int, j: int, dx: float, dy: float):
    return lib().orc_flux_term(ctypes.byref(m), _ptr(nodes), which, i, j, dx, dy)


def stencil(m: Model, nodes: np.ndarray, which: int, i: int, j: int, dx: float, dy: float):
    return lib().orc_mattsson_density_stencil(ctypes.byref(m), _ptr(nodes), which, i, j, dx, dy)


def cg_density_stencil(m: Model, nodes: np.ndarray, which: int, i: int, j: int, dx: float,
                       dy: float):
    """Cockburn and Gau stencil `which` (cases as mattsson_density_stencil) of
    the mass density at (i, j), 1-based."""
    return lib().orc_cg_density_stencil(ctypes.byref(m), _ptr(nodes), which, i, j, dx, dy)


def coordinates(nx, ny, x_min, x_max, y_min, y_max):
    xm = np.zeros(nx)
    ym = np.zeros(ny)
    lib().orc_coordinates(nx, ny, x_min, x_max, y_min, y_max, _ptr(xm), _ptr(ym))
    return xm, ym


def time_dev(m: Model, nodes, x_map, y_map, t=0.0):
    out = np.empty_like(nodes)
    lib().orc_compute_time_dev(ctypes.byref(m), t, _ptr(nodes), _ptr(x_map), _ptr(y_map), _ptr(out))
    return out


def apply_bc(m: Model, nodes, x_map, y_map, nodes_tmp=None, t=0.0):
    if nodes_tmp is None:
        nodes_tmp = nodes.copy()
    lib().orc_apply_bc_on_nodes(ctypes.byref(m), t, _ptr(x_map), _ptr(y_map), _ptr(nodes_tmp),
                                _ptr(nodes))
    return nodes


def run(m: Model, nodes, x_map, y_map, dts, t0=0.0):
    """Advance `nodes` in place by len(dts) steps of the model's integrator;
    returns the failing step or 0."""
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    return lib().orc_run(ctypes.byref(m), t0, len(dts), _ptr(dts), _ptr(x_map), _ptr(y_map),
                         _ptr(nodes))


def openbc_timedev(m: Model, nodes, i, j, ydir, right, grad_type, dn, pe=None):
    """Characteristic time derivative of an open boundary point (1-based i, j);
    pe = [rho, ux, uy, P, c] given explicitly, or None for the point's own."""
    out = np.zeros(4)
    pe_arr = None if pe is None else np.ascontiguousarray(pe, dtype=np.float64)
    lib().orc_openbc_timedev(ctypes.byref(m), _ptr(nodes), i, j, ydir, right, grad_type, dn,
                             None if pe_arr is None else pe_arr.ctypes.data, _ptr(out))
    return out


def lsrk3_amplification(z):
    return lib().orc_lsrk3_amplification(float(z))


def cfl_dt(m: Model, nodes, dx, dy, cfl):
    out = np.zeros(4)
    lib().orc_cfl_dt(ctypes.byref(m), _ptr(nodes), dx, dy, cfl, _ptr(out))
    return out


def openbc_undermined_point(m: Model, v) -> bool:
    v = np.ascontiguousarray(v, dtype=np.float64)
    return bool(lib().orc_openbc_undermined_point(ctypes.byref(m), _ptr(v)))


def openbc_undermined_at(m: Model, nodes, i: int, j: int, dx, dy, md_threshold=-1.0) -> bool:
    """The criterion at interior point (i, j), 1-based."""
    return bool(lib().orc_openbc_undermined_at(ctypes.byref(m), _ptr(nodes), i, j, dx, dy,
                                               md_threshold))


def openbc_undermined(m: Model, nodes, dist: int, dx=1.0, dy=1.0, md_threshold=-1.0):
    """Undermined points within `dist` interior columns/rows of each edge (W, E,
    S, N); md_threshold < 0: the default criterion only."""
    out = (ctypes.c_int * 4)()
    lib().orc_openbc_undermined(ctypes.byref(m), _ptr(nodes), dist, dx, dy, md_threshold, out)
    return np.array(out[:], dtype=np.int32)


def decrease_bands(nx: int, ny: int, growth, dist: int):
    """The check band of each edge that has grown since creation (growth = the
    W, E, S, N points added) in the decrease of the domain extension: the
    reference's get_check_line_param (bf_decrease_module.f:252-330) around the
    edge's place at creation (align_W/E/S/N, parameters_bf_layer.f:45-49),
    the interior and buffer-layer parts (check_layer_interior :448-519,
    check_layer_bf :522-600) taken together and clipped to the points whose
    3x3 neighbourhood lies in the grid.  0-based inclusive (i0, i1, j0, j1)
    per edge, or None."""
    gw, ge, gs, gn = (int(g) for g in growth)
    aw, ae, as_, an = 1 + gw, nx - 2 - ge, 1 + gs, ny - 2 - gn
    d = int(dist)
    raw = [(aw - d + 1, aw + d, as_ - d, an + d - 1),       # W
           (ae - d, ae + d - 1, as_ - d, an + d - 1),       # E
           (aw - d + 1, ae + d - 1, as_ - d + 1, as_ + d),  # S
           (aw - d + 1, ae + d - 1, an - d, an + d - 1)]    # N
    out = []
    for k, (i0, i1, j0, j1) in enumerate(raw):
        i0, i1, j0, j1 = max(i0, 1), min(i1, nx - 2), max(j0, 1), min(j1, ny - 2)
        out.append((i0, i1, j0, j1) if int(growth[k]) > 0 and i0 <= i1 and j0 <= j1 else None)
    return out


def openbc_undermined_decrease(m: Model, nodes, growth, dist: int, dx=1.0, dy=1.0,
                               md_threshold=-1.0):
    """Undermined points (orc_openbc_undermined_at) in each decrease band."""
    counts = np.zeros(4, dtype=np.int32)
    for k, band in enumerate(decrease_bands(m.nx, m.ny, growth, dist)):
        if band is None:
            continue
        i0, i1, j0, j1 = band
        for j in range(j0, j1 + 1):
            for i in range(i0, i1 + 1):
                counts[k] += openbc_undermined_at(m, nodes, i + 1, j + 1, dx, dy, md_threshold)
    return counts


def extend(nodes, w: int, e: int, s: int, n: int):
    """The grid grown by w/e columns and s/n rows, every new point repeating
    the nearest old point (the domain extension's copy rule)."""
    return np.ascontiguousarray(np.pad(nodes, ((0, 0), (s, n), (w, e)), mode="edge"))


def sums(m: Model, nodes):
    out = np.zeros(4)
    lib().orc_sums(ctypes.byref(m), _ptr(nodes), _ptr(out))
    return out


def _ints(v):
    return (ctypes.c_int * len(v))(*[int(q) for q in v])


class _Grid:
    """Keeps the arrays of an orc_bf_grid alive.  nodes: (4, ny, nx)."""

    def __init__(self, nodes, x_map, y_map, align=(0, 0)):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.x_map = np.ascontiguousarray(x_map, dtype=np.float64)
        self.y_map = np.ascontiguousarray(y_map, dtype=np.float64)
        ne, ny, nx = self.nodes.shape
        assert ne == 4 and self.x_map.size == nx and self.y_map.size == ny
        self.c = BfGrid(nx, ny, int(align[0]), int(align[1]), _ptr(self.x_map),
                        _ptr(self.y_map), _ptr(self.nodes))


def bf_grid(nodes, x_map, y_map, align=(0, 0)):
    """A grid for the new-point functions: nodes (4, ny, nx), its maps and its
    alignment (bf_align(1,1), bf_align(2,1))."""
    return _Grid(nodes, x_map, y_map, align)


def interp_coeff_1d(x_map, nodes):
    """nodes (2 points, 4 variables) -> coefficients (2, 4): slope, intercept."""
    out = np.zeros((2, 4))
    lib().orc_interp_coeff_1d(_ptr(np.ascontiguousarray(x_map, dtype=np.float64)),
                              _ptr(np.ascontiguousarray(nodes, dtype=np.float64)), _ptr(out))
    return out


def interpolate_1d(x, coeff):
    out = np.zeros(4)
    lib().orc_interpolate_1d(float(x), _ptr(np.ascontiguousarray(coeff, dtype=np.float64)),
                             _ptr(out))
    return out


def interp_coeff_2d(x_map, y_map, nodes):
    out = np.zeros((3, 4))
    lib().orc_interp_coeff_2d(_ptr(np.ascontiguousarray(x_map, dtype=np.float64)),
                              _ptr(np.ascontiguousarray(y_map, dtype=np.float64)),
                              _ptr(np.ascontiguousarray(nodes, dtype=np.float64)), _ptr(out))
    return out


def interpolate_2d(x, y, coeff):
    out = np.zeros(4)
    lib().orc_interpolate_2d(float(x), float(y),
                             _ptr(np.ascontiguousarray(coeff, dtype=np.float64)), _ptr(out))
    return out


def newton_cotes(d0, d1, dt):
    out = np.zeros(4)
    lib().orc_newton_cotes(_ptr(np.ascontiguousarray(d0, dtype=np.float64)),
                           _ptr(np.ascontiguousarray(d1, dtype=np.float64)), float(dt), _ptr(out))
    return out


def _ff(far_field):
    return np.ascontiguousarray(far_field, dtype=np.float64)


def newgrdpt_x(m: Model, far_field, dt, g0, g1, i1, j1, right, grad, eigen_lin=False):
    out = np.zeros(4)
    ff = _ff(far_field)
    lib().orc_newgrdpt_x(ctypes.byref(m), _ptr(ff), int(eigen_lin), float(dt),
                         ctypes.byref(g0.c), ctypes.byref(g1.c), i1, j1, int(right), grad,
                         _ptr(out))
    return out


def newgrdpt_y(m: Model, far_field, dt, g0, g1, i1, j1, right, grad, eigen_lin=False):
    out = np.zeros(4)
    ff = _ff(far_field)
    lib().orc_newgrdpt_y(ctypes.byref(m), _ptr(ff), int(eigen_lin), float(dt),
                         ctypes.byref(g0.c), ctypes.byref(g1.c), i1, j1, int(right), grad,
                         _ptr(out))
    return out


def newgrdpt_xy(m: Model, far_field, dt, g0, g1, i1, j1, n_dir, right, grads, eigen_idx,
                inter_idx1, eigen_lin=False):
    """grads: 6 gradient procedures (x1, y1, x2, y2, x3, y3); eigen_idx (i, j);
    inter_idx1: three (i, j) pairs, all 1-based in grid 1."""
    out = np.zeros(4)
    ff = _ff(far_field)
    flat = [q for pair in inter_idx1 for q in pair]
    lib().orc_newgrdpt_xy(ctypes.byref(m), _ptr(ff), int(eigen_lin), float(dt),
                          ctypes.byref(g0.c), ctypes.byref(g1.c), i1, j1, n_dir, int(right),
                          _ints(grads), _ints(eigen_idx), _ints(flat), _ptr(out))
    return out


def newgrdpt_local(m: Model, far_field, dt, g0, g1, i1, j1, procedure, grad=GRAD_I,
                   eigen_lin=False):
    out = np.zeros(4)
    ff = _ff(far_field)
    lib().orc_newgrdpt_local(ctypes.byref(m), _ptr(ff), int(eigen_lin), float(dt),
                             ctypes.byref(g0.c), ctypes.byref(g1.c), i1, j1, procedure, grad,
                             _ptr(out))
    return out


def newgrdpt_config_procedure(config_id):
    """(procedure, gradient) of a configuration, or None."""
    p = ctypes.c_int()
    g = ctypes.c_int()
    if lib().orc_newgrdpt_config_procedure(config_id, ctypes.byref(p), ctypes.byref(g)):
        return None
    return p.value, g.value


def extend_newgrdpt(m: Model, far_field, dt, nodes0, nodes1, x_map, y_map, w, e, s, n):
    """The grid grown by one layer at the edges with w/e/s/n = 1; new points
    from the characteristic extrapolation (state nodes0 at t - dt, nodes1 at t)."""
    nodes0 = np.ascontiguousarray(nodes0, dtype=np.float64)
    nodes1 = np.ascontiguousarray(nodes1, dtype=np.float64)
    x_map = np.ascontiguousarray(x_map, dtype=np.float64)
    y_map = np.ascontiguousarray(y_map, dtype=np.float64)
    out = np.zeros((4, m.ny + s + n, m.nx + w + e))
    ff = _ff(far_field)
    st = lib().orc_extend_newgrdpt(ctypes.byref(m), _ptr(ff), float(dt), _ptr(nodes0),
                                   _ptr(nodes1), _ptr(x_map), _ptr(y_map), _ints([w, e, s, n]),
                                   _ptr(out))
    if st:
        raise ValueError("unexpected new-point configuration")
    return out
