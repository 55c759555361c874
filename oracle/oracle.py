"""ctypes binding of the CPU oracle (oracle/dim2d_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker.  Arrays follow the reference's Fortran
layout nodes(nx, ny, ne); in numpy they are held as C arrays of shape
(4, ny, nx) (variable, j, i), which is the same memory.
"""
from __future__ import annotations

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


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
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
        _lib = L
    return _lib


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


def extend(nodes, w: int, e: int, s: int, n: int):
    """The grid grown by w/e columns and s/n rows, every new point repeating
    the nearest old point (the domain extension's copy rule)."""
    return np.ascontiguousarray(np.pad(nodes, ((0, 0), (s, n), (w, e)), mode="edge"))


def sums(m: Model, nodes):
    out = np.zeros(4)
    lib().orc_sums(ctypes.byref(m), _ptr(nodes), _ptr(out))
    return out
