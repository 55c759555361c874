"""Host-side mirror of the reference's field / td_operators / td_integrator /
bc_operators interface, running on the HIP library through its C ABI.

Reference interface mirrored here (jdesmarais/dim2D_finite_volume):
  field%integrate(dt)                       field/field_class.f:61-72
  td_integrator%integrate(field, dt)        td_integrator/rungekutta3rdtvd/td_integrator_class.f:103-161
  td_operators%compute_time_dev(...)        td_operators/finitevolume/td_operators_class.f:119-235
  bc_operators_gen%apply_bc_on_nodes(...)   boundary_conditions/bc_operators_gen_class.f:238-381
  field%ini coordinates and borders         field/field_abstract_class.f:244-315

Arrays are the reference's nodes(nx, ny, ne) in Fortran order, held in numpy
as C arrays of shape (4, ny, nx) -- the same bytes.  Errors raise
capi.Dim2dError (the reference prints and stops).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field as dc_field, replace

import numpy as np

from . import capi

BC_SIZE = 2


@dataclass
class ModelParams:
    """Compile-time parameters of the reference (parameters_input.f,
    dim2d_parameters.f) as run-time values.  Defaults are the reference's
    test physics (dim2d_parameters.f:182-187) on [-0.5, 0.5]^2."""
    nx: int
    ny: int
    bc: tuple = (capi.BC_REFLECTION,) * 4  # (W, E, S, N)
    precision: int = capi.PREC_FP64
    x_min: float = -0.5
    x_max: float = 0.5
    y_min: float = -0.5
    y_max: float = 0.5
    re: float = 5.0
    we: float = 10.0
    pr: float = 20.0
    viscous_r: float = -1.5
    cv_r: float = 2.5
    gravity: float = 0.0
    wall_contact_angle: float = 90.0
    wall_heat_flux: float = 0.0
    integrator: int = capi.INT_RK3TVD
    operators: int = capi.OPS_MATTSSON  # sd_choice: Mattsson or Cockburn-Gau face stencils

    def to_c(self) -> capi.Params:
        return capi.Params(
            self.nx, self.ny, BC_SIZE, self.bc[0], self.bc[1], self.bc[2], self.bc[3],
            self.precision, int(self.gravity != 0.0), self.x_min, self.x_max, self.y_min,
            self.y_max, self.re, self.we, self.pr, self.viscous_r, self.cv_r, self.gravity,
            self.wall_contact_angle, self.wall_heat_flux, self.integrator, self.operators)

    def coordinates(self):
        """x_map, y_map (field_abstract_class.f:286-315)."""
        dx = (self.x_max - self.x_min) / (self.nx - 1 - 2 * BC_SIZE)
        dy = (self.y_max - self.y_min) / (self.ny - 1 - 2 * BC_SIZE)
        x_map = np.array([self.x_min + (i - 1 - BC_SIZE) * dx for i in range(1, self.nx + 1)])
        y_map = np.array([self.y_min + (j - 1 - BC_SIZE) * dy for j in range(1, self.ny + 1)])
        return x_map, y_map

    def spacing(self):
        """dx, dy as td_operators_class.f:156-157 takes them."""
        x_map, y_map = self.coordinates()
        return x_map[1] - x_map[0], y_map[1] - y_map[0]


def _dptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Field:
    """One simulation resident in HBM (field_abstract + its operators)."""

    def __init__(self, params: ModelParams, nodes: np.ndarray | None = None, time: float = 0.0):
        self.params = params
        self.x_map, self.y_map = params.coordinates()
        self.time = time
        self._lib = capi.load()
        self._ctx = ctypes.c_void_p()
        cp = params.to_c()
        capi.check(self._lib.dim2d_ctx_create(ctypes.byref(cp), ctypes.byref(self._ctx)),
                   "dim2d_ctx_create")
        if nodes is not None:
            self.upload(nodes)

    # -- state transfer -------------------------------------------------
    def _check_shape(self, a: np.ndarray):
        shape = (4, self.params.ny, self.params.nx)
        if a.shape != shape:
            raise ValueError(f"nodes must have shape {shape} (nodes(nx,ny,ne) in Fortran order)")

    def upload(self, nodes: np.ndarray) -> None:
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self._check_shape(nodes)
        capi.check(self._lib.dim2d_upload(self._ctx, _dptr(nodes)), "dim2d_upload")

    def download(self) -> np.ndarray:
        out = np.empty((4, self.params.ny, self.params.nx))
        capi.check(self._lib.dim2d_download(self._ctx, _dptr(out)), "dim2d_download")
        return out

    @property
    def nodes(self) -> np.ndarray:
        return self.download()

    # -- operators ----------------------------------------------------------
    def compute_time_dev(self) -> np.ndarray:
        """td_operators%compute_time_dev of the current state."""
        out = np.empty((4, self.params.ny, self.params.nx))
        capi.check(self._lib.dim2d_time_dev(self._ctx, self.time, _dptr(out)), "dim2d_time_dev")
        return out

    def apply_bc_on_nodes(self) -> None:
        capi.check(self._lib.dim2d_apply_bc(self._ctx, self.time), "dim2d_apply_bc")

    def write_data(self, path: str, dt: float | None = None, t_max: float | None = None,
                   attrs: dict | None = None, ic=None, detail_print: float | None = None,
                   md_threshold: float | None = None) -> None:
        """field%write_data (field_abstract_class.f:597-610): the current state
        as one output file in the reference's NetCDF layout, with the header
        of nf90_write_header (io_netcdf; ic: an io_netcdf.InitialCondition)."""
        from . import io_netcdf
        io_netcdf.write_data(path, self.download(), self.x_map, self.y_map, self.time,
                             self.params, dt=dt, t_max=t_max, attrs=attrs, ic=ic,
                             detail_print=detail_print, md_threshold=md_threshold)

    @classmethod
    def restart(cls, params: ModelParams | None, path: str) -> "Field":
        """A field whose state and time come from an output file: the
        reference's restart (-r, io_operators_class.f:183-236).  The file's
        grid must be the one params describe; with params None they come
        from the file's header (io_netcdf.header_params)."""
        from . import io_netcdf
        nodes, x_map, y_map, time, attrs = io_netcdf.read_data(path)
        if params is None:
            params, _ = io_netcdf.header_params(attrs, nodes.shape[2], nodes.shape[1])
        if nodes.shape != (4, params.ny, params.nx):
            raise ValueError(f"{path}: grid {nodes.shape[2]}x{nodes.shape[1]} is not "
                             f"{params.nx}x{params.ny}")
        xm, ym = params.coordinates()
        h = min(params.spacing())
        if np.abs(xm - x_map).max() > 1e-9 * h or np.abs(ym - y_map).max() > 1e-9 * h:
            raise ValueError(f"{path}: coordinates differ from the parameters' domain")
        return cls(params, nodes, time=time)

    def check_steady_state(self, limit: float = 1e-12):
        """field%check_steady_state (field_abstract_class.f:629-667): returns
        (steady, max |time_dev| per variable over the cells with a derivative)."""
        out = np.zeros(4)
        steady = ctypes.c_int32()
        capi.check(self._lib.dim2d_check_steady_state(self._ctx, limit, _dptr(out),
                                                      ctypes.byref(steady)),
                   "dim2d_check_steady_state")
        return bool(steady.value), out

    # -- domain extension at open edges -------------------------------------
    EDGES = {"W": 1, "E": 2, "S": 4, "N": 8}

    def _refresh_params(self) -> None:
        q = capi.Params()
        capi.check(self._lib.dim2d_get_params(self._ctx, ctypes.byref(q)), "dim2d_get_params")
        self.params = replace(self.params, nx=q.nx, ny=q.ny, x_min=q.x_min, x_max=q.x_max,
                              y_min=q.y_min, y_max=q.y_max)
        self.x_map, self.y_map = self.params.coordinates()

    def openbc_undermined(self, distance: int = 5, md_threshold: float | None = None) -> np.ndarray:
        """Undermined points within `distance` points of each open edge (W, E,
        S, N): are_openbc_undermined, pmodel_eq_class.f:1991-2070; an imaginary
        sound speed, or with md_threshold (the reference's
        bf_openbc_md_threshold) a density inside the two-phase band."""
        out = (ctypes.c_int32 * 4)()
        thr = -1.0 if md_threshold is None else float(md_threshold)
        capi.check(self._lib.dim2d_openbc_undermined(self._ctx, distance, thr, out),
                   "dim2d_openbc_undermined")
        return np.array(out[:], dtype=np.int32)

    def set_far_field(self, far_field) -> None:
        """The conservative far-field state of the open edges (the initial
        condition's get_far_field), or None.  With it, extend() computes new
        points by the reference's characteristic extrapolation
        (bf_newgrdpt_class.f) from the last step, one layer per edge and call."""
        if far_field is None:
            capi.check(self._lib.dim2d_set_far_field(self._ctx, None), "dim2d_set_far_field")
            return
        ff = np.ascontiguousarray(far_field, dtype=np.float64)
        if ff.shape != (4,):
            raise ValueError("far_field must hold 4 conservative values")
        capi.check(self._lib.dim2d_set_far_field(self._ctx, ff.ctypes.data),
                   "dim2d_set_far_field")

    def extend(self, w: int = 0, e: int = 0, s: int = 0, n: int = 0) -> None:
        """Grow the grid at open edges by whole columns/rows (dx, dy kept);
        with a far field set, by one layer with extrapolated new points."""
        capi.check(self._lib.dim2d_extend(self._ctx, w, e, s, n), "dim2d_extend")
        self._refresh_params()

    def growth(self) -> np.ndarray:
        """Columns/rows added per edge (W, E, S, N) since creation."""
        out = (ctypes.c_int32 * 4)()
        capi.check(self._lib.dim2d_get_growth(self._ctx, out), "dim2d_get_growth")
        return np.array(out[:], dtype=np.int32)

    def shrink(self, w: int = 0, e: int = 0, s: int = 0, n: int = 0) -> None:
        """Remove columns/rows added earlier (at most growth() per edge)."""
        capi.check(self._lib.dim2d_shrink(self._ctx, w, e, s, n), "dim2d_shrink")
        self._refresh_params()

    def openbc_undermined_decrease(self, distance: int = 5,
                                   md_threshold: float | None = None) -> np.ndarray:
        """Undermined points in each grown edge's decrease check band
        (dim2d_openbc_undermined_decrease)."""
        out = (ctypes.c_int32 * 4)()
        thr = -1.0 if md_threshold is None else float(md_threshold)
        capi.check(self._lib.dim2d_openbc_undermined_decrease(self._ctx, distance, thr, out),
                   "dim2d_openbc_undermined_decrease")
        return np.array(out[:], dtype=np.int32)

    def adapt_domain(self, edges: str = "WESN", distance: int = 5, grow: int = 32,
                     md_threshold: float | None = None, shrink: str = "") -> np.ndarray:
        """field_extended%adapt_domain (field_extended_class.f:427-452): each
        edge listed in `shrink` that has grown and whose decrease band holds
        no undermined point goes back to its place at creation; then each edge
        listed in `edges` whose band of `distance` points is undermined moves
        out by `grow` points.  Returns the points added (negative: removed)
        per edge (W, E, S, N)."""
        mask = sum(self.EDGES[c] for c in set(edges.upper()))
        mask += 16 * sum(self.EDGES[c] for c in set(shrink.upper()))
        out = (ctypes.c_int32 * 4)()
        thr = -1.0 if md_threshold is None else float(md_threshold)
        capi.check(self._lib.dim2d_adapt_domain(self._ctx, mask, distance, thr, grow, out),
                   "dim2d_adapt_domain")
        added = np.array(out[:], dtype=np.int32)
        self._refresh_params()
        return added

    def integrate(self, dt: float) -> None:
        """field%integrate: one RK3-TVD step, then time += dt (field_class.f:61-72)."""
        capi.check(self._lib.dim2d_rk3_step(self._ctx, self.time, dt), "dim2d_rk3_step")
        self.time += dt

    def run(self, nsteps: int, dt: float = 0.0, cfl: float = 0.0, sync: bool = True) -> None:
        """nsteps steps with a fixed dt, or with the CFL rule when cfl > 0."""
        capi.check(self._lib.dim2d_integrate(self._ctx, self.time, dt, cfl, nsteps),
                   "dim2d_integrate")
        if sync:
            self.synchronize()
            self.time = self.device_time()[0]

    def synchronize(self) -> None:
        capi.check(self._lib.dim2d_synchronize(self._ctx), "dim2d_synchronize")

    def device_time(self):
        out = np.zeros(2)
        capi.check(self._lib.dim2d_time(self._ctx, _dptr(out)), "dim2d_time")
        return out

    def cfl_dt(self, cfl: float):
        out = np.zeros(4)
        capi.check(self._lib.dim2d_cfl_dt(self._ctx, cfl, _dptr(out)), "dim2d_cfl_dt")
        return out

    def diagnostics(self):
        out = np.zeros(7)
        capi.check(self._lib.dim2d_diagnostics(self._ctx, _dptr(out)), "dim2d_diagnostics")
        return out

    def set_tiles(self, block_x: int = 0, rows: int = 0) -> None:
        capi.check(self._lib.dim2d_set_tiles(self._ctx, block_x, rows), "dim2d_set_tiles")

    def set_stage_fusion(self, fuse: bool, rows: int = 0) -> None:
        """RK3-TVD stages 1 and 2 in one launch (on by default where the edges
        allow it) or apart; rows: tile height of the fused launch (0 keeps it)."""
        capi.check(self._lib.dim2d_set_stage_fusion(self._ctx, int(bool(fuse)), rows),
                   "dim2d_set_stage_fusion")

    KIND = {"stage1": 1, "stage2": 2, "stage3": 4, "wall": 8}

    def profile(self, enable, kinds=None) -> None:
        """Record HIP events around the kernel launches of the following steps:
        all kinds (enable True), the named ones (kinds, e.g. ["stage3"]), or none."""
        mask = 0
        if enable:
            mask = 15 if kinds is None else sum(self.KIND[k] for k in kinds)
        capi.check(self._lib.dim2d_profile_enable(self._ctx, mask), "dim2d_profile_enable")

    def kernel_times(self) -> dict:
        """Mean duration (ms) and count per kernel since profile(True)."""
        out = np.zeros(8)
        capi.check(self._lib.dim2d_profile_read(self._ctx, _dptr(out)), "dim2d_profile_read")
        names = ("stage1", "stage2", "stage3", "wall")
        return {n: (float(out[2 * i]), int(out[2 * i + 1])) for i, n in enumerate(names)}

    def close(self) -> None:
        if self._ctx:
            self._lib.dim2d_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent simulations advanced together (persistent stage
    kernels over one work list, replayed from a captured HIP graph)."""

    def __init__(self, params: list[ModelParams]):
        self.params = list(params)
        self._lib = capi.load()
        arr = (capi.Params * len(params))(*[p.to_c() for p in params])
        self._b = ctypes.c_void_p()
        capi.check(self._lib.dim2d_batch_create(arr, len(params), ctypes.byref(self._b)),
                   "dim2d_batch_create")

    def upload(self, sim: int, nodes: np.ndarray) -> None:
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        capi.check(self._lib.dim2d_batch_upload(self._b, sim, _dptr(nodes)), "dim2d_batch_upload")

    def download(self, sim: int) -> np.ndarray:
        p = self.params[sim]
        out = np.empty((4, p.ny, p.nx))
        capi.check(self._lib.dim2d_batch_download(self._b, sim, _dptr(out)),
                   "dim2d_batch_download")
        return out

    def integrate(self, dts, nsteps: int, sync: bool = True) -> None:
        d = np.ascontiguousarray(dts, dtype=np.float64)
        capi.check(self._lib.dim2d_batch_integrate(self._b, _dptr(d), nsteps),
                   "dim2d_batch_integrate")
        if sync:
            self.synchronize()

    def synchronize(self) -> None:
        capi.check(self._lib.dim2d_batch_synchronize(self._b), "dim2d_batch_synchronize")

    def set_tiles(self, block_x: int = 0, rows: int = 0, workgroups: int = 0) -> None:
        capi.check(self._lib.dim2d_batch_set_tiles(self._b, block_x, rows, workgroups),
                   "dim2d_batch_set_tiles")

    def set_stage_fusion(self, fuse: bool, rows: int = 0, workgroups: int = 0,
                         pipeline: bool = False) -> None:
        """RK3-TVD stages 1 and 2 in one persistent launch (on by default when
        every simulation's edges allow it); rows and workgroups of its tiles.
        pipeline (with fuse): two groups of simulations half a step apart, one
        group's stage pair running beside the other's stage 3 in each launch
        (workgroups then sizes those launches); the same results bit for bit."""
        mode = 2 if (fuse and pipeline) else int(bool(fuse))
        capi.check(self._lib.dim2d_batch_set_stage_fusion(self._b, mode, rows, workgroups),
                   "dim2d_batch_set_stage_fusion")

    def tiles(self) -> dict:
        """The batch's launch shapes (dim2d_batch_get_tiles)."""
        out = (ctypes.c_int32 * 6)()
        capi.check(self._lib.dim2d_batch_get_tiles(self._b, out), "dim2d_batch_get_tiles")
        return {"stage_fusion": bool(out[0]), "pipelined": out[0] == 2,
                "pair_rows": out[1], "pair_workgroups": out[2],
                "block_x": out[3], "rows": out[4], "workgroups": out[5]}

    def autotune(self, dts, nsteps: int = 10):
        """Pick the tile heights and persistent workgroup counts by timing
        (dim2d_batch_autotune): the stage pair's first when stages 1 and 2 are
        fused, then stage 3's (or all three stages' otherwise); the state is left
        as it was.  Returns (rows, workgroups, ms_per_step) of the last search;
        tiles() gives the rest."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        out = np.zeros(3)
        capi.check(self._lib.dim2d_batch_autotune(self._b, _dptr(d), nsteps, _dptr(out)),
                   "dim2d_batch_autotune")
        return int(out[0]), int(out[1]), float(out[2])

    def close(self) -> None:
        if self._b:
            self._lib.dim2d_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
