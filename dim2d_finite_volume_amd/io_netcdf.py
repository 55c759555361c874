"""Output and restart files in the layout of the reference's NetCDF writer.

The reference writes one file per output time, data<N>.nc, through nf90
(io_operators/nf90_operators/io_operators_class.f:102-236,
nf90_operators_module.f:195-1353) and reads it back for a restart
(nf90_operators_read_module.f:48-196).  These functions write and read the
same dimensions, variables and attributes:

  dimensions  time = 1, x = nx, y = ny (ghost points included)
  variables   time(time), x(x), y(y) and mass, momentum_x, momentum_y,
              energy, each over (time, x, y) in Fortran order, i.e.
              [j, i, 0] = nodes(i, j, k) in C order
  attributes  units, long_name (and axis for x, y) on the coordinates;
              units, long_name, missing_data and _FillValue on the fields;
              title, history, the domain, dt and the model parameters
              (viscous_r, Re, We, Pr, cv_r, gravity) as global attributes

The reference creates NetCDF-4 files (NF90_NETCDF4, an HDF5 container); no
HDF5 library is available to this package, so the files here use the classic
format with 64-bit offsets (CDF-2), which the netCDF library opens the same
way (nf90_open, and the reference's post-processing readers).  The reference's
IC/BC description attributes (nf90_write_header_ic/bc) are not written.
"""
from __future__ import annotations

import datetime

import numpy as np
from scipy.io import netcdf_file

# pmodel_eq_class.f:534-598 (get_var_name, get_var_longname, get_var_unit)
VAR_NAMES = ("mass", "momentum_x", "momentum_y", "energy")
VAR_LONG_NAMES = ("mass density", "momentum density along the x-axis",
                  "momentum density along the y-axis", "total energy density")
VAR_UNITS = ("(kg/m3)/(kg/m3)", "(kg/(m2.s))/(kg/(m2.s))", "(kg/(m2.s))/(kg/(m2.s))",
             "(J/m3)/(J.m3)")
# NF90_FILL_DOUBLE, the fill value of FP64 variables (nf90_def_var_model)
FILL_DOUBLE = 9.9692099683868690e+36


def write_data(path: str, nodes: np.ndarray, x_map: np.ndarray, y_map: np.ndarray, time: float,
               params=None, dt: float | None = None, t_max: float | None = None,
               attrs: dict | None = None) -> None:
    """Write one output time.  nodes: (4, ny, nx), the reference's
    nodes(nx, ny, ne) in C order; params: a ModelParams (domain and model
    parameters become global attributes)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    ne, ny, nx = nodes.shape
    if ne != 4 or x_map.shape != (nx,) or y_map.shape != (ny,):
        raise ValueError("nodes must be (4, ny, nx) with x_map (nx,) and y_map (ny,)")
    f = netcdf_file(path, "w", version=2)
    try:
        now = datetime.datetime.now()
        f.title = b"DIM2D"
        f.history = now.strftime("date %Y/%m/%d time %H:%M:%S").encode()
        if params is not None:
            f.x_min, f.x_max = float(params.x_min), float(params.x_max)
            f.y_min, f.y_max = float(params.y_min), float(params.y_max)
            # get_sim_parameters, pmodel_eq_class.f:644-675
            f.viscous_r = float(params.viscous_r)
            f.Re = float(params.re)
            f.We = float(params.we)
            f.Pr = float(params.pr)
            f.cv_r = float(params.cv_r)
            f.gravity = float(params.gravity)
        if dt is not None:
            f.dt = float(dt)
        if t_max is not None:
            f.t_max = float(t_max)
        for k, v in (attrs or {}).items():
            setattr(f, k, v.encode() if isinstance(v, str) else v)

        f.createDimension("time", 1)
        f.createDimension("x", nx)
        f.createDimension("y", ny)
        t = f.createVariable("time", "d", ("time",))
        x = f.createVariable("x", "d", ("x",))
        y = f.createVariable("y", "d", ("y",))
        t.units, x.units, y.units = b"s/s", b"m/m", b"m/m"
        x.axis, y.axis = b"X", b"Y"
        t.long_name = b"reduced time"
        x.long_name = b"reduced x-coordinate"
        y.long_name = b"reduced y-coordinate"
        t[:] = [time]
        x[:] = x_map
        y[:] = y_map
        for k in range(4):
            v = f.createVariable(VAR_NAMES[k], "d", ("y", "x", "time"))
            v.units = VAR_UNITS[k].encode()
            v.long_name = VAR_LONG_NAMES[k].encode()
            v.missing_data = FILL_DOUBLE
            v._FillValue = FILL_DOUBLE
            v[:, :, 0] = nodes[k]
    finally:
        f.close()


def read_data(path: str):
    """Read a file written in this layout (by write_data or by the
    reference's writer as a classic file): (nodes (4, ny, nx), x_map, y_map,
    time, global attributes)."""
    f = netcdf_file(path, "r", mmap=False)
    try:
        nodes = np.stack([np.array(f.variables[n][:, :, 0], dtype=np.float64) for n in VAR_NAMES])
        x_map = np.array(f.variables["x"][:], dtype=np.float64)
        y_map = np.array(f.variables["y"][:], dtype=np.float64)
        time = float(f.variables["time"][0])
        attrs = {k: (v.decode() if isinstance(v, bytes) else v) for k, v in f._attributes.items()}
    finally:
        f.close()
    return nodes, x_map, y_map, time, attrs
