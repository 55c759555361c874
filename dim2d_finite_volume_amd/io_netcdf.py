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

The global attributes follow nf90_write_header (nf90_operators_module.f:200-
336): title and history, the program description, the domain, t_max, dt and
detail_print, the initial condition and its description
(nf90_write_header_ic_dim2d, :352-597), the boundary conditions and their
open-boundary strategies (nf90_write_header_bc, :613-894), then the model
parameters (nf90_write_sim_parameters, :913-952).  Names and their order are
the reference's; institution, source, prog_commit and references describe
this package.  header_params() reads the model, domain and boundary choice
back from such a header, the reference's or this package's.

The reference creates NetCDF-4 files (NF90_NETCDF4, an HDF5 container); no
HDF5 library is available to this package, so the files here use the classic
format with 64-bit offsets (CDF-2), which the netCDF library opens the same
way (nf90_open, and the reference's post-processing readers).
"""
from __future__ import annotations

import datetime
from dataclasses import dataclass

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

# code tables of parameters_constant.f:59-370
DIM2D_IC_CODES = ("steady_state", "drop_retraction", "bubble_ascending", "homogeneous_liquid",
                  "drop_collision", "phase_separation", "bubble_transported",
                  "bubble_next_to_wall", "bubble_collapse", "bubble_nucleation",
                  "bubble_spherical_cap", "newgrdpt_test", "bubbles_transported")
BC_CODES = ("periodic_xy", "reflection_xy", "hedstrom_xy", "poinsot_xy", "yoolodato_xy",
            "wall_xy", "wall_S_reflection", "wall_S_open", "half_wall_S_open")
FLOW_PROFILE_CODES = ("parabolic_profile", "linear_profile")
PHASE_CODES = ("liquid", "vapor")
WALL_SOURCE_CODES = ("no_heat_source", "constant_heat_source", "gaussian_heat_source")
SURFACE_TYPE_CODES = ("uniform_surface", "surface_with_heaters")
# the initial conditions whose header carries the wall description (:366-369)
WALL_ICS = ("bubble_next_to_wall", "bubble_collapse", "bubble_nucleation",
            "bubble_spherical_cap")
# the boundary codes of the edge sets (W, E, S, N) the solver takes; the
# reference has no code for periodic x with walls in y, written here as
# "periodic_x_wall_y"
_P, _R, _W, _O = 1, 0, 2, 3  # capi.BC_PERIODIC, BC_REFLECTION, BC_WALL, BC_OPEN
BC_OF_EDGES = {(_P, _P, _P, _P): "periodic_xy", (_R, _R, _R, _R): "reflection_xy",
               (_O, _O, _O, _O): "hedstrom_xy", (_W, _W, _W, _W): "wall_xy",
               (_R, _R, _W, _R): "wall_S_reflection", (_O, _O, _W, _O): "wall_S_open",
               (_P, _P, _W, _W): "periodic_x_wall_y", (_W, _W, _R, _R): "wall_x_reflection_y",
               (_W, _W, _P, _P): "wall_x_periodic_y"}
EDGES_OF_BC = {v: k for k, v in BC_OF_EDGES.items()}


@dataclass
class InitialCondition:
    """The description of an initial condition that nf90_write_header_ic_dim2d
    writes (parameters_input.f: ic_choice, T0, flow_velocity, flow_profile,
    phase_at_center, ratio_bubble_interface, ic_perturbation_*).  The wall
    description of the wall initial conditions comes from the model
    parameters (contact angle, constant heat flux)."""
    name: str = "steady_state"
    temperature: float = 0.95
    flow_velocity: float = 0.0
    flow_profile: str = "parabolic_profile"
    phase_at_center: str = "vapor"
    ratio_bubble_radius_interface: float = 2.0
    perturbation_amp: float | None = None

    def __post_init__(self):
        if self.name not in DIM2D_IC_CODES:
            raise ValueError(f"unknown initial condition {self.name!r}")
        if self.flow_profile not in FLOW_PROFILE_CODES:
            raise ValueError(f"unknown flow profile {self.flow_profile!r}")
        if self.phase_at_center not in PHASE_CODES:
            raise ValueError(f"unknown phase {self.phase_at_center!r}")


def header_attributes(params=None, ic: InitialCondition | None = None, dt=None, t_max=None,
                      detail_print=None, md_threshold=None, history=None) -> dict:
    """The global attributes of nf90_write_header in the reference's order.
    params: ModelParams; md_threshold: the open-boundary density threshold of
    the domain extension (bf_openbc_md_threshold), if active."""
    a = {"title": "DIM2D",
         "history": history or datetime.datetime.now().strftime("date %Y/%m/%d time %H:%M:%S"),
         "institution": "dim2d_finite_volume_amd",
         "source": "dim2d_finite_volume_amd (MI355X HIP)",
         "prog_commit": "n/a",
         "references": "DIM2D solver of jdesmarais/dim2D_finite_volume",
         "convention": "cf-1.6"}
    if params is not None:
        a.update(x_min=float(params.x_min), x_max=float(params.x_max),
                 y_min=float(params.y_min), y_max=float(params.y_max))
    if t_max is not None:
        a["t_max"] = float(t_max)
    if dt is not None:
        a["dt"] = float(dt)
    if detail_print is not None:
        a["detail_print"] = float(detail_print)
    if ic is not None:
        a["initial_conditions"] = ic.name
        a["flow_velocity"] = float(ic.flow_velocity)
        a["temperature"] = float(ic.temperature)
        if ic.name in WALL_ICS:
            # the reference writes these codes untrimmed, padded to the
            # length of their tables' entries
            if abs(ic.flow_velocity) > 0.0:
                a["flow_profile"] = ic.flow_profile.ljust(17)
            a["phase_at_center"] = ic.phase_at_center.ljust(6)
            if ic.name == "bubble_collapse":
                a["ratio_bubble_radius_interface"] = float(ic.ratio_bubble_radius_interface)
            a["wall_surface_type"] = "uniform_surface".ljust(20)
            a["wall_micro_contact_angle"] = float(params.wall_contact_angle if params else 90.0)
            q = float(params.wall_heat_flux) if params is not None else 0.0
            if q != 0.0:
                a["wall_heat_source"] = "constant_heat_source"
                a["wall_maximum_heat_source"] = q
        if ic.perturbation_amp is not None:
            a["ic_perturbation_ac"] = "activated"
            a["ic_perturbation_amp"] = float(ic.perturbation_amp)
    if params is not None:
        code = BC_OF_EDGES.get(tuple(params.bc))
        if code is None:
            names = {0: "reflection", 1: "periodic", 2: "wall", 3: "open"}
            code = "_".join(f"{e}:{names[b]}" for e, b in zip("WESN", params.bc))
        a["boundary_conditions"] = code
        if code in ("hedstrom_xy", "wall_S_open"):
            # the open-boundary strategies this solver runs: the reference's
            # defaults (parameters_input.f:649-669)
            a["obc_eigenqties_strategy"] = "obc_eigenqties_bc"
            a["obc_edge_xy_strategy"] = "obc_edge_xy_flux"
            a["obc_edge_flux_strategy"] = "obc_edge_flux_capillarity"
            a["obc_edge_overlap_ac"] = "activated"
            if md_threshold is not None:
                a["openbc_md_threshold_ac"] = "activated"
                a["openbc_md_threshold"] = float(md_threshold)
        # get_sim_parameters, pmodel_eq_class.f:644-675
        a.update(viscous_r=float(params.viscous_r), Re=float(params.re), We=float(params.we),
                 Pr=float(params.pr), cv_r=float(params.cv_r), gravity=float(params.gravity))
    return a


def header_params(attrs: dict, nx: int, ny: int):
    """ModelParams and InitialCondition from a header written by the
    reference's nf90_write_header or by write_data: domain, model parameters,
    boundary code (the reference's codes that the solver takes, and this
    package's compositions), wall angle and constant heat flux.  The phase
    at the centre and the flow profile are in the header of the wall
    conditions only; elsewhere they read as the defaults."""
    from .solver import ModelParams
    code = attrs.get("boundary_conditions", "reflection_xy")
    if code not in EDGES_OF_BC:
        raise ValueError(f"boundary conditions {code!r} have no counterpart in the solver")
    g = float(attrs.get("gravity", 0.0))
    p = ModelParams(nx=nx, ny=ny, bc=EDGES_OF_BC[code],
                    x_min=float(attrs["x_min"]), x_max=float(attrs["x_max"]),
                    y_min=float(attrs["y_min"]), y_max=float(attrs["y_max"]),
                    re=float(attrs["Re"]), we=float(attrs["We"]), pr=float(attrs["Pr"]),
                    viscous_r=float(attrs["viscous_r"]), cv_r=float(attrs["cv_r"]), gravity=g,
                    wall_contact_angle=float(attrs.get("wall_micro_contact_angle", 90.0)),
                    wall_heat_flux=(float(attrs["wall_maximum_heat_source"])
                                    if str(attrs.get("wall_heat_source", "")).strip()
                                    == "constant_heat_source" else 0.0))
    ic = None
    if "initial_conditions" in attrs:
        ic = InitialCondition(
            name=attrs["initial_conditions"].strip(),
            temperature=float(attrs.get("temperature", 0.95)),
            flow_velocity=float(attrs.get("flow_velocity", 0.0)),
            flow_profile=attrs.get("flow_profile", "parabolic_profile").strip(),
            phase_at_center=attrs.get("phase_at_center", "vapor").strip(),
            ratio_bubble_radius_interface=float(attrs.get("ratio_bubble_radius_interface", 2.0)),
            perturbation_amp=(float(attrs["ic_perturbation_amp"])
                              if attrs.get("ic_perturbation_ac") == "activated" else None))
    return p, ic


def write_data(path: str, nodes: np.ndarray, x_map: np.ndarray, y_map: np.ndarray, time: float,
               params=None, dt: float | None = None, t_max: float | None = None,
               attrs: dict | None = None, ic: InitialCondition | None = None,
               detail_print: float | None = None, md_threshold: float | None = None) -> None:
    """Write one output time.  nodes: (4, ny, nx), the reference's
    nodes(nx, ny, ne) in C order; params: a ModelParams and ic an
    InitialCondition, written into the header with dt, t_max, detail_print
    and md_threshold as header_attributes() orders them; attrs: further
    global attributes, written last."""
    nodes = np.asarray(nodes, dtype=np.float64)
    ne, ny, nx = nodes.shape
    if ne != 4 or x_map.shape != (nx,) or y_map.shape != (ny,):
        raise ValueError("nodes must be (4, ny, nx) with x_map (nx,) and y_map (ny,)")
    f = netcdf_file(path, "w", version=2)
    try:
        head = header_attributes(params, ic, dt, t_max, detail_print, md_threshold)
        for k, v in list(head.items()) + list((attrs or {}).items()):
            # the reference's attributes are real(rkind) (NF90_DOUBLE); a
            # Python float would be stored as NC_FLOAT
            if isinstance(v, str):
                v = v.encode()
            elif isinstance(v, float):
                v = np.float64(v)
            setattr(f, k, v)

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
