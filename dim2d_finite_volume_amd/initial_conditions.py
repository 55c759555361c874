"""Seeded initial conditions for the DIM workloads.

Profiles follow the reference's drop/bubble recipe
(physical_models/dim2d/dim2d_ic/dim2d_dropbubble_module.f:180-386): a tanh
density profile between the saturated liquid and vapour densities of the
reference's fits (dim2d_state_eq_module.f:145-227) and the total energy
  E = rho (8/3 cv T0 - 3 rho) + |grad rho|^2 / (2 We) + rho |u|^2 / 2.
Random numbers come from numpy's PCG64 with an explicit seed, so the GPU run
and the oracle start from identical bits.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi
from .solver import ModelParams


def saturated_densities(T0: float):
    """(rho_liquid, rho_vapor) at T0 (dim2d_lowTemperature = .false. fits)."""
    s = math.sqrt(1.0 - T0)
    return 1.0 + 2.06 * s, 1.0 - 1.91 * s


def _grid(p: ModelParams):
    x_map, y_map = p.coordinates()
    X, Y = np.meshgrid(x_map, y_map)  # shape (ny, nx)
    return X, Y


def circles_field(p: ModelParams, T0: float, centers, radius: float, phase_at_center: str,
                  width: float, flow=(0.0, 0.0), noise: float = 0.0, seed: int = 0) -> np.ndarray:
    """Drops ("liquid" at the centre) or bubbles ("vapor") of one radius.
    The profile of the nearest centre is used at every point."""
    rho_l, rho_v = saturated_densities(T0)
    X, Y = _grid(p)
    d = np.full(X.shape, np.inf)
    for (xc, yc) in centers:
        d = np.minimum(d, np.sqrt((X - xc) ** 2 + (Y - yc) ** 2))
    sign = -1.0 if phase_at_center == "liquid" else 1.0
    mid, half = 0.5 * (rho_l + rho_v), 0.5 * (rho_l - rho_v)
    th = np.tanh(sign * 2.0 * (d - radius) / width)
    rho = mid + half * th
    grad2 = (half * 2.0 / width) ** 2 * (1.0 - th ** 2) ** 2
    if noise > 0.0:
        rng = np.random.default_rng(seed)
        rho = rho + rng.uniform(-noise, noise, size=rho.shape)
    u0, v0 = flow
    nodes = np.empty((4, p.ny, p.nx))
    nodes[0] = rho
    nodes[1] = rho * u0
    nodes[2] = rho * v0
    nodes[3] = rho * (8.0 / 3.0 * p.cv_r * T0 - 3.0 * rho) + grad2 / (2.0 * p.we) + \
        0.5 * rho * (u0 * u0 + v0 * v0)
    return nodes


def drop64(precision: int = capi.PREC_FP64, n: int = 64, seed: int = 1):
    """64^2 liquid drop (radius 0.15, at rest) in vapour at T0 = 0.95, interface
    width 2 sqrt(We) dx, uniform density noise 1e-6, reflection on all edges."""
    p = ModelParams(nx=n + 4, ny=n + 4, precision=precision)
    dx, _ = p.spacing()
    nodes = circles_field(p, 0.95, [(0.0, 0.0)], 0.15, "liquid", 2.0 * math.sqrt(p.we) * dx,
                          noise=1e-6, seed=seed)
    return p, nodes


def two_bubbles(n: int = 4096, precision: int = capi.PREC_FP64, seed: int = 2,
                bc=(capi.BC_PERIODIC, capi.BC_PERIODIC, capi.BC_WALL, capi.BC_WALL)):
    """Two vapour bubbles of radius 0.08 at seeded centres in liquid at
    T0 = 0.95 with mean flow u = 0.1; periodic in x, walls in y."""
    p = ModelParams(nx=n + 4, ny=n + 4, bc=bc, precision=precision)
    dx, _ = p.spacing()
    rng = np.random.default_rng(seed)
    r = 0.08
    while True:
        c = rng.uniform(-0.5 + 0.12, 0.5 - 0.12, size=(2, 2))
        if np.hypot(*(c[0] - c[1])) > 2.0 * r + 0.05:
            break
    nodes = circles_field(p, 0.95, [tuple(c[0]), tuple(c[1])], r, "vapor",
                          2.0 * math.sqrt(p.we) * dx, flow=(0.1, 0.0))
    return p, nodes


def open_bubble(n: int = 8192, precision: int = capi.PREC_MIXED, center=(0.0, 0.0),
                radius: float = 0.1, flow=(0.2, 0.05)):
    """A single vapour bubble (default radius 0.1, centred) in liquid at
    T0 = 0.95 transported by the mean flow (default (u, v) = (0.2, 0.05)),
    Hedstrom open boundaries on all four edges (bc_choice hedstrom_xy_choice)."""
    p = ModelParams(nx=n + 4, ny=n + 4, bc=(capi.BC_OPEN,) * 4, precision=precision)
    dx, _ = p.spacing()
    nodes = circles_field(p, 0.95, [tuple(center)], radius, "vapor", 2.0 * math.sqrt(p.we) * dx,
                          flow=tuple(flow))
    return p, nodes


def far_field(p: ModelParams, T0: float = 0.95, flow=(0.0, 0.0), phase: str = "liquid"):
    """The conservative far-field state of the bubble conditions: the
    surrounding phase at rest temperature T0 carried by the mean flow
    (get_far_field of dim2d_ic/bubble_transported/ic_class.f with the
    saturation densities of this module)."""
    rho_l, rho_v = saturated_densities(T0)
    rho = rho_l if phase == "liquid" else rho_v
    u0, v0 = flow
    return np.array([rho, rho * u0, rho * v0,
                     0.5 * rho * (u0 * u0 + v0 * v0) + rho * (8.0 / 3.0 * p.cv_r * T0 - 3.0 * rho)])


def batch_configs(nsims: int = 256, seed: int = 4, precision: int = capi.PREC_FP64,
                  nmin: int = 64, nmax: int = 512):
    """Independent drop/bubble simulations: square grids n = round(exp(U[ln nmin, ln nmax])),
    T0 ~ U[0.90, 0.99], We ~ U[5, 20], radius ~ U[0.05, 0.2], seeded centres,
    reflection on all edges."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nsims):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        T0 = rng.uniform(0.90, 0.99)
        we = rng.uniform(5.0, 20.0)
        radius = rng.uniform(0.05, 0.2)
        c = rng.uniform(-0.5 + radius, 0.5 - radius, size=2)
        phase = "liquid" if rng.uniform() < 0.5 else "vapor"
        p = ModelParams(nx=n + 4, ny=n + 4, we=we, precision=precision)
        dx, _ = p.spacing()
        nodes = circles_field(p, T0, [tuple(c)], radius, phase, 2.0 * math.sqrt(we) * dx)
        out.append((p, nodes))
    return out


def sound_speed(p: ModelParams, nodes: np.ndarray) -> np.ndarray:
    """speed_of_sound (dim2d_prim_module.f:1179-1215) of every cell."""
    rho, qx, qy, E = nodes
    P = 3.0 / ((3.0 - rho) * p.cv_r) * (E - 0.5 * rho * ((qx / rho) ** 2 + (qy / rho) ** 2)
                                        + 3.0 * rho ** 2) - 3.0 * rho ** 2
    e = P + rho ** 2 * (-3.0 + 2.0 * rho)
    b = np.sqrt((P + 3.0 * rho ** 2) / (p.cv_r * e))
    a = 1.0 / np.sqrt(1.0 + b ** 2)
    d = np.sqrt(rho ** 3) * np.sqrt((3.0 - rho) / e)
    return math.sqrt(3.0) * rho / (a * d)


def cfl_sound_speed(p: ModelParams, nodes: np.ndarray) -> np.ndarray:
    """Sound speed of the time step rule: the reduced form
    c^2 = 3 (e + (P + 3 rho^2)/cv) / (rho (3 - rho)) of speed_of_sound, with
    c = 0 where c^2 <= 0 (spinodal part of an interface)."""
    rho, qx, qy, E = nodes
    P = 3.0 / ((3.0 - rho) * p.cv_r) * (E - 0.5 * rho * ((qx / rho) ** 2 + (qy / rho) ** 2)
                                        + 3.0 * rho ** 2) - 3.0 * rho ** 2
    e = P + rho ** 2 * (-3.0 + 2.0 * rho)
    c2 = 3.0 * (e + (P + 3.0 * rho ** 2) / p.cv_r) / (rho * (3.0 - rho))
    return np.sqrt(np.maximum(c2, 0.0))


def cfl_dt(p: ModelParams, nodes: np.ndarray, cfl: float) -> float:
    """The library's time step rule (dim2d_cfl_dt) evaluated on the host."""
    it = nodes[:, 2:-2, 2:-2]
    rho = it[0]
    speed = np.sqrt((it[1] / rho) ** 2 + (it[2] / rho) ** 2) + cfl_sound_speed(p, it)
    dx, dy = p.spacing()
    h = min(dx, dy)
    dt = min(h / speed.max(), p.re * rho.min() * h * h, math.sqrt(p.we / rho.max()) * h * h)
    return cfl * dt
