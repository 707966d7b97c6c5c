"""Point-mass gravity and gas sinks: the reference's BlackHole bodies
(src/BinarySupernova.jl:130-171) as sources of the Euler RHS.

``add_bh_gravity_source_`` and ``add_sink_sources_`` mirror
add_bh_gravity_source! (src/gravity_bh.jl:58-78) and add_sink_sources!
(src/sinks.jl:32-58) on a dU array.  ``set_point_masses`` / ``set_sinks``
attach the bodies to a grid so that every later step adds them to each
stage's RHS inside the fused update kernel, in the order of the reference's
coupled step (scripts/run_sn50_fiducial.jl:772-816): flux divergence,
point-mass gravity, the phi source, sinks.

``gas_force_on_bh`` (src/gravity_bh.jl:91-113) and ``accrete_``
(src/sinks.jl:73-109) read the gas on the device: only the sums come back to
the host, where the reference copies the grid to the host every call.  The
N-body update stays with the caller.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import BSN_HOST, call
from .euler3d import Grid, _as_grid, _ptr

MAX_BODIES = 8


@dataclass
class BlackHole:
    """src/BinarySupernova.jl:130-137."""
    pos: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    vel: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    mass: float = 0.0
    eps: float = 0.0
    c_code: float = 1.0
    r_floor: float = 0.0


def r_sink(bh: BlackHole) -> float:
    """max(6 M / c^2, r_floor)  (src/BinarySupernova.jl:159)."""
    a = 6.0 * bh.mass / (bh.c_code * bh.c_code)
    if a != a or bh.r_floor != bh.r_floor:      # Julia's max propagates NaN
        return math.nan
    return a if a > bh.r_floor else float(bh.r_floor)


def t_sink(bh: BlackHole, f_sink: float = 1.0) -> float:
    """f_sink * r_sink / sqrt(2 M / r_sink)  (src/BinarySupernova.jl:166-171)."""
    rs = r_sink(bh)
    v_ff = math.sqrt(2.0 * bh.mass / rs)
    return f_sink * rs / v_ff


def _arr(vals):
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).ravel())
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _masses(bhs):
    if len(bhs) > MAX_BODIES:
        raise ValueError(f"at most {MAX_BODIES} bodies")
    pos = _arr([c for b in bhs for c in b.pos])
    mass = _arr([b.mass for b in bhs])
    eps = _arr([b.eps for b in bhs])
    return pos, mass, eps


def _sinks(bhs, f_sink):
    if len(bhs) > MAX_BODIES:
        raise ValueError(f"at most {MAX_BODIES} bodies")
    pos = _arr([c for b in bhs for c in b.pos])
    vel = _arr([c for b in bhs for c in b.vel])
    rs = _arr([r_sink(b) for b in bhs])
    ts = _arr([t_sink(b, f_sink) for b in bhs])
    return pos, vel, rs, ts


def add_bh_gravity_source_(dU, U, nx, ny, nz, dx, dy, dz, bhs, x0, y0, z0):
    """add_bh_gravity_source! (src/gravity_bh.jl:58-78) on a host dU."""
    g, _ = _as_grid(U, nx, ny, nz)
    pos, mass, eps = _masses(bhs)
    call("bsn_add_bh_gravity_source", g.handle, _ptr(dU), BSN_HOST, float(dx), float(dy),
         float(dz), len(bhs), pos[1], mass[1], eps[1], float(x0), float(y0), float(z0))


def add_sink_sources_(dU, U, nx, ny, nz, dx, dy, dz, bhs, x0, y0, z0, *, f_sink=1.0,
                      torque_free=True, rho_sink_min=0.0):
    """add_sink_sources! (src/sinks.jl:32-58) on a host dU."""
    g, _ = _as_grid(U, nx, ny, nz)
    pos, vel, rs, ts = _sinks(bhs, f_sink)
    call("bsn_add_sink_sources", g.handle, _ptr(dU), BSN_HOST, float(dx), float(dy), float(dz),
         len(bhs), pos[1], vel[1], rs[1], ts[1], int(bool(torque_free)), float(rho_sink_min),
         float(x0), float(y0), float(z0))


def set_point_masses(grid: Grid, bhs, x0, y0, z0) -> None:
    """Point-mass gravity of `bhs` in every later step of `grid`."""
    pos, mass, eps = _masses(bhs)
    call("bsn_grid_set_point_masses", grid.handle, len(bhs), pos[1], mass[1], eps[1], float(x0),
         float(y0), float(z0))


def set_sinks(grid: Grid, bhs, x0, y0, z0, *, f_sink=1.0, torque_free=True,
              rho_sink_min=0.0) -> None:
    """Sink drains of `bhs` in every later step of `grid`."""
    pos, vel, rs, ts = _sinks(bhs, f_sink)
    call("bsn_grid_set_sinks", grid.handle, len(bhs), pos[1], vel[1], rs[1], ts[1],
         int(bool(torque_free)), float(rho_sink_min), float(x0), float(y0), float(z0))


def gas_forces_on_bhs(U, nx, ny, nz, dx, dy, dz, bhs, x0, y0, z0) -> np.ndarray:
    """gas_force_on_bh for every body of `bhs` in one pass over the density:
    an (n, 3) array, row b the force of the gas on bhs[b]."""
    g, _ = _as_grid(U, nx, ny, nz)
    pos, mass, eps = _masses(bhs)
    F = np.zeros((len(bhs), 3))
    if len(bhs):
        call("bsn_gas_force_on_bh", g.handle, float(dx), float(dy), float(dz), len(bhs), pos[1],
             mass[1], eps[1], float(x0), float(y0), float(z0), _ptr(F))
    return F


def gas_force_on_bh(U, nx, ny, nz, dx, dy, dz, bh, x0, y0, z0) -> np.ndarray:
    """gas_force_on_bh (src/gravity_bh.jl:91-113): F = +sum rho grad(Phi_bh) dV."""
    return gas_forces_on_bhs(U, nx, ny, nz, dx, dy, dz, [bh], x0, y0, z0)[0]


def accrete_all_(bhs, U, nx, ny, nz, dx, dy, dz, x0, y0, z0, dt, *, f_sink=1.0,
                 rho_sink_min=0.0) -> np.ndarray:
    """accrete! for every body of `bhs` (each with its own sink radius and
    time, from the state U): updates bh.mass and bh.vel in place and returns
    the accreted masses."""
    g, _ = _as_grid(U, nx, ny, nz)
    n = len(bhs)
    if n > MAX_BODIES:
        raise ValueError(f"at most {MAX_BODIES} bodies")
    pos = _arr([c for b in bhs for c in b.pos])
    vel = np.ascontiguousarray([float(c) for b in bhs for c in b.vel], dtype=np.float64)
    mass = np.ascontiguousarray([float(b.mass) for b in bhs], dtype=np.float64)
    rs = _arr([r_sink(b) for b in bhs])
    ts = _arr([t_sink(b, f_sink) for b in bhs])
    dm = np.zeros(n)
    if n:
        call("bsn_accrete", g.handle, float(dx), float(dy), float(dz), n, pos[1], _ptr(vel),
             _ptr(mass), rs[1], ts[1], float(rho_sink_min), float(x0), float(y0), float(z0),
             float(dt), _ptr(dm))
    for b, bh in enumerate(bhs):
        bh.mass = float(mass[b])
        bh.vel = [float(v) for v in vel[3 * b:3 * b + 3]]
    return dm


def accrete_(bh, U, nx, ny, nz, dx, dy, dz, x0, y0, z0, dt, *, f_sink=1.0,
             rho_sink_min=0.0) -> None:
    """accrete! (src/sinks.jl:73-109): bh gains the mass and the full momentum
    of the gas within r_sink(bh) drained over dt."""
    accrete_all_([bh], U, nx, ny, nz, dx, dy, dz, x0, y0, z0, dt, f_sink=f_sink,
                 rho_sink_min=rho_sink_min)

