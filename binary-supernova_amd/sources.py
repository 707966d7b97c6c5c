"""Point-mass gravity and gas sinks: the reference's BlackHole bodies
(src/BinarySupernova.jl:130-171) as sources of the Euler RHS.

``add_bh_gravity_source_`` and ``add_sink_sources_`` mirror
add_bh_gravity_source! (src/gravity_bh.jl:58-78) and add_sink_sources!
(src/sinks.jl:32-58) on a dU array.  ``set_point_masses`` / ``set_sinks``
attach the bodies to a grid so that every later step adds them to each
stage's RHS inside the fused update kernel, in the order of the reference's
coupled step (scripts/run_sn50_fiducial.jl:772-816): flux divergence,
point-mass gravity, the phi source, sinks.  The N-body update and accrete!
stay with the caller.
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
