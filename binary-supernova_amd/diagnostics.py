"""Scalar diagnostics of a run (src/diagnostics.jl): the gas sums run on the
device in one pass over the state (bsn_gas_diagnostics), the body sums on the
host as in the reference."""
from __future__ import annotations

import numpy as np

from ._lib import call
from .euler3d import _as_grid, _ptr
from .sources import MAX_BODIES, _masses

FIELDS = ("gas_energy_total", "gas_kinetic_total", "gas_momentum_total",
          "gas_angular_momentum_total", "bound_gas_mass")


def gas_diagnostics(U, nx, ny, nz, dx, dy, dz, x0=0.0, y0=0.0, z0=0.0, bhs=(),
                    gamma=5.0 / 3.0) -> dict:
    """Every gas diagnostic of src/diagnostics.jl:21-175 from one pass:
    {gas_energy_total, gas_kinetic_total, gas_momentum_total (3),
    gas_angular_momentum_total (3, about the origin), bound_gas_mass (with
    the potential of `bhs`)}."""
    g, _ = _as_grid(U, nx, ny, nz)
    if len(bhs) > MAX_BODIES:
        raise ValueError(f"at most {MAX_BODIES} bodies")
    pos, mass, eps = _masses(bhs)
    out = np.zeros(9)
    call("bsn_gas_diagnostics", g.handle, float(dx), float(dy), float(dz), float(x0), float(y0),
         float(z0), len(bhs), pos[1], mass[1], eps[1], float(gamma), _ptr(out))
    return {"gas_energy_total": float(out[0]), "gas_kinetic_total": float(out[1]),
            "gas_momentum_total": out[2:5].copy(), "gas_angular_momentum_total": out[5:8].copy(),
            "bound_gas_mass": float(out[8])}


def gas_energy_total(U, nx, ny, nz, dx, dy, dz) -> float:
    """src/diagnostics.jl:21-26."""
    return gas_diagnostics(U, nx, ny, nz, dx, dy, dz)["gas_energy_total"]


def gas_kinetic_total(U, nx, ny, nz, dx, dy, dz) -> float:
    """src/diagnostics.jl:33-44."""
    return gas_diagnostics(U, nx, ny, nz, dx, dy, dz)["gas_kinetic_total"]


def gas_momentum_total(U, nx, ny, nz, dx, dy, dz) -> np.ndarray:
    """src/diagnostics.jl:51-59."""
    return gas_diagnostics(U, nx, ny, nz, dx, dy, dz)["gas_momentum_total"]


def gas_angular_momentum_total(U, nx, ny, nz, dx, dy, dz, x0, y0, z0) -> np.ndarray:
    """src/diagnostics.jl:67-94 (about the origin)."""
    return gas_diagnostics(U, nx, ny, nz, dx, dy, dz, x0, y0, z0)["gas_angular_momentum_total"]


def bound_gas_mass(U, nx, ny, nz, dx, dy, dz, x0, y0, z0, bhs, gamma) -> float:
    """src/diagnostics.jl:137-175: mass of the cells with e_kin + e_int + Phi_BH <= 0."""
    return gas_diagnostics(U, nx, ny, nz, dx, dy, dz, x0, y0, z0, bhs, gamma)["bound_gas_mass"]


def bh_kinetic_total(bhs) -> float:
    """src/diagnostics.jl:97-104."""
    KE = 0.0
    for bh in bhs:
        v = bh.vel
        KE += 0.5 * bh.mass * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return KE


def bh_angular_momentum_total(bhs) -> np.ndarray:
    """src/diagnostics.jl:111-122: sum of M r x v about the origin."""
    Lx = Ly = Lz = 0.0
    for bh in bhs:
        M = bh.mass
        x, y, z = bh.pos[0], bh.pos[1], bh.pos[2]
        vx, vy, vz = bh.vel[0], bh.vel[1], bh.vel[2]
        Lx += M * (y * vz - z * vy)
        Ly += M * (z * vx - x * vz)
        Lz += M * (x * vy - y * vx)
    return np.array([Lx, Ly, Lz])
