"""Gas self-gravity on the device: the isolated Poisson solve
(solve_poisson_isolated, src/self_gravity.jl:35-76) and the per-stage
self-gravity source of the coupled step (add_self_gravity_source!,
src/self_gravity.jl:93-115)."""
from __future__ import annotations

import numpy as np

from ._lib import BSN_HOST, call
from .euler3d import Grid, _cache, _ptr


def solve_poisson_isolated(rho, nx, ny, nz, dx, dy, dz) -> np.ndarray:
    """Phi (same shape as rho, (nz, ny, nx) C-order = (nx, ny, nz) column-major)
    of the active-cell density rho (solve_poisson_isolated,
    src/self_gravity.jl:35-76), computed on the GPU: pruned register FFTs
    when every zero-padded axis is a power of two, hipFFT otherwise."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    if rho.shape != (nz, ny, nx):
        raise ValueError(f"rho shape {rho.shape} != {(nz, ny, nx)}")
    key = ("poisson", nx, ny, nz)
    g = _cache.get(key)
    if g is None:
        g = _cache[key] = Grid(nx, ny, nz)
    phi = np.empty_like(rho)
    call("bsn_solve_poisson_isolated", g.handle, _ptr(rho), BSN_HOST, _ptr(phi), BSN_HOST,
         float(dx), float(dy), float(dz))
    return phi


def set_self_gravity(grid: Grid, on: bool = True) -> None:
    """Every later step of `grid` solves phi from each stage's density and
    adds the self-gravity source (replaces a phi set with Grid.set_phi)."""
    call("bsn_grid_set_self_gravity", grid.handle, int(bool(on)))
