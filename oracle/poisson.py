"""Isolated-boundary Poisson solve, restated from solve_poisson_isolated
(src/self_gravity.jl:35-76) with numpy's complex FFTs (the reference uses
FFTW's complex fft/ifft the same way).  TEST INFRASTRUCTURE ONLY: the checker
of the device solver and of the self-gravity step.

Arrays are (nz, ny, nx) C-order, the same bytes as the reference's (nx, ny,
nz) column-major arrays.  Bit parity with FFTW is not pinned: FFT rounding
differs between libraries, so the device result is compared with a tolerance.
"""
from __future__ import annotations

import numpy as np


def green(nx, ny, nz, dx, dy, dz):
    """K = -1/sqrt(r^2 + eps^2) on the (2nx, 2ny, 2nz) grid, folded so that
    index 0 is the origin (src/self_gravity.jl:44-57)."""
    Nx, Ny, Nz = 2 * nx, 2 * ny, 2 * nz
    eps2 = (dx * dy * dz) ** (2.0 / 3.0)
    fi = np.array([i if i < nx else i - Nx for i in range(Nx)], dtype=np.float64)
    fj = np.array([j if j < ny else j - Ny for j in range(Ny)], dtype=np.float64)
    fk = np.array([k if k < nz else k - Nz for k in range(Nz)], dtype=np.float64)
    r2 = ((fi * dx) ** 2)[None, None, :] + ((fj * dy) ** 2)[None, :, None] + \
        ((fk * dz) ** 2)[:, None, None]
    return -1.0 / np.sqrt(r2 + eps2)


def solve_poisson_isolated(rho, dx, dy, dz):
    nz, ny, nx = rho.shape
    K = green(nx, ny, nz, dx, dy, dz)
    pad = np.zeros((2 * nz, 2 * ny, 2 * nx))
    pad[:nz, :ny, :nx] = rho
    phi = np.real(np.fft.ifftn(np.fft.fftn(K) * np.fft.fftn(pad))) * (dx * dy * dz)
    return np.ascontiguousarray(phi[:nz, :ny, :nx])


def self_gravity_step(U, dx, dy, dz, dt, gamma, bc="outflow", rho_floor=0.0, P_floor=0.0):
    """euler3d_step! with add_self_gravity_source! after the RHS of every
    stage (scripts/run_sn50_fiducial.jl:772-816 with self_gravity on and no
    bodies): phi is solved from the stage's floored density."""
    from . import oracle as O
    NG = 3
    Un = U.copy()
    for stage in (1, 2, 3):
        dU = O.euler3d_rhs(U, dx, dy, dz, gamma, bc, rho_floor, P_floor)[0]
        rho = np.ascontiguousarray(U[NG:-NG, NG:-NG, NG:-NG, 0])
        phi = solve_poisson_isolated(rho, dx, dy, dz)
        O.apply_sg_force(dU, U, phi, dx, dy, dz)
        O.rk3_combine(stage, U, Un, dU, dt)
        O.apply_floors(U, rho_floor, P_floor, gamma)
