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

NG = 3


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


# ---------------------------------------------------------------------------
# FMR self-gravity (src/fmr_gravity.jl): composite density on the root level,
# one isolated solve there, trilinear prolongation of phi down the levels.

def restrict_density_to_coarse(Uc, Uf, lo_hi, ratio):
    """restrict_density_to_coarse! (src/fmr_gravity.jl:40-60): the coarse
    density inside the fine box becomes the 1/r^3 mean of the fine densities,
    summed in the reference's order (sx outermost, sz innermost), then scaled
    by 1/r^3.  Uc and Uf are (Nz, Ny, Nx, 5) states; only Uc[..., 0] changes."""
    ci_lo, ci_hi, cj_lo, cj_hi, ck_lo, ck_hi = (int(v) for v in lo_hi)
    r = int(ratio)
    n_ic, n_jc, n_kc = ci_hi - ci_lo + 1, cj_hi - cj_lo + 1, ck_hi - ck_lo + 1
    Af = Uf[NG:NG + r * n_kc, NG:NG + r * n_jc, NG:NG + r * n_ic, 0]
    val = np.zeros((n_kc, n_jc, n_ic))
    for sx in range(r):
        for sy in range(r):
            for sz in range(r):
                val = val + Af[sz::r, sy::r, sx::r]
    inv_r3 = 1.0 / float(r ** 3)
    Uc[NG + ck_lo - 1:NG + ck_hi, NG + cj_lo - 1:NG + cj_hi, NG + ci_lo - 1:NG + ci_hi, 0] = \
        val * inv_r3


def prolong_scalar_trilinear(phi_c, lo, ratio, nf):
    """_prolong_scalar_trilinear! (src/fmr_gravity.jl:70-106) with the
    reference's operation order: fine cell centre in coarse active-cell
    coordinates xc = ((lo - 1) + (i - 0.5)/r) + 0.5, base cell clamped to
    [1, n_c - 1], eight weights ((1-fx)(1-fy))(1-fz) ..., summed left to right.
    phi_c is (nzc, nyc, nxc); lo = (ci_lo, cj_lo, ck_lo) 1-based; nf =
    (nxf, nyf, nzf).  Returns (nzf, nyf, nxf)."""
    ncz, ncy, ncx = phi_c.shape
    axes = []
    for c_lo, n_f, n_c in zip(lo, nf, (ncx, ncy, ncz)):
        pos = (float(c_lo - 1) + (np.arange(1, n_f + 1, dtype=np.float64) - 0.5) / ratio) + 0.5
        ic = np.clip(np.floor(pos).astype(np.int64), 1, n_c - 1)
        axes.append((ic, pos - ic))
    (ix, fx), (iy, fy), (iz, fz) = axes
    fx, fy, fz = fx[None, None, :], fy[None, :, None], fz[:, None, None]
    kx, ky, kz = (ix - 1)[None, None, :], (iy - 1)[None, :, None], (iz - 1)[:, None, None]

    def P(a, b, c):
        return phi_c[kz + c, ky + b, kx + a]

    w000 = (1 - fx) * (1 - fy) * (1 - fz)
    w100 = fx * (1 - fy) * (1 - fz)
    w010 = (1 - fx) * fy * (1 - fz)
    w110 = fx * fy * (1 - fz)
    w001 = (1 - fx) * (1 - fy) * fz
    w101 = fx * (1 - fy) * fz
    w011 = (1 - fx) * fy * fz
    w111 = fx * fy * fz
    out = (w000 * P(0, 0, 0) + w100 * P(1, 0, 0) + w010 * P(0, 1, 0) + w110 * P(1, 1, 0)
           + w001 * P(0, 0, 1) + w101 * P(1, 0, 1) + w011 * P(0, 1, 1) + w111 * P(1, 1, 1))
    return np.ascontiguousarray(out)


def _active_dims(U):
    Nz, Ny, Nx, _ = U.shape
    return Nx - 2 * NG, Ny - 2 * NG, Nz - 2 * NG


def solve_fmr_poisson(Us, dxs, lo_his, ratios):
    """solve_fmr_poisson (src/fmr_gravity.jl:119-133) over a chain of nested
    levels (root first).  The density is restricted from the finest level up
    (each restriction as restrict_density_to_coarse!, which changes the parent
    density inside the child's box), the root is solved, and phi is prolonged
    trilinearly down the chain.  With two levels this is the reference's
    function.  Returns the list of per-level phi, (nz, ny, nx) each."""
    nlev = len(Us)
    for l in range(nlev - 1, 0, -1):
        restrict_density_to_coarse(Us[l - 1], Us[l], lo_his[l - 1], ratios[l - 1])
    dx, dy, dz = dxs[0]
    phis = [solve_poisson_isolated(np.ascontiguousarray(Us[0][NG:-NG, NG:-NG, NG:-NG, 0]),
                                   dx, dy, dz)]
    for l in range(1, nlev):
        lh = lo_his[l - 1]
        phis.append(prolong_scalar_trilinear(phis[-1], (lh[0], lh[2], lh[4]), ratios[l - 1],
                                             _active_dims(Us[l])))
    return phis


def fmr3d_step_self_gravity(Uc, Uf, dxc, lo_hi, ratio, gamma, dt_c, bc="outflow",
                            rho_floor=0.0, P_floor=0.0, G_grav=1.0):
    """fmr3d_step!(G, dt_c; self_gravity=true, G_grav) (src/fmr3d.jl:451-473):
    one composite solve at the start of the step, phi scaled by G_grav when it
    is not 1, then the two-level step with that frozen phi on both levels."""
    from . import oracle as O
    d = (dxc,) * 3 if np.isscalar(dxc) else tuple(dxc)
    phi_c, phi_f = solve_fmr_poisson([Uc, Uf], [d], [lo_hi], [ratio])
    if G_grav != 1.0:
        phi_c = phi_c * float(G_grav)
        phi_f = phi_f * float(G_grav)
    O.fmr3d_step(Uc, Uf, d, lo_hi, ratio, gamma, dt_c, bc, rho_floor, P_floor, phi_c, phi_f)
    return phi_c, phi_f
