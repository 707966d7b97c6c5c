"""Initial conditions and phi producers for the benchmark workloads.

These are the "reference path" inputs: they run once on the host and are not
timed.  Host arrays use the reference layout -- a state of the reference's
shape ``(5, Nx, Ny, Nz)`` column-major is held as a C-contiguous numpy array of
shape ``(Nz, Ny, Nx, 5)`` (identical bytes); a potential of the reference's
shape ``(nx, ny, nz)`` is held as ``(nz, ny, nx)``.

Restated from:
  sedov_ic_3d!            src/euler3d.jl:275-316
  Sod set-up              test/test_sod3d.jl:79-104
  lane_emden              src/stellar_ic.jl:34-63
  polytrope_ic_3d!        src/stellar_ic.jl:87-178
  bh_potential            src/gravity_bh.jl:28-31
  solve_poisson_isolated  src/self_gravity.jl:35-76   (numpy FFT, not FFTW)
  restrict_density_to_coarse!  src/fmr_gravity.jl:40-60
  _prolong_scalar_trilinear!  src/fmr_gravity.jl:70-106
"""
from __future__ import annotations

import math

import numpy as np

NG = 3


def empty_state(nx: int, ny: int, nz: int) -> np.ndarray:
    return np.zeros((nz + 2 * NG, ny + 2 * NG, nx + 2 * NG, 5), dtype=np.float64)


def active(U: np.ndarray) -> np.ndarray:
    """View of the active cells, shape (nz, ny, nx, 5)."""
    return U[NG:-NG, NG:-NG, NG:-NG, :]


def _centres(n: int, d: float, offset: float) -> np.ndarray:
    # (i - ng - 0.5) * d + offset for the 1-based padded index i of active cells
    return (np.arange(1, n + 1, dtype=np.float64) - 0.5) * d + offset


# ---------------------------------------------------------------------------
# Sod tube (test/test_sod3d.jl:79-104), generalised to any left/right states.

def sod_3d(direction: str = "x", n_normal: int = 128, n_trans: int = 4, gamma: float = 1.4,
           left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), length: float = 1.0):
    """Returns (U, (nx, ny, nz), (dx, dy, dz))."""
    axis = "xyz".index(direction)
    n = [n_trans, n_trans, n_trans]
    n[axis] = n_normal
    nx, ny, nz = n
    d = [length / nx, length / ny, length / nz]
    U = empty_state(nx, ny, nz)
    coord = (np.arange(1, n_normal + 1, dtype=np.float64) - 0.5) * d[axis] - 0.5 * length
    prof = np.empty((n_normal, 5))
    for a, c in enumerate(coord):
        r, v, p = left if c < 0.0 else right
        prof[a] = (r, 0.0, 0.0, 0.0, p / (gamma - 1) + (0.5 * r) * (v * v))
        prof[a, 1 + axis] = r * v
    A = active(U)
    shape = [1, 1, 1]
    shape[2 - axis] = n_normal
    A[...] = prof.reshape(shape + [5])
    return U, (nx, ny, nz), tuple(d)


# ---------------------------------------------------------------------------
# Sedov (src/euler3d.jl:275-316).  Note that the reference deposits
# E_blast*(gamma-1) of energy; pass E_blast = E0/(gamma-1) to deposit E0.

def sedov_ic_3d(nx, ny, nz, dx, dy, dz, gamma, E_blast=1.0, r_inject=None, rho_bg=1.0,
                P_floor=1e-5, x_offset=0.0, y_offset=0.0, z_offset=0.0):
    U = empty_state(nx, ny, nz)
    r_inj = 2.5 * min(dx, dy, dz) if r_inject is None else float(r_inject)
    A = active(U)
    E_bg = P_floor / (gamma - 1)
    A[..., 0] = rho_bg
    A[..., 4] = E_bg
    xc = _centres(nx, dx, x_offset)[None, None, :]
    yc = _centres(ny, dy, y_offset)[None, :, None]
    zc = _centres(nz, dz, z_offset)[:, None, None]
    r = np.sqrt(xc * xc + yc * yc + zc * zc)
    inside = r <= r_inj
    V_inj = 0.0
    dV = dx * dy * dz
    for _ in range(int(inside.sum())):      # sequential accumulation, as the reference
        V_inj += dV
    dE = E_blast * (gamma - 1) / V_inj if V_inj > 0.0 else 0.0
    A[..., 4] = np.where(inside, E_bg + dE, E_bg)
    return U


# ---------------------------------------------------------------------------
# Turbulent periodic box: rho = exp(s), s a Gaussian random field with
# P(k) ~ k^(-11/3) for 2 <= |k| <= 32, sigma(s) = 0.5; velocity components are
# independent fields with the same spectrum scaled to rms Mach 2; p = 1.

def gaussian_field(n: int, seed: int, kmin: float = 2.0, kmax: float = 32.0,
                   slope: float = -11.0 / 3.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz = np.fft.rfftfreq(n, d=1.0 / n)
    kk = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    band = (kk >= kmin) & (kk <= kmax)
    amp[band] = kk[band] ** (slope / 2.0)
    spec = (rng.standard_normal(kk.shape) + 1j * rng.standard_normal(kk.shape)) * amp
    f = np.fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2))
    f -= f.mean()
    f /= f.std()
    return f


def turbulent_box(n: int = 256, gamma: float = 5.0 / 3.0, mach: float = 2.0, sigma_s: float = 0.5,
                  seed: int = 1234, p0: float = 1.0, kmin: float = 2.0, kmax: float = 32.0):
    """Returns (U, (n,n,n), (dx,dx,dx)) on the unit periodic cube."""
    s = sigma_s * gaussian_field(n, seed, kmin, kmax)
    rho = np.exp(s)
    vel = [gaussian_field(n, seed + 1 + a, kmin, kmax) for a in range(3)]
    # scale so that sqrt(<|v|^2 rho / (gamma p)>) = mach
    v2 = vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2
    scale = mach / math.sqrt(float(np.mean(v2 * rho / (gamma * p0))))
    U = empty_state(n, n, n)
    A = active(U)
    A[..., 0] = rho
    KE = np.zeros_like(rho)
    for a in range(3):
        v = vel[a] * scale
        A[..., 1 + a] = rho * v
        KE += v * v
    A[..., 4] = p0 / (gamma - 1) + 0.5 * rho * KE
    dx = 1.0 / n
    return U, (n, n, n), (dx, dx, dx)


# ---------------------------------------------------------------------------
# Lane-Emden polytrope (src/stellar_ic.jl:34-178)

def lane_emden(n: float, dxi: float = 1e-3):
    nf = float(n)
    xi0 = dxi
    th = 1.0 - xi0 ** 2 / 6.0 + nf * xi0 ** 4 / 120.0
    dth = -xi0 / 3.0 + nf * xi0 ** 3 / 30.0
    xis, ths, dths = [xi0], [th], [dth]
    xi = xi0

    def f(x, t, d):
        return -(max(t, 0.0)) ** nf - 2.0 * d / x

    while th > 0.0:
        k1t = dxi * dth
        k1d = dxi * f(xi, th, dth)
        k2t = dxi * (dth + k1d / 2)
        k2d = dxi * f(xi + dxi / 2, th + k1t / 2, dth + k1d / 2)
        k3t = dxi * (dth + k2d / 2)
        k3d = dxi * f(xi + dxi / 2, th + k2t / 2, dth + k2d / 2)
        k4t = dxi * (dth + k3d)
        k4d = dxi * f(xi + dxi, th + k3t, dth + k3d)
        th += (k1t + 2 * k2t + 2 * k3t + k4t) / 6
        dth += (k1d + 2 * k2d + 2 * k3d + k4d) / 6
        xi += dxi
        xis.append(xi)
        ths.append(th)
        dths.append(dth)
    return np.array(xis), np.array(ths), np.array(dths)


def polytrope_scales(gamma: float, R_star: float, rho_c: float):
    """(M_star, r_scale, K, xi_1) of the polytrope with central density rho_c."""
    n = 1.0 / (gamma - 1.0)
    xis, ths, dths = lane_emden(n)
    xi1, dth1 = xis[-1], dths[-1]
    omega = -xi1 ** 2 * dth1
    r_scale = R_star / xi1
    M_star = rho_c * 4 * math.pi * r_scale ** 3 * omega
    K = 4 * math.pi * r_scale ** 2 * rho_c ** (1.0 - 1.0 / n) / (n + 1.0)
    return M_star, r_scale, K, xi1


def polytrope_ic_3d(nx, ny, nz, dx, dy, dz, gamma, R_star=1.0, rho_c=1.0, x0=0.0, y0=0.0,
                    z0=0.0, rho_floor=1e-10, P_floor=1e-12, M_star=None):
    """polytrope_ic_3d! (src/stellar_ic.jl:87-178) without a hollow core:
    rho = rho_c theta^n, P = K rho^(1+1/n) inside R_star, floors outside, at
    rest.  The star is given by its central density rho_c, or, as in the
    reference, by its mass M_star (then rho_c = M_star/(4 pi r_scale^3 omega_n)).
    x0, y0, z0 are the left edge of the active domain; the star sits at the
    origin."""
    n = 1.0 / (gamma - 1.0)
    xis, ths, _ = lane_emden(n)
    xi1 = xis[-1]
    if M_star is not None:
        omega = -xi1 ** 2 * lane_emden(n)[2][-1]
        rho_c = float(M_star) / (4 * math.pi * (R_star / xi1) ** 3 * omega)
    M_star, r_scale, K, _ = polytrope_scales(gamma, R_star, rho_c)
    dxi = xis[1] - xis[0]
    xc = _centres(nx, dx, x0)[None, None, :]
    yc = _centres(ny, dy, y0)[None, :, None]
    zc = _centres(nz, dz, z0)[:, None, None]
    r = np.sqrt(xc * xc + yc * yc + zc * zc)
    xq = r / r_scale
    i_f = xq / dxi
    i = np.clip(np.floor(i_f).astype(np.int64), 1, len(xis) - 1)
    a = i_f - i
    th = (1.0 - a) * ths[i - 1] + a * ths[np.minimum(i, len(ths) - 1)]
    th = np.where(xq <= 0.0, 1.0, np.where(xq >= xi1, 0.0, th))
    inside = th > 0.0
    rho = np.where(inside, rho_c * np.maximum(th, 0.0) ** n, rho_floor)
    P = np.where(inside, K * rho ** (1.0 + 1.0 / n), P_floor)
    U = empty_state(nx, ny, nz)
    A = active(U)
    A[..., 0] = rho
    A[..., 4] = P / (gamma - 1.0)
    return U


# ---------------------------------------------------------------------------
# Potentials

def point_mass_phi(nx, ny, nz, dx, dy, dz, x0, y0, z0, masses):
    """Sum of Plummer potentials -M/sqrt(r^2 + eps^2) (src/gravity_bh.jl:28-31)
    sampled at cell centres; masses = [(M, (px,py,pz), eps), ...]."""
    xc = _centres(nx, dx, x0)[None, None, :]
    yc = _centres(ny, dy, y0)[None, :, None]
    zc = _centres(nz, dz, z0)[:, None, None]
    phi = np.zeros((nz, ny, nx))
    for M, (px, py, pz), eps in masses:
        ddx, ddy, ddz = xc - px, yc - py, zc - pz
        phi -= M / np.sqrt(ddx ** 2 + ddy ** 2 + ddz ** 2 + eps ** 2)
    return phi


def solve_poisson_isolated(rho: np.ndarray, dx: float, dy: float, dz: float) -> np.ndarray:
    """Hockney-Eastwood zero-padded convolution with K = -1/sqrt(r^2+eps^2),
    eps^2 = (dx dy dz)^(2/3) (src/self_gravity.jl:35-76).  rho is (nz,ny,nx)."""
    nz, ny, nx = rho.shape
    Nx, Ny, Nz = 2 * nx, 2 * ny, 2 * nz
    dV = dx * dy * dz
    eps2 = (dx * dy * dz) ** (2.0 / 3.0)
    fi = np.where(np.arange(Nx) < nx, np.arange(Nx), np.arange(Nx) - Nx).astype(np.float64)
    fj = np.where(np.arange(Ny) < ny, np.arange(Ny), np.arange(Ny) - Ny).astype(np.float64)
    fk = np.where(np.arange(Nz) < nz, np.arange(Nz), np.arange(Nz) - Nz).astype(np.float64)
    r2 = ((fk * dz) ** 2)[:, None, None] + ((fj * dy) ** 2)[None, :, None] + \
         ((fi * dx) ** 2)[None, None, :]
    K = -1.0 / np.sqrt(r2 + eps2)
    pad = np.zeros((Nz, Ny, Nx))
    pad[:nz, :ny, :nx] = rho
    phi = np.fft.irfftn(np.fft.rfftn(K) * np.fft.rfftn(pad), s=(Nz, Ny, Nx), axes=(0, 1, 2)) * dV
    return np.ascontiguousarray(phi[:nz, :ny, :nx])


def restrict_density_to_coarse(rho_c: np.ndarray, rho_f: np.ndarray, box: tuple,
                               ratio: int) -> None:
    """restrict_density_to_coarse! (src/fmr_gravity.jl:40-60) on active-cell
    density arrays (z, y, x): the coarse cells of box = (ci_lo, ci_hi, cj_lo,
    cj_hi, ck_lo, ck_hi) (1-based) become the sum of their r^3 fine cells,
    taken in the reference's order (x offset outermost, z innermost), times
    1/r^3."""
    ci_lo, ci_hi, cj_lo, cj_hi, ck_lo, ck_hi = (int(v) for v in box)
    r = int(ratio)
    n_i, n_j, n_k = ci_hi - ci_lo + 1, cj_hi - cj_lo + 1, ck_hi - ck_lo + 1
    f = rho_f[:r * n_k, :r * n_j, :r * n_i]
    val = np.zeros((n_k, n_j, n_i))
    for sx in range(r):
        for sy in range(r):
            for sz in range(r):
                val = val + f[sz::r, sy::r, sx::r]
    rho_c[ck_lo - 1:ck_hi, cj_lo - 1:cj_hi, ci_lo - 1:ci_hi] = val * (1.0 / float(r ** 3))


def prolong_scalar_trilinear(phi_c: np.ndarray, lo: tuple, ratio: int, nf: tuple) -> np.ndarray:
    """Trilinear prolongation of an active-cell scalar onto a nested box
    (src/fmr_gravity.jl:70-106).  lo = (ci_lo, cj_lo, ck_lo), 1-based;
    nf = (nxf, nyf, nzf).  Arrays are (z, y, x)."""
    ncz, ncy, ncx = phi_c.shape
    out = []
    for axis, (c_lo, n_f, n_c) in enumerate(zip(lo, nf, (ncx, ncy, ncz))):
        pos = (c_lo - 1) + (np.arange(1, n_f + 1) - 0.5) / ratio + 0.5
        ic = np.clip(np.floor(pos).astype(np.int64), 1, n_c - 1)
        out.append((ic, pos - ic))
    (ix, fx), (iy, fy), (iz, fz) = out
    fx = fx[None, None, :]
    fy = fy[None, :, None]
    fz = fz[:, None, None]
    P = lambda a, b, c: phi_c[(iz + c - 1)[:, None, None], (iy + b - 1)[None, :, None],
                              (ix + a - 1)[None, None, :]]
    return ((1 - fx) * (1 - fy) * (1 - fz) * P(0, 0, 0) + fx * (1 - fy) * (1 - fz) * P(1, 0, 0)
            + (1 - fx) * fy * (1 - fz) * P(0, 1, 0) + fx * fy * (1 - fz) * P(1, 1, 0)
            + (1 - fx) * (1 - fy) * fz * P(0, 0, 1) + fx * (1 - fy) * fz * P(1, 0, 1)
            + (1 - fx) * fy * fz * P(0, 1, 1) + fx * fy * fz * P(1, 1, 1))


# ---------------------------------------------------------------------------
# Batch of independent problems of uneven sizes

BATCH_SIZES = (32, 48, 64, 96, 128)


def batch_problems(count: int = 48, seed: int = 4321):
    """`count` independent problems on the unit cube: per-axis sizes drawn
    from BATCH_SIZES; a Sod tube along a random axis (left/right rho and p
    log-uniform on [0.1, 10], at rest) or a Sedov blast (E0 log-uniform on
    [0.1, 10] deposited within 2.5 dx of the centre, ambient rho = 1,
    p = 1e-5); gamma 1.4 or 5/3; outflow boundaries.
    Returns a list of dicts (U, dims, d, gamma, kind, bc, floors)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny, nz = (int(v) for v in rng.choice(BATCH_SIZES, size=3))
        gamma = float(rng.choice([1.4, 5.0 / 3.0]))
        d = (1.0 / nx, 1.0 / ny, 1.0 / nz)
        if rng.random() < 0.5:
            axis = int(rng.integers(3))
            rL, rR, pL, pR = (float(v) for v in 10.0 ** rng.uniform(-1, 1, size=4))
            U = empty_state(nx, ny, nz)
            A = active(U)
            n_axis = (nx, ny, nz)[axis]
            c = (np.arange(1, n_axis + 1) - 0.5) * d[axis] - 0.5
            left = c < 0.0
            rho = np.where(left, rL, rR)
            E = np.where(left, pL, pR) / (gamma - 1)
            shp = [1, 1, 1]
            shp[2 - axis] = n_axis
            A[..., 0] = rho.reshape(shp)
            A[..., 4] = E.reshape(shp)
            kind = f"sod-{'xyz'[axis]}"
        else:
            E0 = float(10.0 ** rng.uniform(-1, 1))
            U = sedov_ic_3d(nx, ny, nz, *d, gamma, E_blast=E0 / (gamma - 1), rho_bg=1.0,
                            P_floor=1e-5, x_offset=-0.5, y_offset=-0.5, z_offset=-0.5)
            kind = "sedov"
        out.append(dict(U=U, dims=(nx, ny, nz), d=d, gamma=gamma, kind=kind, bc="outflow",
                        floors=(1e-8, 1e-8)))
    return out


# ---------------------------------------------------------------------------
# Three-level FMR star: n = 3 polytrope (rho_c = 1, R = 1, gamma = 4/3) in a
# 192^3 base on [-8, 8]^3, nested 96^3 (dx/2) and 128x128x96 (dx/4) boxes,
# atmosphere rho = 1e-10, p = 1e-12, static phi from two Plummer point masses
# (0.5 at (+-2, 0, 0), softening 0.1) plus the star's self-potential.

def fmr_star(base: int = 192, half: float = 8.0, gamma: float = 4.0 / 3.0, rho_floor: float = 1e-10,
             P_floor: float = 1e-12, self_gravity: bool = True):
    """Returns dict(levels=[(U, dims, d, x0)], boxes, ratios, phis, gamma, floors)."""
    dx0 = 2 * half / base
    lev = [((base,) * 3, (dx0,) * 3, (-half,) * 3)]
    # level 1: [-2, 2]^3 of the base, ratio 2
    c = base // 2
    w1 = int(round(2.0 / dx0))                      # base cells per half-width
    b1 = (c - w1 + 1, c + w1, c - w1 + 1, c + w1, c - w1 + 1, c + w1)
    n1 = 2 * (2 * w1)
    dx1 = dx0 / 2
    lev.append(((n1,) * 3, (dx1,) * 3, (-w1 * dx0,) * 3))
    # level 2: (2/3 base) x (2/3 base) x (1/2 base) cells at dx/4 centred in
    # level 1 (128 x 128 x 96 for the 192^3 base)
    c1 = n1 // 2
    n2 = (2 * base // 3, 2 * base // 3, base // 2)
    hx, hz = n2[0] // 4, n2[2] // 4                  # level-1 cells per half-width
    b2 = (c1 - hx + 1, c1 + hx, c1 - hx + 1, c1 + hx, c1 - hz + 1, c1 + hz)
    dx2 = dx1 / 2
    lev.append((n2, (dx2,) * 3, (-hx * dx1, -hx * dx1, -hz * dx1)))
    levels = []
    for dims, d, x0 in lev:
        U = polytrope_ic_3d(*dims, *d, gamma, R_star=1.0, rho_c=1.0, x0=x0[0], y0=x0[1],
                            z0=x0[2], rho_floor=rho_floor, P_floor=P_floor)
        levels.append((U, dims, d, x0))
    masses = [(0.5, (2.0, 0.0, 0.0), 0.1), (0.5, (-2.0, 0.0, 0.0), 0.1)]
    phis = [point_mass_phi(*dims, *d, *x0, masses) for (_, dims, d, x0) in levels]
    if self_gravity:
        # composite density restricted to the base, solved once, prolonged down
        rho = [active(U)[..., 0].copy() for (U, _, _, _) in levels]
        for l in (2, 1):
            restrict_density_to_coarse(rho[l - 1], rho[l], (b1, b2)[l - 1], 2)
        phi_c = solve_poisson_isolated(rho[0], dx0, dx0, dx0)
        phi_1 = prolong_scalar_trilinear(phi_c, (b1[0], b1[2], b1[4]), 2, lev[1][0])
        phi_2 = prolong_scalar_trilinear(phi_1, (b2[0], b2[2], b2[4]), 2, lev[2][0])
        for p, s in zip(phis, (phi_c, phi_1, phi_2)):
            p += s
    return dict(levels=levels, boxes=[b1, b2], ratios=[2, 2], phis=phis, gamma=gamma,
                floors=(rho_floor, P_floor))
