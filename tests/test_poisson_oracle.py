"""The oracle's isolated Poisson solve (src/self_gravity.jl:35-76) against the
reference's own checks (test/test_self_gravity.jl): exterior potential of a
uniform sphere scales as 1/r (1 %), self-energy -3M^2/(5R) (10 %), the 1/r
error falls with resolution, and a spherically symmetric density gives no net
momentum source."""
import numpy as np

from oracle import poisson as PO
from oracle import oracle as O

NG = 3


def sphere(n, R, rho0=1.0):
    dx = 1.0 / n
    c = (np.arange(1, n + 1) - 0.5) * dx - 0.5
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    return np.where(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) < R, rho0, 0.0), dx


def ratio_error(n):
    # test/test_self_gravity.jl:30-73: Phi(rA) / Phi(rB) against rB / rA
    rho, dx = sphere(n, 0.15)
    phi = PO.solve_poisson_isolated(rho, dx, dx, dx)
    # round(Int, .) of the reference: nearest, ties to even (Python's round)
    jc = round(0.5 / dx + 0.5)                       # 1-based cell indices
    iA = round((0.5 + 0.25) / dx + 0.5)
    iB = round((0.5 + 0.35) / dx + 0.5)
    r = lambda i: np.sqrt(((i - 0.5) * dx - 0.5) ** 2 + 2 * ((jc - 0.5) * dx - 0.5) ** 2)
    num = phi[jc - 1, jc - 1, iA - 1] / phi[jc - 1, jc - 1, iB - 1]
    return abs(num - r(iB) / r(iA)) / (r(iB) / r(iA))


def test_exterior_potential_scales_as_one_over_r():
    assert ratio_error(32) < 0.01


def test_self_energy_of_uniform_sphere():
    rho, dx = sphere(32, 0.15)
    M = rho.sum() * dx ** 3
    phi = PO.solve_poisson_isolated(rho, dx, dx, dx)
    E = 0.5 * (rho * phi).sum() * dx ** 3
    Ea = -3.0 * M ** 2 / (5.0 * 0.15)
    assert abs((E - Ea) / Ea) < 0.10


def test_convergence():
    assert ratio_error(32) < ratio_error(16)


def test_symmetric_source_has_no_net_momentum():
    n = 16
    rho, dx = sphere(n, 0.3)
    U = np.zeros((n + 6, n + 6, n + 6, 5))
    U[NG:-NG, NG:-NG, NG:-NG, 0] = rho + 0.1
    U[NG:-NG, NG:-NG, NG:-NG, 4] = 1.0
    phi = PO.solve_poisson_isolated(np.ascontiguousarray(U[NG:-NG, NG:-NG, NG:-NG, 0]), dx, dx, dx)
    dU = np.zeros_like(U)
    O.apply_sg_force(dU, U, phi, dx, dx, dx)
    s = dU[NG:-NG, NG:-NG, NG:-NG, 1:4].sum((0, 1, 2))
    assert np.abs(s).max() < 1e-10 * np.abs(dU[..., 1:4]).max()


def test_equals_direct_sum_on_a_ragged_grid():
    # The zero-padded (2n) circular convolution is the linear convolution
    # exactly: phi_i = dV * sum_j rho_j * (-1 / sqrt(|x_i - x_j|^2 + eps^2)),
    # eps^2 = (dx dy dz)^(2/3) (src/self_gravity.jl:44-45), by brute force.
    # A dropped fold, a transposed axis or a wrong scale fails this.
    nx, ny, nz = 7, 5, 4
    dx, dy, dz = 0.1, 0.13, 0.07
    rho = np.random.default_rng(5).uniform(0.0, 2.0, (nz, ny, nx))
    eps2 = (dx * dy * dz) ** (2.0 / 3.0)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, Z = (i * dx).ravel(), (j * dy).ravel(), (k * dz).ravel()
    r2 = (X[:, None] - X[None]) ** 2 + (Y[:, None] - Y[None]) ** 2 + (Z[:, None] - Z[None]) ** 2
    direct = (-1.0 / np.sqrt(r2 + eps2)) @ rho.ravel() * (dx * dy * dz)
    phi = PO.solve_poisson_isolated(rho, dx, dy, dz)
    assert np.abs(phi.ravel() - direct).max() <= 1e-13 * np.abs(direct).max()
