"""The oracle's phi gravity source (orc_apply_sg_force, after
_sg_gradient_kernel!, src/gpu_kernels.jl:291-319) and its FMR self-gravity
composition (src/fmr_gravity.jl:40-133, src/fmr3d.jl:451-473), pinned by
closed forms that a flipped sign, a wrong edge difference, a dropped or wrong
energy term or a transposed gradient component all fail, and
by the reference's own FMR self-gravity tests (test/test_fmr_gravity.jl:
58-152): the regression with self_gravity off, the composite solve against a
manual restriction and solve, the symmetric sphere's zero net momentum and
the infall of a self-gravitating sphere.

Every GPU phi test compares the device with orc_apply_sg_force bit for bit, so
these checks are what tie the device phi source to the reference.

The edge cells: the reference takes ip = min(ii + 1, n), im = max(ii - 1, 1)
and the factor inv2d when ip > im, 2 inv2d otherwise (src/gpu_kernels.jl:
300-306).  At ii = 1 and ii = n (n >= 2) ip > im holds, so the edge gradient
is the one-sided difference times inv2d = 1/(2d): half the one-sided slope.
The doubled factor is taken only when n = 1, where the difference is zero.
The closed forms below state that behaviour; the seemingly intended doubling
at the edges is one of the planted mistakes they catch."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import poisson as PO

NG = 3


def centres(n, d, x0):
    return x0 + (np.arange(1, n + 1) - 0.5) * d


def random_state(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    U = np.zeros((nz + 2 * NG, ny + 2 * NG, nx + 2 * NG, 5))
    a = U[NG:-NG, NG:-NG, NG:-NG]
    a[..., 0] = rng.uniform(0.5, 2.0, a.shape[:3])
    a[..., 1:4] = rng.uniform(-1.0, 1.0, a.shape[:3] + (3,))
    a[..., 4] = rng.uniform(2.0, 3.0, a.shape[:3])
    return U


def grid_case(phi_fn, nx=7, ny=5, nz=4, d=(0.1, 0.13, 0.07), x0=(-0.3, 0.2, -0.15), seed=1):
    """(U, phi, coordinates) for a potential given as a function of the cell
    centres; the state is random so that every source term is nonzero."""
    U = random_state(nx, ny, nz, seed)
    x = centres(nx, d[0], x0[0])[None, None, :]
    y = centres(ny, d[1], x0[1])[None, :, None]
    z = centres(nz, d[2], x0[2])[:, None, None]
    phi = np.ascontiguousarray(np.broadcast_to(phi_fn(x, y, z), (nz, ny, nx)), dtype=np.float64)
    return U, phi, (x, y, z)


def sg_source_errors(apply_fn, d=(0.1, 0.13, 0.07)):
    """Largest relative deviations of a phi-source implementation from the
    closed forms.  Linear phi = a x + b y + c z: the centred differences give
    exactly (a, b, c) inside, and the edge cells half of it, so dU_m = -rho g
    and dU_E = -(m . g) with g = (a, b, c) inside and halved per axis at that
    axis's edges.  Quadratic phi = x^2 + 2 y^2 - z^2: centred differences are
    exact inside (2x), and the edge difference (phi_2 - phi_1)/(2 dx) is
    x_1 + dx/2 at the low edge and x_n - dx/2 at the high edge."""
    errs = {}
    a, b, c = 0.7, -1.3, 2.1
    U, phi, _ = grid_case(lambda x, y, z: a * x + b * y + c * z, d=d)
    dU = np.zeros_like(U)
    apply_fn(dU, U, phi, *d)
    A = U[NG:-NG, NG:-NG, NG:-NG]
    D = dU[NG:-NG, NG:-NG, NG:-NG]
    nz, ny, nx = phi.shape
    G = np.empty((nz, ny, nx, 3))
    G[..., 0], G[..., 1], G[..., 2] = a, b, c
    G[:, :, [0, -1], 0] *= 0.5
    G[:, [0, -1], :, 1] *= 0.5
    G[[0, -1], :, :, 2] *= 0.5
    want_m = -A[..., 0:1] * G
    want_E = -(A[..., 1:4] * G).sum(-1)
    errs["linear_m"] = np.abs(D[..., 1:4] - want_m).max() / np.abs(want_m).max()
    errs["linear_E"] = np.abs(D[..., 4] - want_E).max() / np.abs(want_E).max()
    errs["linear_rho"] = np.abs(D[..., 0]).max()

    U, phi, (x, y, z) = grid_case(lambda x, y, z: x * x + 2 * y * y - z * z, d=d)
    dU = np.zeros_like(U)
    apply_fn(dU, U, phi, *d)
    A = U[NG:-NG, NG:-NG, NG:-NG]
    D = dU[NG:-NG, NG:-NG, NG:-NG]
    nz, ny, nx = phi.shape

    def one_sided(c, h, coef):
        gc = 2 * coef * c
        gc = gc.copy()
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        ax = [i for i in range(3) if c.shape[i] > 1][0]
        lo[ax], hi[ax] = 0, -1
        gc[tuple(lo)] = coef * (c[tuple(lo)] + 0.5 * h)
        gc[tuple(hi)] = coef * (c[tuple(hi)] - 0.5 * h)
        return gc

    gx = np.broadcast_to(one_sided(x, d[0], 1.0), (nz, ny, nx))
    gy = np.broadcast_to(one_sided(y, d[1], 2.0), (nz, ny, nx))
    gz = np.broadcast_to(one_sided(z, d[2], -1.0), (nz, ny, nx))
    G = np.stack([gx, gy, gz], -1)
    want_m = -A[..., 0:1] * G
    want_E = -(A[..., 1:4] * G).sum(-1)
    errs["quadratic_m"] = np.abs(D[..., 1:4] - want_m).max() / np.abs(want_m).max()
    errs["quadratic_E"] = np.abs(D[..., 4] - want_E).max() / np.abs(want_E).max()
    return errs


def planted(sign=1.0, edge=1.0, energy=True, swap_yz=False):
    """A numpy phi source with one planted mistake, to show that the checks
    above see it: sign of the source, the edge factor (2: the one-sided slope
    instead of the reference's half of it), the energy term, or the y and z
    gradient components exchanged.  The defaults are the reference's
    behaviour."""
    def apply(dU, U, phi, dx, dy, dz):
        nz, ny, nx = phi.shape
        A = U[NG:-NG, NG:-NG, NG:-NG]
        D = dU[NG:-NG, NG:-NG, NG:-NG]

        def grad(axis, h, n):
            p = np.moveaxis(phi, axis, 0)
            g = np.empty_like(p)
            if n == 1:
                g[:] = 0.0
            else:
                g[1:-1] = (p[2:] - p[:-2]) * (0.5 / h)
                g[0] = (p[1] - p[0]) * (edge * 0.5 / h)
                g[-1] = (p[-1] - p[-2]) * (edge * 0.5 / h)
            return np.moveaxis(g, 0, axis)

        gx, gy, gz = grad(2, dx, nx), grad(1, dy, ny), grad(0, dz, nz)
        if swap_yz:
            gy, gz = gz, gy
        r = A[..., 0]
        D[..., 1] -= sign * r * gx
        D[..., 2] -= sign * r * gy
        D[..., 3] -= sign * r * gz
        if energy:
            D[..., 4] -= sign * r * ((A[..., 1] * gx + A[..., 2] * gy + A[..., 3] * gz) / r)
    return apply


TOL = 1e-12


def test_oracle_sg_source_matches_closed_forms():
    errs = sg_source_errors(O.apply_sg_force)
    assert errs["linear_rho"] == 0.0
    for k, v in errs.items():
        assert v <= TOL, (k, v)


def test_correct_numpy_restatement_passes_the_same_checks():
    # The planted-mistake harness itself is right when nothing is planted.
    errs = sg_source_errors(planted())
    for k, v in errs.items():
        assert v <= TOL, (k, v)


@pytest.mark.parametrize("mistake,keys", [
    (dict(sign=-1.0), ("linear_m", "linear_E", "quadratic_m", "quadratic_E")),
    (dict(edge=2.0), ("linear_m", "linear_E", "quadratic_m", "quadratic_E")),
    (dict(edge=0.0), ("linear_m", "linear_E", "quadratic_m", "quadratic_E")),
    (dict(energy=False), ("linear_E", "quadratic_E")),
    (dict(swap_yz=True), ("linear_m", "linear_E", "quadratic_m")),
])
def test_planted_mistakes_fail_the_closed_forms(mistake, keys):
    errs = sg_source_errors(planted(**mistake))
    for k in keys:
        assert errs[k] > 1e-3, (mistake, k, errs[k])


def test_single_cell_axis_has_no_gradient():
    # ip = min(ii + 1, 1) = im = max(ii - 1, 1): the difference vanishes.
    U, phi, _ = grid_case(lambda x, y, z: x + 3 * y + 5 * z, nx=6, ny=1, nz=3)
    dU = np.zeros_like(U)
    O.apply_sg_force(dU, U, phi, 0.1, 0.13, 0.07)
    A = U[NG:-NG, NG:-NG, NG:-NG]
    D = dU[NG:-NG, NG:-NG, NG:-NG]
    assert np.array_equal(D[..., 2], np.zeros_like(D[..., 2]))
    gx = np.array([0.5, 1, 1, 1, 1, 0.5])[None, None, :]
    gz = np.array([0.5, 1, 0.5])[:, None, None] * 5.0
    np.testing.assert_allclose(D[..., 1], -A[..., 0] * gx, rtol=1e-12)
    np.testing.assert_allclose(D[..., 3], -A[..., 0] * gz, rtol=1e-12)


def test_source_adds_to_the_existing_rhs():
    # dU[q] -= ... (src/gpu_kernels.jl:314-317), not dU[q] = -...
    U, phi, _ = grid_case(lambda x, y, z: x * y + z)
    base = np.random.default_rng(3).normal(size=U.shape)
    dU0 = np.zeros_like(U)
    O.apply_sg_force(dU0, U, phi, 0.1, 0.13, 0.07)
    dU = base.copy()
    O.apply_sg_force(dU, U, phi, 0.1, 0.13, 0.07)
    np.testing.assert_allclose(dU - base, dU0, rtol=0, atol=1e-13)
    assert np.array_equal(dU[..., 0], base[..., 0])


# ---------------------------------------------------------------------------
# FMR self-gravity: the reference's _make_sg_grid (test/test_fmr_gravity.jl:
# 16-56), a uniform sphere on an 8^3 coarse level over [-L, L]^3 with a 4:1
# fine patch over coarse cells 3..6.

def make_sg_grid(nc=8, L=0.5, ratio=4, R=0.2, rho0=1.0, rho_bg=1e-4, gamma=5 / 3, e_int=1e-3):
    dxc = 2 * L / nc
    Uc = np.zeros((nc + 2 * NG,) * 3 + (5,))
    xc = centres(nc, dxc, -L)
    Z, Y, X = np.meshgrid(xc, xc, xc, indexing="ij")
    rho = np.where(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) < R, rho0, rho_bg)
    Uc[NG:-NG, NG:-NG, NG:-NG, 0] = rho
    Uc[NG:-NG, NG:-NG, NG:-NG, 4] = rho * e_int
    O.fill_ghost(Uc, "outflow")
    lo, hi = nc // 4 + 1, 3 * (nc // 4)
    lo_hi = (lo, hi, lo, hi, lo, hi)
    nf = (hi - lo + 1) * ratio
    dxf = dxc / ratio
    xf = centres(nf, dxf, (lo - 1) * dxc - L)
    Z, Y, X = np.meshgrid(xf, xf, xf, indexing="ij")
    Uf = np.zeros((nf + 2 * NG,) * 3 + (5,))
    rho = np.where(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) < R, rho0, rho_bg)
    Uf[NG:-NG, NG:-NG, NG:-NG, 0] = rho
    Uf[NG:-NG, NG:-NG, NG:-NG, 4] = rho * e_int
    O.prolong_fine_ghosts(Uf, Uc, lo_hi, ratio)
    return dict(Uc=Uc, Uf=Uf, dxc=dxc, dxf=dxf, lo_hi=lo_hi, ratio=ratio, gamma=gamma,
                xf=xf, L=L)


FLOORS = (1e-12, 1e-12)


def test_self_gravity_off_is_the_plain_step():
    # test/test_fmr_gravity.jl:58-70
    G1, G2 = make_sg_grid(), make_sg_grid()
    O.fmr3d_step(G1["Uc"], G1["Uf"], G1["dxc"], G1["lo_hi"], 4, G1["gamma"], 1e-4, "outflow",
                 *FLOORS)
    O.fmr3d_step(G2["Uc"], G2["Uf"], G2["dxc"], G2["lo_hi"], 4, G2["gamma"], 1e-4, "outflow",
                 *FLOORS, None, None)
    assert np.array_equal(G1["Uc"], G2["Uc"]) and np.array_equal(G1["Uf"], G2["Uf"])


def test_composite_solve_equals_manual_restrict_and_solve():
    # test/test_fmr_gravity.jl:72-89
    G = make_sg_grid()
    PO.restrict_density_to_coarse(G["Uc"], G["Uf"], G["lo_hi"], 4)
    d = G["dxc"]
    phi_ref = PO.solve_poisson_isolated(np.ascontiguousarray(G["Uc"][NG:-NG, NG:-NG, NG:-NG, 0]),
                                        d, d, d)
    G2 = make_sg_grid()
    phi_c, phi_f = PO.solve_fmr_poisson([G2["Uc"], G2["Uf"]], [(d, d, d)], [G2["lo_hi"]], [4])
    assert np.array_equal(phi_c, phi_ref)
    assert phi_f.shape == (16, 16, 16)


def test_symmetric_sphere_keeps_zero_net_momentum():
    # test/test_fmr_gravity.jl:91-120
    G = make_sg_grid()
    Uc, Uf = G["Uc"], G["Uf"]
    assert np.abs(Uc[NG:-NG, NG:-NG, NG:-NG, 1:4].sum((0, 1, 2))).max() < 1e-14
    phi_c, _ = PO.fmr3d_step_self_gravity(Uc, Uf, G["dxc"], G["lo_hi"], 4, G["gamma"], 1e-4,
                                          "outflow", *FLOORS)
    M_c = Uc[NG:-NG, NG:-NG, NG:-NG, 0].sum()
    Pc = Uc[NG:-NG, NG:-NG, NG:-NG, 1:4].sum((0, 1, 2))
    Pf = Uf[NG:-NG, NG:-NG, NG:-NG, 1:4].sum((0, 1, 2))
    assert np.all(np.abs(Pc) / M_c < 1e-10)
    assert np.all(np.abs(Pf) / M_c < 1e-10)
    # ... and the source did act: the step moved momentum cell by cell.
    assert np.abs(Uf[NG:-NG, NG:-NG, NG:-NG, 1:4]).max() > 1e-8
    assert phi_c.min() < 0.0


def radial_momentum(G):
    xf = G["xf"]
    Z, Y, X = np.meshgrid(xf, xf, xf, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    A = G["Uf"][NG:-NG, NG:-NG, NG:-NG]
    keep = r >= 1e-12
    rad = (A[..., 1] * X + A[..., 2] * Y + A[..., 3] * Z) / np.where(keep, r, 1.0)
    return (rad * keep).sum() * G["dxf"] ** 3


def test_self_gravitating_sphere_falls_inward():
    # test/test_fmr_gravity.jl:122-152: after one step the fine level's
    # integral of rho v . r_hat is negative.
    G = make_sg_grid(R=0.2, rho0=1.0, rho_bg=1e-6, e_int=1e-4)
    PO.fmr3d_step_self_gravity(G["Uc"], G["Uf"], G["dxc"], G["lo_hi"], 4, G["gamma"], 5e-4,
                               "outflow", *FLOORS)
    rm = radial_momentum(G)
    assert rm < 0.0
    # Without gravity the cold sphere's pressure pushes outward: the sign above
    # comes from the source, not from the hydrodynamics.
    H = make_sg_grid(R=0.2, rho0=1.0, rho_bg=1e-6, e_int=1e-4)
    O.fmr3d_step(H["Uc"], H["Uf"], H["dxc"], H["lo_hi"], 4, H["gamma"], 5e-4, "outflow", *FLOORS)
    assert radial_momentum(H) > 0.0


def test_G_grav_scales_the_potential():
    # fmr3d_step!(...; G_grav) multiplies both frozen potentials (src/fmr3d.jl:469-472).
    G = make_sg_grid()
    d = G["dxc"]
    phi1 = PO.solve_fmr_poisson([G["Uc"].copy(), G["Uf"]], [(d, d, d)], [G["lo_hi"]], [4])
    G2 = make_sg_grid()
    phi_c, phi_f = PO.fmr3d_step_self_gravity(G2["Uc"], G2["Uf"], d, G2["lo_hi"], 4, G2["gamma"],
                                              1e-4, "outflow", *FLOORS, G_grav=2.5)
    assert np.array_equal(phi_c, phi1[0] * 2.5)
    assert np.array_equal(phi_f, phi1[1] * 2.5)


# ---------------------------------------------------------------------------
# The two operators of the composite solve.

def test_restrict_density_is_the_r3_mean_and_touches_only_rho():
    rng = np.random.default_rng(11)
    nc, r = 6, 2
    lo_hi = (2, 4, 3, 5, 2, 3)
    Uc = rng.uniform(0.5, 1.5, (nc + 6, nc + 6, nc + 6, 5))
    nfx, nfy, nfz = 3 * r, 3 * r, 2 * r
    Uf = rng.uniform(0.5, 1.5, (nfz + 6, nfy + 6, nfx + 6, 5))
    before = Uc.copy()
    PO.restrict_density_to_coarse(Uc, Uf, lo_hi, r)
    Af = Uf[NG:-NG, NG:-NG, NG:-NG, 0]
    mean = Af.reshape(nfz // r, r, nfy // r, r, nfx // r, r).mean(axis=(1, 3, 5))
    box = (slice(NG + 1, NG + 3), slice(NG + 2, NG + 5), slice(NG + 1, NG + 4))
    np.testing.assert_allclose(Uc[box + (0,)], mean, rtol=1e-15)
    # mass: r^3 times the coarse box sum is the fine sum
    assert abs(Uc[box + (0,)].sum() * r ** 3 - Af.sum()) <= 1e-13 * Af.sum()
    changed = Uc != before
    assert not changed[..., 1:].any()
    outside = changed[..., 0].copy()
    outside[box] = False
    assert not outside.any()


def test_trilinear_prolongation_is_exact_on_multilinear_fields():
    # Trilinear interpolation reproduces a + bx + cy + dz + exy + fxyz exactly,
    # including the clamped edge cells, where it extrapolates from the last pair.
    ncx, ncy, ncz = 7, 6, 5
    f = lambda x, y, z: 0.3 + 1.1 * x - 0.7 * y + 0.4 * z + 0.25 * x * y - 0.05 * x * y * z
    x = np.arange(1, ncx + 1, dtype=np.float64)[None, None, :]
    y = np.arange(1, ncy + 1, dtype=np.float64)[None, :, None]
    z = np.arange(1, ncz + 1, dtype=np.float64)[:, None, None]
    phi_c = np.ascontiguousarray(np.broadcast_to(f(x, y, z), (ncz, ncy, ncx)))
    for lo, r, nf in (((1, 1, 1), 2, (14, 12, 10)), ((2, 3, 1), 4, (12, 8, 16)),
                      ((3, 2, 2), 2, (8, 6, 6))):
        out = PO.prolong_scalar_trilinear(phi_c, lo, r, nf)
        xf = (lo[0] - 1) + (np.arange(1, nf[0] + 1) - 0.5) / r + 0.5
        yf = (lo[1] - 1) + (np.arange(1, nf[1] + 1) - 0.5) / r + 0.5
        zf = (lo[2] - 1) + (np.arange(1, nf[2] + 1) - 0.5) / r + 0.5
        want = f(xf[None, None, :], yf[None, :, None], zf[:, None, None])
        assert out.shape == (nf[2], nf[1], nf[0])
        np.testing.assert_allclose(out, want, rtol=0, atol=1e-13)
