"""GPU: FMR self-gravity (src/fmr_gravity.jl, fmr3d_step!(...; self_gravity,
G_grav) of src/fmr3d.jl:451-473) through the C ABI against the oracle
(oracle/poisson.py).

The density restriction and the trilinear prolongation are compared bit for
bit.  The potential comes from the device FFT, which differs from the numpy
FFT in the last bits (about 1e-15 of max |phi|), so the potential is held to
1e-12 of its maximum and the stepped states to rtol 1e-12 on smooth states
(the reference's HLLC switches wave-speed estimates when a face pressure is
clamped to zero, a jump a last-bit change can flip; on a smooth state that
does not happen).  The reference's own physical checks (zero net momentum of
a symmetric sphere, infall of a self-gravitating sphere) run on the device
path as well.
"""
import numpy as np
import pytest

import binary_supernova_amd as bsn
from oracle import oracle as O
from oracle import poisson as PO

pytestmark = pytest.mark.gpu
NG = 3
FLOORS = (1e-12, 1e-12)


def centres(n, d, x0):
    return x0 + (np.arange(1, n + 1) - 0.5) * d


def sg_grid(nc=8, L=0.5, ratio=4, R=0.2, rho0=1.0, rho_bg=1e-4, gamma=5 / 3, e_int=1e-3,
            smooth=False):
    """The reference's _make_sg_grid (test/test_fmr_gravity.jl:16-56): a sphere
    on an nc^3 coarse level over [-L, L]^3 with a ratio:1 patch over coarse
    cells nc/4+1 .. 3nc/4.  smooth: a Gaussian of width R instead of a
    top hat."""
    dxc = 2 * L / nc

    def fill(U, xs):
        Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
        r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
        rho = rho_bg + rho0 * np.exp(-(r / R) ** 2) if smooth else np.where(r < R, rho0, rho_bg)
        U[NG:-NG, NG:-NG, NG:-NG, 0] = rho
        U[NG:-NG, NG:-NG, NG:-NG, 4] = rho * e_int

    Uc = np.zeros((nc + 2 * NG,) * 3 + (5,))
    fill(Uc, centres(nc, dxc, -L))
    O.fill_ghost(Uc, "outflow")
    lo, hi = nc // 4 + 1, 3 * (nc // 4)
    lo_hi = (lo, hi, lo, hi, lo, hi)
    nf = (hi - lo + 1) * ratio
    dxf = dxc / ratio
    xf = centres(nf, dxf, (lo - 1) * dxc - L)
    Uf = np.zeros((nf + 2 * NG,) * 3 + (5,))
    fill(Uf, xf)
    O.prolong_fine_ghosts(Uf, Uc, lo_hi, ratio)
    return dict(Uc=Uc, Uf=Uf, dxc=dxc, dxf=dxf, lo_hi=lo_hi, ratio=ratio, gamma=gamma, xf=xf,
                nc=nc, nf=nf)


def device(G, form="fused", floors=FLOORS):
    d, r = G["dxc"], G["ratio"]
    H = bsn.FMRHierarchy([(G["nc"],) * 3, (G["nf"],) * 3], [(d,) * 3, (d / r,) * 3],
                         [G["lo_hi"]], [r], G["gamma"], "outflow", *floors)
    H.set_mode(form == "inplace")
    H.levels[0].set_state(G["Uc"])
    H.levels[1].set_state(G["Uf"])
    return H


def phi_close(a, b):
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())


def test_restrict_density_and_prolong_scalar_bitwise():
    print("fmr gravity operators", flush=True)
    G = sg_grid(nc=12, ratio=2)
    rng = np.random.default_rng(3)
    G["Uf"][NG:-NG, NG:-NG, NG:-NG, 0] *= rng.uniform(0.5, 1.5, (G["nf"],) * 3)
    H = device(G)
    H.restrict_density(1)
    Uc = G["Uc"].copy()
    PO.restrict_density_to_coarse(Uc, G["Uf"], G["lo_hi"], 2)
    np.testing.assert_array_equal(H.levels[0].get_state(), Uc)
    phi_c = rng.normal(size=(12, 12, 12))
    want = PO.prolong_scalar_trilinear(phi_c, G["lo_hi"][0::2], 2, (G["nf"],) * 3)
    np.testing.assert_array_equal(H.prolong_scalar(1, phi_c), want)
    out = np.empty_like(want)
    bsn._prolong_scalar_trilinear_(out, phi_c, H)
    np.testing.assert_array_equal(out, want)


def test_solve_fmr_poisson_matches_oracle():
    # test/test_fmr_gravity.jl:72-89 on the device: the composite solve is the
    # restriction followed by the isolated solve of the coarse density
    print("fmr composite solve", flush=True)
    G = sg_grid()
    H = device(G)
    phis = bsn.solve_fmr_poisson(H)
    d = G["dxc"]
    Uc = G["Uc"].copy()
    ref = PO.solve_fmr_poisson([Uc, G["Uf"]], [(d, d, d)], [G["lo_hi"]], [4])
    np.testing.assert_array_equal(H.levels[0].get_state(), Uc)    # density restricted in place
    assert phis[1].shape == (16, 16, 16)
    phi_close(phis[0], ref[0])
    phi_close(phis[1], ref[1])
    # the fine phi is the device's own prolongation of the device's coarse phi
    np.testing.assert_array_equal(phis[1], H.prolong_scalar(1, phis[0]))


@pytest.mark.parametrize("form", ["fused", "inplace"])
@pytest.mark.parametrize("G_grav", [1.0, 2.5])
def test_self_gravity_step_matches_oracle(form, G_grav):
    print(f"fmr self-gravity step {form} G={G_grav}", flush=True)
    G = sg_grid(smooth=True, R=0.15, e_int=1e-2)
    H = device(G, form)
    Uc, Uf = G["Uc"].copy(), G["Uf"].copy()
    for dt in (2e-3, 1.5e-3):
        bsn.fmr3d_step_(H, dt, self_gravity=True, G_grav=G_grav)
        PO.fmr3d_step_self_gravity(Uc, Uf, G["dxc"], G["lo_hi"], 4, G["gamma"], dt, "outflow",
                                   *FLOORS, G_grav=G_grav)
    np.testing.assert_allclose(H.levels[0].get_state(), Uc, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(H.levels[1].get_state(), Uf, rtol=1e-12, atol=1e-15)
    # and the source acts: without it the fine momentum stays at rest longer
    V = sg_grid(smooth=True, R=0.15, e_int=1e-2)
    for dt in (2e-3, 1.5e-3):
        O.fmr3d_step(V["Uc"], V["Uf"], V["dxc"], V["lo_hi"], 4, V["gamma"], dt, "outflow", *FLOORS)
    assert np.abs(V["Uf"] - Uf).max() > 1e-6


def test_self_gravity_off_is_the_plain_step_and_level_phi_returns():
    # test/test_fmr_gravity.jl:58-70, and a static phi set on the levels is
    # theirs again after a self-gravity step
    print("fmr self-gravity off", flush=True)
    G = sg_grid(smooth=True, R=0.15, e_int=1e-2)
    H = device(G)
    bsn.fmr3d_step_(H, 1e-3, self_gravity=False)
    Uc, Uf = G["Uc"].copy(), G["Uf"].copy()
    O.fmr3d_step(Uc, Uf, G["dxc"], G["lo_hi"], 4, G["gamma"], 1e-3, "outflow", *FLOORS)
    np.testing.assert_array_equal(H.levels[0].get_state(), Uc)
    np.testing.assert_array_equal(H.levels[1].get_state(), Uf)
    rng = np.random.default_rng(5)
    phis = [0.01 * rng.normal(size=(8, 8, 8)), 0.01 * rng.normal(size=(16, 16, 16))]
    H.set_phis(phis)
    bsn.fmr3d_step_(H, 1e-3, self_gravity=True)
    PO.fmr3d_step_self_gravity(Uc, Uf, G["dxc"], G["lo_hi"], 4, G["gamma"], 1e-3, "outflow", *FLOORS)
    Uc = H.levels[0].get_state()
    Uf = H.levels[1].get_state()
    bsn.fmr3d_step_(H, 1e-3)
    O.fmr3d_step(Uc, Uf, G["dxc"], G["lo_hi"], 4, G["gamma"], 1e-3, "outflow", *FLOORS, *phis)
    np.testing.assert_array_equal(H.levels[0].get_state(), Uc)
    np.testing.assert_array_equal(H.levels[1].get_state(), Uf)


def test_symmetric_sphere_and_infall_on_the_device():
    # test/test_fmr_gravity.jl:91-152 through the device path
    print("fmr self-gravity sphere", flush=True)
    G = sg_grid()
    H = device(G)
    bsn.fmr3d_step_(H, 1e-4, self_gravity=True)
    Uc, Uf = H.levels[0].get_state(), H.levels[1].get_state()
    M_c = Uc[NG:-NG, NG:-NG, NG:-NG, 0].sum()
    for U in (Uc, Uf):
        P = U[NG:-NG, NG:-NG, NG:-NG, 1:4].sum((0, 1, 2))
        assert np.all(np.abs(P) / M_c < 1e-10)
    assert np.abs(Uf[NG:-NG, NG:-NG, NG:-NG, 1:4]).max() > 1e-8
    G = sg_grid(R=0.2, rho0=1.0, rho_bg=1e-6, e_int=1e-4)
    H = device(G)
    bsn.fmr3d_step_(H, 5e-4, self_gravity=True)
    A = H.levels[1].get_state()[NG:-NG, NG:-NG, NG:-NG]
    xf = G["xf"]
    Z, Y, X = np.meshgrid(xf, xf, xf, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    rad = (A[..., 1] * X + A[..., 2] * Y + A[..., 3] * Z) / np.where(r > 0, r, 1.0)
    assert rad.sum() < 0.0


def test_three_levels_and_run_match_oracle():
    # a chain of three levels (ratio 2 then 2): density restricted from the
    # finest up, phi prolonged down; then the device-dt run with the solve at
    # every root step against the oracle's host loop
    print("fmr self-gravity three levels", flush=True)
    gm, L, n0 = 5 / 3, 0.5, 16
    d0 = 2 * L / n0
    xs0 = centres(n0, d0, -L)

    def level(n, d, x0):
        U = np.zeros((n + 2 * NG,) * 3 + (5,))
        xs = centres(n, d, x0)
        Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
        rho = 1e-3 + np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 0.15 ** 2)
        U[NG:-NG, NG:-NG, NG:-NG, 0] = rho
        U[NG:-NG, NG:-NG, NG:-NG, 4] = rho * 1e-2
        return U

    b1, b2 = (5, 12, 5, 12, 5, 12), (5, 12, 5, 12, 5, 12)
    Us = [level(16, d0, -L), level(16, d0 / 2, -L + 4 * d0), level(16, d0 / 4, -L + 4 * d0 + 4 * d0 / 2)]
    O.fill_ghost(Us[0], "outflow")
    O.prolong_fine_ghosts(Us[1], Us[0], b1, 2)
    O.prolong_fine_ghosts(Us[2], Us[1], b2, 2)
    dxs = [(d0,) * 3, (d0 / 2,) * 3, (d0 / 4,) * 3]
    for form in ("fused", "inplace"):
        H = bsn.FMRHierarchy([(16,) * 3] * 3, dxs, [b1, b2], [2, 2], gm, "outflow", *FLOORS)
        H.set_mode(form == "inplace")
        for g, U in zip(H.levels, Us):
            g.set_state(U)
        V = [U.copy() for U in Us]
        for dt in (2e-3, 1e-3):
            H.step(dt, self_gravity=True, G_grav=1.5)
            phis = PO.solve_fmr_poisson(V, dxs, [b1, b2], [2, 2])
            O.fmr_multilevel_step(V, dxs, [b1, b2], [2, 2], gm, dt, "outflow", *FLOORS,
                                  phis=[p * 1.5 for p in phis])
        for g, U in zip(H.levels, V):
            np.testing.assert_allclose(g.get_state(), U, rtol=1e-12, atol=1e-15)
    # device dt (cfl_dt_fmr3d of the step's start state) with the solve per
    # root step, two levels
    G = sg_grid(smooth=True, R=0.15, e_int=1e-2)
    H = device(G)
    H.set_self_gravity(True, 1.0)
    H.run(0.4, 2)
    Uc, Uf = G["Uc"].copy(), G["Uf"].copy()
    t = 0.0
    for _ in range(2):
        dt = O.cfl_dt_fmr3d(Uc, Uf, G["dxc"], G["lo_hi"], 4, G["gamma"], 0.4)
        PO.fmr3d_step_self_gravity(Uc, Uf, G["dxc"], G["lo_hi"], 4, G["gamma"], dt, "outflow",
                                   *FLOORS)
        t += dt
    np.testing.assert_allclose(H.time()[0], t, rtol=1e-12)
    np.testing.assert_allclose(H.levels[0].get_state(), Uc, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(H.levels[1].get_state(), Uf, rtol=1e-12, atol=1e-15)
