"""The oracle against the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py): a change of the oracle's arithmetic
shows here as a bit difference.  The Sod fixture is also held to the published
star-region values of the exact Riemann solution (Toro, Riemann Solvers and
Numerical Methods for Fluid Dynamics, Table 4.2, test 1)."""
import os

import numpy as np

from oracle import analytic
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_weno_and_hllc_vectors():
    z = load("weno_hllc.npz")
    a = z["weno_in"]
    wl = np.array([O.weno5_left(*r[0:5]) for r in a])
    wr = np.array([O.weno5_right(*r[1:6]) for r in a])
    np.testing.assert_array_equal(wl, z["weno_left"])
    np.testing.assert_array_equal(wr, z["weno_right"])
    g = float(z["hllc_gamma"])
    for i in range(len(z["hllc_left"])):
        for d in range(3):
            F = O.hllc_flux(d, z["hllc_left"][i], z["hllc_right"][i], g)
            np.testing.assert_array_equal(F, z["hllc_flux"][i, d])


def test_step_and_cfl_vectors():
    z = load("step_7x5x4.npz")
    d, dt = tuple(z["d"]), float(z["dt"])
    for name, bc in (("outflow", "outflow"), ("periodic", "periodic"),
                     ("mixed", ("periodic", "outflow", "periodic"))):
        V = z[f"{name}_in"].copy()
        O.euler3d_step(V, *d, dt, 1.4, bc, 1e-4, 1e-4)
        np.testing.assert_array_equal(V, z[f"{name}_out"])
    assert O.cfl_dt_3d(z["cfl_in"].copy(), *d, 1.4, 0.4) == float(z["cfl_dt"])


def test_fmr_vectors():
    z = load("fmr_6cube_r2.npz")
    Uc, Uf = z["coarse_in"].copy(), z["fine_in"].copy()
    lo_hi, r, dxc, gm = tuple(int(v) for v in z["lo_hi"]), int(z["ratio"]), float(z["dxc"]), float(z["gamma"])
    dt = O.cfl_dt_fmr3d(Uc, Uf, dxc, lo_hi, r, gm, 0.4)
    assert dt == float(z["dt"])
    O.fmr3d_step(Uc, Uf, dxc, lo_hi, r, gm, dt, "outflow", 1e-10, 1e-5)
    np.testing.assert_array_equal(Uc, z["coarse_out"])
    np.testing.assert_array_equal(Uf, z["fine_out"])


def test_sod_exact_profile_and_star_state():
    z = load("sod_exact_t0.2.npz")
    rho, v, p = analytic.sod_exact(z["x"] - 0.5, float(z["t"]))
    np.testing.assert_array_equal(rho, z["rho"])
    np.testing.assert_array_equal(v, z["v"])
    np.testing.assert_array_equal(p, z["p"])
    # star region: between the contact (x - 0.5 = u* t) and the shock
    xs = z["x"] - 0.5
    t = float(z["t"])
    left_star = (xs > 0.05) & (xs < 0.92745 * t - 0.01)
    right_star = (xs > 0.92745 * t + 0.01) & (xs < 1.75 * t - 0.01)
    assert left_star.any() and right_star.any()
    assert np.allclose(z["p"][left_star | right_star], 0.30313, atol=5e-5)
    assert np.allclose(z["v"][left_star | right_star], 0.92745, atol=5e-5)
    assert np.allclose(z["rho"][left_star], 0.42632, atol=5e-5)
    assert np.allclose(z["rho"][right_star], 0.26557, atol=5e-5)
