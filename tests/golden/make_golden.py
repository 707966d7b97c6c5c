"""Writes the golden fixtures of this directory (run from the repository root:
python tests/golden/make_golden.py).

The reference ships no data fixtures (its tests use tolerances only, e.g.
test/test_weno5.jl, test/test_sod3d.jl, test/test_fmr3d.jl) and its Julia
solver cannot run in this environment, so the vectors here are outputs of the
CPU oracle on seeded inputs, plus the exact Riemann solution of the Sod tube
(restated from test/test_sod3d.jl:13-75).  They pin the oracle against
regressions (tests/test_golden.py) and give the GPU tests committed data to
match bit for bit.  All files are .npz without pickled objects.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import analytic  # noqa: E402
from oracle import oracle as O  # noqa: E402

NG = 3


def _pkg():
    import importlib.util
    name = "binary_supernova_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(ROOT, "binary-supernova_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def random_state(nx, ny, nz, seed, g=1.4):
    """Random primitives over every padded cell; ~3 % of the cells carry
    states that drive the guarded branches (negative or tiny density,
    |v| > 100, negative pressure)."""
    rng = np.random.default_rng(seed)
    shp = (nz + 2 * NG, ny + 2 * NG, nx + 2 * NG)
    rho = rng.uniform(0.2, 2.0, shp)
    v = rng.normal(0.0, 1.0, shp + (3,))
    p = rng.uniform(0.1, 2.0, shp)
    m = rng.random(shp)
    rho = np.where(m < 0.005, -rng.uniform(0, 1, shp), rho)
    rho = np.where((m >= 0.005) & (m < 0.01), 1e-7, rho)
    v = np.where(((m >= 0.01) & (m < 0.02))[..., None], v * 300.0, v)
    p = np.where((m >= 0.02) & (m < 0.03), -0.5, p)
    U = np.zeros(shp + (5,))
    U[..., 0] = rho
    U[..., 1:4] = rho[..., None] * v
    U[..., 4] = p / (g - 1) + 0.5 * rho * (v * v).sum(-1)
    return U


def weno_hllc_vectors():
    rng = np.random.default_rng(7)
    n = 96
    a = rng.normal(0.0, 1.0, (n, 6))
    a[n // 3: 2 * n // 3, 3:] += 5.0            # a jump inside the stencil
    a[2 * n // 3:] = np.cumsum(rng.uniform(0, 1, (n - 2 * n // 3, 6)), axis=1) ** 2
    wl = np.array([O.weno5_left(*r[0:5]) for r in a])
    wr = np.array([O.weno5_right(*r[1:6]) for r in a])
    # HLLC states: random, plus supersonic, strong density ratio, vacuum-ish
    g = 1.4
    m = 64
    prim = np.empty((m, 2, 5))
    prim[:, :, 0] = rng.uniform(0.1, 2.0, (m, 2))
    prim[:, :, 1:4] = rng.normal(0, 1, (m, 2, 3))
    prim[:, :, 4] = rng.uniform(0.05, 2.0, (m, 2))
    prim[8:16, :, 1] += 6.0                       # supersonic to the right
    prim[16:24, :, 1] -= 6.0                      # supersonic to the left
    prim[24:32, 0, 0] *= 10.0                     # density ratio > 3 (HLL branch)
    prim[32:36, :, 0] = -0.1                      # non-positive density reset
    prim[36:40, :, 1] *= 400.0                    # velocity cap
    prim[40:44, :, 4] = -0.2                      # E below KE
    cons = np.empty_like(prim)
    cons[..., 0] = prim[..., 0]
    cons[..., 1:4] = prim[..., 0:1] * prim[..., 1:4]
    cons[..., 4] = prim[..., 4] / (g - 1) + 0.5 * prim[..., 0] * (prim[..., 1:4] ** 2).sum(-1)
    F = np.array([[O.hllc_flux(d, cons[i, 0], cons[i, 1], g) for d in range(3)] for i in range(m)])
    return dict(weno_in=a, weno_left=wl, weno_right=wr, hllc_left=cons[:, 0], hllc_right=cons[:, 1],
                hllc_gamma=np.float64(g), hllc_flux=F)


def step_vectors():
    out = {}
    nx, ny, nz = 7, 5, 4
    d = (0.1, 0.12, 0.09)
    for name, bc, seed in (("outflow", "outflow", 11), ("periodic", "periodic", 12),
                           ("mixed", ("periodic", "outflow", "periodic"), 13)):
        U = random_state(nx, ny, nz, seed)
        V = U.copy()
        dt = 0.004
        O.euler3d_step(V, *d, dt, 1.4, bc, 1e-4, 1e-4)
        out[f"{name}_in"] = U
        out[f"{name}_out"] = V
    out["d"] = np.array(d)
    out["dt"] = np.float64(0.004)
    # one CFL dt of a wild state
    U = random_state(nx, ny, nz, 14)
    out["cfl_in"] = U
    out["cfl_dt"] = np.float64(O.cfl_dt_3d(U, *d, 1.4, 0.4))
    return out


def fmr_vectors():
    P = _pkg().problems
    gm, nc, L, r = 5 / 3, 6, 0.5, 2
    dxc = 2 * L / nc
    lo_hi = (2, 4, 2, 4, 3, 5)
    Uc = P.sedov_ic_3d(nc, nc, nc, dxc, dxc, dxc, gm, E_blast=1.0, r_inject=0.2, rho_bg=1.0,
                       P_floor=1e-5, x_offset=-L, y_offset=-L, z_offset=-L)
    O.fill_ghost(Uc, "outflow")
    Uf = P.empty_state(6, 6, 6)
    O.prolong_all(Uf, Uc, lo_hi, r)
    out = dict(coarse_in=Uc.copy(), fine_in=Uf.copy(), lo_hi=np.array(lo_hi), ratio=np.int64(r),
               dxc=np.float64(dxc), gamma=np.float64(gm))
    dt = O.cfl_dt_fmr3d(Uc, Uf, dxc, lo_hi, r, gm, 0.4)
    O.fmr3d_step(Uc, Uf, dxc, lo_hi, r, gm, dt, "outflow", 1e-10, 1e-5)
    out.update(dt=np.float64(dt), coarse_out=Uc, fine_out=Uf)
    return out


def sod_exact_profile():
    n = 128
    x = (np.arange(n) + 0.5) / n
    rho, v, p = analytic.sod_exact(x - 0.5, 0.2)
    # star-region values the exact solver must reproduce (Toro, Table 4.2,
    # test 1: p* = 0.30313, u* = 0.92745, rho*L = 0.42632, rho*R = 0.26557)
    return dict(x=x, t=np.float64(0.2), rho=rho, v=v, p=p)


FILES = {
    "weno_hllc.npz": weno_hllc_vectors,
    "step_7x5x4.npz": step_vectors,
    "fmr_6cube_r2.npz": fmr_vectors,
    "sod_exact_t0.2.npz": sod_exact_profile,
}


def main():
    for name, fn in FILES.items():
        path = os.path.join(HERE, name)
        np.savez_compressed(path, **fn())
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
