"""ctypes front end of the CPU FP64 oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module, and only as the checker.  The product package never does.

Arrays follow the reference layout: a conserved state U of the reference's
shape ``(5, Nx, Ny, Nz)`` column-major is held here as a C-contiguous numpy
array of shape ``(Nz, Ny, Nx, 5)`` -- the same bytes.  phi is ``(nz, ny, nx)``
C-order, i.e. the reference's ``(nx, ny, nz)`` column-major.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsn_oracle.so")
NG = 3

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_d = C.c_double
_i64 = C.c_int64
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_pi64 = C.POINTER(C.c_int64)


def _declare(L):
    L.orc_wenoz_weights.argtypes = [_d] * 6 + [_pd] * 3
    L.orc_weno5_left.argtypes = [_d] * 5
    L.orc_weno5_left.restype = _d
    L.orc_weno5_right.argtypes = [_d] * 5
    L.orc_weno5_right.restype = _d
    L.orc_hllc_normal.argtypes = [_d] * 11 + [_pd]
    L.orc_hllc_flux.argtypes = [C.c_int, _pd, _pd, _d, _pd]
    L.orc_cell_pressure.argtypes = [_d] * 6
    L.orc_cell_pressure.restype = _d
    L.orc_EL_from_P.argtypes = [_d] * 6
    L.orc_EL_from_P.restype = _d
    L.orc_fill_ghost.argtypes = [_pd, _i64, _i64, _i64, _pi]
    L.orc_apply_floors.argtypes = [_pd, _i64, _i64, _i64, _d, _d, _d]
    L.orc_weno_fluxes.argtypes = [C.c_int, _pd, _pd, _i64, _i64, _i64, _d]
    L.orc_flux_divergence.argtypes = [_pd, _pd, _pd, _pd, _i64, _i64, _i64, _d, _d, _d]
    L.orc_apply_sg_force.argtypes = [_pd, _pd, _pd, _i64, _i64, _i64, _d, _d, _d]
    L.orc_cfl_dt_3d.argtypes = [_pd, _i64, _i64, _i64, _d, _d, _d, _d, _d]
    L.orc_cfl_dt_3d.restype = _d
    L.orc_euler3d_rhs.argtypes = [_pd, _pd, _pd, _pd, _pd, _i64, _i64, _i64, _d, _d, _d,
                                  _d, _pi, _d, _d, C.c_int]
    L.orc_rk3_combine.argtypes = [C.c_int, _pd, _pd, _pd, _d, _i64]
    L.orc_euler3d_step.argtypes = [_pd, _i64, _i64, _i64, _d, _d, _d, _d, _d, _pi, _d, _d, _pd]
    L.orc_euler3d_step.restype = C.c_int
    L.orc_conserved_totals.argtypes = [_pd, _i64, _i64, _i64, _d, _pd]
    L.orc_lag5_weights.argtypes = [C.c_int, _pd]
    L.orc_prolong_all.argtypes = [_pd, _pd, _i64, _i64, _i64, _pi64, C.c_int]
    L.orc_prolong_fine_ghosts.argtypes = [_pd, _pd, _i64, _i64, _i64, _pi64, C.c_int]
    L.orc_restrict.argtypes = [_pd, _pd, _i64, _i64, _i64, _pi64, C.c_int]
    L.orc_fmr3d_step.argtypes = [_pd, _pd, _i64, _i64, _i64, _d, _d, _d, _pi64, C.c_int,
                                 _d, _pi, _d, _d, _d, _pd, _pd]
    L.orc_fmr3d_step.restype = C.c_int
    L.orc_cfl_dt_fmr3d.argtypes = [_pd, _pd, _i64, _i64, _i64, _d, _d, _d, _pi64, C.c_int,
                                   _d, _d]
    L.orc_cfl_dt_fmr3d.restype = _d
    L.orc_fmr_multilevel_step.argtypes = [C.c_int, C.POINTER(_pd), _pi64, _pd, _pi64, _pi,
                                          _d, _pi, _d, _d, _d, C.POINTER(_pd)]
    L.orc_fmr_multilevel_step.restype = C.c_int
    L.orc_bh_gravity_source.argtypes = [_pd, _pd, _i64, _i64, _i64, _d, _d, _d, _d, _d, _d,
                                        C.c_int, _pd, _pd, _pd, _pd, _pd]
    L.orc_sink_sources.argtypes = [_pd, _pd, _i64, _i64, _i64, _d, _d, _d, _d, _d, _d, C.c_int,
                                   _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, C.c_int, _d]
    L.orc_coupled_step.argtypes = [_pd, _i64, _i64, _i64, _d, _d, _d, _d, _d, _pi, _d, _d, _pd,
                                   _d, _d, _d, C.c_int, _pd, _pd, _pd, _pd, C.c_int, _pd, _pd,
                                   C.c_int, _d]
    L.orc_coupled_step.restype = C.c_int
    L.orc_gas_force_on_bh.argtypes = [_pd, _i64, _i64, _i64] + [_d] * 11 + [_pd]
    L.orc_accrete_sums.argtypes = [_pd, _i64, _i64, _i64] + [_d] * 13 + [_pd]
    L.orc_gas_diagnostics.argtypes = [_pd, _i64, _i64, _i64] + [_d] * 6 + [C.c_int] + \
        [_pd] * 5 + [_d, _pd]


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_pd)


def _bc(bc):
    if isinstance(bc, str):
        bc = [bc] * 3
    return (C.c_int * 3)(*[1 if b in ("periodic", 1) else 0 for b in bc])


def _dims(U):
    Nz, Ny, Nx, five = U.shape
    assert five == 5
    return Nx - 2 * NG, Ny - 2 * NG, Nz - 2 * NG


# -- scalar functions -------------------------------------------------------

def weno5_left(*v):
    return lib().orc_weno5_left(*[float(x) for x in v])


def weno5_right(*v):
    return lib().orc_weno5_right(*[float(x) for x in v])


def wenoz_weights(b0, b1, b2, d0, d1, d2):
    w = [C.c_double() for _ in range(3)]
    lib().orc_wenoz_weights(b0, b1, b2, d0, d1, d2, *[C.byref(x) for x in w])
    return tuple(x.value for x in w)


def hllc_flux(direction, L, R, gamma):
    d = {"x": 0, "y": 1, "z": 2}[direction] if isinstance(direction, str) else direction
    Lp = (C.c_double * 5)(*L)
    Rp = (C.c_double * 5)(*R)
    F = (C.c_double * 5)()
    lib().orc_hllc_flux(d, Lp, Rp, gamma, F)
    return np.array(list(F))


def cell_pressure(r, mx, my, mz, E, g):
    return lib().orc_cell_pressure(r, mx, my, mz, E, g)


# -- grid functions (in place on U) -----------------------------------------

def fill_ghost(U, bc="outflow"):
    nx, ny, nz = _dims(U)
    lib().orc_fill_ghost(_p(U), nx, ny, nz, _bc(bc))


def apply_floors(U, rho_floor, P_floor, gamma):
    nx, ny, nz = _dims(U)
    lib().orc_apply_floors(_p(U), nx, ny, nz, rho_floor, P_floor, gamma)


def flux_shapes(nx, ny, nz):
    Nx, Ny, Nz = nx + 2 * NG, ny + 2 * NG, nz + 2 * NG
    return ((Nz, Ny, Nx + 1, 5), (Nz, Ny + 1, Nx, 5), (Nz + 1, Ny, Nx, 5))


def euler3d_rhs(U, dx, dy, dz, gamma, bc="outflow", rho_floor=0.0, P_floor=0.0,
                fill_ghosts=True):
    """Returns (dU, Fx, Fy, Fz); U is modified in place (ghosts, floors)."""
    nx, ny, nz = _dims(U)
    sx, sy, sz = flux_shapes(nx, ny, nz)
    dU = np.zeros_like(U)
    Fx, Fy, Fz = np.zeros(sx), np.zeros(sy), np.zeros(sz)
    lib().orc_euler3d_rhs(_p(dU), _p(Fx), _p(Fy), _p(Fz), _p(U), nx, ny, nz, dx, dy, dz,
                          gamma, _bc(bc), rho_floor, P_floor, int(bool(fill_ghosts)))
    return dU, Fx, Fy, Fz


def euler3d_step(U, dx, dy, dz, dt, gamma, bc="outflow", rho_floor=0.0, P_floor=0.0,
                 phi=None):
    nx, ny, nz = _dims(U)
    if phi is not None:
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert phi.shape == (nz, ny, nx)
    rc = lib().orc_euler3d_step(_p(U), nx, ny, nz, dx, dy, dz, dt, gamma, _bc(bc),
                                rho_floor, P_floor, _p(phi))
    if rc:
        raise MemoryError("oracle euler3d_step allocation failed")


def cfl_dt_3d(U, dx, dy, dz, gamma, cfl):
    nx, ny, nz = _dims(U)
    return lib().orc_cfl_dt_3d(_p(U), nx, ny, nz, dx, dy, dz, gamma, cfl)


def apply_sg_force(dU, U, phi, dx, dy, dz):
    nx, ny, nz = _dims(U)
    lib().orc_apply_sg_force(_p(dU), _p(U), _p(np.ascontiguousarray(phi)), nx, ny, nz,
                             dx, dy, dz)


def rk3_combine(stage, U, Un, dU, dt):
    lib().orc_rk3_combine(stage, _p(U), _p(Un), _p(dU), dt, U.size)


def conserved_totals(U, dV):
    nx, ny, nz = _dims(U)
    out = (C.c_double * 5)()
    lib().orc_conserved_totals(_p(U), nx, ny, nz, dV, out)
    return np.array(list(out))


# -- FMR ---------------------------------------------------------------------

def _lohi(lo_hi):
    return (C.c_int64 * 6)(*[int(v) for v in lo_hi])


def lag5_weights(ratio):
    w = np.zeros((ratio, 5))
    lib().orc_lag5_weights(ratio, _p(w))
    return w


def prolong_all(Uf, Uc, lo_hi, ratio):
    nx, ny, nz = _dims(Uc)
    lib().orc_prolong_all(_p(Uf), _p(Uc), nx, ny, nz, _lohi(lo_hi), ratio)


def prolong_fine_ghosts(Uf, Uc, lo_hi, ratio):
    nx, ny, nz = _dims(Uc)
    lib().orc_prolong_fine_ghosts(_p(Uf), _p(Uc), nx, ny, nz, _lohi(lo_hi), ratio)


def restrict(Uc, Uf, lo_hi, ratio):
    nx, ny, nz = _dims(Uc)
    lib().orc_restrict(_p(Uc), _p(Uf), nx, ny, nz, _lohi(lo_hi), ratio)


def fmr3d_step(Uc, Uf, dxc, lo_hi, ratio, gamma, dt_c, bc="outflow", rho_floor=0.0,
               P_floor=0.0, phi_c=None, phi_f=None):
    nx, ny, nz = _dims(Uc)
    dx, dy, dz = (dxc, dxc, dxc) if np.isscalar(dxc) else dxc
    rc = lib().orc_fmr3d_step(_p(Uc), _p(Uf), nx, ny, nz, dx, dy, dz, _lohi(lo_hi), ratio,
                              gamma, _bc(bc), rho_floor, P_floor, dt_c, _p(phi_c), _p(phi_f))
    if rc:
        raise MemoryError("oracle fmr3d_step allocation failed")


def cfl_dt_fmr3d(Uc, Uf, dxc, lo_hi, ratio, gamma, cfl):
    nx, ny, nz = _dims(Uc)
    dx, dy, dz = (dxc, dxc, dxc) if np.isscalar(dxc) else dxc
    return lib().orc_cfl_dt_fmr3d(_p(Uc), _p(Uf), nx, ny, nz, dx, dy, dz, _lohi(lo_hi),
                                  ratio, gamma, cfl)


def fmr_multilevel_step(Us, dxs, lo_his, ratios, gamma, dt, bc="outflow", rho_floor=0.0,
                        P_floor=0.0, phis=None):
    """Us: list of level states (root first); dxs: list of (dx,dy,dz);
    lo_his[l]: box of level l+1 in level l's 1-based active indices."""
    nlev = len(Us)
    dims = (C.c_int64 * (3 * nlev))(*[d for U in Us for d in _dims(U)])
    dx = (C.c_double * (3 * nlev))(*[float(v) for t in dxs for v in t])
    lh = (C.c_int64 * max(1, 6 * (nlev - 1)))(*[int(v) for t in lo_his for v in t])
    rt = (C.c_int * max(1, nlev - 1))(*[int(r) for r in ratios])
    Up = (_pd * nlev)(*[_p(U) for U in Us])
    if phis is None:
        ph = None
    else:
        ph = (_pd * nlev)(*[_p(p) if p is not None else None for p in phis])
    rc = lib().orc_fmr_multilevel_step(nlev, Up, dims, dx, lh, rt, gamma, _bc(bc),
                                       rho_floor, P_floor, dt, ph)
    if rc:
        raise MemoryError("oracle multilevel step failed")


# -- point masses and sinks (src/gpu_kernels.jl:199-284) ---------------------

def _v(vals):
    a = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).ravel())
    return a


def bh_gravity_source(dU, U, dx, dy, dz, x0, y0, z0, pos, mass, eps):
    """_bh_gravity_kernel! over the active cells; pos is (n, 3)."""
    nx, ny, nz = _dims(U)
    p = _v(pos).reshape(-1, 3)
    px, py, pz = (np.ascontiguousarray(p[:, a]) for a in range(3))
    m, e = _v(mass), _v(eps)
    lib().orc_bh_gravity_source(_p(dU), _p(U), nx, ny, nz, dx, dy, dz, x0, y0, z0, len(m),
                                _p(px), _p(py), _p(pz), _p(m), _p(e))


def sink_sources(dU, U, dx, dy, dz, x0, y0, z0, pos, vel, rs, ts, torque_free=True,
                 rho_sink_min=0.0):
    """_sink_sources_kernel! over the active cells; pos, vel are (n, 3)."""
    nx, ny, nz = _dims(U)
    p = _v(pos).reshape(-1, 3)
    v = _v(vel).reshape(-1, 3)
    cols = [np.ascontiguousarray(a[:, c]) for a in (p, v) for c in range(3)]
    r, t = _v(rs), _v(ts)
    lib().orc_sink_sources(_p(dU), _p(U), nx, ny, nz, dx, dy, dz, x0, y0, z0, len(r),
                           *[_p(c) for c in cols], _p(r), _p(t), int(bool(torque_free)),
                           rho_sink_min)


def coupled_step(U, dx, dy, dz, dt, gamma, bc="outflow", rho_floor=0.0, P_floor=0.0, phi=None,
                 origin=(0.0, 0.0, 0.0), pos=(), vel=(), mass=(), eps=(), nsink=0, rs=(), ts=(),
                 torque_free=True, rho_sink_min=0.0):
    """The hydrodynamic part of coupled_step! (scripts/run_sn50_fiducial.jl:
    772-816): gravity from every body, sinks from the first nsink."""
    nx, ny, nz = _dims(U)
    p, v, m, e = _v(pos), _v(vel), _v(mass), _v(eps)
    r, t = _v(rs), _v(ts)
    rc = lib().orc_coupled_step(_p(U), nx, ny, nz, dx, dy, dz, dt, gamma, _bc(bc), rho_floor,
                                P_floor, _p(phi), *origin, len(m), _p(p), _p(v), _p(m), _p(e),
                                int(nsink), _p(r), _p(t), int(bool(torque_free)), rho_sink_min)
    if rc:
        raise ValueError(f"oracle coupled_step failed ({rc})")


def gas_force_on_bh(U, dx, dy, dz, pos, mass, eps, x0, y0, z0):
    """Force of the gas on one body (src/gravity_bh.jl:91-113), sequential sum."""
    nx, ny, nz = _dims(U)
    F = np.zeros(3)
    lib().orc_gas_force_on_bh(_p(U), nx, ny, nz, dx, dy, dz, float(pos[0]), float(pos[1]),
                              float(pos[2]), float(mass), float(eps), x0, y0, z0, _p(F))
    return F


def accrete_sums(U, dx, dy, dz, x0, y0, z0, pos, rs, ts, dt, rho_sink_min=0.0):
    """(dm, dPx, dPy, dPz) of accrete! for one body (src/sinks.jl:86-103)."""
    nx, ny, nz = _dims(U)
    out = np.zeros(4)
    lib().orc_accrete_sums(_p(U), nx, ny, nz, dx, dy, dz, x0, y0, z0, float(pos[0]),
                           float(pos[1]), float(pos[2]), float(rs), float(ts), float(dt),
                           float(rho_sink_min), _p(out))
    return out


def accrete(mass, vel, sums):
    """The body update of accrete! (src/sinks.jl:104-108): (mass, vel) after
    accreting `sums` = (dm, dPx, dPy, dPz)."""
    dm = sums[0]
    vel = [float(v) for v in vel]
    if dm > 0.0:
        M_new = mass + dm
        vel = [(mass * vel[c] + sums[1 + c]) / M_new for c in range(3)]
        mass = M_new
    return mass, vel


def gas_diagnostics(U, dx, dy, dz, x0, y0, z0, pos=(), mass=(), eps=(), gamma=5.0 / 3.0):
    """(E, KE, Px, Py, Pz, Lx, Ly, Lz, M_bound) of src/diagnostics.jl:21-175,
    sequential sums; pos holds 3 values per body."""
    nx, ny, nz = _dims(U)
    n = len(mass)
    pos = np.asarray(pos, dtype=np.float64).reshape(n, 3) if n else np.zeros((1, 3))
    px, py, pz = (np.ascontiguousarray(pos[:, c]) for c in range(3))
    m = np.ascontiguousarray(mass, dtype=np.float64) if n else np.zeros(1)
    e = np.ascontiguousarray(eps, dtype=np.float64) if n else np.zeros(1)
    out = np.zeros(9)
    lib().orc_gas_diagnostics(_p(U), nx, ny, nz, dx, dy, dz, x0, y0, z0, n, _p(px), _p(py),
                              _p(pz), _p(m), _p(e), float(gamma), _p(out))
    return out

