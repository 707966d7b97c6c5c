"""Host-side mirror of the reference's uniform-grid Euler entry points.

Same names (Julia's trailing ``!`` becomes a trailing ``_``), same argument
meaning and order, same in-place semantics as src/euler3d.jl, src/rk3.jl and
src/fmr_gravity.jl:146.  The work runs in libbsn_hip on the GPU.

``U`` may be
  * a :class:`Grid` -- device-resident state, the analogue of a device array
    in the reference (no host transfer per call), or
  * a numpy array of the reference layout (``(Nz, Ny, Nx, 5)`` C-order, the
    bytes of the reference's ``(5, Nx, Ny, Nz)`` column-major array), which
    is uploaded, advanced on the GPU and written back in place.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BSN_DEVICE, BSN_HOST, PREC, call

NG = 3

_BC = {"outflow": 0, ":outflow": 0, "periodic": 7, ":periodic": 7}


def bc_mask(bc) -> int:
    """bc as the reference's symbol ("outflow"/"periodic"), a per-axis
    sequence of those, or an int bit mask (bit a = periodic along axis a)."""
    if isinstance(bc, (int, np.integer)):
        return int(bc)
    if isinstance(bc, str):
        if bc not in _BC:
            raise ValueError(f"unknown bc {bc!r}")
        return _BC[bc]
    m = 0
    for a, b in enumerate(bc):
        if bc_mask(b) == 7:
            m |= 1 << a
    return m


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class Context:
    """One HIP device and stream."""

    _default = {}

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("bsn_ctx_create", device, C.byref(h))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def info(self) -> str:
        buf = C.create_string_buffer(256)
        call("bsn_ctx_info", self.handle, buf, 256)
        return buf.value.decode()

    def synchronize(self) -> None:
        call("bsn_ctx_synchronize", self.handle)


class Grid:
    """A uniform grid resident on the device (the reference's U plus its
    nx, ny, nz; cf. FMRLevel3D, src/fmr3d.jl:57-71)."""

    def __init__(self, nx: int, ny: int, nz: int, precision: str = "fp64",
                 ctx: Context | None = None):
        self.ctx = ctx or Context.default()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.precision = precision
        h = C.c_void_p()
        call("bsn_grid_create", self.ctx.handle, self.nx, self.ny, self.nz, PREC[precision],
             C.byref(h))
        self.handle = h
        self._phi = None

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().bsn_grid_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    @property
    def shape(self):
        return (self.nz + 2 * NG, self.ny + 2 * NG, self.nx + 2 * NG, 5)

    def set_state(self, U: np.ndarray) -> "Grid":
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.shape:
            raise ValueError(f"state shape {U.shape} != {self.shape}")
        call("bsn_grid_set_state", self.handle, _ptr(U), BSN_HOST)
        return self

    def get_state(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.shape)
        if out.shape != self.shape or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array of the state shape")
        call("bsn_grid_get_state", self.handle, _ptr(out), BSN_HOST)
        return out

    def set_phi(self, phi) -> "Grid":
        if phi is None:
            call("bsn_grid_set_phi", self.handle, None, BSN_HOST)
            self._phi = None
            return self
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if phi.shape != (self.nz, self.ny, self.nx):
            raise ValueError(f"phi shape {phi.shape} != {(self.nz, self.ny, self.nx)}")
        call("bsn_grid_set_phi", self.handle, _ptr(phi), BSN_HOST)
        self._phi = phi
        return self

    def set_params(self, dx, dy, dz, gamma, bc="outflow", rho_floor=0.0, P_floor=0.0) -> "Grid":
        call("bsn_grid_set_params", self.handle, float(dx), float(dy), float(dz), float(gamma),
             bc_mask(bc), float(rho_floor), float(P_floor))
        return self

    def conserved_totals(self, dV: float) -> np.ndarray:
        out = (C.c_double * 5)()
        call("bsn_conserved_totals", self.handle, float(dV), out)
        return np.array(list(out))


# ---------------------------------------------------------------------------
# numpy arrays go through a cached device grid per shape

_cache: dict = {}


def _as_grid(U, nx, ny, nz, precision="fp64"):
    if isinstance(U, Grid):
        if (U.nx, U.ny, U.nz) != (nx, ny, nz):
            raise ValueError("grid dims do not match nx, ny, nz")
        return U, False
    if not isinstance(U, np.ndarray) or U.dtype != np.float64 or not U.flags.c_contiguous:
        raise TypeError("U must be a Grid or a C-contiguous float64 numpy array")
    key = (nx, ny, nz, precision)
    g = _cache.get(key)
    if g is None:
        g = _cache[key] = Grid(nx, ny, nz, precision)
    g.set_state(U)
    return g, True


def _writeback(g: Grid, U, host: bool):
    if host:
        g.get_state(U)


def fill_ghost_3d_outflow_(U, nx, ny, nz):
    """fill_ghost_3d_outflow! (src/euler3d.jl:33-44)"""
    g, host = _as_grid(U, nx, ny, nz)
    call("bsn_fill_ghost_3d", g.handle, 0)
    _writeback(g, U, host)


def fill_ghost_3d_periodic_(U, nx, ny, nz):
    """fill_ghost_3d_periodic! (src/euler3d.jl:51-62)"""
    g, host = _as_grid(U, nx, ny, nz)
    call("bsn_fill_ghost_3d", g.handle, 7)
    _writeback(g, U, host)


def apply_floors_3d_(U, nx, ny, nz, rho_floor, P_floor, gamma):
    """apply_floors_3d! (src/euler3d.jl:73-80)"""
    g, host = _as_grid(U, nx, ny, nz)
    call("bsn_apply_floors_3d", g.handle, float(rho_floor), float(P_floor), float(gamma))
    _writeback(g, U, host)


def flux_shapes(nx, ny, nz):
    Nx, Ny, Nz = nx + 2 * NG, ny + 2 * NG, nz + 2 * NG
    return (Nz, Ny, Nx + 1, 5), (Nz, Ny + 1, Nx, 5), (Nz + 1, Ny, Nx, 5)


def euler3d_rhs_(dU, Fx, Fy, Fz, U, nx, ny, nz, dx, dy, dz, gamma, *, bc="outflow",
                 rho_floor=0.0, P_floor=0.0, fill_ghosts=True):
    """euler3d_rhs! (src/euler3d.jl:182-205): fills dU and Fx, Fy, Fz (host
    arrays of the reference shapes, any may be None) and leaves U with its
    ghosts and floors."""
    g, host = _as_grid(U, nx, ny, nz)
    outs = []
    for a, shp in zip((dU, Fx, Fy, Fz), ((g.shape,) + flux_shapes(nx, ny, nz))):
        if a is None:
            outs.append(None)
            continue
        if a.shape != shp or a.dtype != np.float64 or not a.flags.c_contiguous:
            raise ValueError(f"output of shape {a.shape}, expected {shp}")
        outs.append(_ptr(a))
    call("bsn_euler3d_rhs", g.handle, *outs, BSN_HOST, float(dx), float(dy), float(dz),
         float(gamma), bc_mask(bc), float(rho_floor), float(P_floor), int(bool(fill_ghosts)))
    _writeback(g, U, host)


def euler3d_step_(U, nx, ny, nz, dx, dy, dz, dt, gamma, *, bc="outflow", rho_floor=0.0,
                  P_floor=0.0):
    """euler3d_step! (src/euler3d.jl:220-260)"""
    g, host = _as_grid(U, nx, ny, nz)
    call("bsn_euler3d_step", g.handle, float(dx), float(dy), float(dz), float(dt), float(gamma),
         bc_mask(bc), float(rho_floor), float(P_floor))
    _writeback(g, U, host)


def cfl_dt_3d(U, nx, ny, nz, dx, dy, dz, gamma, cfl) -> float:
    """cfl_dt_3d (src/euler3d.jl:157-169)"""
    g, _ = _as_grid(U, nx, ny, nz)
    out = C.c_double()
    call("bsn_cfl_dt_3d", g.handle, float(dx), float(dy), float(dz), float(gamma), float(cfl),
         C.byref(out))
    return out.value


def _check_state_array(a, g: Grid, name: str) -> None:
    if (not isinstance(a, np.ndarray) or a.shape != g.shape or a.dtype != np.float64
            or not a.flags.c_contiguous):
        raise ValueError(f"{name} must be a C-contiguous float64 array of shape {g.shape}")


_sg_cache: dict = {}


def _apply_sg_force_level_(dU, U, phi, nx, ny, nz, dx, dy, dz):
    """_apply_sg_force_level! (src/fmr_gravity.jl:146-156) on a host dU.

    A numpy U goes through a grid of its own (not the per-shape grid the step
    functions share), so the phi set here never reaches a later step; a Grid
    U gets its previous phi (or none) back before returning."""
    if isinstance(U, Grid):
        g = U
        if (g.nx, g.ny, g.nz) != (nx, ny, nz):
            raise ValueError("grid dims do not match nx, ny, nz")
        own, prev = False, g._phi
    else:
        if not isinstance(U, np.ndarray) or U.dtype != np.float64 or not U.flags.c_contiguous:
            raise TypeError("U must be a Grid or a C-contiguous float64 numpy array")
        key = (nx, ny, nz)
        g = _sg_cache.get(key)
        if g is None:
            g = _sg_cache[key] = Grid(nx, ny, nz)
        g.set_state(U)
        own, prev = True, None
    _check_state_array(dU, g, "dU")
    g.set_phi(phi)
    try:
        call("bsn_apply_sg_force", g.handle, _ptr(dU), BSN_HOST, float(dx), float(dy), float(dz))
    finally:
        if not own:
            g.set_phi(prev)


def rk3_step_(U: Grid, L_, dt, tmp1, tmp2=None):
    """rk3_step! (src/rk3.jl:14-31) on a device grid: L_(tmp1, U) must fill the
    host array tmp1 with L(U); the base state Un lives on the device."""
    if not isinstance(U, Grid):
        raise TypeError("rk3_step_ advances a Grid")
    _check_state_array(tmp1, U, "tmp1")
    call("bsn_rk3_save", U.handle)
    for stage in (1, 2, 3):
        L_(tmp1, U)
        call("bsn_rk3_combine", U.handle, stage, _ptr(tmp1), BSN_HOST, float(dt))


def conserved_totals(U, nx, ny, nz, dV):
    """(sum rho, sum m, sum E) * dV over active cells (src/diagnostics.jl:21-58)."""
    g, _ = _as_grid(U, nx, ny, nz)
    return g.conserved_totals(dV)
