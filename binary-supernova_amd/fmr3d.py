"""Host-side mirror of the reference's fixed-mesh-refinement API
(src/fmr3d.jl): FMRGrid3D, fmr3d_step!, cfl_dt_fmr3d, the prolongation and
restriction helpers, and a multi-level generalisation (levels nested
recursively, each advanced by the two-level algorithm of fmr3d_step!).
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from ._lib import BSN_HOST, call
from .euler3d import Context, Grid, bc_mask

NG = 3


class FMRHierarchy:
    """Levels 0..L-1 on the device; level l+1 covers the 1-based active box
    boxes[l] = (ci_lo, ci_hi, cj_lo, cj_hi, ck_lo, ck_hi) of level l with
    refinement ratios[l]."""

    def __init__(self, dims, dxs, boxes, ratios, gamma, bc="outflow", rho_floor=0.0,
                 P_floor=0.0, precision="fp64", ctx: Context | None = None):
        self.ctx = ctx or Context.default()
        self.levels = []
        for (nx, ny, nz), (dx, dy, dz) in zip(dims, dxs):
            g = Grid(nx, ny, nz, precision, self.ctx)
            g.set_params(dx, dy, dz, gamma, bc, rho_floor, P_floor)
            self.levels.append(g)
        self.dims = [tuple(int(v) for v in d) for d in dims]
        self.dxs = [tuple(float(v) for v in d) for d in dxs]
        self.boxes = [tuple(int(v) for v in b) for b in boxes]
        self.ratios = [int(r) for r in ratios]
        self.gamma, self.bc = float(gamma), bc_mask(bc)
        self.rho_floor, self.P_floor = float(rho_floor), float(P_floor)
        n = len(self.levels)
        arr = (C.c_void_p * n)(*[g.handle for g in self.levels])
        lh = (C.c_int64 * max(1, 6 * (n - 1)))(*[v for b in self.boxes for v in b])
        rt = (C.c_int * max(1, n - 1))(*self.ratios)
        h = C.c_void_p()
        call("bsn_fmr_create", self.ctx.handle, n, arr, lh, rt, C.byref(h))
        self.handle = h

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                from ._lib import load
                load().bsn_fmr_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_mode(self, inplace: bool) -> None:
        """Level step form: False (default) fuses the z sweep with each
        stage update; True steps every level in place with all three flux
        arrays materialised.  Both give the same bits."""
        call("bsn_fmr_set_mode", self.handle, int(bool(inplace)))

    def autotune(self) -> list:
        """Tile and kernel shapes of every level timed on its current state
        (the fastest kept); returns the per-level timing reports."""
        buf = C.create_string_buffer(16384)
        call("bsn_fmr_autotune", self.handle, buf, 16384)
        return json.loads(buf.value.decode())

    def prolong_all(self, level: int) -> None:
        """_prolong_all! (src/fmr3d.jl:206-214)"""
        call("bsn_fmr_prolong_all", self.handle, level)

    def prolong_fine_ghosts(self, level: int) -> None:
        """_prolong_fine_ghosts! (src/fmr3d.jl:223-255)"""
        call("bsn_fmr_prolong_fine_ghosts", self.handle, level)

    def restrict(self, level: int) -> None:
        """_restrict! (src/fmr3d.jl:266-285)"""
        call("bsn_fmr_restrict", self.handle, level)

    def step(self, dt_c: float, self_gravity: bool = False, G_grav: float = 1.0) -> None:
        """fmr3d_step!(G, dt_c; self_gravity, G_grav): with self_gravity the
        composite potential is solved once and frozen for the root step."""
        call("bsn_fmr3d_step_gravity", self.handle, float(dt_c), int(bool(self_gravity)),
             float(G_grav))

    def set_self_gravity(self, on: bool = True, G_grav: float = 1.0) -> None:
        """The frozen-phi self-gravity solve at every root step of run()."""
        call("bsn_fmr_set_self_gravity", self.handle, int(bool(on)), float(G_grav))

    def solve_poisson(self) -> list:
        """solve_fmr_poisson: per-level active-cell phi, (nz, ny, nx) each;
        restricts the density into the parents first."""
        outs = [np.empty((nz, ny, nx)) for (nx, ny, nz) in self.dims]
        arr = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
        call("bsn_solve_fmr_poisson", self.handle, arr, BSN_HOST)
        return outs

    def restrict_density(self, level: int) -> None:
        """restrict_density_to_coarse! of `level` into its parent"""
        call("bsn_fmr_restrict_density", self.handle, level)

    def prolong_scalar(self, level: int, phi_c: np.ndarray) -> np.ndarray:
        """_prolong_scalar_trilinear! of level-1's phi (nz, ny, nx) onto `level`"""
        pc = self.dims[level - 1]
        phi_c = np.ascontiguousarray(phi_c, dtype=np.float64)
        if phi_c.shape != (pc[2], pc[1], pc[0]):
            raise ValueError(f"phi_c shape {phi_c.shape} != {(pc[2], pc[1], pc[0])}")
        nx, ny, nz = self.dims[level]
        out = np.empty((nz, ny, nx))
        call("bsn_fmr_prolong_scalar", self.handle, level, C.c_void_p(phi_c.ctypes.data), BSN_HOST,
             C.c_void_p(out.ctypes.data), BSN_HOST)
        return out

    def cfl_dt(self, cfl: float) -> float:
        out = C.c_double()
        call("bsn_cfl_dt_fmr3d", self.handle, float(cfl), C.byref(out))
        return out.value

    def run(self, cfl: float, nsteps: int, t_end: float = 0.0, graph: bool = True) -> None:
        call("bsn_fmr_run", self.handle, float(cfl), float(t_end), int(nsteps), int(bool(graph)))

    def time(self):
        t, dt = C.c_double(), C.c_double()
        call("bsn_fmr_time", self.handle, C.byref(t), C.byref(dt))
        return t.value, dt.value

    def set_phis(self, phis) -> None:
        for g, p in zip(self.levels, phis):
            g.set_phi(p)

    def zone_updates_per_step(self) -> int:
        """sum over levels of active cells x steps at that level per root step"""
        tot, mult = 0, 1
        for l, (nx, ny, nz) in enumerate(self.dims):
            tot += nx * ny * nz * mult
            if l < len(self.ratios):
                mult *= self.ratios[l]
        return tot


class FMRGrid3D(FMRHierarchy):
    """Two-level grid with the reference's constructor (src/fmr3d.jl:131-151):
    the fine level is initialised by prolongation of all its cells."""

    def __init__(self, coarse_U: np.ndarray, nx, ny, nz, dx, dy, dz, ci_lo, ci_hi, cj_lo, cj_hi,
                 ck_lo, ck_hi, ratio, gamma, *, bc="outflow", rho_floor=0.0, P_floor=0.0,
                 precision="fp64", ctx=None):
        nf = ((ci_hi - ci_lo + 1) * ratio, (cj_hi - cj_lo + 1) * ratio, (ck_hi - ck_lo + 1) * ratio)
        super().__init__([(nx, ny, nz), nf], [(dx, dy, dz), (dx / ratio, dy / ratio, dz / ratio)],
                         [(ci_lo, ci_hi, cj_lo, cj_hi, ck_lo, ck_hi)], [ratio], gamma, bc,
                         rho_floor, P_floor, precision, ctx)
        self.ratio = ratio
        self.coarse, self.fine = self.levels
        self.coarse.set_state(coarse_U)
        self.prolong_all(1)


def fmr3d_step_(G: FMRHierarchy, dt_c: float, *, self_gravity: bool = False,
                G_grav: float = 1.0) -> None:
    """fmr3d_step!(G, dt_c; self_gravity=false, G_grav=1.0) (src/fmr3d.jl:
    451-604).  Without self_gravity a level's phi, when set, is a static
    potential; with it the composite potential of solve_fmr_poisson, times
    G_grav, is frozen across the root step."""
    G.step(dt_c, self_gravity, G_grav)


def solve_fmr_poisson(G: FMRHierarchy) -> list:
    """solve_fmr_poisson(G) (src/fmr_gravity.jl:119-133): per-level phi."""
    return G.solve_poisson()


def restrict_density_to_coarse_(G: FMRHierarchy, level: int = 1) -> None:
    """restrict_density_to_coarse!(G) (src/fmr_gravity.jl:40-60)"""
    G.restrict_density(level)


def _prolong_scalar_trilinear_(phi_f: np.ndarray, phi_c: np.ndarray, G: FMRHierarchy,
                               level: int = 1) -> None:
    """_prolong_scalar_trilinear!(phi_f, phi_c, G) (src/fmr_gravity.jl:70-106)"""
    phi_f[...] = G.prolong_scalar(level, phi_c)


def cfl_dt_fmr3d(G: FMRHierarchy, cfl: float) -> float:
    """cfl_dt_fmr3d (src/fmr3d.jl:616-622)"""
    return G.cfl_dt(cfl)
