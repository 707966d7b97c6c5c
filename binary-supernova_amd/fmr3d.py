"""Host-side mirror of the reference's fixed-mesh-refinement API
(src/fmr3d.jl): FMRGrid3D, fmr3d_step!, cfl_dt_fmr3d, the prolongation and
restriction helpers, and a multi-level generalisation (levels nested
recursively, each advanced by the two-level algorithm of fmr3d_step!).
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from ._lib import call
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

    def step(self, dt_c: float) -> None:
        call("bsn_fmr3d_step", self.handle, float(dt_c))

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


def fmr3d_step_(G: FMRHierarchy, dt_c: float) -> None:
    """fmr3d_step! (src/fmr3d.jl:451-604); a level's phi, when set, is the
    frozen potential of the self_gravity branch."""
    G.step(dt_c)


def cfl_dt_fmr3d(G: FMRHierarchy, cfl: float) -> float:
    """cfl_dt_fmr3d (src/fmr3d.jl:616-622)"""
    return G.cfl_dt(cfl)
