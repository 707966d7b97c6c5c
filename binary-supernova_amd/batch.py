"""Batches of independent problems of uneven sizes: one persistent launch per
kernel sweeps every grid through a device-resident tile work-list; a whole
RK3 step (dt computed on the device per problem) can be captured in a
hipGraph; tile shapes are chosen by an autotuner that times each candidate."""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from ._lib import call
from .euler3d import Context, Grid


class Batch:
    def __init__(self, grids: list[Grid], ctx: Context | None = None):
        self.ctx = ctx or grids[0].ctx
        self.grids = list(grids)
        arr = (C.c_void_p * len(grids))(*[g.handle for g in grids])
        h = C.c_void_p()
        call("bsn_batch_create", self.ctx.handle, len(grids), arr, C.byref(h))
        self.handle = h

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                from ._lib import load
                load().bsn_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def autotune(self) -> dict:
        buf = C.create_string_buffer(4096)
        call("bsn_batch_autotune", self.handle, buf, 4096)
        return json.loads(buf.value.decode())

    def set_shapes(self, sx: int, sy: int, sz: int) -> None:
        call("bsn_batch_set_shapes", self.handle, sx, sy, sz)

    def set_update_shape(self, shape: int) -> None:
        """Fused z-sweep + update kernel shape 0..11 = 3 * variant +
        segmentation: variant 0 plain, 1 operands loaded one cell ahead (FP64)
        or variant 3 without a register cap (split), 2 LDS ring of the
        window's raw cells, 3 ring + early window load; segmentation 0/1 = z
        runs of 16/32 cells, 2 = balanced run lengths."""
        call("bsn_batch_set_update_shape", self.handle, int(shape))

    def get_shapes(self) -> list:
        """[x, y, z flux shape, fused-kernel shape] ids in use."""
        out = (C.c_int * 4)()
        call("bsn_batch_get_shapes", self.handle, out)
        return list(out)

    def set_mode(self, inplace: bool) -> None:
        """True: in-place step with every flux materialised; False (default):
        out-of-place step with the z sweep fused into the stage update."""
        call("bsn_batch_set_mode", self.handle, int(bool(inplace)))

    def step(self, dts) -> None:
        d = (C.c_double * len(self.grids))(*[float(v) for v in dts])
        call("bsn_batch_step", self.handle, d)

    def run(self, cfl: float, nsteps: int, t_end: float = 0.0, graph: bool = True) -> None:
        call("bsn_batch_run", self.handle, float(cfl), float(t_end), int(nsteps), int(bool(graph)))

    def times(self):
        n = len(self.grids)
        t, dt = (C.c_double * n)(), (C.c_double * n)()
        call("bsn_batch_times", self.handle, t, dt)
        return np.array(list(t)), np.array(list(dt))

    def profile(self, nsteps: int):
        """Runs nsteps device-dt steps with HIP events around every flux and
        update launch; returns (ms per flux launch [x, y, z], ms per update
        launch, ms per step)."""
        f3 = (C.c_double * 3)()
        u, s = C.c_double(), C.c_double()
        call("bsn_batch_profile", self.handle, int(nsteps), f3, C.byref(u), C.byref(s))
        return list(f3), u.value, s.value

    def zone_updates_per_step(self) -> int:
        return sum(g.nx * g.ny * g.nz for g in self.grids)
