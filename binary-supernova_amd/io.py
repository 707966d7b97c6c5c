"""Snapshot and checkpoint files (HDF5, the reference's layout) streamed
between a GPU-resident grid and disk through libbsn_io.so.

Mirrors write_snapshot / read_snapshot (src/io.jl:29-84), save_checkpoint /
load_checkpoint of the two-level FMR grid (src/checkpoint.jl:1-172),
save_checkpoint_uniform / load_checkpoint_uniform (src/checkpoint.jl:244-318)
and the body trajectory file (src/io.jl:86-161).
A Grid argument streams its state through pinned slabs (the device packs the
file layout); a numpy array in the reference layout is written from the host.
"""
from __future__ import annotations

import ctypes as C
import os
from collections import namedtuple

import numpy as np

from ._lib import BSN_OK, BsnError, load
from .euler3d import NG, Grid, _ptr
from .sources import BlackHole

IO_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbsn_io.so")
EXPORTS = [
    "bsn_io_last_error", "bsn_write_snapshot", "bsn_write_snapshot_host", "bsn_snapshot_info",
    "bsn_read_snapshot", "bsn_read_snapshot_host", "bsn_save_checkpoint_uniform",
    "bsn_checkpoint_uniform_info", "bsn_load_checkpoint_uniform", "bsn_save_checkpoint_fmr",
    "bsn_checkpoint_fmr_info", "bsn_load_checkpoint_fmr", "bsn_init_trajectory_file",
    "bsn_append_trajectory", "bsn_trajectory_info", "bsn_read_trajectory",
]

_io = None
_d, _i, _i64, _p, _s = C.c_double, C.c_int, C.c_int64, C.c_void_p, C.c_char_p
_pd, _pi64 = C.POINTER(C.c_double), C.POINTER(C.c_int64)

GridParams = namedtuple("GridParams", "nx ny nz dx dy dz gamma x0 y0 z0 rho_floor P_floor")
CheckpointMeta = namedtuple("CheckpointMeta", "t step dt_last t_snap t_bh2_sink_on format_version")
FMRCheckpointMeta = namedtuple("FMRCheckpointMeta", "t step dt_last format_version")
FMRGridInfo = namedtuple("FMRGridInfo", "dims dxs box ratio gamma bc rho_floor P_floor")
Trajectory = namedtuple("Trajectory", "mass rsink pos vel")


def load_io():
    """Load libbsn_io.so (raises if it has not been built)."""
    global _io
    if _io is not None:
        return _io
    load()                                   # libbsn_hip.so first: libbsn_io links it
    if not os.path.exists(IO_LIB_PATH):
        raise BsnError(f"{IO_LIB_PATH} is missing: build it with binary-supernova_amd/build.py")
    L = C.CDLL(IO_LIB_PATH)
    sig = {
        "bsn_io_last_error": ([], C.c_char_p),
        "bsn_write_snapshot": ([_p, _s, _d, _d, _d, _d, _d], _i),
        "bsn_write_snapshot_host": ([_s, _p, _i64, _i64, _i64, _d, _d, _d, _d, _d], _i),
        "bsn_snapshot_info": ([_s, _pi64, _pd], _i),
        "bsn_read_snapshot": ([_p, _s], _i),
        "bsn_read_snapshot_host": ([_s, _p], _i),
        "bsn_save_checkpoint_uniform": ([_p, _p, _i64, _i64, _i64, _s, _i64, _pd, _pd, _i, _p, _p,
                                         _p], _i),
        "bsn_checkpoint_uniform_info": ([_s, _pi64, C.POINTER(_i)], _i),
        "bsn_load_checkpoint_uniform": ([_p, _p, _s, _pi64, _pd, _pd, _i, C.POINTER(_i), _p, _p,
                                         _p], _i),
        "bsn_save_checkpoint_fmr": ([_p, _p, _p, _p, _pi64, _pd, _s, _i64, _pd, _pi64, _i, _pd, _s,
                                     _i, _p, _p, _p], _i),
        "bsn_checkpoint_fmr_info": ([_s, _pi64, _pd, _pi64, C.POINTER(_i), _pd, C.c_char_p, _i,
                                     C.POINTER(_i)], _i),
        "bsn_load_checkpoint_fmr": ([_p, _p, _p, _p, _s, _pi64, _pd, _i, C.POINTER(_i), _p, _p,
                                     _p], _i),
        "bsn_init_trajectory_file": ([_s, _i], _i),
        "bsn_append_trajectory": ([_s, _d, _i, _p, _p, _p, _p], _i),
        "bsn_trajectory_info": ([_s, C.POINTER(_i), _pi64], _i),
        "bsn_read_trajectory": ([_s, _p, _p, _p, _p, _p], _i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _io = L
    return L


def _call(name, *args):
    rc = getattr(load_io(), name)(*args)
    if rc != BSN_OK:
        msg = load_io().bsn_io_last_error().decode(errors="replace")
        raise BsnError(f"{name}: status {rc}: {msg}")


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def _state(U, nx, ny, nz):
    if isinstance(U, Grid):
        if (U.nx, U.ny, U.nz) != (nx, ny, nz):
            raise ValueError("grid dims do not match nx, ny, nz")
        return U, None
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.shape != (nz + 2 * NG, ny + 2 * NG, nx + 2 * NG, 5):
        raise ValueError(f"U shape {U.shape} does not match nx, ny, nz with ghosts")
    return None, U


def write_snapshot(filename, U, nx, ny, nz, dx, dy, dz, t, gamma) -> None:
    """write_snapshot (src/io.jl:29-51): /gas/{rho, rhovx, rhovy, rhovz, E}
    of the active cells and the attributes t, nx, ny, nz, dx, dy, dz, gamma."""
    g, A = _state(U, nx, ny, nz)
    if g is not None:
        _call("bsn_write_snapshot", g.handle, _path(filename), float(dx), float(dy), float(dz),
              float(t), float(gamma))
    else:
        _call("bsn_write_snapshot_host", _path(filename), _ptr(A), nx, ny, nz, float(dx),
              float(dy), float(dz), float(t), float(gamma))


def snapshot_info(filename):
    n = (C.c_int64 * 3)()
    meta = (C.c_double * 5)()
    _call("bsn_snapshot_info", _path(filename), n, meta)
    return tuple(n), tuple(meta)


def read_snapshot(filename):
    """read_snapshot (src/io.jl:60-84): (U_active, nx, ny, nz, dx, dy, dz, t,
    gamma); U_active is the Julia (5, nx, ny, nz) array, here of shape
    (nz, ny, nx, 5)."""
    (nx, ny, nz), (dx, dy, dz, t, gamma) = snapshot_info(filename)
    U = np.empty((nz, ny, nx, 5))
    _call("bsn_read_snapshot_host", _path(filename), _ptr(U))
    return U, nx, ny, nz, dx, dy, dz, t, gamma


def read_snapshot_(grid: Grid, filename) -> None:
    """The active cells of `grid` from a snapshot, streamed onto the device
    (its ghost cells are left as they are)."""
    _call("bsn_read_snapshot", grid.handle, _path(filename))


def save_checkpoint_uniform(path, U, bhs, *, t, step, dt_last, t_snap, t_bh2_sink_on, nx, ny,
                            nz, dx, dy, dz, gamma, x0, y0, z0, rho_floor, P_floor):
    """save_checkpoint_uniform (src/checkpoint.jl:244-283)."""
    g, A = _state(U, nx, ny, nz)
    meta = (C.c_double * 4)(t, dt_last, t_snap, t_bh2_sink_on)
    params = (C.c_double * 9)(dx, dy, dz, gamma, x0, y0, z0, rho_floor, P_floor)
    n = len(bhs)
    pos = np.ascontiguousarray([[float(c) for c in b.pos] for b in bhs], dtype=np.float64).reshape(n, 3)
    vel = np.ascontiguousarray([[float(c) for c in b.vel] for b in bhs], dtype=np.float64).reshape(n, 3)
    props = np.ascontiguousarray([[b.mass, b.eps, b.c_code, b.r_floor] for b in bhs],
                                 dtype=np.float64).reshape(n, 4)
    _call("bsn_save_checkpoint_uniform", g.handle if g is not None else None,
          _ptr(A) if A is not None else None, nx, ny, nz, _path(path), int(step), meta, params, n,
          _ptr(pos), _ptr(vel), _ptr(props))
    return path


def checkpoint_uniform_info(path):
    n = (C.c_int64 * 3)()
    nb = C.c_int()
    _call("bsn_checkpoint_uniform_info", _path(path), n, C.byref(nb))
    return tuple(n), nb.value


def load_checkpoint_uniform(path, grid: Grid | None = None):
    """load_checkpoint_uniform (src/checkpoint.jl:285-318): (U, bhs, grid
    params, meta).  With `grid` the state is streamed into it and U is the
    grid; otherwise U is a host array in the reference layout."""
    (nx, ny, nz), nb = checkpoint_uniform_info(path)
    U = None
    if grid is None:
        U = np.empty((nz + 2 * NG, ny + 2 * NG, nx + 2 * NG, 5))
    step = C.c_int64()
    meta = (C.c_double * 5)()
    params = (C.c_double * 9)()
    got = C.c_int()
    pos, vel, props = np.zeros((max(nb, 1), 3)), np.zeros((max(nb, 1), 3)), np.zeros((max(nb, 1), 4))
    _call("bsn_load_checkpoint_uniform", grid.handle if grid is not None else None,
          _ptr(U) if U is not None else None, _path(path), C.byref(step), meta, params, max(nb, 1),
          C.byref(got), _ptr(pos), _ptr(vel), _ptr(props))
    bhs = [BlackHole(pos=list(pos[i]), vel=list(vel[i]), mass=float(props[i, 0]),
                     eps=float(props[i, 1]), c_code=float(props[i, 2]),
                     r_floor=float(props[i, 3])) for i in range(got.value)]
    gp = GridParams(nx, ny, nz, *params)
    m = CheckpointMeta(meta[0], step.value, meta[1], meta[2], meta[3], int(meta[4]))
    return (grid if grid is not None else U), bhs, gp, m


def _bodies(bhs):
    n = len(bhs)
    pos = np.ascontiguousarray([[float(c) for c in b.pos] for b in bhs], dtype=np.float64).reshape(n, 3)
    vel = np.ascontiguousarray([[float(c) for c in b.vel] for b in bhs], dtype=np.float64).reshape(n, 3)
    props = np.ascontiguousarray([[b.mass, b.eps, b.c_code, b.r_floor] for b in bhs],
                                 dtype=np.float64).reshape(n, 4)
    return n, pos, vel, props


def _bc_name(mask: int) -> str:
    if mask not in (0, 7):
        raise ValueError("the two-level checkpoint stores one bc for all axes")
    return "periodic" if mask == 7 else "outflow"


def save_checkpoint(path, G, bhs, *, t=0.0, step=0, dt_last=0.0):
    """save_checkpoint(path, G, bhs; t, step, dt_last) (src/checkpoint.jl:33-117)
    of a two-level FMRHierarchy: both levels' states stream from the device
    into /grid/coarse/U and /grid/fine/U."""
    from .fmr3d import FMRHierarchy
    if not isinstance(G, FMRHierarchy) or len(G.levels) != 2:
        raise ValueError("the checkpoint format holds a two-level FMR grid")
    n = (C.c_int64 * 6)(*G.dims[0], *G.dims[1])
    d = (C.c_double * 6)(*G.dxs[0], *G.dxs[1])
    box = (C.c_int64 * 6)(*G.boxes[0])
    params = (C.c_double * 3)(G.gamma, G.rho_floor, G.P_floor)
    meta = (C.c_double * 2)(float(t), float(dt_last))
    nb, pos, vel, props = _bodies(bhs)
    _call("bsn_save_checkpoint_fmr", G.levels[0].handle, G.levels[1].handle, None, None, n, d,
          _path(path), int(step), meta, box, int(G.ratios[0]), params,
          _bc_name(G.bc).encode(), nb, _ptr(pos), _ptr(vel), _ptr(props))
    return path


def checkpoint_fmr_info(path):
    """(FMRGridInfo, number of bodies) of a two-level checkpoint."""
    n = (C.c_int64 * 6)()
    d = (C.c_double * 6)()
    box = (C.c_int64 * 6)()
    ratio = C.c_int()
    params = (C.c_double * 3)()
    bc = C.create_string_buffer(32)
    nb = C.c_int()
    _call("bsn_checkpoint_fmr_info", _path(path), n, d, box, C.byref(ratio), params, bc, 32,
          C.byref(nb))
    info = FMRGridInfo(dims=[tuple(n[:3]), tuple(n[3:])], dxs=[tuple(d[:3]), tuple(d[3:])],
                       box=tuple(box), ratio=ratio.value, gamma=params[0],
                       bc=bc.value.decode(), rho_floor=params[1], P_floor=params[2])
    return info, nb.value


def load_checkpoint(path, G=None, *, precision="fp64"):
    """load_checkpoint(path) (src/checkpoint.jl:119-172): (G, bhs, meta).
    Without G a two-level FMRHierarchy of the stored shape is built first;
    both levels' saved states (the fine one as saved, not prolonged again)
    are streamed onto the device."""
    from .fmr3d import FMRHierarchy
    info, nb = checkpoint_fmr_info(path)
    if G is None:
        G = FMRHierarchy(info.dims, info.dxs, [info.box], [info.ratio], info.gamma, info.bc,
                         info.rho_floor, info.P_floor, precision=precision)
    elif len(G.levels) != 2:
        raise ValueError("the checkpoint format holds a two-level FMR grid")
    step = C.c_int64()
    meta = (C.c_double * 3)()
    got = C.c_int()
    pos, vel, props = np.zeros((max(nb, 1), 3)), np.zeros((max(nb, 1), 3)), np.zeros((max(nb, 1), 4))
    _call("bsn_load_checkpoint_fmr", G.levels[0].handle, G.levels[1].handle, None, None,
          _path(path), C.byref(step), meta, max(nb, 1), C.byref(got), _ptr(pos), _ptr(vel),
          _ptr(props))
    bhs = [BlackHole(pos=list(pos[i]), vel=list(vel[i]), mass=float(props[i, 0]),
                     eps=float(props[i, 1]), c_code=float(props[i, 2]),
                     r_floor=float(props[i, 3])) for i in range(got.value)]
    return G, bhs, FMRCheckpointMeta(meta[0], step.value, meta[1], int(meta[2]))


def save_checkpoint_host(path, Uc, Uf, bhs, *, dims, dxs, box, ratio, gamma, bc="outflow",
                         rho_floor=0.0, P_floor=0.0, t=0.0, step=0, dt_last=0.0):
    """The same file from host arrays in the reference layout (coarse and
    fine states with ghosts)."""
    (nxc, nyc, nzc), (nxf, nyf, nzf) = dims
    Uc = np.ascontiguousarray(Uc, dtype=np.float64)
    Uf = np.ascontiguousarray(Uf, dtype=np.float64)
    if Uc.shape != (nzc + 2 * NG, nyc + 2 * NG, nxc + 2 * NG, 5) or \
            Uf.shape != (nzf + 2 * NG, nyf + 2 * NG, nxf + 2 * NG, 5):
        raise ValueError("state shapes do not match dims with ghosts")
    n = (C.c_int64 * 6)(nxc, nyc, nzc, nxf, nyf, nzf)
    d = (C.c_double * 6)(*dxs[0], *dxs[1])
    nb, pos, vel, props = _bodies(bhs)
    _call("bsn_save_checkpoint_fmr", None, None, _ptr(Uc), _ptr(Uf), n, d, _path(path), int(step),
          (C.c_double * 2)(float(t), float(dt_last)), (C.c_int64 * 6)(*box), int(ratio),
          (C.c_double * 3)(gamma, rho_floor, P_floor), bc.encode(), nb, _ptr(pos), _ptr(vel),
          _ptr(props))
    return path


def load_checkpoint_host(path):
    """(Uc, Uf, bhs, FMRGridInfo, meta) of a two-level checkpoint, as host
    arrays in the reference layout."""
    info, nb = checkpoint_fmr_info(path)
    (nxc, nyc, nzc), (nxf, nyf, nzf) = info.dims
    Uc = np.empty((nzc + 2 * NG, nyc + 2 * NG, nxc + 2 * NG, 5))
    Uf = np.empty((nzf + 2 * NG, nyf + 2 * NG, nxf + 2 * NG, 5))
    step = C.c_int64()
    meta = (C.c_double * 3)()
    got = C.c_int()
    pos, vel, props = np.zeros((max(nb, 1), 3)), np.zeros((max(nb, 1), 3)), np.zeros((max(nb, 1), 4))
    _call("bsn_load_checkpoint_fmr", None, None, _ptr(Uc), _ptr(Uf), _path(path), C.byref(step),
          meta, max(nb, 1), C.byref(got), _ptr(pos), _ptr(vel), _ptr(props))
    bhs = [BlackHole(pos=list(pos[i]), vel=list(vel[i]), mass=float(props[i, 0]),
                     eps=float(props[i, 1]), c_code=float(props[i, 2]),
                     r_floor=float(props[i, 3])) for i in range(got.value)]
    return Uc, Uf, bhs, info, FMRCheckpointMeta(meta[0], step.value, meta[1], int(meta[2]))


def init_trajectory_file(filename, n_bh: int) -> None:
    """init_trajectory_file (src/io.jl:94-108)."""
    _call("bsn_init_trajectory_file", _path(filename), int(n_bh))


def append_trajectory(filename, t, bhs) -> None:
    """append_trajectory (src/io.jl:115-133): one record of every body, with
    its r_sink (src/BinarySupernova.jl:159)."""
    from .sources import r_sink
    n = len(bhs)
    mass = np.ascontiguousarray([float(b.mass) for b in bhs], dtype=np.float64)
    rs = np.ascontiguousarray([r_sink(b) for b in bhs], dtype=np.float64)
    _, pos, vel, _ = _bodies(bhs)
    _call("bsn_append_trajectory", _path(filename), float(t), n, _ptr(mass), _ptr(rs), _ptr(pos),
          _ptr(vel))


def read_trajectory(filename):
    """read_trajectory (src/io.jl:144-161): (t, [Trajectory(mass, rsink,
    pos (N x 3), vel (N x 3)) per body])."""
    nb, n = C.c_int(), C.c_int64()
    _call("bsn_trajectory_info", _path(filename), C.byref(nb), C.byref(n))
    N, B = n.value, nb.value
    t = np.zeros(N)
    mass, rs = np.zeros((max(B, 1), N)), np.zeros((max(B, 1), N))
    pos, vel = np.zeros((max(B, 1), N, 3)), np.zeros((max(B, 1), N, 3))
    _call("bsn_read_trajectory", _path(filename), _ptr(t), _ptr(mass), _ptr(rs), _ptr(pos),
          _ptr(vel))
    return t, [Trajectory(mass[b], rs[b], pos[b], vel[b]) for b in range(B)]
