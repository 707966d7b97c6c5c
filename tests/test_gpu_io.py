"""Snapshot and checkpoint files streamed from and to a GPU-resident grid
(bsn_grid_export_slabs / bsn_grid_import_slabs through pinned slabs): every
value must arrive bit for bit, in the layout the host-array path writes (whose
layout tests/test_io_host.py checks with h5dump)."""
import ctypes as C

import numpy as np
import pytest

import binary_supernova_amd as bsn
from binary_supernova_amd import _lib
from binary_supernova_amd import io as bio

pytestmark = pytest.mark.gpu
NG = 3


def grid_with_random_state(nx, ny, nz, seed):
    U = np.random.default_rng(seed).normal(size=(nz + 2 * NG, ny + 2 * NG, nx + 2 * NG, 5))
    return bsn.Grid(nx, ny, nz).set_state(U), U


def export(g, layout, slab_k):
    """Collects every slab the device streams into one array per part."""
    nx, ny, nz = g.nx, g.ny, g.nz
    if layout == _lib.LAYOUT_FIELDS:
        out = np.full((5, nz, ny * nx), np.nan)
    else:
        out = np.full((1, nz + 2 * NG, (ny + 2 * NG) * (nx + 2 * NG) * 5), np.nan)
    seen = []

    def sink(user, part, k0, nk, data, count):
        a = np.ctypeslib.as_array(data, shape=(count,))
        out[part, k0:k0 + nk] = a.reshape(nk, -1)
        seen.append((part, k0, nk))
        return 0

    cb = _lib.SLAB_FN(sink)
    _lib.call("bsn_grid_export_slabs", g.handle, layout, slab_k, cb, None)
    return out, seen


@pytest.mark.parametrize("slab_k", [1, 3, 0])
def test_export_slabs_in_both_layouts(slab_k):
    print(f"export slabs slab_k={slab_k}", flush=True)
    nx, ny, nz = 13, 9, 7
    g, U = grid_with_random_state(nx, ny, nz, 5)
    F, seen = export(g, _lib.LAYOUT_FIELDS, slab_k)
    A = U[NG:-NG, NG:-NG, NG:-NG]
    for q in range(5):
        np.testing.assert_array_equal(F[q].reshape(nz, ny, nx), A[..., q])
    sk = slab_k if slab_k > 0 else nz
    assert seen[0] == (0, 0, min(sk, nz)) and len(seen) == 5 * -(-nz // sk)
    R, seen = export(g, _lib.LAYOUT_REF, slab_k)
    np.testing.assert_array_equal(R[0].reshape(U.shape), U)


@pytest.mark.parametrize("layout", [_lib.LAYOUT_FIELDS, _lib.LAYOUT_REF])
def test_import_slabs_round_trip(layout):
    print(f"import slabs layout={layout}", flush=True)
    nx, ny, nz = 11, 6, 9
    g, U = grid_with_random_state(nx, ny, nz, 6)
    V = np.random.default_rng(7).normal(size=U.shape)

    def source(user, part, k0, nk, data, count):
        a = np.ctypeslib.as_array(data, shape=(count,))
        if layout == _lib.LAYOUT_FIELDS:
            a[:] = V[NG + k0:NG + k0 + nk, NG:-NG, NG:-NG, part].ravel()
        else:
            a[:] = V[k0:k0 + nk].ravel()
        return 0

    _lib.call("bsn_grid_import_slabs", g.handle, layout, 2, _lib.SLAB_FN(source), None)
    W = g.get_state()
    if layout == _lib.LAYOUT_FIELDS:
        expect = U.copy()
        expect[NG:-NG, NG:-NG, NG:-NG] = V[NG:-NG, NG:-NG, NG:-NG]   # ghosts untouched
    else:
        expect = V
    np.testing.assert_array_equal(W, expect)


def test_a_failing_consumer_stops_the_stream():
    g, _ = grid_with_random_state(5, 5, 5, 1)
    calls = []

    def sink(user, part, k0, nk, data, count):
        calls.append(part)
        return 1 if part == 2 else 0

    with pytest.raises(bsn.BsnError, match="consumer failed"):
        _lib.call("bsn_grid_export_slabs", g.handle, _lib.LAYOUT_FIELDS, 0, _lib.SLAB_FN(sink), None)
    assert calls == [0, 1, 2]
    g.get_state()                              # the grid is still usable


def test_snapshot_from_and_to_the_device(tmp_path):
    print("snapshot device", flush=True)
    nx, ny, nz = 100, 90, 80
    g, U = grid_with_random_state(nx, ny, nz, 8)
    dev, host = tmp_path / "dev.h5", tmp_path / "host.h5"
    bio.write_snapshot(dev, g, nx, ny, nz, 0.1, 0.2, 0.3, 4.5, 5 / 3)
    bio.write_snapshot(host, U, nx, ny, nz, 0.1, 0.2, 0.3, 4.5, 5 / 3)
    a, b = bio.read_snapshot(dev), bio.read_snapshot(host)
    assert a[1:] == b[1:] == (nx, ny, nz, 0.1, 0.2, 0.3, 4.5, 5 / 3)
    np.testing.assert_array_equal(a[0], U[NG:-NG, NG:-NG, NG:-NG])
    np.testing.assert_array_equal(b[0], a[0])
    g2, U2 = grid_with_random_state(nx, ny, nz, 9)
    bio.read_snapshot_(g2, dev)
    expect = U2.copy()
    expect[NG:-NG, NG:-NG, NG:-NG] = U[NG:-NG, NG:-NG, NG:-NG]
    np.testing.assert_array_equal(g2.get_state(), expect)
    g3 = bsn.Grid(nx, ny, nz + 1)
    with pytest.raises(bsn.BsnError, match="grid is"):
        bio.read_snapshot_(g3, dev)


def test_uniform_checkpoint_from_and_to_the_device(tmp_path):
    print("checkpoint device", flush=True)
    nx, ny, nz = 100, 90, 80                       # 86 padded k-planes: two slabs
    g, U = grid_with_random_state(nx, ny, nz, 10)
    B = bsn.BlackHole
    bhs = [B(pos=[0.1, 0.2, 0.3], vel=[1.0, 2.0, 3.0], mass=0.7, eps=0.01, c_code=30.0,
             r_floor=0.02)]
    kw = dict(t=1.5, step=77, dt_last=1e-3, t_snap=2.0, t_bh2_sink_on=0.5, nx=nx, ny=ny, nz=nz,
              dx=0.1, dy=0.1, dz=0.1, gamma=1.4, x0=-5.0, y0=-4.5, z0=-4.0, rho_floor=1e-8,
              P_floor=1e-10)
    p = tmp_path / "ckpt.h5"
    bio.save_checkpoint_uniform(p, g, bhs, **kw)
    Uh, bhs_h, grid_h, meta_h = bio.load_checkpoint_uniform(p)
    np.testing.assert_array_equal(Uh, U)
    assert meta_h == (1.5, 77, 1e-3, 2.0, 0.5, 1)
    assert bhs_h[0].pos == bhs[0].pos and bhs_h[0].c_code == 30.0
    g2 = bsn.Grid(nx, ny, nz)
    _, bhs_d, grid_d, meta_d = bio.load_checkpoint_uniform(p, grid=g2)
    np.testing.assert_array_equal(g2.get_state(), U)
    assert (grid_d, meta_d) == (grid_h, meta_h)
    # the host-array writer gives the same file contents
    q = tmp_path / "ckpt_host.h5"
    bio.save_checkpoint_uniform(q, U, bhs, **kw)
    np.testing.assert_array_equal(bio.load_checkpoint_uniform(q)[0], Uh)


def test_fmr_checkpoint_from_and_to_the_device(tmp_path):
    # save_checkpoint / load_checkpoint of a two-level grid (src/checkpoint.jl
    # :33-172) streamed from and to the device: a stepped hierarchy, saved and
    # loaded into a hierarchy built from the file, holds the same bits, and
    # both step on to the same bits (the restart is faithful)
    from oracle import oracle as O
    from binary_supernova_amd import problems as P
    print("fmr checkpoint", flush=True)
    gm, nc, L, r = 5 / 3, 8, 0.5, 2
    dxc = 2 * L / nc
    box = (3, 6, 3, 6, 3, 6)
    Uc = P.sedov_ic_3d(nc, nc, nc, dxc, dxc, dxc, gm, E_blast=1.0, r_inject=0.15, rho_bg=1.0,
                       P_floor=1e-5, x_offset=-L, y_offset=-L, z_offset=-L)
    O.fill_ghost(Uc, "outflow")
    G = bsn.FMRGrid3D(Uc, nc, nc, nc, dxc, dxc, dxc, *box, r, gm, bc="outflow",
                      rho_floor=1e-10, P_floor=1e-5)
    for _ in range(2):
        G.step(G.cfl_dt(0.4))
    bhs = [bsn.BlackHole(pos=[0.1, 0.0, 0.0], vel=[0.0, 0.2, 0.0], mass=0.3, eps=1e-3,
                         c_code=20.0, r_floor=1e-3)]
    p = tmp_path / "fmr.h5"
    bio.save_checkpoint(p, G, bhs, t=0.5, step=2, dt_last=1e-3)
    Uc_f, Uf_f, _, info, _ = bio.load_checkpoint_host(p)
    np.testing.assert_array_equal(Uc_f, G.levels[0].get_state())
    np.testing.assert_array_equal(Uf_f, G.levels[1].get_state())
    H, bhs2, meta = bio.load_checkpoint(p)
    assert meta == (0.5, 2, 1e-3, 1) and len(bhs2) == 1 and bhs2[0].pos == bhs[0].pos
    assert H.boxes == [box] and H.ratios == [r] and H.dims == G.dims
    for a, b in zip(G.levels, H.levels):
        np.testing.assert_array_equal(b.get_state(), a.get_state())
    for _ in range(2):
        dt = G.cfl_dt(0.4)
        assert H.cfl_dt(0.4) == dt
        G.step(dt)
        H.step(dt)
    for a, b in zip(G.levels, H.levels):
        np.testing.assert_array_equal(b.get_state(), a.get_state())
