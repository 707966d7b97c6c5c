"""Snapshot and checkpoint files from host arrays (no GPU): the layout is
checked with HDF5's own h5dump, independently of libbsn_io's reader, against
what the reference writes (src/io.jl:29-51, src/checkpoint.jl:244-283; a Julia
(a, b, c) array is an HDF5 dataset of dims (c, b, a)), and the reference's
round-trip test (test/test_checkpoint.jl:84-137) is repeated."""
import os
import re
import subprocess

import numpy as np
import pytest

import binary_supernova_amd as bsn
from binary_supernova_amd import io as bio

NG = 3
H5DUMP = "/opt/conda/bin/h5dump"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def h5dump(*args):
    r = subprocess.run([H5DUMP, *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def h5_array(path, name, shape, tmp_path):
    out = tmp_path / (name.strip("/").replace("/", "_") + ".bin")
    h5dump("-d", name, "-b", "LE", "-o", str(out), path)
    return np.fromfile(out, dtype="<f8").reshape(shape)


def header(path):
    """{object path: (kind, datatype, dims)} from h5dump -H."""
    out = h5dump("-H", path)
    items, stack = {}, []
    cur = None
    for line in out.splitlines():
        m = re.match(r"\s*(GROUP|DATASET|ATTRIBUTE) \"([^\"]+)\" \{", line)
        if m:
            kind, name = m.groups()
            stack.append(name)
            cur = "/".join(s for s in stack if s is not None and s != "/")
            items[cur] = [kind, None, None]
            continue
        m = re.match(r"\s*DATATYPE\s+(\S+)", line)
        if m and cur:
            items[cur][1] = m.group(1)
        m = re.match(r"\s*DATASPACE\s+(SCALAR|SIMPLE \{ \( ([^)]*) \))", line)
        if m and cur:
            items[cur][2] = () if m.group(1) == "SCALAR" else tuple(int(x) for x in m.group(2).split(","))
        elif line.rstrip().endswith("{"):
            stack.append(None)          # an unnamed block, such as a string datatype
        if re.match(r"\s*\}\s*$", line) and stack:
            stack.pop()
            named = [s for s in stack if s is not None]
            cur = "/".join(s for s in named if s != "/") if named else None
    return {k: tuple(v) for k, v in items.items()}


def ref_state(nx, ny, nz, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(nz + 2 * NG, ny + 2 * NG, nx + 2 * NG, 5))


def test_io_library_exports_every_declared_symbol():
    import ctypes
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "bsn_io.h")).read(), flags=re.S)
    declared = sorted(set(re.findall(r"\b(bsn_[a-z0-9_]+)\s*\(", text)))
    assert declared == sorted(bio.EXPORTS)
    bio.load_io()
    lib = ctypes.CDLL(bio.IO_LIB_PATH)
    assert all(hasattr(lib, f) for f in declared)


def test_snapshot_layout_and_values(tmp_path):
    nx, ny, nz = 7, 5, 4
    U = ref_state(nx, ny, nz, 1)
    p = str(tmp_path / "snap.h5")
    bio.write_snapshot(p, U, nx, ny, nz, 0.1, 0.2, 0.3, 1.25, 5.0 / 3.0)
    h = header(p)
    for a in ("t", "dx", "dy", "dz", "gamma"):
        assert h[a] == ("ATTRIBUTE", "H5T_IEEE_F64LE", ())
    for a in ("nx", "ny", "nz"):
        assert h[a] == ("ATTRIBUTE", "H5T_STD_I64LE", ())
    A = U[NG:-NG, NG:-NG, NG:-NG]
    for q, name in enumerate(("rho", "rhovx", "rhovy", "rhovz", "E")):
        assert h["gas/" + name] == ("DATASET", "H5T_IEEE_F64LE", (nz, ny, nx))
        np.testing.assert_array_equal(h5_array(p, "/gas/" + name, (nz, ny, nx), tmp_path), A[..., q])
    out = h5dump("-a", "/ny", p)
    assert re.search(r"\(0\): 5\b", out)


def test_snapshot_round_trip(tmp_path):
    nx, ny, nz = 6, 4, 5
    U = ref_state(nx, ny, nz, 2)
    p = tmp_path / "snap.h5"
    bio.write_snapshot(p, U, nx, ny, nz, 0.05, 0.06, 0.07, 3.5, 1.4)
    A, *rest = bio.read_snapshot(p)
    assert tuple(rest) == (nx, ny, nz, 0.05, 0.06, 0.07, 3.5, 1.4)
    np.testing.assert_array_equal(A, U[NG:-NG, NG:-NG, NG:-NG])


def _reference_case():
    # test/test_checkpoint.jl:84-105: distinct values including ghosts
    nx = ny = nz = 6
    U = 0.5 + 1.0e-3 * np.arange(1, 5 * (nx + 6) * (ny + 6) * (nz + 6) + 1, dtype=np.float64)
    U = U.reshape(nz + 6, ny + 6, nx + 6, 5)     # Julia (5, Nx, Ny, Nz), linear index order
    B = bsn.BlackHole
    bhs = [B(pos=[-0.3, 0.0, 0.0], vel=[0.0, -0.7, 0.0], mass=0.6, eps=2e-3, c_code=41.2,
             r_floor=2e-3),
           B(pos=[0.3, 0.05, -0.1], vel=[0.01, 0.7, 0.0], mass=0.4, eps=2e-3, c_code=41.2,
             r_floor=2e-3)]
    kw = dict(t=7.5, step=12345, dt_last=2.5e-3, t_snap=8.0, t_bh2_sink_on=5.0, nx=nx, ny=ny,
              nz=nz, dx=0.2, dy=0.2, dz=0.2, gamma=4 / 3, x0=-1.0, y0=-1.0, z0=-1.0,
              rho_floor=3e-4, P_floor=1e-5)
    return U, bhs, kw


def test_checkpoint_uniform_round_trip(tmp_path):
    U, bhs, kw = _reference_case()
    p = tmp_path / "ckpt_uniform.h5"
    bio.save_checkpoint_uniform(p, U, bhs, **kw)
    U2, bhs2, grid, meta = bio.load_checkpoint_uniform(p)
    assert U2.shape == U.shape and np.array_equal(U2, U)        # bit for bit, ghosts too
    assert meta == (7.5, 12345, 2.5e-3, 8.0, 5.0, 1)
    assert (grid.nx, grid.ny, grid.nz) == (6, 6, 6)
    assert (grid.dx, grid.gamma) == (0.2, 4 / 3)
    assert (grid.x0, grid.y0, grid.z0) == (-1.0, -1.0, -1.0)
    assert (grid.rho_floor, grid.P_floor) == (3e-4, 1e-5)
    assert len(bhs2) == 2
    for b, b2 in zip(bhs, bhs2):
        assert (b2.pos, b2.vel, b2.mass, b2.eps, b2.c_code, b2.r_floor) == \
            (b.pos, b.vel, b.mass, b.eps, b.c_code, b.r_floor)
    # no bodies at all
    bio.save_checkpoint_uniform(p, U, [], **kw)
    assert bio.load_checkpoint_uniform(p)[1] == []


def test_checkpoint_uniform_layout(tmp_path):
    U, bhs, kw = _reference_case()
    p = str(tmp_path / "ckpt.h5")
    bio.save_checkpoint_uniform(p, U, bhs, **kw)
    h = header(p)
    F, I = "H5T_IEEE_F64LE", "H5T_STD_I64LE"
    for a, t in (("format_version", I), ("t", F), ("step", I), ("dt_last", F), ("t_snap", F),
                 ("t_bh2_sink_on", F)):
        assert h["manifest/" + a] == ("ATTRIBUTE", t, ()), a
    assert h["manifest/kind"][0] == "ATTRIBUTE" and h["manifest/kind"][2] == ()
    assert '"uniform"' in h5dump("-a", "/manifest/kind", p)
    for a in ("nx", "ny", "nz"):
        assert h["grid/" + a] == ("ATTRIBUTE", I, ())
    for a in ("dx", "dy", "dz", "gamma", "x0", "y0", "z0", "rho_floor", "P_floor"):
        assert h["grid/" + a] == ("ATTRIBUTE", F, ())
    assert h["grid/U"] == ("DATASET", F, (12, 12, 12, 5))
    np.testing.assert_array_equal(h5_array(p, "/grid/U", U.shape, tmp_path), U)
    assert h["bhs/n"] == ("ATTRIBUTE", I, ())
    assert h["bhs/pos"] == ("DATASET", F, (2, 3)) and h["bhs/vel"] == ("DATASET", F, (2, 3))
    for d in ("mass", "eps", "c_code", "r_floor"):
        assert h["bhs/" + d] == ("DATASET", F, (2,))
    np.testing.assert_array_equal(h5_array(p, "/bhs/pos", (2, 3), tmp_path),
                                  [b.pos for b in bhs])


def test_a_snapshot_is_not_a_uniform_checkpoint(tmp_path):
    # test/test_checkpoint.jl:135-136: the kinds are not cross-loadable
    nx, ny, nz = 3, 3, 3
    p = tmp_path / "snap.h5"
    bio.write_snapshot(p, ref_state(nx, ny, nz), nx, ny, nz, 0.1, 0.1, 0.1, 0.0, 1.4)
    with pytest.raises(bsn.BsnError, match="manifest"):
        bio.load_checkpoint_uniform(p)
    with pytest.raises(bsn.BsnError, match="cannot open"):
        bio.read_snapshot(tmp_path / "missing.h5")


def _fmr_reference_case():
    # test/test_checkpoint.jl:4-34: coarse 8^3 (dx = 1/8) with rho = 1 + 0.1 r^2
    # and E = 1/(gamma - 1) on the active cells, fine region = coarse cells
    # 4:5 in each axis at ratio 4, prolonged and then scaled by 1.01 so the
    # saved fine state differs from a re-prolongation
    from oracle import oracle as O
    g = 5 / 3
    n, r, box = 8, 4, (4, 5, 4, 5, 4, 5)
    Uc = np.zeros((n + 2 * NG, n + 2 * NG, n + 2 * NG, 5))
    c = (np.arange(n) + 0.5) / n - 0.5
    r2 = c[:, None, None] ** 2 + c[None, :, None] ** 2 + c[None, None, :] ** 2
    Uc[NG:-NG, NG:-NG, NG:-NG, 0] = 1.0 + 0.1 * r2
    Uc[NG:-NG, NG:-NG, NG:-NG, 4] = 1.0 / (g - 1.0)
    Uf = np.zeros_like(Uc)
    O.prolong_all(Uf, Uc, box, r)
    Uf *= 1.01
    B = bsn.BlackHole
    bhs = [B(pos=[-0.5, 0.0, 0.0], vel=[0.0, -0.5, 0.0], mass=0.5, eps=1e-3, c_code=20.0,
             r_floor=1e-3),
           B(pos=[0.5, 0.1, 0.0], vel=[0.0, 0.5, 0.0], mass=0.5, eps=1e-3, c_code=20.0,
             r_floor=1e-3)]
    kw = dict(dims=[(n, n, n), (8, 8, 8)], dxs=[(1 / n,) * 3, (1 / (n * r),) * 3], box=box,
              ratio=r, gamma=g, bc="outflow", rho_floor=1e-12, P_floor=1e-12)
    return Uc, Uf, bhs, kw


def test_checkpoint_fmr_round_trip(tmp_path):
    # test/test_checkpoint.jl:36-63: every field back bit for bit
    Uc, Uf, bhs, kw = _fmr_reference_case()
    p = tmp_path / "ckpt.h5"
    bio.save_checkpoint_host(p, Uc, Uf, bhs, t=0.25, step=42, dt_last=0.001, **kw)
    Uc2, Uf2, bhs2, info, meta = bio.load_checkpoint_host(p)
    assert meta == (0.25, 42, 0.001, 1)
    assert (info.ratio, info.gamma, info.bc, info.rho_floor, info.P_floor) == \
        (4, 5 / 3, "outflow", 1e-12, 1e-12)
    assert info.box == (4, 5, 4, 5, 4, 5)
    assert info.dims == [(8, 8, 8), (8, 8, 8)] and info.dxs == kw["dxs"]
    assert np.array_equal(Uc2, Uc) and np.array_equal(Uf2, Uf)
    assert len(bhs2) == 2
    for b, b2 in zip(bhs, bhs2):
        assert (b2.pos, b2.vel, b2.mass, b2.eps, b2.c_code, b2.r_floor) == \
            (b.pos, b.vel, b.mass, b.eps, b.c_code, b.r_floor)


def test_checkpoint_fmr_with_no_bodies(tmp_path):
    # test/test_checkpoint.jl:66-80: an empty body list, default t and step
    Uc, Uf, _, kw = _fmr_reference_case()
    p = tmp_path / "ckpt_empty.h5"
    bio.save_checkpoint_host(p, Uc, Uf, [], **kw)
    _, _, bhs2, _, meta = bio.load_checkpoint_host(p)
    assert bhs2 == [] and meta.t == 0.0 and meta.step == 0


def test_checkpoint_fmr_layout(tmp_path):
    # the layout of src/checkpoint.jl:3-28, read with h5dump
    Uc, Uf, bhs, kw = _fmr_reference_case()
    p = str(tmp_path / "ckpt.h5")
    bio.save_checkpoint_host(p, Uc, Uf, bhs, t=0.25, step=42, dt_last=0.001, **kw)
    h = header(p)
    F, I = "H5T_IEEE_F64LE", "H5T_STD_I64LE"
    for a, t in (("format_version", I), ("t", F), ("step", I), ("dt_last", F)):
        assert h["manifest/" + a] == ("ATTRIBUTE", t, ()), a
    assert "manifest/kind" not in h
    for a in ("ratio", "ci_lo", "ci_hi", "cj_lo", "cj_hi", "ck_lo", "ck_hi"):
        assert h["grid/" + a] == ("ATTRIBUTE", I, ()), a
    for a in ("gamma", "rho_floor", "P_floor"):
        assert h["grid/" + a] == ("ATTRIBUTE", F, ()), a
    assert '"outflow"' in h5dump("-a", "/grid/bc", p)
    for lev, U in (("coarse", Uc), ("fine", Uf)):
        for a in ("nx", "ny", "nz"):
            assert h[f"grid/{lev}/{a}"] == ("ATTRIBUTE", I, ())
        for a in ("dx", "dy", "dz"):
            assert h[f"grid/{lev}/{a}"] == ("ATTRIBUTE", F, ())
        assert h[f"grid/{lev}/U"] == ("DATASET", F, (14, 14, 14, 5))
        np.testing.assert_array_equal(h5_array(p, f"/grid/{lev}/U", U.shape, tmp_path), U)
    assert h["bhs/pos"] == ("DATASET", F, (2, 3))


def test_fmr_and_uniform_checkpoints_are_not_cross_loadable(tmp_path):
    Uc, Uf, bhs, kw = _fmr_reference_case()
    pf = tmp_path / "fmr.h5"
    bio.save_checkpoint_host(pf, Uc, Uf, bhs, **kw)
    with pytest.raises(bsn.BsnError, match="kind"):
        bio.load_checkpoint_uniform(pf)
    U, bhs_u, kwu = _reference_case()
    pu = tmp_path / "uniform.h5"
    bio.save_checkpoint_uniform(pu, U, bhs_u, **kwu)
    with pytest.raises(bsn.BsnError, match="two-level"):
        bio.load_checkpoint_host(pu)
    # a fine level that is not the box times the ratio is refused
    bad = dict(kw, box=(4, 6, 4, 5, 4, 5))
    with pytest.raises(bsn.BsnError, match="ratio"):
        bio.save_checkpoint_host(tmp_path / "bad.h5", Uc, Uf, bhs, **bad)


def test_trajectory_file_round_trip_and_layout(tmp_path):
    # init_trajectory_file / append_trajectory / read_trajectory
    # (src/io.jl:94-161): three records of two bodies
    B = bsn.BlackHole
    p = str(tmp_path / "traj.h5")
    bio.init_trajectory_file(p, 2)
    recs = []
    for k in range(3):
        bhs = [B(pos=[0.1 * k, 0.2, -0.3], vel=[0.0, 0.5 + k, 0.0], mass=0.5 + 0.01 * k, eps=1e-3,
                 c_code=20.0, r_floor=1e-3),
               B(pos=[-0.1 * k, 0.0, 0.4], vel=[0.3, 0.0, -k], mass=0.25, eps=1e-3, c_code=20.0,
                 r_floor=2e-2)]
        bio.append_trajectory(p, 0.5 * k, bhs)
        recs.append(bhs)
    t, data = bio.read_trajectory(p)
    np.testing.assert_array_equal(t, [0.0, 0.5, 1.0])
    assert len(data) == 2
    for i in range(2):
        np.testing.assert_array_equal(data[i].mass, [r[i].mass for r in recs])
        np.testing.assert_array_equal(data[i].rsink, [bsn.r_sink(r[i]) for r in recs])
        np.testing.assert_array_equal(data[i].pos, [r[i].pos for r in recs])
        np.testing.assert_array_equal(data[i].vel, [r[i].vel for r in recs])
    # r_sink = max(6 M / c^2, r_floor): the floor of the second body wins
    assert data[1].rsink[0] == 2e-2
    h = header(p)
    F = "H5T_IEEE_F64LE"
    assert h["time"] == ("DATASET", F, (3,))
    for i in (1, 2):
        assert h[f"bh{i}/mass"] == ("DATASET", F, (3,))
        assert h[f"bh{i}/rsink"] == ("DATASET", F, (3,))
        # Julia's (N, 3) arrays appear as (3, N)
        assert h[f"bh{i}/pos"] == ("DATASET", F, (3, 3))
        np.testing.assert_array_equal(h5_array(p, f"/bh{i}/pos", (3, 3), tmp_path),
                                      np.array([r[i - 1].pos for r in recs]).T)
    assert "H5S_UNLIMITED" in h5dump("-H", p)
    # an empty file reads back empty
    bio.init_trajectory_file(str(tmp_path / "empty.h5"), 1)
    t0, d0 = bio.read_trajectory(str(tmp_path / "empty.h5"))
    assert len(t0) == 0 and len(d0) == 1 and len(d0[0].mass) == 0
