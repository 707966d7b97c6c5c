#!/usr/bin/env python
"""Benchmark of the Euler RHS + SSP-RK3 step on one MI355X.

Metric: zone-updates/s per RK3 step (sum over levels of active cells x RK3
steps at that level / wall seconds, CFL reduction and ghost fills included),
split FP32/FP64 against plain FP64, with the L1 density error of the split
form against the FP64 path and the relative mass / energy drift.

Default workload (turb256): 256^3 periodic turbulent box, gamma = 5/3,
rho = exp(s) with s a Gaussian field, P(k) ~ k^-11/3 for 2 <= k <= 32,
sigma(s) = 0.5, rms Mach 2, p = 1, CFL 0.4, K RK3 steps (default 50) with dt
computed on the device each step.  The timed region is the split form; the
FP64 form is run and timed the same way for the speed-up and the L1 error.

Other workloads (--config): sod128, sedov128, batch48, fmr3, selfgrav128 (the
coupled step with an FFT Poisson solve every RK3 stage), binary128 (the coupled
step with two bodies: gas forces, sinks, point masses, self-gravity, accretion),
io256 (snapshot and checkpoint files streamed from and to the device).

Prints ONE JSON line.  Inputs are resident in HBM before the timed region.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
LEAN = False
# precision -> [x, y, z, fused] shape ids to use instead of the autotuner
# (--shapes), so that a profiled run holds only the timed launches
SHAPES: dict = {}

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
VALU_FP64_TFLOPS = 78.6
VALU_FP32_TFLOPS = 157.3

# Algorithmic work per interface of one flux sweep (counted from the
# reference formulas, shared subexpressions counted once; a division or square
# root counts as one flop): WENO5-Z pair of 5 variables 620, two interface
# energies 24, HLLC 130, cell pressure 10.
FLOPS_PER_INTERFACE = 784
# Algorithmic HBM bytes per cell of the fused z-sweep + stage-update kernel
# (k_zupdate), averaged over the three launches of a step: every stage reads
# its input state (5 doubles), the x and y fluxes (5 words each: the face
# shared by two cells counts once) and writes the new state (5 doubles);
# stages 2 and 3 also read the RK3 base state (5 doubles), which in stage 1
# is the input itself.  The z fluxes never leave registers.
def update_bytes_per_cell(prec: str) -> float:
    w = 8 if prec == "fp64" else 4
    return 2 * 5 * 8 + 2 * 5 * w + (2.0 / 3.0) * 5 * 8


def profiled_traffic(kernel: str, prec: str):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) for the
    default 256^3 workload, as recorded in the committed profile summary
    (lines 'traffic <kernel> <prec> <bytes> [variant]'); None if absent."""
    f = ROOT / "profiles" / "k_zupdate_turb256_hbm.md"
    try:
        for line in f.read_text().splitlines():
            w = line.strip().strip("`").split()
            if len(w) >= 4 and w[0] == "traffic" and w[1] == kernel and w[2] == prec:
                return float(w[3])
    except OSError:
        pass
    return None


def load_package():
    name = "binary_supernova_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = ROOT / "binary-supernova_amd"
    spec = importlib.util.spec_from_file_location(name, pkg_dir / "__init__.py",
                                                  submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------

def run_uniform(bsn, U, dims, d, gamma, bc, floors, prec, steps, warmup, cfl=0.4):
    """Warm-up, then the timed region: `steps` device-dt RK3 steps with HIP
    events around every kernel launch.  Returns timings and the final state."""
    nx, ny, nz = dims
    g = bsn.Grid(nx, ny, nz, prec).set_state(U)
    g.set_params(*d, gamma, bc, *floors)
    b = bsn.Batch([g])
    if prec in SHAPES:
        sx, sy, sz, su = SHAPES[prec]
        b.set_shapes(sx, sy, sz)
        b.set_update_shape(su)
        tune = {"fixed": SHAPES[prec]}
    else:
        tune = b.autotune()
    shape_ids = b.get_shapes()
    if warmup:
        b.run(cfl, warmup, graph=True)
        g.set_state(U)
    g.ctx.synchronize()
    t0 = time.perf_counter()
    ms_flux, ms_upd, ms_step = b.profile(steps)
    g.ctx.synchronize()
    wall = time.perf_counter() - t0
    out = g.get_state()
    # the same steps replayed as one captured hipGraph per step
    wall_graph = float("nan")
    if not LEAN:
        g.set_state(U)
        g.ctx.synchronize()
        t0 = time.perf_counter()
        b.run(cfl, steps, graph=True)
        wall_graph = time.perf_counter() - t0
    zones = nx * ny * nz * steps
    return dict(wall=wall, wall_graph=wall_graph, zones=zones, ms_flux=ms_flux, ms_update=ms_upd,
                ms_step=ms_step, tune=tune, shape_ids=shape_ids, state=out, grid=g, batch=b)


def totals(U, dV):
    a = U[3:-3, 3:-3, 3:-3, :]
    return np.array([a[..., q].sum(dtype=np.float64) for q in range(5)]) * dV


def host_cpu() -> dict:
    """The host the CPU baseline ran on: CPU model and logical CPUs."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(model=model, nproc=os.cpu_count() or 1)


def cpu_baseline(U, dims, d, gamma, bc, floors, nsteps=1):
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as O
    V = U.copy()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        dt = O.cfl_dt_3d(V, *d, gamma, 0.4)
        O.euler3d_step(V, *d, dt, gamma, bc, *floors)
    sec = time.perf_counter() - t0
    nx, ny, nz = dims
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=nx * ny * nz * nsteps / sec, unit="zone-updates/s", cores=cores,
                kind="port", label="reference-restatement CPU baseline", host=host_cpu(),
                sample=f"{nsteps} RK3 step(s) (cfl_dt_3d + euler3d_step) of the same "
                       f"{nx}x{ny}x{nz} state, FP64 C/OpenMP restatement, {sec:.1f} s")


def roofline(res, prec, interfaces_per_sweep, cells, n=None):
    """Roofline positions.  The dominant kernel is the fused z-sweep +
    stage-update kernel (k_zupdate, the largest share of the step): HBM-bound
    by its algorithmic bytes, with the VALU rate of its z-flux work beside
    it.  The x and y flux sweeps are VALU-bound."""
    fl = [m for m in res["ms_flux"] if m > 0]
    flux_ms = sum(fl) / len(fl)
    upd_ms = res["ms_update"]
    peak = VALU_FP64_TFLOPS if prec == "fp64" else VALU_FP32_TFLOPS
    flux = dict(kernel="x and y flux sweeps (k_fxwalk, k_fwalk or k_flux as tuned), avg", bound="valu",
                achieved=FLOPS_PER_INTERFACE * interfaces_per_sweep / (flux_ms * 1e-3) / 1e12,
                peak=peak, unit="TFLOP/s", traffic=None)
    flux["frac"] = flux["achieved"] / peak
    nbytes = update_bytes_per_cell(prec) * cells
    upd_gbs = nbytes / (upd_ms * 1e-3) / 1e9
    traffic = profiled_traffic("k_zupdate", prec) if n == 256 else None
    upd = dict(kernel="k_zupdate (z sweep + flux divergence + RK3 stage + floors + CFL)",
               bound="hbm", achieved=upd_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
               frac=upd_gbs / HBM_PEAK_GBS, traffic=traffic, algorithmic_bytes=nbytes,
               ms_per_launch=upd_ms,
               valu=dict(achieved=FLOPS_PER_INTERFACE * interfaces_per_sweep / (upd_ms * 1e-3) / 1e12,
                         peak=peak, unit="TFLOP/s"))
    upd["valu"]["frac"] = upd["valu"]["achieved"] / peak
    flux["ms_per_launch"] = flux_ms
    return flux, upd


def standalone_update_roofline(res, prec, cells, steps=10):
    """The stand-alone update kernel, k_update: divergence of the three
    materialised flux arrays, RK3 stage, floors and CFL, in place (the step
    form of the FMR levels).  Timed with HIP events in the in-place form of
    the same batch and state.  Algorithmic bytes per cell, averaged over the
    stages: U read and written (80), the x, y and z fluxes (3 x 5 words) and
    the base state in stages 2 and 3 (2/3 x 40)."""
    b = res["batch"]
    b.set_mode(True)
    try:
        _, ms, _ = b.profile(steps)
    finally:
        b.set_mode(False)
    w = 8 if prec == "fp64" else 4
    per_cell = 2 * 5 * 8 + 3 * 5 * w + (2.0 / 3.0) * 5 * 8
    nbytes = per_cell * cells
    gbs = nbytes / (ms * 1e-3) / 1e9
    return dict(kernel="k_update (in-place form: flux divergence of three flux arrays + RK3 stage "
                       "+ floors + CFL)", bound="hbm", achieved=gbs, peak=HBM_PEAK_GBS,
                unit="GB/s", frac=gbs / HBM_PEAK_GBS, algorithmic_bytes=nbytes,
                bytes_per_cell=per_cell, ms_per_launch=ms, steps=steps)


def step_roofline(zone_updates_per_s, bytes_per_zone_update=320):
    """Whole-step HBM position: 320 B is the algorithmic FP64-state traffic of
    one SSP-RK3 step with perfect on-chip stencil reuse (stage 1 reads U and
    writes U1; stages 2 and 3 read Un and U(k) and write U(k+1): 40 words)."""
    a = zone_updates_per_s * bytes_per_zone_update / 1e9
    return dict(bound="hbm", achieved=a, peak=HBM_PEAK_GBS, unit="GB/s", frac=a / HBM_PEAK_GBS,
                bytes_per_zone_update=bytes_per_zone_update)


def bench_turb(bsn, args):
    n = args.n or 256
    U, dims, d = bsn.problems.turbulent_box(n, seed=1234)
    gamma, bc, floors = 5.0 / 3.0, "periodic", (0.0, 0.0)
    dV = d[0] * d[1] * d[2]
    t0 = totals(U, dV)
    log(f"turbulent box {n}^3: fp64 run")
    r64 = run_uniform(bsn, U, dims, d, gamma, bc, floors, "fp64", args.steps, args.warmup)
    log(f"turbulent box {n}^3: split run")
    r32 = run_uniform(bsn, U, dims, d, gamma, bc, floors, "split", args.steps, args.warmup)
    # calibration for HBM counters: one plain copy of the state (U -> Un),
    # 5 * field_stride doubles read and written (k_copy_state)
    import ctypes as C
    fs = C.c_int64()
    bsn.load().bsn_grid_device_state(r32["grid"].handle, None, C.byref(fs), None)
    bsn.load().bsn_rk3_save(r32["grid"].handle)
    copy_bytes = 5 * fs.value * 8
    A, B = r64["state"], r32["state"]
    ra, rb = A[3:-3, 3:-3, 3:-3, 0], B[3:-3, 3:-3, 3:-3, 0]
    L1 = float(np.abs(ra - rb).sum() / np.abs(ra).sum())
    tA, tB = totals(A, dV), totals(B, dV)
    v64 = r64["zones"] / r64["wall"]
    v32 = r32["zones"] / r32["wall"]
    cells = n ** 3
    interfaces = (n + 1) * n * n
    flux, upd = roofline(r32, "split", interfaces, cells, n)
    flux64, upd64 = roofline(r64, "fp64", interfaces, cells, n)
    out = dict(
        metric="zone-updates/s per RK3 step (split FP32/FP64)",
        value=v32, unit="zone-updates/s",
        ms_per_step=1e3 * r32["wall"] / args.steps,
        dtype="f32-flux/f64-state",
        config=dict(workload=f"turbulent periodic box {n}^3, gamma 5/3, rms Mach 2, "
                             f"P(k)~k^-11/3 (2<=k<=32), CFL 0.4, device dt",
                    grid=[n, n, n], precision="split"),
        fp64=dict(value=v64, ms_per_step=1e3 * r64["wall"] / args.steps,
                  value_graph=r64["zones"] / r64["wall_graph"], roofline=upd64,
                  roofline_flux=flux64, tiles=r64["tune"], shape_ids=r64["shape_ids"]),
        split_over_fp64=v32 / v64,
        value_graph=r32["zones"] / r32["wall_graph"],
        l1_density_split_vs_fp64=L1,
        mass_drift_split=float(abs(tB[0] - t0[0]) / t0[0]),
        energy_drift_split=float(abs(tB[4] - t0[4]) / t0[4]),
        mass_drift_fp64=float(abs(tA[0] - t0[0]) / t0[0]),
        energy_drift_fp64=float(abs(tA[4] - t0[4]) / t0[4]),
        roofline=upd, roofline_flux=flux, tiles=r32["tune"], shape_ids=r32["shape_ids"],
        kernel_ms=dict(flux_xyz=r32["ms_flux"], update=r32["ms_update"],
                       fp64_flux_xyz=r64["ms_flux"], fp64_update=r64["ms_update"]),
        copy_calibration_bytes=copy_bytes,
        roofline_step=step_roofline(v32),
    )
    log("stand-alone update kernel (in-place form)")
    out["roofline_update"] = standalone_update_roofline(r32, "split", cells)
    out["fp64"]["roofline_update"] = standalone_update_roofline(r64, "fp64", cells)
    if not LEAN:
        out["pcie_inclusive"] = host_array_rate(bsn, U, dims, d, gamma, r32["tune"])
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(U, dims, d, gamma, bc, floors, 4)
    return out


def host_array_rate(bsn, U, dims, d, gamma, tune, nsteps=3):
    """The host-array drop-in (bsn_euler3d_step_host, split precision): every
    call uploads the host state, steps it with a fixed dt and downloads it, so
    its rate includes both PCIe transfers and the layout conversions."""
    import ctypes as C
    lib = bsn.load()
    nx, ny, nz = dims
    V = np.ascontiguousarray(U.copy())
    dt = 1e-4
    def step():
        rc = lib.bsn_euler3d_step_host(C.c_void_p(V.ctypes.data), nx, ny, nz, *d, dt, gamma,
                                       7, 0.0, 0.0, 1)
        if rc != 0:
            raise RuntimeError(lib.bsn_last_error().decode())
    step()                                        # creates the cached grid
    t0 = time.perf_counter()
    for _ in range(nsteps):
        step()
    sec = time.perf_counter() - t0
    return dict(value=nx * ny * nz * nsteps / sec, unit="zone-updates/s",
                ms_per_step=1e3 * sec / nsteps,
                note="host-array drop-in: upload + step + download per call (PCIe both ways)")


def bench_sedov(bsn, args):
    n = args.n or 128
    g = 5.0 / 3.0
    dx = 1.0 / n
    U = bsn.problems.sedov_ic_3d(n, n, n, dx, dx, dx, g, E_blast=1.0 / (g - 1), rho_bg=1.0,
                                 P_floor=1e-5, x_offset=-0.5, y_offset=-0.5, z_offset=-0.5)
    return bench_uniform_pair(bsn, args, U, (n, n, n), (dx, dx, dx), g, "outflow", (1e-8, 1e-8),
                              f"Sedov-Taylor {n}^3 on [-0.5,0.5]^3, gamma 5/3, E0=1 within 2.5 dx, "
                              f"rho=1, p=1e-5, CFL 0.4, device dt", steps_default=200,
                              headline="fp64")


def bench_sod(bsn, args):
    """The reference's own driver loop (test/test_sod3d.jl:107-113): host dt
    from cfl_dt_3d, then euler3d_step!, on a device-resident grid."""
    n = args.n or 128
    U, dims, d = bsn.problems.sod_3d("x", n, 8, 1.4)
    nx, ny, nz = dims
    res = {}
    for prec in ("fp64", "split"):
        g = bsn.Grid(nx, ny, nz, prec).set_state(U)
        g.set_params(*d, 1.4, "outflow", 0.0, 0.0)
        g.ctx.synchronize()
        t0 = time.perf_counter()
        t, steps = 0.0, 0
        while t < 0.2:
            dt = min(bsn.cfl_dt_3d(g, nx, ny, nz, *d, 1.4, 0.4), 0.2 - t)
            bsn.euler3d_step_(g, nx, ny, nz, *d, dt, 1.4, bc="outflow")
            t += dt
            steps += 1
        wall = time.perf_counter() - t0
        res[prec] = dict(wall=wall, steps=steps, state=g.get_state())
    A, B = res["fp64"]["state"], res["split"]["state"]
    ra, rb = A[3:-3, 3:-3, 3:-3, 0], B[3:-3, 3:-3, 3:-3, 0]
    dV = d[0] * d[1] * d[2]
    t0, tA, tB = totals(U, dV), totals(A, dV), totals(B, dV)
    v = {p: nx * ny * nz * res[p]["steps"] / res[p]["wall"] for p in res}
    args.steps = res["fp64"]["steps"]
    out = dict(metric="zone-updates/s per RK3 step (fp64)", value=v["fp64"], unit="zone-updates/s",
               ms_per_step=1e3 * res["fp64"]["wall"] / res["fp64"]["steps"], dtype="f64",
               config=dict(workload=f"Sod tube {nx}x{ny}x{nz}, gamma 1.4, outflow, CFL 0.4 to "
                                    f"t=0.2 ({res['fp64']['steps']} steps), host dt loop",
                           grid=list(dims), precision="fp64"),
               split_value=v["split"], split_over_fp64=v["split"] / v["fp64"],
               l1_density_split_vs_fp64=float(np.abs(ra - rb).sum() / np.abs(ra).sum()),
               mass_drift_fp64=float(abs(tA[0] - t0[0]) / t0[0]),
               energy_drift_fp64=float(abs(tA[4] - t0[4]) / t0[4]),
               mass_drift_split=float(abs(tB[0] - t0[0]) / t0[0]),
               energy_drift_split=float(abs(tB[4] - t0[4]) / t0[4]))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(U, dims, d, 1.4, "outflow", (0.0, 0.0),
                                           res["fp64"]["steps"])
    return out


def bench_uniform_pair(bsn, args, U, dims, d, gamma, bc, floors, workload, steps_default,
                       headline):
    nx, ny, nz = dims
    dV = d[0] * d[1] * d[2]
    t0 = totals(U, dV)
    res = {}
    steps = args.steps or steps_default
    for prec in ("fp64", "split"):
        log(f"{workload}: {prec}")
        res[prec] = run_uniform(bsn, U, dims, d, gamma, bc, floors, prec, steps, args.warmup)
    A, B = res["fp64"]["state"], res["split"]["state"]
    ra, rb = A[3:-3, 3:-3, 3:-3, 0], B[3:-3, 3:-3, 3:-3, 0]
    tA, tB = totals(A, dV), totals(B, dV)
    v = {p: res[p]["zones"] / res[p]["wall"] for p in res}
    r = res[headline]
    flux, upd = roofline(r, headline, (nx + 1) * ny * nz, nx * ny * nz)
    out = dict(metric=f"zone-updates/s per RK3 step ({headline})", value=v[headline],
               unit="zone-updates/s", ms_per_step=1e3 * r["wall"] / steps,
               dtype="f64" if headline == "fp64" else "f32-flux/f64-state",
               config=dict(workload=workload, grid=list(dims), precision=headline),
               fp64_value=v["fp64"], split_value=v["split"], split_over_fp64=v["split"] / v["fp64"],
               l1_density_split_vs_fp64=float(np.abs(ra - rb).sum() / np.abs(ra).sum()),
               mass_drift_fp64=float(abs(tA[0] - t0[0]) / t0[0]),
               energy_drift_fp64=float(abs(tA[4] - t0[4]) / t0[4]),
               mass_drift_split=float(abs(tB[0] - t0[0]) / t0[0]),
               energy_drift_split=float(abs(tB[4] - t0[4]) / t0[4]),
               roofline=upd, roofline_flux=flux, tiles=r["tune"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(U, dims, d, gamma, bc, floors, 30)
    return out


def bench_batch(bsn, args):
    probs = bsn.problems.batch_problems(48, seed=4321)
    steps = args.steps or 20
    res = {}
    for prec in ("fp64", "split"):
        log(f"batch of 48: {prec}")
        grids = []
        for p in probs:
            g = bsn.Grid(*p["dims"], prec).set_state(p["U"])
            g.set_params(*p["d"], p["gamma"], p["bc"], *p["floors"])
            grids.append(g)
        b = bsn.Batch(grids)
        tune = b.autotune()
        b.run(0.4, args.warmup, graph=True)
        for g, p in zip(grids, probs):
            g.set_state(p["U"])
        grids[0].ctx.synchronize()
        t0 = time.perf_counter()
        b.run(0.4, steps, graph=True)
        wall = time.perf_counter() - t0
        states = [g.get_state() for g in grids]
        res[prec] = dict(wall=wall, zones=b.zone_updates_per_step() * steps, tune=tune, states=states)
    l1n = l1d = 0.0
    mass = {"fp64": [0.0, 0.0], "split": [0.0, 0.0]}
    energy = {"fp64": [0.0, 0.0], "split": [0.0, 0.0]}
    for i, p in enumerate(probs):
        A, B = res["fp64"]["states"][i], res["split"]["states"][i]
        ra, rb = A[3:-3, 3:-3, 3:-3, 0], B[3:-3, 3:-3, 3:-3, 0]
        l1n += float(np.abs(ra - rb).sum())
        l1d += float(np.abs(ra).sum())
        dV = p["d"][0] * p["d"][1] * p["d"][2]
        t0 = totals(p["U"], dV)
        for prec, S in (("fp64", A), ("split", B)):
            t1 = totals(S, dV)
            mass[prec][0] += t1[0] - t0[0]
            mass[prec][1] += t0[0]
            energy[prec][0] += t1[4] - t0[4]
            energy[prec][1] += t0[4]
    v = {pr: res[pr]["zones"] / res[pr]["wall"] for pr in res}
    cells = sum(p["dims"][0] * p["dims"][1] * p["dims"][2] for p in probs)
    return dict(metric="zone-updates/s per RK3 step (split, batched)", value=v["split"],
                unit="zone-updates/s", ms_per_step=1e3 * res["split"]["wall"] / steps,
                dtype="f32-flux/f64-state",
                config=dict(workload=f"batch of 48 uneven Sod/Sedov problems ({cells} cells), "
                                     f"persistent tile work-list, hipGraph step, device dt per problem",
                            precision="split"),
                fp64_value=v["fp64"], split_over_fp64=v["split"] / v["fp64"],
                l1_density_split_vs_fp64=l1n / l1d,
                mass_drift_fp64=abs(mass["fp64"][0]) / mass["fp64"][1],
                energy_drift_fp64=abs(energy["fp64"][0]) / energy["fp64"][1],
                mass_drift_split=abs(mass["split"][0]) / mass["split"][1],
                energy_drift_split=abs(energy["split"][0]) / energy["split"][1],
                tiles=res["split"]["tune"], tiles_fp64=res["fp64"]["tune"])


def bench_fmr(bsn, args):
    log("building the 3-level star")
    st = bsn.problems.fmr_star()
    steps = args.steps or 100
    dims = [lv[1] for lv in st["levels"]]
    dxs = [lv[2] for lv in st["levels"]]

    def hierarchy(prec):
        H = bsn.FMRHierarchy(dims, dxs, st["boxes"], st["ratios"], st["gamma"], "outflow",
                             *st["floors"], precision=prec)
        for g, lv in zip(H.levels, st["levels"]):
            g.set_state(lv[0])
        H.prolong_fine_ghosts(1)
        H.prolong_fine_ghosts(2)
        H.set_phis(st["phis"])
        return H

    res = {}
    for prec in ("fp64", "split"):
        log(f"fmr 3 levels: {prec}")
        H = hierarchy(prec)
        init = [g.get_state() for g in H.levels]
        tune = H.autotune()
        H.run(0.4, args.warmup, graph=True)
        for g, U in zip(H.levels, init):
            g.set_state(U)
        H.levels[0].ctx.synchronize()
        t0 = time.perf_counter()
        H.run(0.4, steps, graph=True)
        wall = time.perf_counter() - t0
        res[prec] = dict(wall=wall, zones=H.zone_updates_per_step() * steps,
                         states=[g.get_state() for g in H.levels], init=init, t=H.time()[0],
                         tune=[r["zupdate"]["best"] for r in tune])
        del H
    v = {pr: res[pr]["zones"] / res[pr]["wall"] for pr in res}
    A, B = res["fp64"]["states"][0], res["split"]["states"][0]
    dV = np.prod(dxs[0])
    t0 = totals(res["fp64"]["init"][0], dV)
    tA, tB = totals(A, dV), totals(B, dV)
    # The split error at equal times: with device dt each form takes its own
    # CFL steps, and on this problem the dt sequence itself is sensitive to
    # round-off (atmosphere floors), so after 100 steps the two runs stand at
    # different times.  Both forms are stepped again with the FP64 form's dt.
    log("fmr 3 levels: split vs fp64 on one dt sequence")
    H64, H32 = hierarchy("fp64"), hierarchy("split")
    for _ in range(steps):
        dt = H64.cfl_dt(0.4)
        H64.step(dt)
        H32.step(dt)
    ra = H64.levels[0].get_state()[3:-3, 3:-3, 3:-3, 0]
    rb = H32.levels[0].get_state()[3:-3, 3:-3, 3:-3, 0]
    l1 = float(np.abs(ra - rb).sum() / np.abs(ra).sum())
    del H64, H32
    return dict(metric="zone-updates/s per RK3 step (split, 3-level FMR)", value=v["split"],
                unit="zone-updates/s", ms_per_step=1e3 * res["split"]["wall"] / steps,
                dtype="f32-flux/f64-state",
                config=dict(workload="3-level FMR (192^3 + 96^3 + 128x128x96, ratio 2) n=3 "
                                     "polytrope + two Plummer point masses + self-potential phi",
                            precision="split"),
                fp64_value=v["fp64"], split_over_fp64=v["split"] / v["fp64"],
                l1_density_split_vs_fp64_base=l1,
                l1_note=f"root density after {steps} steps of one dt sequence (the FP64 form's CFL dt)",
                device_dt_end_times=dict(fp64=res["fp64"]["t"], split=res["split"]["t"]),
                fused_kernel_shapes=dict(fp64=res["fp64"]["tune"], split=res["split"]["tune"]),
                mass_drift_fp64_base=float(abs(tA[0] - t0[0]) / t0[0]),
                mass_drift_split_base=float(abs(tB[0] - t0[0]) / t0[0]),
                energy_drift_fp64_base=float(abs(tA[4] - t0[4]) / t0[4]),
                energy_drift_split_base=float(abs(tB[4] - t0[4]) / t0[4]))


def bench_selfgrav(bsn, args):
    """A gamma = 5/3 polytrope star (R = 1, rho_c = 1) on n^3 over [-2, 2]^3
    with gas self-gravity: every RK3 stage solves phi from its density with
    the zero-padded hipFFT convolution (a (2n)^3 real transform pair) and adds
    the source in the update kernel, as the coupled step does with
    add_self_gravity_source! (scripts/run_sn50_fiducial.jl:772-816)."""
    n = args.n or 128
    gamma, L = 5.0 / 3.0, 4.0
    d = (L / n,) * 3
    floors = (1e-10, 1e-12)
    U = bsn.problems.polytrope_ic_3d(n, n, n, *d, gamma, x0=-L / 2, y0=-L / 2, z0=-L / 2,
                                     rho_floor=floors[0], P_floor=floors[1])
    steps = args.steps
    res = {}
    for prec in ("split", "fp64"):
        g = bsn.Grid(n, n, n, prec).set_state(U)
        g.set_params(*d, gamma, "outflow", *floors)
        b = bsn.Batch([g])
        b.autotune()
        bsn.set_self_gravity(g, True)
        b.run(0.4, args.warmup, graph=False)
        g.set_state(U)
        g.ctx.synchronize()
        t0 = time.perf_counter()
        ms_flux, ms_upd, ms_step = b.profile(steps)
        g.ctx.synchronize()
        wall = time.perf_counter() - t0
        # the same grid and steps without the solve: the difference is the
        # cost of three Poisson solves per step
        bsn.set_self_gravity(g, False)
        g.set_state(U)
        b.profile(2)
        g.set_state(U)
        g.ctx.synchronize()
        _, _, ms_plain = b.profile(steps)
        res[prec] = dict(wall=wall, ms_step=ms_step, ms_plain=ms_plain, ms_update=ms_upd)
        log(f"self-gravity {prec}: {ms_step:.3f} ms/step, {ms_plain:.3f} ms/step without")
    # the host-array Poisson drop-in (upload rho, solve, download phi)
    rho = np.ascontiguousarray(U[3:-3, 3:-3, 3:-3, 0])
    bsn.solve_poisson_isolated(rho, n, n, n, *d)
    t0 = time.perf_counter()
    for _ in range(5):
        bsn.solve_poisson_isolated(rho, n, n, n, *d)
    ms_dropin = 1e3 * (time.perf_counter() - t0) / 5
    r32, r64 = res["split"], res["fp64"]
    cells = n ** 3
    out = dict(
        metric="zone-updates/s per RK3 step with per-stage self-gravity (split FP32/FP64)",
        value=cells * steps / r32["wall"], unit="zone-updates/s",
        ms_per_step=1e3 * r32["wall"] / steps, dtype="f32-flux/f64-state",
        config=dict(workload=f"gamma 5/3 polytrope star on {n}^3 over [-2,2]^3, outflow, "
                             f"phi from an isolated FFT Poisson solve every RK3 stage, "
                             f"CFL 0.4, device dt", grid=[n, n, n], precision="split"),
        fp64=dict(value=cells * steps / r64["wall"], ms_per_step=1e3 * r64["wall"] / steps),
        poisson=dict(ms_per_solve=(r32["ms_step"] - r32["ms_plain"]) / 3,
                     fft=(f"{2 * n}^3 zero-padded FP64 convolution: pruned register four-step "
                          f"transforms for power-of-two padded axes, hipFFT otherwise"),
                     step_share=(r32["ms_step"] - r32["ms_plain"]) / r32["ms_step"],
                     ms_step_without=r32["ms_plain"],
                     dropin_ms_per_solve_host_arrays=ms_dropin),
    )
    if not args.no_cpu:
        log("cpu baseline (one coupled step, numpy FFT Poisson)")
        sys.path.insert(0, str(ROOT))
        from oracle import poisson as PO
        V = U.copy()
        nsteps = 3                                   # about 10 s of CPU work
        t0 = time.perf_counter()
        for _ in range(nsteps):
            PO.self_gravity_step(V, *d, 1e-4, gamma, "outflow", *floors)
        sec = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = dict(
            value=cells * nsteps / sec, unit="zone-updates/s", cores=cores, kind="port",
            label="reference-restatement CPU baseline", host=host_cpu(),
            sample=f"{nsteps} RK3 steps, each with three numpy FFT Poisson solves, from the same "
                   f"{n}^3 state, FP64 C/OpenMP restatement, {sec:.1f} s")
    return out


def bench_binary(bsn, args):
    """The coupled step of the binary run (coupled_step!,
    scripts/run_sn50_fiducial.jl:756-819) with every gas operation on the
    device: a gamma = 5/3 polytrope star on n^3 over [-2, 2]^3 with two bodies
    (one at the centre, one at x = 1.2 on a circular orbit), torque-free
    sinks, point-mass gravity and self-gravity in every RK3 stage.  Per step:
    gas_force_on_bh for both bodies, a kick-drift of the bodies on the host
    (the N-body integrator is not part of this build), the step, accrete! for
    both bodies.  The gas diagnostics run once per step as well."""
    import copy
    n = args.n or 128
    gamma, L = 5.0 / 3.0, 4.0
    d = (L / n,) * 3
    o = (-L / 2,) * 3
    floors = (1e-10, 1e-12)
    U = bsn.problems.polytrope_ic_3d(n, n, n, *d, gamma, x0=o[0], y0=o[1], z0=o[2],
                                     rho_floor=floors[0], P_floor=floors[1])
    B = bsn.BlackHole
    bhs0 = [B(pos=[0.0, 0.0, 0.0], vel=[0.0, 0.0, 0.0], mass=0.5, eps=d[0], c_code=1e3,
              r_floor=2.5 * d[0]),
            B(pos=[1.2, 0.0, 0.0], vel=[0.0, 0.9, 0.0], mass=0.3, eps=d[0], c_code=1e3,
              r_floor=2.5 * d[0])]
    steps = args.steps
    g = bsn.Grid(n, n, n, "split").set_state(U)
    g.set_params(*d, gamma, "outflow", *floors)
    b = bsn.Batch([g])
    b.autotune()
    bsn.set_self_gravity(g, True)
    dt = 0.2 * bsn.cfl_dt_3d(g, n, n, n, *d, gamma, 0.4)

    def run(k, timed):
        bhs = copy.deepcopy(bhs0)
        g.set_state(U)
        t = dict(force=0.0, step=0.0, accrete=0.0, diag=0.0)
        g.ctx.synchronize()
        t_all = time.perf_counter()
        for _ in range(k):
            t0 = time.perf_counter()
            F = bsn.gas_forces_on_bhs(g, n, n, n, *d, bhs, *o)
            t1 = time.perf_counter()
            for bh, f in zip(bhs, F):              # kick-drift with the gas force only
                for c in range(3):
                    bh.vel[c] += dt * f[c] / bh.mass
                    bh.pos[c] += dt * bh.vel[c]
            bsn.set_point_masses(g, bhs, *o)
            bsn.set_sinks(g, bhs, *o, f_sink=1.0, torque_free=True, rho_sink_min=1e-6)
            t2 = time.perf_counter()
            b.step([dt])
            g.ctx.synchronize()
            t3 = time.perf_counter()
            bsn.accrete_all_(bhs, g, n, n, n, *d, *o, dt, rho_sink_min=1e-6)
            t4 = time.perf_counter()
            bsn.diagnostics.gas_diagnostics(g, n, n, n, *d, *o, bhs, gamma)
            t5 = time.perf_counter()
            t["force"] += t1 - t0
            t["step"] += t3 - t2
            t["accrete"] += t4 - t3
            t["diag"] += t5 - t4
        wall = time.perf_counter() - t_all
        return wall, t, bhs

    run(args.warmup, False)
    wall, t, bhs = run(steps, True)
    log(f"binary {n}^3: {1e3 * wall / steps:.3f} ms per coupled step")
    cells = n ** 3
    out = dict(
        metric="zone-updates/s per coupled RK3 step (point masses, sinks, self-gravity, "
               "gas forces, accretion; split FP32/FP64)",
        value=cells * steps / wall, unit="zone-updates/s",
        ms_per_step=1e3 * wall / steps, dtype="f32-flux/f64-state",
        config=dict(workload=f"gamma 5/3 polytrope star on {n}^3 over [-2,2]^3 with two bodies, "
                             f"torque-free sinks, point-mass gravity and self-gravity every RK3 "
                             f"stage, gas forces and accretion per step, fixed dt",
                    grid=[n, n, n], precision="split", bodies=2),
        ms_per_call={k: 1e3 * v / steps for k, v in t.items()},
        accreted_mass=float(sum(bh.mass for bh in bhs) - sum(bh.mass for bh in bhs0)),
    )
    if not args.no_cpu:
        log("cpu baseline (one coupled step on the host: FP64 restatement + numpy FFT)")
        sec, csteps = cpu_coupled_steps(bsn, U, n, d, o, gamma, floors, bhs0, dt, 3)
        out["cpu_baseline"] = dict(
            value=cells * csteps / sec, unit="zone-updates/s", cores=os.cpu_count() or 1,
            kind="port", label="reference-restatement CPU baseline", host=host_cpu(),
            sample=f"{csteps} coupled step(s) of the same {n}^3 state: per RK3 stage the FP64 "
                   f"C/OpenMP right-hand side, a numpy FFT isolated Poisson solve (single "
                   f"thread), the self-gravity, point-mass and sink sources, the stage "
                   f"combine and floors; gas forces and accretion sums of both bodies per "
                   f"step; {sec:.1f} s")
        out["gpu_ms_force_and_accrete"] = out["ms_per_call"]["force"] + out["ms_per_call"]["accrete"]
    return out


def cpu_coupled_steps(bsn, U0, n, d, o, gamma, floors, bhs0, dt, nsteps):
    """The binary run's coupled step (scripts/run_sn50_fiducial.jl:756-819) on
    the host with the oracle, for the CPU baseline: per stage the right-hand
    side, then the sources in the step's order (point masses, self-gravity
    from the stage's density, sinks), the stage combine and the floors; the
    bodies' gas forces and accretion sums once per step.  Returns (seconds,
    steps)."""
    import copy
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as O
    from oracle import poisson as PO
    NG = 3
    U = U0.copy()
    bhs = copy.deepcopy(bhs0)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        for bh in bhs:
            O.gas_force_on_bh(U, *d, bh.pos, bh.mass, bh.eps, *o)
        pos = [bh.pos for bh in bhs]
        vel = [bh.vel for bh in bhs]
        Un = U.copy()
        for stage in (1, 2, 3):
            dU = O.euler3d_rhs(U, *d, gamma, "outflow", *floors)[0]
            O.bh_gravity_source(dU, U, *d, *o, pos, [bh.mass for bh in bhs], [bh.eps for bh in bhs])
            phi = PO.solve_poisson_isolated(np.ascontiguousarray(U[NG:-NG, NG:-NG, NG:-NG, 0]), *d)
            O.apply_sg_force(dU, U, phi, *d)
            O.sink_sources(dU, U, *d, *o, pos, vel, [bsn.r_sink(bh) for bh in bhs],
                           [bsn.t_sink(bh) for bh in bhs], True, 1e-6)
            O.rk3_combine(stage, U, Un, dU, dt)
            O.apply_floors(U, *floors, gamma)
        for bh in bhs:
            O.accrete_sums(U, *d, *o, bh.pos, bsn.r_sink(bh), bsn.t_sink(bh), dt, 1e-6)
    return time.perf_counter() - t0, nsteps


def bench_io(bsn, args):
    """Snapshot and checkpoint files of the n^3 turbulent box (src/io.jl:29-51,
    src/checkpoint.jl:244-283) written from and read into the GPU-resident
    grid through pinned slabs, against the reference's way (copy the whole
    state to the host, then write), into a temporary directory (page cache:
    no fsync).  value = snapshot bytes written per second from the device."""
    import ctypes as C
    import shutil
    import tempfile
    from binary_supernova_amd import _lib
    from binary_supernova_amd import io as bio
    n = args.n or 256
    U, dims, d = bsn.problems.turbulent_box(n, seed=1234)
    g = bsn.Grid(n, n, n).set_state(U)
    reps = args.steps
    tmp = tempfile.mkdtemp(prefix="bsn_io_", dir=os.environ.get("BSN_IO_DIR", "/tmp"))
    snap_bytes = 5 * n ** 3 * 8
    ckpt_bytes = 5 * (n + 6) ** 3 * 8
    kw = dict(t=1.0, step=10, dt_last=1e-3, t_snap=1.1, t_bh2_sink_on=0.0, nx=n, ny=n, nz=n,
              dx=d[0], dy=d[1], dz=d[2], gamma=5 / 3, x0=0.0, y0=0.0, z0=0.0, rho_floor=0.0,
              P_floor=0.0)

    def timed(fn):
        fn()                                          # warm-up (allocations, page cache)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    def noop(user, part, k0, nk, data, count):
        return 0
    cb = _lib.SLAB_FN(noop)
    try:
        snap, ckpt = os.path.join(tmp, "snap.h5"), os.path.join(tmp, "ckpt.h5")
        t_pipe = timed(lambda: _lib.call("bsn_grid_export_slabs", g.handle, _lib.LAYOUT_FIELDS, 0,
                                         cb, None))
        t_snap = timed(lambda: bio.write_snapshot(snap, g, n, n, n, *d, 1.0, 5 / 3))
        t_read = timed(lambda: bio.read_snapshot_(g, snap))
        t_ckpt = timed(lambda: bio.save_checkpoint_uniform(ckpt, g, [], **kw))
        t_load = timed(lambda: bio.load_checkpoint_uniform(ckpt, grid=g))
        host = np.empty(g.shape)

        def ref_snap():
            g.get_state(host)
            bio.write_snapshot(snap, host, n, n, n, *d, 1.0, 5 / 3)
        t_ref = timed(ref_snap)

        def ref_ckpt():
            g.get_state(host)
            bio.save_checkpoint_uniform(ckpt, host, [], **kw)
        t_refc = timed(ref_ckpt)
        ok = np.array_equal(g.get_state(), U)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    gbs = lambda b, t: b / t / 1e9
    return dict(
        metric="snapshot GB/s written from the GPU-resident state (HDF5, page cache)",
        value=gbs(snap_bytes, t_snap), unit="GB/s", ms_per_step=1e3 * t_snap, dtype="f64",
        config=dict(workload=f"snapshot and uniform checkpoint of the turbulent box {n}^3 "
                             f"(HDF5 files of src/io.jl and src/checkpoint.jl), median of {reps}",
                    grid=[n, n, n]),
        snapshot=dict(bytes=snap_bytes, write_ms=1e3 * t_snap, read_ms=1e3 * t_read,
                      write_gbs=gbs(snap_bytes, t_snap), read_gbs=gbs(snap_bytes, t_read),
                      reference_way_ms=1e3 * t_ref,
                      reference_way="whole state to the host, then write (src/io.jl:36-40)"),
        checkpoint=dict(bytes=ckpt_bytes, save_ms=1e3 * t_ckpt, load_ms=1e3 * t_load,
                        save_gbs=gbs(ckpt_bytes, t_ckpt), load_gbs=gbs(ckpt_bytes, t_load),
                        reference_way_ms=1e3 * t_refc),
        slab_pipeline=dict(ms=1e3 * t_pipe, gbs=gbs(snap_bytes, t_pipe),
                           note="pack + PCIe copy of the five fields into pinned slabs, "
                                "with a consumer that does nothing"),
        state_restored_bitwise=bool(ok),
    )


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="turb256",
                    choices=["turb256", "sedov128", "sod128", "batch48", "fmr3", "selfgrav128",
                             "binary128", "io256"])
    ap.add_argument("--n", type=int, default=None, help="override the grid size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lean", action="store_true", help="skip the hipGraph replay pass and "
                    "the host-array rate (for profiler runs)")
    ap.add_argument("--shapes", action="append", default=[], metavar="PREC=X,Y,Z,U",
                    help="use these shape ids (as a line's shape_ids reports them) for one "
                         "precision (fp64 or split) instead of the autotuner")
    args = ap.parse_args()
    global LEAN
    LEAN = args.lean
    for spec in args.shapes:
        prec, ids = spec.split("=")
        SHAPES[prec] = [int(v) for v in ids.split(",")]
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    bsn = load_package()
    ctx = bsn.Context.default()
    log("device:", ctx.info())
    if args.config == "turb256":
        args.steps = args.steps or 50
        out = bench_turb(bsn, args)
    elif args.config == "sedov128":
        out = bench_sedov(bsn, args)
        args.steps = args.steps or 200
    elif args.config == "batch48":
        out = bench_batch(bsn, args)
        args.steps = args.steps or 20
    elif args.config == "fmr3":
        out = bench_fmr(bsn, args)
        args.steps = args.steps or 100
    elif args.config == "selfgrav128":
        args.steps = args.steps or 20
        out = bench_selfgrav(bsn, args)
    elif args.config == "binary128":
        args.steps = args.steps or 20
        out = bench_binary(bsn, args)
    elif args.config == "io256":
        args.steps = args.steps or 3
        out = bench_io(bsn, args)
    else:
        args.steps = args.steps or 100
        out = bench_sod(bsn, args)
    line = dict(metric=out.pop("metric"), value=out.pop("value"), unit=out.pop("unit"), n_gpus=1,
                steps=args.steps, warmup=args.warmup, ms_per_step=out.pop("ms_per_step"),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype=out.pop("dtype"),
                data="synthetic", config=out.pop("config"))
    line.update(out)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
