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

Other workloads (--config): sod128, sedov128, batch48, fmr3.

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

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
VALU_FP64_TFLOPS = 78.6
VALU_FP32_TFLOPS = 157.3

# Algorithmic work per interface of one flux sweep (counted from the
# reference formulas, shared subexpressions counted once; a division or square
# root counts as one flop): WENO5-Z pair of 5 variables 620, two interface
# energies 24, HLLC 130, cell pressure 10.
FLOPS_PER_INTERFACE = 784
# Algorithmic HBM bytes per cell of the fused update kernel: read 15 flux
# words, U and Un (5 words each), write U (5 words).
def update_bytes_per_cell(prec: str) -> int:
    return 15 * (8 if prec == "fp64" else 4) + 15 * 8


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
    tune = b.autotune()
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
    g.set_state(U)
    g.ctx.synchronize()
    t0 = time.perf_counter()
    b.run(cfl, steps, graph=True)
    wall_graph = time.perf_counter() - t0
    zones = nx * ny * nz * steps
    return dict(wall=wall, wall_graph=wall_graph, zones=zones, ms_flux=ms_flux, ms_update=ms_upd,
                ms_step=ms_step, tune=tune, state=out, grid=g, batch=b)


def totals(U, dV):
    a = U[3:-3, 3:-3, 3:-3, :]
    return np.array([a[..., q].sum(dtype=np.float64) for q in range(5)]) * dV


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
                kind="port",
                sample=f"{nsteps} RK3 step(s) (cfl_dt_3d + euler3d_step) of the same "
                       f"{nx}x{ny}x{nz} state, FP64 C/OpenMP restatement, {sec:.1f} s")


def roofline(res, prec, interfaces_per_sweep, cells):
    """Dominant kernel of the step and its roofline position."""
    f = res["ms_flux"]
    flux_ms = sum(f) / 3.0
    upd_ms = res["ms_update"]
    peak = VALU_FP64_TFLOPS if prec == "fp64" else VALU_FP32_TFLOPS
    flux = dict(kernel="k_flux (x/y/z sweeps, avg)", bound="valu",
                achieved=FLOPS_PER_INTERFACE * interfaces_per_sweep / (flux_ms * 1e-3) / 1e12,
                peak=peak, unit="TFLOP/s", traffic=None)
    flux["frac"] = flux["achieved"] / peak
    upd_gbs = update_bytes_per_cell(prec) * cells / (upd_ms * 1e-3) / 1e9
    upd = dict(kernel="k_update (fused divergence + RK3 stage + floors)", bound="hbm",
               achieved=upd_gbs, peak=HBM_PEAK_GBS, unit="GB/s", frac=upd_gbs / HBM_PEAK_GBS,
               traffic=None)
    return flux, upd


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
    A, B = r64["state"], r32["state"]
    ra, rb = A[3:-3, 3:-3, 3:-3, 0], B[3:-3, 3:-3, 3:-3, 0]
    L1 = float(np.abs(ra - rb).sum() / np.abs(ra).sum())
    tA, tB = totals(A, dV), totals(B, dV)
    v64 = r64["zones"] / r64["wall"]
    v32 = r32["zones"] / r32["wall"]
    cells = n ** 3
    interfaces = (n + 1) * n * n
    flux, upd = roofline(r32, "split", interfaces, cells)
    flux64, upd64 = roofline(r64, "fp64", interfaces, cells)
    out = dict(
        metric="zone-updates/s per RK3 step (split FP32/FP64)",
        value=v32, unit="zone-updates/s",
        ms_per_step=1e3 * r32["wall"] / args.steps,
        dtype="f32-flux/f64-state",
        config=dict(workload=f"turbulent periodic box {n}^3, gamma 5/3, rms Mach 2, "
                             f"P(k)~k^-11/3 (2<=k<=32), CFL 0.4, device dt",
                    grid=[n, n, n], precision="split"),
        fp64=dict(value=v64, ms_per_step=1e3 * r64["wall"] / args.steps,
                  value_graph=r64["zones"] / r64["wall_graph"], roofline=flux64,
                  roofline_update=upd64, tiles=r64["tune"]),
        split_over_fp64=v32 / v64,
        value_graph=r32["zones"] / r32["wall_graph"],
        l1_density_split_vs_fp64=L1,
        mass_drift_split=float(abs(tB[0] - t0[0]) / t0[0]),
        energy_drift_split=float(abs(tB[4] - t0[4]) / t0[4]),
        mass_drift_fp64=float(abs(tA[0] - t0[0]) / t0[0]),
        energy_drift_fp64=float(abs(tA[4] - t0[4]) / t0[4]),
        roofline=flux, roofline_update=upd, tiles=r32["tune"],
        kernel_ms=dict(flux_xyz=r32["ms_flux"], update=r32["ms_update"],
                       fp64_flux_xyz=r64["ms_flux"], fp64_update=r64["ms_update"]),
    )
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(U, dims, d, gamma, bc, floors, 1)
    return out


def bench_sedov(bsn, args):
    n = args.n or 128
    g = 5.0 / 3.0
    dx = 1.0 / n
    U = bsn.problems.sedov_ic_3d(n, n, n, dx, dx, dx, g, E_blast=1.0 / (g - 1), rho_bg=1.0,
                                 P_floor=1e-5, x_offset=-0.5, y_offset=-0.5, z_offset=-0.5)
    floors = (0.0, 0.0)
    steps = args.steps or 200
    res = {}
    for prec in ("fp64", "split"):
        log(f"sedov {n}^3 {prec}")
        res[prec] = run_uniform(bsn, U, (n, n, n), (dx, dx, dx), g, "outflow", floors, prec, steps,
                                args.warmup)
    v64 = res["fp64"]["zones"] / res["fp64"]["wall"]
    v32 = res["split"]["zones"] / res["split"]["wall"]
    A, B = res["fp64"]["state"], res["split"]["state"]
    ra, rb = A[3:-3, 3:-3, 3:-3, 0], B[3:-3, 3:-3, 3:-3, 0]
    flux, upd = roofline(res["fp64"], "fp64", (n + 1) * n * n, n ** 3)
    out = dict(metric="zone-updates/s per RK3 step (FP64)", value=v64, unit="zone-updates/s",
               ms_per_step=1e3 * res["fp64"]["wall"] / steps, dtype="f64",
               config=dict(workload=f"Sedov-Taylor {n}^3, gamma 5/3, E0=1, {steps} steps, CFL 0.4",
                           grid=[n, n, n], precision="fp64"),
               split=dict(value=v32, l1_density_vs_fp64=float(np.abs(ra - rb).sum() / np.abs(ra).sum())),
               roofline=flux, roofline_update=upd)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(U, (n, n, n), (dx, dx, dx), g, "outflow", floors, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="turb256", choices=["turb256", "sedov128"])
    ap.add_argument("--n", type=int, default=None, help="override the grid size")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
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
    else:
        out = bench_sedov(bsn, args)
        args.steps = args.steps or 200
    line = dict(metric=out.pop("metric"), value=out.pop("value"), unit=out.pop("unit"), n_gpus=1,
                steps=args.steps, warmup=args.warmup, ms_per_step=out.pop("ms_per_step"),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype=out.pop("dtype"),
                data="synthetic", config=out.pop("config"))
    line.update(out)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
