"""binary_supernova_amd -- MI355X-native Euler right-hand side, SSP-RK3 step,
CFL reduction and fixed-mesh refinement of Binary-Supernova.

The compute runs in libbsn_hip.so (hand-written HIP for gfx950, C ABI in
include/bsn.h); this package is the host-side mirror of the reference's
Julia entry points (src/euler3d.jl, src/rk3.jl, src/fmr3d.jl,
src/fmr_gravity.jl, src/gravity_bh.jl, src/sinks.jl, src/self_gravity.jl,
src/diagnostics.jl, src/io.jl, src/checkpoint.jl) plus batching and the benchmark
workloads.
"""
from ._lib import BsnError, EXPORTS, LIB_PATH, load
from .euler3d import (NG, Context, Grid, apply_floors_3d_, bc_mask, cfl_dt_3d,
                      conserved_totals, euler3d_rhs_, euler3d_step_, fill_ghost_3d_outflow_,
                      fill_ghost_3d_periodic_, flux_shapes, rk3_step_, _apply_sg_force_level_)
from .fmr3d import (FMRGrid3D, FMRHierarchy, cfl_dt_fmr3d, fmr3d_step_, solve_fmr_poisson,
                    restrict_density_to_coarse_, _prolong_scalar_trilinear_)
from .batch import Batch
from .gravity import set_self_gravity, solve_poisson_isolated
from .sources import (BlackHole, accrete_, accrete_all_, add_bh_gravity_source_,
                      add_sink_sources_, gas_force_on_bh, gas_forces_on_bhs, r_sink,
                      set_point_masses, set_sinks, t_sink)
from . import diagnostics
from . import io
from . import problems

__all__ = [
    "BsnError", "EXPORTS", "LIB_PATH", "load", "NG", "Context", "Grid", "apply_floors_3d_",
    "bc_mask", "cfl_dt_3d", "conserved_totals", "euler3d_rhs_", "euler3d_step_",
    "fill_ghost_3d_outflow_", "fill_ghost_3d_periodic_", "flux_shapes", "rk3_step_",
    "_apply_sg_force_level_", "FMRGrid3D", "FMRHierarchy", "cfl_dt_fmr3d", "fmr3d_step_",
    "solve_fmr_poisson", "restrict_density_to_coarse_", "_prolong_scalar_trilinear_",
    "Batch", "problems", "BlackHole", "add_bh_gravity_source_", "add_sink_sources_", "r_sink",
    "set_point_masses", "set_sinks", "t_sink", "set_self_gravity", "solve_poisson_isolated",
    "accrete_", "accrete_all_", "gas_force_on_bh", "gas_forces_on_bhs", "diagnostics",
    "io",
]
