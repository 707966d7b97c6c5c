// bsn_internal.hpp -- device data layout shared by the kernels and the host
// orchestration of libbsn_hip.
//
// Layout in HBM (one grid = one uniform patch: a problem or an FMR level):
//   U, Un, Uout  five FP64 fields each, structure of arrays (same layout).  Cell (q, i, j, k) with
//           padded indices i in [0, Nx), j in [0, Ny), k in [0, Nz) lives at
//           q*fs + (k*Ny + j)*pitch + XOFF + i.  XOFF puts the first active
//           cell (i = 3) on a 128-byte boundary; pitch is a multiple of 16
//           doubles, so every row starts on a cache line.
//   F[0..2] interface fluxes, five fields each, FP64 (plain form) or FP32
//           (split form), active-range interfaces only:
//             Fx(q, f, j, k) = Fx[q*ffs[0] + (k*ny + j)*fp[0] + f],     f in [0, nx]
//             Fy(q, i, f, k) = Fy[q*ffs[1] + (k*(ny+1) + f)*fp[1] + i], f in [0, ny]
//             Fz(q, i, j, f) = Fz[q*ffs[2] + (f*ny + j)*fp[2] + i],     f in [0, nz]
//           Interface f lies between active cells f-1 and f (0-based); it is
//           the reference's F?[q, NG+f, ...] (src/gpu_kernels.jl:51-154).
//   phi     (nz, ny, nx) row-major = the reference's (nx, ny, nz) column-major.
//   scal    per-grid device scalars (see SCAL_*).
#pragma once
#include <stdint.h>

#define BSN_NG 3
#define BSN_XOFF 13
#define BSN_MAXLEV 8

enum {
    SCAL_DT = 0,      // dt of the current step (FP64)
    SCAL_SMAX = 1,    // running max wave speed, bit pattern as uint64
    SCAL_T = 2,       // simulated time
    SCAL_TEND = 3,    // optional end time (dt is clipped to t_end - t when > 0)
    SCAL_CFL = 4,     // CFL number used by the device-side dt
    SCAL_TOT = 8,     // 5 conserved totals
    SCAL_COUNT = 16
};

#define BSN_MAX_BH 8

// Point masses and gas sinks of a grid (the reference's BlackHole list as
// passed to add_bh_gravity_source! / add_sink_sources!): positions and
// velocities, Plummer softening, sink radius and removal time per body, and
// the left edge of the active domain each source uses for cell centres.
struct SrcDev {
    int nbh, nsink, torque_free, pad;
    double dx, dy, dz;
    double gx0[3], sx0[3];
    double px[BSN_MAX_BH], py[BSN_MAX_BH], pz[BSN_MAX_BH], mass[BSN_MAX_BH], eps[BSN_MAX_BH];
    double spx[BSN_MAX_BH], spy[BSN_MAX_BH], spz[BSN_MAX_BH];
    double svx[BSN_MAX_BH], svy[BSN_MAX_BH], svz[BSN_MAX_BH];
    double rs[BSN_MAX_BH], ts[BSN_MAX_BH];
    double rho_sink_min;
};

struct GridDev {
    double* U;         // state the stage reads (stencils, ghost fills)
    double* Un;        // RK3 base state of the step
    double* Uout;      // where the stage update writes (== U in the in-place path)
    void* F[3];
    const double* phi;
    const SrcDev* src;  // point masses / sinks added to every stage's RHS, or null
    double* scal;
    int64_t fs;
    int64_t ffs[3];
    int nx, ny, nz;
    int Ny, Nz;
    int pitch;
    int fp[3];
    int bc;            // bit a set: periodic along axis a, else outflow
    int prec;          // 0 = exact FP64, 1 = split FP32/FP64
    int lmode;         // how a stage reads U: bit 0 = a ghost cell reads the
                       // active cell its index maps to under bc (no ghost fill
                       // needed); bit 1 = active cells are floored on load
    double idx, idy, idz;
    double gamma, gfac, rho_floor, P_floor;
    float gf32[3];     // gamma, gfac and 1/(gamma - 1) rounded to FP32 (split flux)
    int nzkeep;        // z interfaces whose fluxes the fused z-sweep + update kernel
    int zkeep[4];      // stores to F[2] (the FMR flux registers read them)
};

// A tile of work: which grid, and three integer coordinates whose meaning is
// fixed by the kernel that consumes the list.
struct Tile {
    int g, a, b, c;
};

// Column tiles (the y/z walks and the fused z-sweep + update) flag in a a
// half-width tile: 32 lanes along x and 8 rows per workgroup instead of 64 and
// 4, for the last x-tile of a grid whose rows end 32 cells or fewer into it.
#define COL_HALF (1 << 30)

// Fused z-sweep + update tiles pack the first cell and the column length in c.
#define ZT_K0_MASK 0x1FFFFF
#define ZT_LEN_SHIFT 21
#define ZT_MAX_LEN 1023

// FMR coupling between a parent grid and the nested child grid.
struct FmrDev {
    int parent, child;        // grid indices into the GridDev table
    int ratio;
    int lo[3], hi[3];         // child box in parent 1-based active indices
    double w[4 * 5];          // Lagrange weights (src/fmr3d.jl:32-46), [s][j]
    double* frc;              // coarse-side flux register, 6 faces
    double* frf;              // fine-side flux register
    int64_t froff[7];         // face offsets into a register (in doubles)
    double* Ucn;              // parent state at the start of its step
    double* Ucnp1;            // parent state at the end of its RK3 step
    double* Uci;              // parent state interpolated in time (parent layout)
    double* Ucit;             // the same over the plo..phi box, y fastest, then z, then x
    int plo[3], phi[3];       // padded parent cells any prolongation stencil reads
};
