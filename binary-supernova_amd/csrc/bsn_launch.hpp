// bsn_launch.hpp -- host-side launch helpers for the kernels in bsn_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "bsn_internal.hpp"

#define UPDATE_KZ 8   // cells per thread along z in the update and CFL kernels

namespace bsn {

struct LaunchCtx {
    hipStream_t stream;
    int cus;          // compute units of the device
};

// Flux sweep shapes: x 0..2 tiled, then FLUX_XWALK_SHAPE0 .. +2 short walks
// of 2 or 4 interfaces per lane (fluxes stored through LDS) and of 4 (stored
// directly); y and z 0..4 tiled, then FLUX_WALK_SHAPE0 .. +2 register walks
// of 16, 32 or 64 interfaces.
#define FLUX_XWALK_SHAPE0 3
#define FLUX_WALK_SHAPE0 5
void flux_shape_dims(int dir, int shape, int& W, int& H);
int flux_shape_count(int dir);
int flux_walk_len(int dir, int shape);
// Workgroups of a flux sweep shape that stay resident on the device (the
// persistent launch grid before it is clipped to the tile count).
int flux_slots(int dir, int prec, int shape, int cus);

hipError_t launch_flux(const LaunchCtx& L, int dir, int prec, int shape, const GridDev* grids,
                       const Tile* tiles, int ntiles, int max_wg);
hipError_t launch_update(const LaunchCtx& L, int prec, const GridDev* grids, const Tile* tiles,
                         int ntiles, int stage, int do_cfl, double* dU_out);
hipError_t launch_cfl(const LaunchCtx& L, const GridDev* grids, const Tile* tiles, int ntiles);
int zupdate_shape_count();
int zupdate_seg(int shape);
int zupdate_slots(int prec, int shape, int cus);
hipError_t launch_zupdate(const LaunchCtx& L, int prec, int shape, const GridDev* grids,
                          const Tile* tiles, int ntiles, int stage, int do_cfl);
hipError_t launch_ghost_fill(const LaunchCtx& L, const GridDev* grids, const int64_t* off,
                             int ngrids, int64_t total, int copy_un);
hipError_t launch_ghost_combine(const LaunchCtx& L, const GridDev* grids, const int64_t* off,
                                int ngrids, int64_t total, int stage);
hipError_t launch_floors(const LaunchCtx& L, const GridDev* grids, const int64_t* off, int ngrids,
                         int64_t total, int copy_un);
hipError_t launch_dt(const LaunchCtx& L, const GridDev* grids, int ngrids);
hipError_t launch_advance_t(const LaunchCtx& L, const GridDev* grids, int ngrids);
hipError_t launch_set_scal(const LaunchCtx& L, double* scal, int idx, double v);
hipError_t launch_fill_u64(const LaunchCtx& L, uint64_t* p, int64_t n, uint64_t v);
hipError_t launch_aos_to_soa(const LaunchCtx& L, const GridDev& G, const double* aos, double* dst);
hipError_t launch_soa_to_aos(const LaunchCtx& L, const GridDev& G, const double* src, double* aos);
hipError_t launch_flux_export(const LaunchCtx& L, const GridDev& G, int dir, double* out);
hipError_t launch_sg_force(const LaunchCtx& L, const GridDev& G, double* dU);
hipError_t launch_src(const LaunchCtx& L, const GridDev& G, double* dU, int which);
hipError_t launch_combine_all(const LaunchCtx& L, const GridDev& G, int stage, const double* dU,
                              double dt);
hipError_t launch_copy_state(const LaunchCtx& L, const GridDev& G, const double* src, double* dst);
// Blocks of one conserved-totals launch (bsn_conserved_totals and its
// allocation-free variant use the same count, hence the same sums), and the
// size of the variant's fixed scratch: the partials, then the five sums.
#define TOTALS_BLOCKS 1024
#define TOTALS_SCRATCH_DOUBLES (5 * TOTALS_BLOCKS + 8)
#define BSN_MAX_DEVICES 64
hipError_t totals_scratch_setup(int device, double** ptr);
hipError_t launch_totals(const LaunchCtx& L, const GridDev& G, double* part, int nblocks, double dV,
                         double* out);
// Buffers (child layout) a prolongation writes; none = the child's state.
struct ProlongDst {
    int n;
    double* p[3];
};
// Ucat: the interpolated parent box transposed (FmrDev.Ucit) for the x-slab
// ghost cells, or null to read Uca for every cell.
hipError_t launch_prolong(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr, int64_t total,
                          const double* Uca, const double* Ucat, int all_cells,
                          ProlongDst dst = ProlongDst{0, {}});
hipError_t launch_time_interp(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr,
                              int64_t total, double alpha);
// _restrict! (nfields = 5) or restrict_density_to_coarse! (nfields = 1)
hipError_t launch_restrict(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr, int64_t total,
                           int nfields = 5);
// _prolong_scalar_trilinear! of active-cell phi (x fastest), box at
// coarse active cell lo (1-based)
hipError_t launch_prolong_scalar(const LaunchCtx& L, const double* phc, const int nc[3], double* phf,
                                 const int nf[3], const int lo[3], int ratio);
hipError_t launch_scale(const LaunchCtx& L, double* a, int64_t n, double s);
hipError_t launch_fr_zero(const LaunchCtx& L, const FmrDev* fmr, int64_t total);
hipError_t launch_fr_coarse(const LaunchCtx& L, int prec, const GridDev* grids, const FmrDev* fmr,
                            int64_t nelem, double weight);
hipError_t launch_fr_fine(const LaunchCtx& L, int prec, const GridDev* grids, const FmrDev* fmr,
                          int64_t nelem, double weight);
hipError_t launch_bc_correct(const LaunchCtx& L, const GridDev* grids, const FmrDev* fmr, int64_t nelem);
hipError_t launch_child_dt(const LaunchCtx& L, const GridDev* grids, int parent, int child, int ratio);
hipError_t launch_fold_dt(const LaunchCtx& L, const GridDev* grids, int root, int child, double factor);

}  // namespace bsn
