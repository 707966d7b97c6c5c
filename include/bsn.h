/*
 * bsn.h -- C ABI of libbsn_hip: the Euler right-hand side, the SSP-RK3 step,
 * the CFL time step and the fixed-mesh-refinement step of Binary-Supernova,
 * on AMD Instinct MI355X (gfx950).
 *
 * Every entry point is extern "C", takes plain pointers and scalars, returns
 * an int status (BSN_OK = 0) and never throws; bsn_last_error() describes the
 * last failure of the calling thread.  A call is synchronous (the device work
 * it enqueues has finished when it returns) unless its name ends in _async.
 *
 * Arrays crossing the boundary use the reference's Julia layout:
 *   state U   (5, nx+6, ny+6, nz+6) Float64, column-major, field index fastest
 *             (src/euler3d.jl:1-8), ghost cells included;
 *   fluxes    Fx (5, Nx+1, Ny, Nz), Fy (5, Nx, Ny+1, Nz), Fz (5, Nx, Ny, Nz+1)
 *             (src/euler3d.jl:231-233);
 *   phi       (nx, ny, nz) Float64 column-major, active cells only
 *             (src/fmr_gravity.jl:146).
 * `where` says whether such a pointer is host memory (BSN_HOST) or device
 * memory on the grid's GPU (BSN_DEVICE).
 *
 * Boundary conditions are a per-axis bit mask: bit a set = periodic along
 * axis a (x = bit 0), otherwise outflow.  The reference's bc=:outflow is 0
 * and bc=:periodic is 7 (src/euler3d.jl:185-194).
 */
#ifndef BSN_H
#define BSN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSN_OK 0
#define BSN_ERR_ARG 1        /* invalid argument or shape */
#define BSN_ERR_HIP 2        /* a HIP runtime call failed */
#define BSN_ERR_NONFINITE 3  /* non-finite value detected where checked */
#define BSN_ERR_STATE 4      /* call not valid in the object's state */
#define BSN_ERR_IO 5         /* a file could not be read or written, or holds the wrong data */

#define BSN_BC_OUTFLOW 0
#define BSN_BC_PERIODIC 7

#define BSN_PREC_FP64 0      /* all-FP64, bit-identical to the reference */
#define BSN_PREC_SPLIT 1     /* FP32 WENO/HLLC, FP64 state, combine, reductions */

#define BSN_HOST 0
#define BSN_DEVICE 1

typedef struct bsn_ctx bsn_ctx;
typedef struct bsn_grid bsn_grid;
typedef struct bsn_batch bsn_batch;
typedef struct bsn_fmr bsn_fmr;

/* Library and context ---------------------------------------------------- */
int bsn_version(void);
const char* bsn_last_error(void);
int bsn_ctx_create(int device, bsn_ctx** out);
int bsn_ctx_destroy(bsn_ctx* ctx);
int bsn_ctx_synchronize(bsn_ctx* ctx);
int bsn_ctx_info(bsn_ctx* ctx, char* buf, int len);

/* A uniform grid (one problem or one FMR level) resident on the device.
 * Replaces the (U, nx, ny, nz) argument group of the reference's functions
 * and the FMRLevel3D struct (src/fmr3d.jl:57-71).  One x-y plane of one
 * field, ghosts included, must stay under 2 GiB (nx, ny up to about 16000):
 * the kernels address a plane through 32-bit offsets; BSN_ERR_ARG otherwise. */
int bsn_grid_create(bsn_ctx* ctx, int64_t nx, int64_t ny, int64_t nz, int precision,
                    bsn_grid** out);
int bsn_grid_destroy(bsn_grid* g);
int bsn_grid_set_state(bsn_grid* g, const double* U, int where);
int bsn_grid_get_state(bsn_grid* g, double* U, int where);
/* phi for the gravity source (_apply_sg_force_level!, src/fmr_gravity.jl:146);
 * NULL removes it.  The grid keeps its own device copy.  BSN_ERR_STATE while
 * self-gravity is on (the step solves phi itself). */
int bsn_grid_set_phi(bsn_grid* g, const double* phi, int where);
/* Geometry and physics of the grid, used by the step functions below. */
int bsn_grid_set_params(bsn_grid* g, double dx, double dy, double dz, double gamma, int bc,
                        double rho_floor, double P_floor);
int bsn_grid_device_state(bsn_grid* g, double** U_soa, int64_t* field_stride, int64_t* pitch);

/* Drop-ins for the reference's uniform-grid entry points ----------------- */
/* fill_ghost_3d_outflow! / fill_ghost_3d_periodic!  (src/euler3d.jl:33,51) */
int bsn_fill_ghost_3d(bsn_grid* g, int bc);
/* apply_floors_3d!  (src/euler3d.jl:73) */
int bsn_apply_floors_3d(bsn_grid* g, double rho_floor, double P_floor, double gamma);
/* euler3d_rhs!(dU, Fx, Fy, Fz, U, nx, ny, nz, dx, dy, dz, gamma; bc, rho_floor,
 * P_floor, fill_ghosts)  (src/euler3d.jl:182).  dU, Fx, Fy, Fz receive the
 * reference-layout arrays (any may be NULL); U (the grid) gets its ghosts and
 * floors as in the reference. */
int bsn_euler3d_rhs(bsn_grid* g, double* dU, double* Fx, double* Fy, double* Fz, int where,
                    double dx, double dy, double dz, double gamma, int bc, double rho_floor,
                    double P_floor, int fill_ghosts);
/* euler3d_step!(U, nx, ny, nz, dx, dy, dz, dt, gamma; bc, rho_floor, P_floor)
 * (src/euler3d.jl:220).  If the grid has a phi, its source is added after the
 * flux divergence of every stage, as fmr3d_step! does (src/fmr3d.jl:509-513). */
int bsn_euler3d_step(bsn_grid* g, double dx, double dy, double dz, double dt, double gamma,
                     int bc, double rho_floor, double P_floor);
/* cfl_dt_3d(U, nx, ny, nz, dx, dy, dz, gamma, cfl) -> dt  (src/euler3d.jl:157) */
int bsn_cfl_dt_3d(bsn_grid* g, double dx, double dy, double dz, double gamma, double cfl,
                  double* dt_out);
/* _apply_sg_force_level!(dU, U, phi, ...)  (src/fmr_gravity.jl:146): dU
 * (reference layout) is updated in place with the grid's phi and state. */
int bsn_apply_sg_force(bsn_grid* g, double* dU, int where, double dx, double dy, double dz);
/* rk3_step!(U, L!, dt, tmp1, tmp2) building blocks (src/rk3.jl:14-31):
 * bsn_rk3_save copies U to the grid's base state (Un .= U); bsn_rk3_combine
 * applies stage 1, 2 or 3 of the Shu-Osher combination with dU (reference
 * layout) over the whole array. */
int bsn_rk3_save(bsn_grid* g);
int bsn_rk3_combine(bsn_grid* g, int stage, const double* dU, int where, double dt);
/* Conservation totals sum(U[q, active]) * dV, q = 1..5 (src/diagnostics.jl:21-58). */
int bsn_conserved_totals(bsn_grid* g, double dV, double out[5]);
/* No counterpart in the reference: bsn_conserved_totals for callers that
 * cannot allocate device memory (code under hipGraph capture, serving loops
 * that forbid hipMalloc).  The 1024 block partial sums live in a fixed
 * library array of 5 x 1024 + 8 doubles (41 KB) on each device, held until
 * the process ends, instead of a buffer allocated on the first call; the sums
 * are those of bsn_conserved_totals, bit for bit.  The scratch does not grow
 * with the grid, so no grid is too large for it.  out: 5 doubles in host
 * memory (where = BSN_HOST; the call waits for the stream) or device memory
 * (BSN_DEVICE; nothing waits, so the call can be captured).  Calls that
 * overlap on one device race on the array. */
int bsn_conserved_totals_noalloc(bsn_grid* g, double dV, double* out, int where);

/* Point masses and gas sinks ---------------------------------------------- */
/* The bodies of the reference's BlackHole list (src/BinarySupernova.jl:130-137)
 * as plain arrays: pos and vel hold 3 values per body (x, y, z); n <= 8.
 * x0, y0, z0 are the left edge of the active domain, so active cell ii
 * (1-based) has centre x0 + (ii - 0.5)*dx.  Sink radius and removal time are
 * the caller's r_sink(bh) and t_sink(bh, f_sink) (src/BinarySupernova.jl:159-171). */
#define BSN_MAX_BH 8
/* add_bh_gravity_source!(dU, U, nx, ny, nz, dx, dy, dz, bhs, x0, y0, z0)
 * (src/gravity_bh.jl:58-78): adds point-mass gravity from the grid's state to
 * dU (reference layout). */
int bsn_add_bh_gravity_source(bsn_grid* g, double* dU, int where, double dx, double dy, double dz,
                              int n, const double* pos, const double* mass, const double* eps,
                              double x0, double y0, double z0);
/* add_sink_sources!(dU, U, ...; f_sink, torque_free, rho_sink_min)
 * (src/sinks.jl:32-58): adds the sink drains of the n bodies to dU. */
int bsn_add_sink_sources(bsn_grid* g, double* dU, int where, double dx, double dy, double dz,
                         int n, const double* pos, const double* vel, const double* r_sink,
                         const double* t_sink, int torque_free, double rho_sink_min, double x0,
                         double y0, double z0);
/* Sources every later step of the grid adds to each stage's RHS, in the order
 * of the reference's coupled step (scripts/run_sn50_fiducial.jl:772-816):
 * flux divergence, point-mass gravity, the phi source, sinks.  n = 0 removes
 * them.  The step uses the grid's cell sizes. */
int bsn_grid_set_point_masses(bsn_grid* g, int n, const double* pos, const double* mass,
                              const double* eps, double x0, double y0, double z0);
int bsn_grid_set_sinks(bsn_grid* g, int n, const double* pos, const double* vel,
                       const double* r_sink, const double* t_sink, int torque_free,
                       double rho_sink_min, double x0, double y0, double z0);

/* Gas self-gravity ---------------------------------------------------------- */
/* solve_poisson_isolated(rho, nx, ny, nz, dx, dy, dz) -> phi
 * (src/self_gravity.jl:35-76): Hockney-Eastwood zero-padded FFT convolution
 * with the softened kernel -1/sqrt(r^2 + eps^2), eps^2 = (dx dy dz)^(2/3).  rho
 * and phi are (nx, ny, nz) column-major arrays of the grid's active shape. */
int bsn_solve_poisson_isolated(bsn_grid* g, const double* rho, int where_rho, double* phi,
                               int where_phi, double dx, double dy, double dz);
/* on = 1: every later step solves phi from the density of each stage's input
 * (after its floors) and adds the self-gravity source with it, as the
 * reference's coupled step does with add_self_gravity_source!
 * (src/self_gravity.jl:93-115).  It replaces a phi set by bsn_grid_set_phi;
 * on = 0 removes it.  Uniform grids and batches; FMR levels keep their own phi. */
int bsn_grid_set_self_gravity(bsn_grid* g, int on);

/* What the bodies and the run's diagnostics read from the gas ---------------
 * Device reductions over the grid's active cells; the host receives only the
 * sums.  Bodies as above (pos: 3 values per body, n <= BSN_MAX_BH).  The
 * per-cell terms follow the reference's operation order; the order of the sums
 * differs (they are deterministic from call to call). */
/* gas_force_on_bh(U, nx, ny, nz, dx, dy, dz, bh, x0, y0, z0) for n bodies in one
 * pass over the density (src/gravity_bh.jl:91-113): F[3*b + c] is the force of
 * the gas on body b, +sum rho M (r - r_b) / (|r - r_b|^2 + eps^2)^(3/2) dV. */
int bsn_gas_force_on_bh(bsn_grid* g, double dx, double dy, double dz, int n, const double* pos,
                        const double* mass, const double* eps, double x0, double y0, double z0,
                        double* F);
/* accrete!(bh, U, nx, ny, nz, dx, dy, dz, x0, y0, z0, dt; f_sink, rho_sink_min)
 * (src/sinks.jl:73-109) for n bodies: the sums over the cells with r < r_sink[b]
 * and rho > rho_sink_min run on the device over each body's sink box, then vel
 * and mass (host arrays, updated in place) change as accrete! changes bh.vel and
 * bh.mass when the accreted mass is > 0.  r_sink and t_sink are the caller's
 * r_sink(bh) and t_sink(bh, f_sink).  dm_out (n values, may be NULL) receives the
 * accreted masses. */
int bsn_accrete(bsn_grid* g, double dx, double dy, double dz, int n, const double* pos,
                double* vel, double* mass, const double* r_sink, const double* t_sink,
                double rho_sink_min, double x0, double y0, double z0, double dt, double* dm_out);
/* The gas diagnostics of src/diagnostics.jl in one pass over the state:
 * out[0] gas_energy_total (:21-26), out[1] gas_kinetic_total (:33-44),
 * out[2..4] gas_momentum_total (:51-59), out[5..7] gas_angular_momentum_total
 * about the origin (:67-94), out[8] bound_gas_mass with the potential of the n
 * bodies and gamma (:137-175). */
int bsn_gas_diagnostics(bsn_grid* g, double dx, double dy, double dz, double x0, double y0,
                        double z0, int n, const double* pos, const double* mass,
                        const double* eps, double gamma, double out[9]);

/* Streaming a grid's state to and from files ------------------------------
 * The state crosses PCIe in slabs of k-planes through two pinned host
 * buffers: for each slab the device packs the file layout, copies it to one
 * buffer and the callback consumes it while the next slab is packed and copied
 * into the other (export); or the callback fills one buffer while the previous
 * slab is copied up and unpacked (import).  Layouts:
 *   BSN_LAYOUT_FIELDS  five parts, part q = field q of the active cells as a
 *                      Julia (nx, ny, nz) array: the /gas datasets of
 *                      write_snapshot (src/io.jl:29-51); k runs over 0..nz-1;
 *   BSN_LAYOUT_REF     one part, the reference state array with ghosts,
 *                      (5, nx+6, ny+6, nz+6): /grid/U of save_checkpoint_uniform
 *                      (src/checkpoint.jl:244-283); k runs over 0..nz+5.
 * A slab holds nk k-planes starting at k0, count values in the layout's order
 * (k slowest).  slab_k <= 0 picks about 32 MB per slab.  A callback returning
 * nonzero stops the stream with BSN_ERR_IO.  data is valid only during the call. */
#define BSN_LAYOUT_FIELDS 0
#define BSN_LAYOUT_REF 1
typedef int (*bsn_slab_fn)(void* user, int part, int64_t k0, int64_t nk, double* data,
                           int64_t count);
int bsn_grid_export_slabs(bsn_grid* g, int layout, int64_t slab_k, bsn_slab_fn sink, void* user);
int bsn_grid_import_slabs(bsn_grid* g, int layout, int64_t slab_k, bsn_slab_fn source,
                          void* user);
/* The grid's active extents (nx, ny, nz). */
int bsn_grid_dims(bsn_grid* g, int64_t n[3]);

/* Host-array convenience drop-ins: the reference signature on a host Julia
 * Array U; a cached per-thread grid uploads, runs and downloads. */
int bsn_euler3d_step_host(double* U, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                          double dz, double dt, double gamma, int bc, double rho_floor,
                          double P_floor, int precision);
int bsn_cfl_dt_3d_host(const double* U, int64_t nx, int64_t ny, int64_t nz, double dx, double dy,
                       double dz, double gamma, double cfl, double* dt_out);

/* Batches: independent grids of any shapes swept by one persistent launch
 * per kernel over a tile work-list.  All grids share one precision and must
 * have parameters set (bsn_grid_set_params). */
int bsn_batch_create(bsn_ctx* ctx, int ngrids, bsn_grid* const* grids, bsn_batch** out);
int bsn_batch_destroy(bsn_batch* b);
/* Times every tile shape of every kernel on the current states and keeps the
 * fastest; writes a JSON report of the choice into report (may be NULL).
 * Without it a batch (and every FMR level and single-grid call) starts from
 * the shapes the autotuner keeps for a 256^3 box: split x 0, y/z 2, fused
 * kernel 11; FP64 x 2, y/z 2, fused kernel 2. */
int bsn_batch_autotune(bsn_batch* b, char* report, int report_len);
/* Sets the flux shape ids explicitly (x: 0..2 tiled, 3..5 walks of 2, 4 and 4
 * interfaces per lane through LDS-staged rows, the last storing its fluxes
 * directly; y/z: 0..4 tiled, 5..7
 * register walks along the sweep of 16, 32 or 64 interfaces per thread). */
int bsn_batch_set_shapes(bsn_batch* b, int sx, int sy, int sz);
/* Sets the shape id (0..11) of the fused z-sweep + stage-update kernel used
 * by the out-of-place step: shape = 3 * variant + segmentation.  Variant 0
 * plain, 1 loads the next cell and the update operands one cell ahead, 2 keeps
 * the window's raw cells in an LDS ring instead of reading them twice, 3 the
 * ring plus an early load of the next window cell.  Segmentation 0 and 1: z
 * runs of 16 or 32 cells per column tile; 2: one run length per grid chosen
 * so that the tiles fill the resident workgroups evenly. */
int bsn_batch_set_update_shape(bsn_batch* b, int shape);
/* The shape ids in use: x, y, z flux shapes and the fused kernel's shape (as
 * the autotuner left them), for bsn_batch_set_shapes / _set_update_shape. */
int bsn_batch_get_shapes(bsn_batch* b, int shapes[4]);
/* Step form: 0 (default) = out of place, z sweep fused with the stage update;
 * 1 = in place with all fluxes materialised.  Both give the same bits. */
int bsn_batch_set_mode(bsn_batch* b, int inplace);
/* One SSP-RK3 step of every grid with per-grid dt (host array of ngrids). */
int bsn_batch_step(bsn_batch* b, const double* dt);
/* nsteps CFL-limited steps: dt = cfl/max wave speed, computed on the device
 * per grid each step (clipped to t_end - t when t_end > 0).  use_graph
 * captures one whole step in a hipGraph and replays it. */
int bsn_batch_run(bsn_batch* b, double cfl, double t_end, int nsteps, int use_graph);
/* Per-grid simulated time and last dt after bsn_batch_run. */
int bsn_batch_times(bsn_batch* b, double* t, double* dt);

/* Fixed mesh refinement ---------------------------------------------------- */
/* levels[0] is the root; level l+1 is nested in level l over the 1-based
 * active box lo_hi[6l .. 6l+5] = (ci_lo, ci_hi, cj_lo, cj_hi, ck_lo, ck_hi) of
 * level l with refinement ratio[l] (FMRGrid3D, src/fmr3d.jl:109-151).  With
 * nlevels = 2 this is the reference's two-level grid. */
int bsn_fmr_create(bsn_ctx* ctx, int nlevels, bsn_grid* const* levels, const int64_t* lo_hi,
                   const int* ratio, bsn_fmr** out);
int bsn_fmr_destroy(bsn_fmr* m);
/* Level step form: 0 (default) = per stage the x and y sweeps and the z sweep
 * fused with the stage update, which also stores the z faces the flux
 * registers read; 1 = in place with all three flux arrays materialised (the
 * order of src/fmr3d.jl:497-530).  Both give the same bits.  The fused form
 * holds two more state buffers per level. */
int bsn_fmr_set_mode(bsn_fmr* m, int inplace);
/* bsn_batch_autotune for every level on its current state; report gets a
 * JSON array of the per-level reports.  The level states are not changed. */
int bsn_fmr_autotune(bsn_fmr* m, char* report, int report_len);
/* _prolong_all! / _prolong_fine_ghosts! of level l >= 1 from its parent's
 * current state (src/fmr3d.jl:206-255). */
int bsn_fmr_prolong_all(bsn_fmr* m, int level);
int bsn_fmr_prolong_fine_ghosts(bsn_fmr* m, int level);
/* _restrict! of level l >= 1 into its parent (src/fmr3d.jl:266-285). */
int bsn_fmr_restrict(bsn_fmr* m, int level);
/* fmr3d_step!(G, dt_c) (src/fmr3d.jl:451); each level's phi (if set) is the
 * frozen potential of the self_gravity branch (src/fmr3d.jl:467-473). */
int bsn_fmr3d_step(bsn_fmr* m, double dt_c);
/* cfl_dt_fmr3d(G, cfl) (src/fmr3d.jl:616). */
int bsn_cfl_dt_fmr3d(bsn_fmr* m, double cfl, double* dt_out);
/* nsteps CFL-limited root steps with dt on the device; use_graph as above
 * (ignored while self-gravity is on: the Poisson solve is launched eagerly). */
int bsn_fmr_run(bsn_fmr* m, double cfl, double t_end, int nsteps, int use_graph);
int bsn_fmr_time(bsn_fmr* m, double* t, double* dt);

/* FMR self-gravity (src/fmr_gravity.jl) ------------------------------------
 * fmr3d_step!(G, dt_c; self_gravity, G_grav) (src/fmr3d.jl:451-473): with
 * self_gravity = 1 the composite Poisson problem is solved once at the start
 * of the root step (solve_fmr_poisson, which restricts the density into the
 * parents first), every level's phi is scaled by G_grav and frozen across all
 * stages and sub-cycles of the step; the levels' own phi (bsn_grid_set_phi)
 * is not used during the step and is theirs again afterwards.  More than two
 * levels: the density is restricted from the finest level up and phi is
 * prolonged down the chain.  self_gravity = 0 is bsn_fmr3d_step. */
int bsn_fmr3d_step_gravity(bsn_fmr* m, double dt_c, int self_gravity, double G_grav);
/* The same frozen-phi solve at every root step of bsn_fmr_run (on = 0 stops). */
int bsn_fmr_set_self_gravity(bsn_fmr* m, int on, double G_grav);
/* solve_fmr_poisson(G) (src/fmr_gravity.jl:119-133): phi[l] receives level
 * l's active-cell potential (nx*ny*nz, x fastest), host or device arrays
 * (where); the parents' density inside each child box is restricted first. */
int bsn_solve_fmr_poisson(bsn_fmr* m, double* const* phi, int where);
/* restrict_density_to_coarse!(G) (src/fmr_gravity.jl:40-60) for level l >= 1
 * into its parent (density only). */
int bsn_fmr_restrict_density(bsn_fmr* m, int level);
/* _prolong_scalar_trilinear!(phi_f, phi_c, G) (src/fmr_gravity.jl:70-106)
 * from level l-1's active-cell array to level l's. */
int bsn_fmr_prolong_scalar(bsn_fmr* m, int level, const double* phi_c, int where_c,
                           double* phi_f, int where_f);

/* Kernel timing of the last bsn_batch_run / bsn_fmr_run with profile=1:
 * average milliseconds per launch of the flux (x, y, z) and update kernels. */
int bsn_batch_profile(bsn_batch* b, int nsteps, double* ms_flux3, double* ms_update,
                      double* ms_step);

#ifdef __cplusplus
}
#endif
#endif
