/*
 * bsn_io.h -- C ABI of libbsn_io: the snapshot and checkpoint files of
 * Binary-Supernova (HDF5, the reference's layout) written from and read into
 * a grid resident on the GPU, streaming through pinned slabs
 * (bsn_grid_export_slabs / bsn_grid_import_slabs in bsn.h).
 *
 * Files are the reference's: a Julia array of size (a, b, c) is an HDF5
 * dataset of dims (c, b, a) in float64 little endian; integer attributes are
 * int64, real ones float64, strings fixed-length UTF-8.  Every entry point
 * returns BSN_OK or a BSN_ERR_* status; bsn_io_last_error() describes the last
 * failure of the calling thread.  Host-array forms take the reference layout
 * of bsn.h (state (5, nx+6, ny+6, nz+6) with ghosts).
 */
#ifndef BSN_IO_H
#define BSN_IO_H
#include <stdint.h>
#include "bsn.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* bsn_io_last_error(void);

/* write_snapshot(filename, U, nx, ny, nz, dx, dy, dz, t, γ) (src/io.jl:29-51):
 * /gas/{rho, rhovx, rhovy, rhovz, E} of the active cells and the root
 * attributes t, nx, ny, nz, dx, dy, dz, gamma. */
int bsn_write_snapshot(bsn_grid* g, const char* filename, double dx, double dy, double dz,
                       double t, double gamma);
int bsn_write_snapshot_host(const char* filename, const double* U, int64_t nx, int64_t ny,
                            int64_t nz, double dx, double dy, double dz, double t, double gamma);
/* read_snapshot(filename) (src/io.jl:60-84): n = (nx, ny, nz) and
 * meta = (dx, dy, dz, t, gamma); then the active cells into a grid of that
 * shape (its ghost cells are left as they are), or into a host array U_active
 * of the reference's (5, nx, ny, nz) layout. */
int bsn_snapshot_info(const char* filename, int64_t n[3], double meta[5]);
int bsn_read_snapshot(bsn_grid* g, const char* filename);
int bsn_read_snapshot_host(const char* filename, double* U_active);

/* save_checkpoint_uniform(path, U, bhs; t, step, dt_last, t_snap, t_bh2_sink_on,
 * nx, ny, nz, dx, dy, dz, γ, x0, y0, z0, ρ_floor, P_floor)
 * (src/checkpoint.jl:244-283): the state with ghosts comes from the grid g, or
 * from U_host (reference layout, nx, ny, nz) when g is NULL.
 *   meta   = (t, dt_last, t_snap, t_bh2_sink_on)
 *   params = (dx, dy, dz, gamma, x0, y0, z0, rho_floor, P_floor)
 *   bodies: pos and vel 3 values per body, props 4 per body
 *           (mass, eps, c_code, r_floor), n_bodies of them. */
int bsn_save_checkpoint_uniform(bsn_grid* g, const double* U_host, int64_t nx, int64_t ny,
                                int64_t nz, const char* path, int64_t step, const double meta[4],
                                const double params[9], int n_bodies, const double* pos,
                                const double* vel, const double* props);
/* The shape and body count of a uniform checkpoint. */
int bsn_checkpoint_uniform_info(const char* path, int64_t n[3], int* n_bodies);
/* load_checkpoint_uniform(path) (src/checkpoint.jl:285-318): checks
 * format_version == 1 and kind == "uniform" (BSN_ERR_IO otherwise), reads the
 * state into g (or into U_host when g is NULL) and returns
 *   meta   = (t, dt_last, t_snap, t_bh2_sink_on, format_version)
 *   params as above; up to max_bodies bodies (BSN_ERR_ARG if more are stored). */
int bsn_load_checkpoint_uniform(bsn_grid* g, double* U_host, const char* path, int64_t* step,
                                double meta[5], double params[9], int max_bodies, int* n_bodies,
                                double* pos, double* vel, double* props);

/* save_checkpoint(path, G, bhs; t, step, dt_last) (src/checkpoint.jl:33-117):
 * the two-level FMR grid of the reference (format version 1).  The states
 * with ghosts stream from the grids coarse and fine, or come from the host
 * arrays Uc_host and Uf_host (reference layout) when the grids are NULL.
 *   n      = coarse (nx, ny, nz), fine (nx, ny, nz)
 *   d      = coarse (dx, dy, dz), fine (dx, dy, dz)
 *   meta   = (t, dt_last)
 *   box    = (ci_lo, ci_hi, cj_lo, cj_hi, ck_lo, ck_hi), 1-based coarse cells
 *   params = (gamma, rho_floor, P_floor); bc = "outflow" or "periodic"
 *   bodies as for bsn_save_checkpoint_uniform.
 * BSN_ERR_ARG when the fine shape is not the box times the ratio. */
int bsn_save_checkpoint_fmr(bsn_grid* coarse, bsn_grid* fine, const double* Uc_host,
                            const double* Uf_host, const int64_t n[6], const double d[6],
                            const char* path, int64_t step, const double meta[2],
                            const int64_t box[6], int ratio, const double params[3], const char* bc,
                            int n_bodies, const double* pos, const double* vel,
                            const double* props);
/* The grid description of a two-level checkpoint (shapes and cell sizes as
 * n and d above, box, ratio, params, the bc string into bc[bc_len]) and its
 * body count; BSN_ERR_IO unless format_version == 1 and both levels exist. */
int bsn_checkpoint_fmr_info(const char* path, int64_t n[6], double d[6], int64_t box[6],
                            int* ratio, double params[3], char* bc, int bc_len, int* n_bodies);
/* load_checkpoint(path) (src/checkpoint.jl:119-172): both levels' saved states
 * (the fine one as saved, not prolonged again) into the grids, which must have
 * the stored shapes, or into the host arrays when the grids are NULL;
 *   meta = (t, dt_last, format_version); up to max_bodies bodies. */
int bsn_load_checkpoint_fmr(bsn_grid* coarse, bsn_grid* fine, double* Uc_host, double* Uf_host,
                            const char* path, int64_t* step, double meta[3], int max_bodies,
                            int* n_bodies, double* pos, double* vel, double* props);

/* init_trajectory_file(filename, n_bh) (src/io.jl:94-108): "time" and, per
 * body, bh<i>/{mass, rsink, pos, vel}, extendable along the record axis. */
int bsn_init_trajectory_file(const char* filename, int n_bh);
/* append_trajectory(filename, t, bhs) (src/io.jl:115-133): one record for
 * n_bh bodies; rsink is the caller's r_sink(bh) of each body, pos and vel 3
 * values per body. */
int bsn_append_trajectory(const char* filename, double t, int n_bh, const double* mass,
                          const double* rsink, const double* pos, const double* vel);
/* The body count and the number of records of a trajectory file. */
int bsn_trajectory_info(const char* filename, int* n_bh, int64_t* n_records);
/* read_trajectory(filename) (src/io.jl:144-161): t[N]; per body b,
 * mass[b*N + r], rsink[b*N + r], pos[(b*N + r)*3 + a], vel likewise (the
 * reference's N x 3 arrays, record by record). */
int bsn_read_trajectory(const char* filename, double* t, double* mass, double* rsink,
                        double* pos, double* vel);

#ifdef __cplusplus
}
#endif
#endif
