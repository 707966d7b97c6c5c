/* bsn_oracle.h -- CPU FP64 restatement of the reference hot path (test
 * infrastructure only; see bsn_oracle.c for the per-function citations). */
#ifndef BSN_ORACLE_H
#define BSN_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

void orc_wenoz_weights(double b0, double b1, double b2, double d0, double d1, double d2,
                       double *w0, double *w1, double *w2);
double orc_weno5_left(double vm2, double vm1, double v0, double vp1, double vp2);
double orc_weno5_right(double vm1, double v0, double vp1, double vp2, double vp3);
void orc_hllc_normal(double rL, double mnL, double mt1L, double mt2L, double EL,
                     double rR, double mnR, double mt1R, double mt2R, double ER,
                     double g, double F[5]);
void orc_hllc_flux(int dir, const double L[5], const double R[5], double g, double F[5]);
double orc_cell_pressure(double r, double mx, double my, double mz, double E, double g);
double orc_EL_from_P(double PL, double rL, double mxL, double myL, double mzL, double g);

void orc_fill_ghost(double *U, int64_t nx, int64_t ny, int64_t nz, const int bc[3]);
void orc_apply_floors(double *U, int64_t nx, int64_t ny, int64_t nz,
                      double rho_floor, double P_floor, double g);
void orc_weno_fluxes(int dir, double *F, const double *U, int64_t nx, int64_t ny, int64_t nz,
                     double g);
void orc_flux_divergence(double *dU, const double *Fx, const double *Fy, const double *Fz,
                         int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz);
void orc_apply_sg_force(double *dU, const double *U, const double *phi,
                        int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz);
double orc_cfl_dt_3d(const double *U, int64_t nx, int64_t ny, int64_t nz,
                     double dx, double dy, double dz, double g, double cfl);
void orc_euler3d_rhs(double *dU, double *Fx, double *Fy, double *Fz, double *U,
                     int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz,
                     double g, const int bc[3], double rho_floor, double P_floor,
                     int fill_ghosts);
void orc_rk3_combine(int stage, double *U, const double *Un, const double *dU,
                     double dt, int64_t n);
int orc_euler3d_step(double *U, int64_t nx, int64_t ny, int64_t nz,
                     double dx, double dy, double dz, double dt, double g,
                     const int bc[3], double rho_floor, double P_floor, const double *phi);
void orc_conserved_totals(const double *U, int64_t nx, int64_t ny, int64_t nz,
                          double dV, double out[5]);

void orc_lag5_weights(int ratio, double *w);
void orc_prolong_all(double *Uf, const double *Uc, int64_t ncx, int64_t ncy, int64_t ncz,
                     const int64_t lo_hi[6], int ratio);
void orc_prolong_fine_ghosts(double *Uf, const double *Uc, int64_t ncx, int64_t ncy,
                             int64_t ncz, const int64_t lo_hi[6], int ratio);
void orc_restrict(double *Uc, const double *Uf, int64_t ncx, int64_t ncy, int64_t ncz,
                  const int64_t lo_hi[6], int ratio);
int orc_fmr3d_step(double *Uc, double *Uf,
                   int64_t ncx, int64_t ncy, int64_t ncz, double dxc, double dyc, double dzc,
                   const int64_t lo_hi[6], int ratio, double g, const int bc[3],
                   double rho_floor, double P_floor, double dt_c,
                   const double *phi_c, const double *phi_f);
double orc_cfl_dt_fmr3d(const double *Uc, const double *Uf,
                        int64_t ncx, int64_t ncy, int64_t ncz, double dxc, double dyc, double dzc,
                        const int64_t lo_hi[6], int ratio, double g, double cfl);
int orc_fmr_multilevel_step(int nlev, double **U, const int64_t *dims, const double *dxs,
                            const int64_t *lo_hi, const int *ratio, double g, const int bc[3],
                            double rho_floor, double P_floor, double dt,
                            const double *const *phis);
void orc_bh_gravity_source(double *dU, const double *U, int64_t nx, int64_t ny, int64_t nz,
                           double dx, double dy, double dz, double x0, double y0, double z0,
                           int nbh, const double *px, const double *py, const double *pz,
                           const double *mass, const double *eps);
void orc_sink_sources(double *dU, const double *U, int64_t nx, int64_t ny, int64_t nz,
                      double dx, double dy, double dz, double x0, double y0, double z0,
                      int nbh, const double *px, const double *py, const double *pz,
                      const double *bvx, const double *bvy, const double *bvz,
                      const double *rs, const double *ts, int torque_free, double rho_sink_min);
int orc_coupled_step(double *U, int64_t nx, int64_t ny, int64_t nz,
                     double dx, double dy, double dz, double dt, double g,
                     const int bc[3], double rho_floor, double P_floor, const double *phi,
                     double x0, double y0, double z0, int nbh, const double *pos,
                     const double *vel, const double *mass, const double *eps,
                     int nsink, const double *rs, const double *ts, int torque_free,
                     double rho_sink_min);
void orc_gas_force_on_bh(const double *U, int64_t nx, int64_t ny, int64_t nz,
                         double dx, double dy, double dz, double px, double py, double pz,
                         double mass, double eps, double x0, double y0, double z0,
                         double F[3]);
void orc_accrete_sums(const double *U, int64_t nx, int64_t ny, int64_t nz,
                      double dx, double dy, double dz, double x0, double y0, double z0,
                      double px, double py, double pz, double rs, double ts, double dt,
                      double rho_sink_min, double out[4]);
void orc_gas_diagnostics(const double *U, int64_t nx, int64_t ny, int64_t nz,
                         double dx, double dy, double dz, double x0, double y0, double z0,
                         int nbh, const double *px, const double *py, const double *pz,
                         const double *mass, const double *eps, double g, double out[9]);

#ifdef __cplusplus
}
#endif
#endif
