/*
 * grhd_oracle.h -- CPU ORACLE (test infrastructure only) for the
 * fixed-background GRHD path.  Same argument structs and FArrayBox layout as
 * the product C-ABI in include/grhd_abi.h, but host pointers.  Loaded only
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 */
#ifndef GRHD_ORACLE_H_
#define GRHD_ORACLE_H_

#include "../include/grhd_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

int grhd_oracle_compute_rhs(const grhd_box *b, const grhd_params *o,
                            const grhd_background *bg, double time,
                            double *state, int ncomp, int comp_base,
                            double *rhs, unsigned char *source_mask,
                            grhd_error *err);
int grhd_oracle_recover_primitives(const grhd_box *b, const grhd_params *o,
                                   const grhd_background *bg, double time,
                                   double *state, int ncomp, int comp_base,
                                   unsigned char *mask,
                                   grhd_diagnostics *diag);
int grhd_oracle_update_conserved(const grhd_box *b, const grhd_params *o,
                                 const grhd_background *bg, double *state,
                                 int ncomp, int comp_base, const double *rhs,
                                 double dt, int recover, double recovery_time,
                                 unsigned char *mask, grhd_diagnostics *diag);
int grhd_oracle_compute_max_inverse_dt(const grhd_box *b,
                                       const grhd_params *o,
                                       const grhd_background *bg, double time,
                                       const double *state, int ncomp,
                                       int comp_base, double *out,
                                       grhd_error *err);
int grhd_oracle_compute_stable_dt(const grhd_box *b, const grhd_params *o,
                                  const grhd_background *bg, double time,
                                  const double *state, int ncomp,
                                  int comp_base, double cfl, double *out,
                                  grhd_error *err);
int grhd_oracle_advance_ssprk2(const grhd_box *b, const grhd_params *o,
                               const grhd_background *bg, double time,
                               double dt, double *state, int ncomp,
                               int comp_base, unsigned char *mask,
                               grhd_diagnostics *diag, grhd_error *err);
int grhd_oracle_advance_rk4(const grhd_box *b, const grhd_params *o,
                            const grhd_background *bg, double time, double dt,
                            int nsteps, double *state, int ncomp,
                            int comp_base, unsigned char *mask,
                            grhd_diagnostics *diag, grhd_error *err);

int grhd_oracle_adm_point(const grhd_background *bg, double time,
                          const double x[3], double *out);
int grhd_oracle_cell_geometry_box(const grhd_box *b,
                                  const grhd_background *bg, double time,
                                  double *out);
void grhd_oracle_prim2con(const double *prim, double gamma, const double *g9,
                          double *cons);
int grhd_oracle_con2prim(const double *cons, const grhd_params *o,
                         const double *gu9, double *prim, double *info);
void grhd_oracle_floors(double *prim, const grhd_params *o, const double *g9);
double grhd_oracle_limited_slope(double l, double c, double r, double theta);
double grhd_oracle_minmod(double a, double b);
void grhd_oracle_hlle(const double *pl6, const double *pr6, double gamma,
                      const double *g9, const double *gu9, double lapse,
                      const double *shift3, int d, double *flux5,
                      double *speeds4);
void grhd_oracle_flux(const double *p6, double gamma, const double *g9,
                      double lapse, const double *shift3, int d,
                      double *flux5);
void grhd_oracle_sources_point(const grhd_background *bg, double time,
                               const double x[3], const double *p6,
                               const double *c5, double gamma, int use_K,
                               double *out);
double grhd_oracle_log_sqrt_det_dt(const grhd_background *bg, double time,
                                   const int iv[3], double dx,
                                   const double center[3], double step);
double grhd_oracle_log(double x);
int grhd_oracle_num_threads(void);
void grhd_oracle_set_num_threads(int n);
/* KerrSchild::check_if_excised (KerrSchild.hpp:247-280) */
int grhd_oracle_ks_excised(double M, double a, const double x[3],
                           double inner_buffer, double outer_buffer);

/* Level diagnostics over `nbox` FABs (each ncomp components on its box's
 * storage), interior cells only, in the reference's order: boxes in turn,
 * BoxIterator (i fastest) within a box. */
void grhd_oracle_leveldata_conserved_sum(const grhd_box *boxes, int nbox,
                                         const double *const *fabs, int ncomp,
                                         int comp_base, double sum[5]);
int grhd_oracle_leveldata_conserved_state_is_finite(
    const grhd_box *boxes, int nbox, const double *const *fabs, int ncomp,
    int comp_base);

/* KITPCHK1 payload of a level (write_checkpoint / read_checkpoint_state). */
void grhd_oracle_checkpoint_pack(const grhd_box *boxes, int nbox,
                                 const double *const *fabs, int ncomp,
                                 const int dlo[3], const int dhi[3],
                                 int ghosts, double *values);
void grhd_oracle_checkpoint_unpack(const grhd_box *boxes, int nbox,
                                   double *const *fabs, int ncomp,
                                   const int dlo[3], const int dhi[3],
                                   int ghosts, const double *values);

#ifdef __cplusplus
}
#endif

#endif /* GRHD_ORACLE_H_ */
