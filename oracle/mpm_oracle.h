/*
 * mpm_oracle.h -- FP64 CPU oracle for the depth-averaged MPM time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (amsc_amd/, include/)
 * includes, links or calls this code; only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it, as the checker.
 *
 * Contents (each function cites the reference code it restates):
 *   1. or_*  per-kernel restatements of the 18 gpu_kernels entry points
 *      (main/gpu_kernels.cpp; gpu_offload semantics);
 *   2. or_bin_particles: the quadgrid init_particle_mesh layout (cell of each
 *      particle, CSR by cell, 4-colouring);
 *   3. or_sim_*: the optim_par time step (main/optim_par.cpp:128-682) with
 *      adaptive merge/split inserted at gpu_offload's position
 *      (main/gpu_offload.cpp:282-401, main/merge_split_ops_cmes.h);
 *   4. or_ms_decide / or_ms_execute: merge/split on a state snapshot.
 *
 * Compiled with gcc -O2 -ffp-contract=off: every expression keeps the
 * reference's evaluation order, so results are reproducible bit for bit.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- 1. */
void or_reset_nodal_arrays(int nn, double *Mv, double *mom_vx, double *mom_vy,
                           double *F_ext_vx, double *F_ext_vy,
                           double *F_int_vx, double *F_int_vy,
                           double *Ftot_vx, double *Ftot_vy, double *avx,
                           double *avy, double *vvx, double *vvy, double *vvxL,
                           double *vvyL, double *Fric_x, double *Fric_y,
                           double *FPxv, double *FPyv, double *HV,
                           double *pv_rho, double *pv_vvx, double *pv_vvy,
                           double *pv_avx, double *pv_avy);
void or_scatter_mass_momentum(int np, int nrows, double hx, double hy,
                              const int *cstart, const int *cptcls,
                              const int *ccidx, const int *coloff,
                              const double *x, const double *y,
                              const double *Mp, const double *mom_px,
                              const double *mom_py, double *Mv, double *mom_vx,
                              double *mom_vy);
void or_gradZ_to_particles(int np, int nrows, double hx, double hy,
                           const double *x, const double *y, const int *p2g,
                           const double *Z, double *dZxp, double *dZyp);
void or_gravity_force(int np, double g, const double *hp, const double *H,
                      const double *Mp, const double *dZxp, const double *dZyp,
                      double *Fpx, double *Fpy);
void or_friction_particle_force(int np, const double *Ap, const double *Fb_x,
                                const double *Fb_y, double *Fric_px,
                                double *Fric_py);
void or_scatter_gravity_force(int np, int nrows, double hx, double hy,
                              const int *cstart, const int *cptcls,
                              const int *ccidx, const int *coloff,
                              const double *x, const double *y,
                              const double *Fpx, const double *Fpy,
                              double *FPxv, double *FPyv);
void or_scatter_friction_force(int np, int nrows, double hx, double hy,
                               const int *cstart, const int *cptcls,
                               const int *ccidx, const int *coloff,
                               const double *x, const double *y,
                               const double *Fric_px, const double *Fric_py,
                               double *Fric_x, double *Fric_y);
void or_compute_F_ext(int nn, const double *FPxv, const double *FPyv,
                      const double *Fric_x, const double *Fric_y,
                      double *F_ext_vx, double *F_ext_vy);
void or_scatter_internal_forces(int np, int nrows, double hx, double hy,
                                const int *cstart, const int *cptcls,
                                const int *ccidx, const int *coloff,
                                const double *x, const double *y,
                                const double *Vp, const double *F11,
                                const double *F12, const double *F21,
                                const double *F22, double *F_int_vx,
                                double *F_int_vy);
void or_nodal_force_and_momentum(int nn, double dt, const double *F_ext_vx,
                                 const double *F_ext_vy,
                                 const double *F_int_vx,
                                 const double *F_int_vy, double *Ftot_vx,
                                 double *Ftot_vy, double *mom_vx,
                                 double *mom_vy);
void or_nodal_velocity_and_bc(int nn, int nrows, int ncols, double dt,
                              int bc_flag, const double *Mv,
                              const double *Ftot_vx, const double *Ftot_vy,
                              double *mom_vx, double *mom_vy, double *avx,
                              double *avy, double *vvx, double *vvy,
                              double *vvxL, double *vvyL);
void or_g2p_velocity_and_advect(int np, int nrows, int ncols, double hx,
                                double hy, double dt, const int *p2g,
                                const double *vvxL, const double *vvyL,
                                const double *avx, const double *avy,
                                double *x, double *y, double *vpx, double *vpy,
                                double *apx, double *apy, double *mom_px,
                                double *mom_py, double *vpxL, double *vpyL,
                                const double *Mp);
void or_g2pd_gradients_and_height_update(
    int np, int nrows, double hx, double hy, double dt, const double *x,
    const double *y, const int *p2g, const double *vvxL, const double *vvyL,
    double *vpx_dx, double *vpx_dy, double *vpy_dx, double *vpy_dy,
    double *hp, double *Vp, double *Ap);
void or_voellmy_friction(int np, double rho, double g, double tan_fa,
                         double xi, double fric_on, const double *vpx,
                         const double *vpy, const double *hp, double *Fb_x,
                         double *Fb_y);
void or_bingham_stress_update(int np, double rho, double g, double mu,
                              double tau_Y, double cc, double A_coeff,
                              double C_coeff, const double *vpx,
                              const double *vpy, const double *hp,
                              const double *H, const double *vpx_dx,
                              const double *vpx_dy, const double *vpy_dx,
                              const double *vpy_dy, double *F11, double *F12,
                              double *F21, double *F22);
void or_g2p_height(int np, int nrows, double hx, double hy, const double *x,
                   const double *y, const int *p2g, const double *Z,
                   const double *hp, double *Zp, double *hpZ);
void or_scatter_plotvars(int np, int nrows, double hx, double hy,
                         const int *cstart, const int *cptcls,
                         const int *ccidx, const int *coloff, const double *x,
                         const double *y, const double *Mp, const double *vpx,
                         const double *vpy, const double *apx,
                         const double *apy, const double *hp, double *pv_rho,
                         double *pv_vvx, double *pv_vvy, double *pv_avx,
                         double *pv_avy, double *HV);
void or_normalize_plotvars(int nn, double *pv_rho, double *pv_vvx,
                           double *pv_vvy, double *pv_avx, double *pv_avy);

/* ---------------------------------------------------------------- 2. */
int or_cell_of(double x, double y, int nrows, int ncols, double hx,
               double hy);
void or_bin_particles(int np, const double *x, const double *y, int nrows,
                      int ncols, double hx, double hy, int *p2g, int *cstart,
                      int *cptcls, int *ccidx, int *coloff);

/* ---------------------------------------------------------------- 3. */
typedef struct or_params {
  int nrows, ncols;          /* Ney, Nex */
  double hx, hy;
  double g, rho, mu, tauy, xi, phi; /* phi in degrees */
  double bingham_on, friction_on, cfl, eq_level, dt_fixed, T;
  int bc_flag, nsteps, max_steps;
  int ms_on;
  double ms_alpha, ms_beta, ms_split_hp_min, ms_shear_split, ms_hp_min,
      ms_max_dv, ms_max_ops;
  int ms_min_level, ms_max_level, ms_min_ppc, ms_interval;
  int ms_mode;       /* 0: merge_split_ops_cmes.h, 1: merge_split_ops.h ELFS */
  double ms_stretch; /* ELFS divergence trigger (ms_config::stretch_thresh) */
} or_params;

/* stress_tensor_t (optim_par.cpp:68-120) over np particles. */
void or_stress_optim_par(int np, const or_params *p, const double *vpx,
                         const double *vpy, const double *hp, const double *H,
                         const double *vpx_dx, const double *vpx_dy,
                         const double *vpy_dx, const double *vpy_dy,
                         double *F11, double *F12, double *F21, double *F22);

/* Particle fields held by the oracle state (names used by or_sim_get). */
enum {
  OR_X, OR_Y, OR_MP, OR_AP, OR_VPX, OR_VPY, OR_MOM_PX, OR_MOM_PY, OR_HP,
  OR_VP, OR_H, OR_F11, OR_F12, OR_F21, OR_F22, OR_FB_X, OR_FB_Y, OR_VPX_DX,
  OR_VPX_DY, OR_VPY_DX, OR_VPY_DY, OR_APX, OR_APY, OR_DZXP, OR_DZYP, OR_ZP,
  OR_HPZ, OR_FPX, OR_FPY, OR_FRIC_PX, OR_FRIC_PY, OR_VPXL, OR_VPYL, OR_XP,
  OR_YP, OR_NFIELDS
};

typedef struct or_sim or_sim;

/* Exterior cells and boundary distance of the legacy ELFS merge/split
 * (merge_split_ops.h:86-262) for the occupancy given by cstart. */
void or_elfs_cells(int nrows, int ncols, double h, const int *cstart,
                   char *ext, double *dist);

/* Threads of the per-particle loops (OpenMP builds; no effect otherwise). */
void or_set_threads(int n);
int or_max_threads(void);

or_sim *or_sim_create(const or_params *p, int np, const double *x,
                      const double *y, const double *Mp, const double *Ap,
                      const double *vpx, const double *vpy,
                      const double *mom_px, const double *mom_py,
                      const double *hp, const double *Vp, const double *Z);
void or_sim_destroy(or_sim *s);
/* One time step. Returns 1 if a step was taken, 0 if the run had ended. */
int or_sim_step(or_sim *s);
int or_sim_np(const or_sim *s);
int or_sim_it(const or_sim *s);
double or_sim_t(const or_sim *s);
double or_sim_dt(const or_sim *s);
/* Copy a particle field (see or_field_name) or nodal field into out. */
int or_sim_get(const or_sim *s, const char *name, double *out);
int or_sim_get_int(const or_sim *s, const char *name, int *out);
/* Overwrite a particle field (length np). */
int or_sim_set(or_sim *s, const char *name, const double *in);
/* err_mass, err_mom and t of the last step (optim_par.cpp:395-428). */
void or_sim_conservation(const or_sim *s, double out[3]);
/* Number of merge/split splits and merges done in the last step. */
void or_sim_ms_counts(const or_sim *s, int out[2]);
const char *or_field_name(int f);

/* ---------------------------------------------------------------- 4. */
/* decide_actions (merge_split_ops_cmes.h:116-253) on a snapshot.  Cells come
 * from or_bin_particles of x,y.  Writes action[np] (0 KEEP, 1 SPLIT,
 * 2 MERGE_PRIMARY, 3 MERGE_SECONDARY) and partner[np]; returns
 * n_split in counts[0] and n_merge in counts[1]. */
void or_ms_decide(const or_params *p, int np, double dt, const double *x,
                  const double *y, const double *Ap, const double *hp,
                  const double *vpx_dx, const double *vpx_dy,
                  const double *vpy_dx, const double *vpy_dy,
                  const double *vpx, const double *vpy, const int *level,
                  int *action, int *partner, int counts[2]);
/* execute_merge_split (merge_split_ops_cmes.h:258-427) on arrays: in/out
 * are or_field_count() field arrays in or_field_name() order, out, nlabel,
 * nlevel with room for np + splits; returns the new particle count. */
int or_ms_execute(const or_params *p, int np, const int *act,
                  const int *partner, double *const *in, const int *label,
                  const int *level, double **out, int *nlabel, int *nlevel);
int or_field_count(void);
/* The legacy ELFS rule (ms_mode 1) on a snapshot: decide_actions of
 * merge_split_ops.h:349-510 with the cells of or_bin_particles(x, y) and
 * execute_merge_split of merge_split_ops.h:513-716 (arrays as
 * or_ms_execute, room for np + 3 * splits). */
void or_elfs_decide(const or_params *p, int np, const double *x,
                    const double *y, const double *Ap, const double *hp,
                    const double *vpx_dx, const double *vpy_dy,
                    const double *vpx, const double *vpy, const int *level,
                    int *action, int *partner, int counts[2]);
int or_elfs_execute(const or_params *p, int np, const int *act,
                    const int *partner, double *const *in, const int *label,
                    const int *level, double **out, int *nlabel, int *nlevel);

#ifdef __cplusplus
}
#endif
#endif
