/*
 * amsc_mpm.h -- C ABI of the MI355X (gfx950) depth-averaged MPM time step.
 *
 * Plain C: pointers, sizes and status codes only.  Every function returns
 * AMSC_OK (0) or an AMSC_ERR_* code; amsc_last_error() gives the message.
 * Launches are stream-ordered on the stream set with amsc_set_stream
 * (default: the null stream) and do not synchronise unless stated.
 *
 * Four groups of entry points:
 *   1. the 18 step kernels of the reference's gpu_kernels.h, same parameter
 *      lists (device pointers), same results;
 *   2. device binning (replaces quadgrid's host init_particle_mesh + the
 *      per-step host CSR rebuild) and merge/split decisions on the device;
 *   3. amsc_sim_*: a device-resident time loop (the gpu_offload.cpp driver's
 *      job) built on fused kernels, FP64 with optim_par semantics;
 *   4. amsc_ens_*: the same loop over a batch of independent scenarios
 *      advanced together, in FP64 or mixed FP32/FP64 (no reference
 *      counterpart: the hazard-ensemble form of the time loop).
 */
#ifndef AMSC_MPM_H
#define AMSC_MPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMSC_OK 0
#define AMSC_ERR_INVALID 1   /* bad argument (size, pointer, layout)      */
#define AMSC_ERR_HIP 2       /* a HIP runtime call failed                  */
#define AMSC_ERR_NOMEM 3     /* device allocation failed                   */
#define AMSC_ERR_CAPACITY 4  /* particle capacity exceeded                 */
#define AMSC_ERR_STALE 5     /* field not materialised by the last step    */

const char *amsc_last_error(void);
/* hipStream_t for subsequent launches (NULL = null stream). */
int amsc_set_stream(void *stream);
int amsc_synchronize(void);

/* Device memory on the library's own HIP runtime (for callers that have no
 * other allocator, e.g. Python tests): allocate, free, and synchronous
 * copies (the copy to the host waits for the device first; the
 * device-to-device copy is stream-ordered). */
int amsc_device_alloc(size_t bytes, void **ptr);
int amsc_device_free(void *ptr);
int amsc_copy_to_device(void *dst, const void *src, size_t bytes);
int amsc_copy_on_device(void *dst, const void *src, size_t bytes); /* async */
int amsc_copy_to_host(void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------ 1. */
/* Same parameters and results as gpu_kernels:: in the reference
 * (main/gpu_kernels.h line cited per entry point).  All double and int
 * pointers are device pointers, except `coloff` (int[5], host memory, as in
 * the reference, which dereferences it on the host: gpu_kernels.cpp:105-106)
 * and the diagnostic *_out pointers, which the reference never writes
 * (gpu_kernels.cpp:450-453,539-545) and which are ignored here too.
 * Scatter entry points sum, for every node, the contributions of its (up to
 * 4) cells in colour order, each cell's particles in cptcls order, added to
 * the node's current value: the reference's order.  They require ccidx to
 * list every cell (coloff[4] == nrows * ncols, as quadgrid's colouring does,
 * gpu_offload.cpp:258). */

/* gpu_kernels.h:7-18 */
int amsc_reset_nodal_arrays(int nn, double *d_Mv, double *d_mom_vx,
                            double *d_mom_vy, double *d_F_ext_vx,
                            double *d_F_ext_vy, double *d_F_int_vx,
                            double *d_F_int_vy, double *d_Ftot_vx,
                            double *d_Ftot_vy, double *d_avx, double *d_avy,
                            double *d_vvx, double *d_vvy, double *d_vvxL,
                            double *d_vvyL, double *d_Fric_x, double *d_Fric_y,
                            double *d_FPxv, double *d_FPyv, double *d_HV,
                            double *d_pv_rho, double *d_pv_vvx,
                            double *d_pv_vvy, double *d_pv_avx,
                            double *d_pv_avy);
/* gpu_kernels.h:23-28 */
int amsc_scatter_mass_momentum(int np, int nrows, double hx, double hy,
                               const int *cstart, const int *cptcls,
                               const int *ccidx, const int *coloff,
                               const double *d_x, const double *d_y,
                               const double *d_Mp, const double *d_mom_px,
                               const double *d_mom_py, double *d_Mv,
                               double *d_mom_vx, double *d_mom_vy);
/* gpu_kernels.h:33-37 */
int amsc_gradZ_to_particles(int np, int nrows, double hx, double hy,
                            const double *d_x, const double *d_y,
                            const int *d_p2g, const double *d_Z,
                            double *d_dZxp, double *d_dZyp);
/* gpu_kernels.h:42-46 */
int amsc_gravity_force(int np, double g_c, const double *d_hp,
                       const double *d_H, const double *d_Mp,
                       const double *d_dZxp, const double *d_dZyp,
                       double *d_Fpx, double *d_Fpy);
/* gpu_kernels.h:51-54 */
int amsc_friction_particle_force(int np, const double *d_Ap,
                                 const double *d_Fb_x, const double *d_Fb_y,
                                 double *d_Fric_px, double *d_Fric_py);
/* gpu_kernels.h:59-64 */
int amsc_scatter_gravity_force(int np, int nrows, double hx, double hy,
                               const int *cstart, const int *cptcls,
                               const int *ccidx, const int *coloff,
                               const double *d_x, const double *d_y,
                               const double *d_Fpx, const double *d_Fpy,
                               double *d_FPxv, double *d_FPyv);
/* gpu_kernels.h:69-74 */
int amsc_scatter_friction_force(int np, int nrows, double hx, double hy,
                                const int *cstart, const int *cptcls,
                                const int *ccidx, const int *coloff,
                                const double *d_x, const double *d_y,
                                const double *d_Fric_px,
                                const double *d_Fric_py, double *d_Fric_x,
                                double *d_Fric_y);
/* gpu_kernels.h:79-83 */
int amsc_compute_F_ext(int nn, const double *d_FPxv, const double *d_FPyv,
                       const double *d_Fric_x, const double *d_Fric_y,
                       double *d_F_ext_vx, double *d_F_ext_vy);
/* gpu_kernels.h:88-94 */
int amsc_scatter_internal_forces(int np, int nrows, double hx, double hy,
                                 const int *cstart, const int *cptcls,
                                 const int *ccidx, const int *coloff,
                                 const double *d_x, const double *d_y,
                                 const double *d_Vp, const double *d_F11,
                                 const double *d_F12, const double *d_F21,
                                 const double *d_F22, double *d_F_int_vx,
                                 double *d_F_int_vy);
/* gpu_kernels.h:99-104 */
int amsc_nodal_force_and_momentum(int nn, double dt, const double *d_F_ext_vx,
                                  const double *d_F_ext_vy,
                                  const double *d_F_int_vx,
                                  const double *d_F_int_vy, double *d_Ftot_vx,
                                  double *d_Ftot_vy, double *d_mom_vx,
                                  double *d_mom_vy);
/* gpu_kernels.h:109-115 (bool bc_flag passed as int) */
int amsc_nodal_velocity_and_bc(int nn, int nrows, int ncols, double dt,
                               int bc_flag, const double *d_Mv,
                               const double *d_Ftot_vx,
                               const double *d_Ftot_vy, double *d_mom_vx,
                               double *d_mom_vy, double *d_avx, double *d_avy,
                               double *d_vvx, double *d_vvy, double *d_vvxL,
                               double *d_vvyL);
/* gpu_kernels.h:121-131 */
int amsc_g2p_velocity_and_advect(int np, int nrows, int ncols, double hx,
                                 double hy, double dt, const int *d_p2g,
                                 const double *d_vvxL, const double *d_vvyL,
                                 const double *d_avx, const double *d_avy,
                                 double *d_x, double *d_y, double *d_vpx,
                                 double *d_vpy, double *d_apx, double *d_apy,
                                 double *d_mom_px, double *d_mom_py,
                                 double *d_vpxL, double *d_vpyL,
                                 const double *d_Mp);
/* gpu_kernels.h:137-144 */
int amsc_g2pd_gradients_and_height_update(
    int np, int nrows, double hx, double hy, double dt, const double *d_x,
    const double *d_y, const int *d_p2g, const double *d_vvxL,
    const double *d_vvyL, double *d_vpx_dx, double *d_vpx_dy,
    double *d_vpy_dx, double *d_vpy_dy, double *d_hp, double *d_Vp,
    double *d_Ap, double *min_sc_out, double *min_hp_out);
/* gpu_kernels.h:149-152 */
int amsc_voellmy_friction(int np, double rho, double g_c, double tan_fa,
                          double xi, double fric_on, const double *d_vpx,
                          const double *d_vpy, const double *d_hp,
                          double *d_Fb_x, double *d_Fb_y);
/* gpu_kernels.h:157-167 */
int amsc_bingham_stress_update(
    int np, double rho, double g_c, double mu, double tau_Y, double cc,
    double A_coeff, double C_coeff, const double *d_vpx, const double *d_vpy,
    const double *d_hp, const double *d_H, const double *d_vpx_dx,
    const double *d_vpx_dy, const double *d_vpy_dx, const double *d_vpy_dy,
    double *d_F11, double *d_F12, double *d_F21, double *d_F22,
    double *max_coeff_out, double *min_hp_out, double *max_nv_out,
    double *max_F11_out, double *max_dev_out, double *max_pres_out,
    double *max_Dxx_out);
/* gpu_kernels.h:172-176 */
int amsc_g2p_height(int np, int nrows, double hx, double hy,
                    const double *d_x, const double *d_y, const int *d_p2g,
                    const double *d_Z, const double *d_hp, double *d_Zp,
                    double *d_hpZ);
/* gpu_kernels.h:181-188 */
int amsc_scatter_plotvars(int np, int nrows, double hx, double hy,
                          const int *cstart, const int *cptcls,
                          const int *ccidx, const int *coloff,
                          const double *d_x, const double *d_y,
                          const double *d_Mp, const double *d_vpx,
                          const double *d_vpy, const double *d_apx,
                          const double *d_apy, const double *d_hp,
                          double *d_pv_rho, double *d_pv_vvx,
                          double *d_pv_vvy, double *d_pv_avx,
                          double *d_pv_avy, double *d_HV);
/* gpu_kernels.h:193-196 */
int amsc_normalize_plotvars(int nn, double *d_pv_rho, double *d_pv_vvx,
                            double *d_pv_vvy, double *d_pv_avx,
                            double *d_pv_avy);

/* The adaptive time step of gpu_offload.cpp:282-301 (optim_par.cpp:285-307)
 * with the maxima taken on the device instead of over host copies of the
 * particle arrays: cel = |max(sqrt(g*hmax) + max|vpx|, sqrt(g*hmax) +
 * max|vpy|)| over np > 0 particles; when it > 0, *dt = cfl*min(hx,hy) /
 * (1e-2 + cel), otherwise *dt is left as it is (the first step keeps its
 * dt, gpu_offload.cpp:255).  *cel_out (may be NULL) receives cel.  Host
 * scalars; synchronous. */
int amsc_cfl_dt(int np, const double *d_vpx, const double *d_vpy,
                const double *d_hp, double g, double cfl, double hx, double hy,
                int it, double *dt, double *cel_out);

/* ------------------------------------------------------------------ 2. */
/* Replaces quadgrid particles_t::init_particle_mesh() (called at
 * optim_par.cpp:341 and gpu_offload.cpp:243,380,549, rebuilt on the host
 * every step at gpu_offload.cpp:546-558).  Device arrays: p2g[np] (cell of
 * each particle, floor(x/hx), floor(y/hy) clamped to the grid, cell id
 * r + c*nrows), cstart[nrows*ncols+1], cptcls[np] (particle ids by cell,
 * ascending id within a cell: a stable sort), ccidx[nrows*ncols] (cells of
 * colour (r&1)+2*(c&1), colours 0..3, ascending within a colour);
 * coloff[5] is host memory.  Any output pointer except p2g/cstart/cptcls may
 * be NULL. */
int amsc_bin_particles(int np, int nrows, int ncols, double hx, double hy,
                       const double *d_x, const double *d_y, int *d_p2g,
                       int *d_cstart, int *d_cptcls, int *d_ccidx,
                       int *coloff);

/* No counterpart in the reference: amsc_bin_particles for callers that
 * cannot allocate device memory (code under graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and
 * completion rule; it calls no hipMalloc, hipFree or allocating library
 * routine.  Its workspace (cell counts, arrival slots, scan tile sums) is a
 * fixed array in the library's own device memory, 32 MiB per GPU, held from
 * the library's load to the end of the process: at most 4194304 cells
 * (nrows*ncols) and 4194304 particles, larger sizes are rejected with
 * AMSC_ERR_INVALID before any device work.  Every call on one GPU shares
 * that GPU's array; calls from one process are serialised, calls that
 * overlap on one GPU from several threads of control outside that rule
 * (e.g. two processes sharing a context) would race. */
int amsc_bin_particles_noalloc(int np, int nrows, int ncols, double hx,
                               double hy, const double *d_x, const double *d_y,
                               int *d_p2g, int *d_cstart, int *d_cptcls,
                               int *d_ccidx, int *coloff);

/* Solver parameters: the DATA.json scalars (main/mpm_data.h:14-56), names in
 * lower case; nrows = Ney, ncols = Nex; phi in degrees; max_steps > 0 caps
 * an adaptive run; ms_* as DATA's merge/split keys. */
typedef struct amsc_params {
  int nrows, ncols;
  double hx, hy;
  double g, rho, mu, tauy, xi, phi;
  double bingham_on, friction_on, cfl, eq_level, dt_fixed, T;
  int bc_flag, nsteps, max_steps;
  int ms_on;
  double ms_alpha, ms_beta, ms_split_hp_min, ms_shear_split, ms_hp_min,
      ms_max_dv, ms_max_ops;
  int ms_min_level, ms_max_level, ms_min_ppc, ms_interval;
  /* merge/split rule: 0 = merge_split_ops_cmes.h (the solvers' rule),
   * 1 = the legacy ELFS rule of merge_split_ops.h (split where the distance
   * to the fluid boundary is below ms_alpha * sqrt(Ap) or the divergence
   * exceeds ms_stretch, merge where it is above ms_beta * sqrt(Ap)) */
  int ms_mode;
  double ms_stretch;
} amsc_params;

/* decide_actions (merge_split_ops_cmes.h:116-253) on the device, for a state
 * already binned by amsc_bin_particles (cstart/cptcls).  Writes
 * action[np] (0 KEEP, 1 SPLIT, 2 MERGE_PRIMARY, 3 MERGE_SECONDARY) and
 * partner[np] (device); counts (host, int[2]) = splits, merges.  Identical
 * integer results to the serial reference for every ms_max_ops. */
int amsc_ms_decide(const amsc_params *p, int np, double dt, const double *d_x,
                   const double *d_y, const double *d_Ap, const double *d_hp,
                   const double *d_vpx_dx, const double *d_vpx_dy,
                   const double *d_vpy_dx, const double *d_vpy_dy,
                   const double *d_vpx, const double *d_vpy,
                   const int *d_level, const int *d_cstart,
                   const int *d_cptcls, int *d_action, int *d_partner,
                   int *counts);

/* execute_merge_split (merge_split_ops_cmes.h:258-427) on the caller's
 * device arrays, after amsc_ms_decide: replaces the host
 * adaptive_merge_split of gpu_offload.cpp:321-401 for a caller that keeps
 * its own particle storage.  nfields particle fields (at most 64) named as
 * the reference's dprops: "x" "y" "Mp" "Vp" "Ap" "hp" "vpx" "vpy" "vpx_dx"
 * "vpx_dy" "vpy_dx" "vpy_dy" are required, "mom_px" "mom_py" "xp" "yp" are
 * recomputed when present, any other field is carried (mass-averaged on a
 * merge, copied on a split; Mp Vp Ap mom_px mom_py are extensive).  d_in[k]
 * holds np values, d_out[k] room for `capacity` values (out of place);
 * labels and levels likewise.  *np_out = the new particle count, in the
 * reference's order (kept, split and merged particles at their scan
 * offsets).  AMSC_ERR_CAPACITY if it would exceed capacity (nothing
 * written).  Synchronous. */
int amsc_ms_execute(const amsc_params *p, int np, const int *d_action,
                    const int *d_partner, int nfields,
                    const char *const *names, const double *const *d_in,
                    double *const *d_out, const int *d_label_in,
                    const int *d_level_in, int *d_label_out,
                    int *d_level_out, int capacity, int *np_out);

/* ------------------------------------------------------------------ 3. */
typedef struct amsc_sim amsc_sim;

/* Create a device-resident simulation from host arrays (DATA.json layout:
 * particle arrays of length np, Z of length (nrows+1)*(ncols+1)).
 * capacity >= np bounds the particle count merge/split may reach (0 = 2*np).
 * Performs the step-0 initialisation of optim_par.cpp:235-264. */
int amsc_sim_create(const amsc_params *p, int np, const double *x,
                    const double *y, const double *Mp, const double *Ap,
                    const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp, const double *Z,
                    int capacity, amsc_sim **out);
/* Same, reading a DATA.json file (main/mpm_data.h schema). */
int amsc_sim_create_from_json(const char *path, int max_steps, int capacity,
                              amsc_sim **out);
int amsc_sim_destroy(amsc_sim *s);
/* Advance up to nsteps steps (fewer if the run ends: it == NSTEPS for a
 * fixed dt, t >= T or it == max_steps for an adaptive one).  *done (may be
 * NULL) receives the number of steps taken.  Synchronises only on
 * merge/split steps and when the end condition needs reading. */
int amsc_sim_step(amsc_sim *s, int nsteps, int *done);
/* 1: every step also writes the output-only fields (Ap, Fb, velocity
 * gradients, apx/apy, Zp, hpZ, dZxp/dZyp, Fpx/Fpy, mom) and the nodal
 * arrays; 0 (default): only when a merge/split step needs them. */
int amsc_sim_set_aux(amsc_sim *s, int on);
/* Record HIP events around every kernel launch (per-kernel timing). */
int amsc_sim_set_timing(amsc_sim *s, int on);
/* Accumulated kernel time: names[i] (static strings), total ms and launch
 * count for up to cap kernels; returns the number of kernels in *n. */
int amsc_sim_kernel_times(amsc_sim *s, const char **names, double *ms,
                          int *count, int cap, int *n);
int amsc_sim_reset_kernel_times(amsc_sim *s);
int amsc_sim_num_particles(amsc_sim *s, int *np);
/* it (steps taken), t, dt of the last step; synchronises. */
int amsc_sim_state(amsc_sim *s, int *it, double *t, double *dt);
/* Copy a field to host memory (n entries at most); synchronises.
 * Particle fields: x y Mp Ap vpx vpy mom_px mom_py hp Vp H F_11 F_12 F_21
 * F_22 Fb_x Fb_y vpx_dx vpx_dy vpy_dx vpy_dy apx apy dZxp dZyp Zp hpZ Fpx Fpy
 * Fric_px Fric_py vpxL vpyL; nodal: Mv mom_vx mom_vy F_ext_vx F_ext_vy
 * F_int_vx F_int_vy Fric_x Fric_y FPxv FPyv Ftot_vx Ftot_vy avx avy vvx vvy
 * vvxL vvyL Z. */
int amsc_sim_get_field(amsc_sim *s, const char *name, double *out, int n);
/* Integer fields: label level p2g cptcls cstart. */
int amsc_sim_get_field_int(amsc_sim *s, const char *name, int *out, int n);
/* Overwrite a particle field from host memory (n == num_particles).
 * AMSC_ERR_STALE for Ap, vxdx, vxdy, vydx and vydy after a step that is
 * followed by a merge/split step of a fixed-dt run under the solvers' rule
 * with aux output off: that step keeps them only for the particles whose
 * strain tests passed. */
int amsc_sim_set_field(amsc_sim *s, const char *name, const double *in,
                       int n);
/* err_mass, err_mom (optim_par.cpp:395-428, FP64 device sums) and t of the
 * last step. */
int amsc_sim_conservation(amsc_sim *s, double out[3]);
/* Parameters of the simulation (scenario 0 of a batch). */
int amsc_sim_params(amsc_sim *s, amsc_params *out);
/* The reference solver's output files (optim_par.cpp:269,311-337,655-665),
 * written from the current state:
 *   particles CSV (nc_particles_<it>.csv): a header of quoted field names,
 *     "x","y", then the particle fields in alphabetical order, then "label"
 *     and "level", one row per particle, values with 17 significant digits.
 *     Output-only fields are written when the last step produced them
 *     (amsc_sim_set_aux(s, 1)).  Quadgrid's own column order is not
 *     available, so readers should select columns by name, as the
 *     reference's scripts do;
 *   grid VTS (GRID_forZ.vts, nc_grid_<it>.vts): a VTK XML StructuredGrid of
 *     the (ncols+1) x (nrows+1) nodes, points (x, y, 0), point data Z and
 *     the nodal fields the last step produced. */
int amsc_sim_write_particles_csv(amsc_sim *s, const char *path);
int amsc_sim_write_grid_vts(amsc_sim *s, const char *path);
/* Particle-steps launched so far: the sum over steps of the particle count
 * at that step (steps issued after the end of a run are not counted once
 * the end has been observed). */
int amsc_sim_particle_steps(amsc_sim *s, long long *out);
/* splits, merges of the last merge/split call. */
int amsc_sim_ms_counts(amsc_sim *s, int out[2]);
/* 1: record every step variant once into a HIP graph and replay it (launch
 * extents then follow the particle capacity; results are unchanged).
 * Ignored while per-kernel timing is on. */
int amsc_sim_set_graph(amsc_sim *s, int on);
/* P2G variant v in 0..5: tile shape v % 3 (0: 16x32 nodes, 1: 8x31 nodes
 * swept by one wavefront column by column, 2: 8x32), tile order v / 3 (0:
 * column-major from the last tile column down, 1: linear); particle-kernel
 * workgroup size 256 or 512.  Results do not depend on either.  Default:
 * variant 2 (8x32, column-major from the last tile column down) and 256. */
#define AMSC_P2G_VARIANTS 6
int amsc_sim_set_tiling(amsc_sim *s, int p2g_shape, int particle_block);
/* Autotuner: after one untimed step, time every P2G variant, then every
 * workgroup size, over steps_per_variant live steps each (one step of each
 * candidate per round) and keep the fastest (in mixed precision a swept
 * variant only if it is 5 % faster than the best tiled one: that swept
 * kernel slows down as the particles of a flow spread unevenly over its
 * cells; the linear tile order
 * only if it is 3 % faster than the other order of the same shape, which
 * reads first the particles the previous step wrote last).  The steps are real steps of
 * the run; *steps_used (may be NULL) receives their number,
 * 1 + steps_per_variant * (AMSC_P2G_VARIANTS + 2); variant_ms (may be
 * NULL, AMSC_P2G_VARIANTS + 2 entries) the average P2G time of each
 * variant, then the particle-kernel time of each size. */
int amsc_sim_autotune(amsc_sim *s, int steps_per_variant, int *p2g_shape,
                      int *particle_block, int *steps_used,
                      double *variant_ms);
/* Average duration (ms) of reps back-to-back P2G launches on the current
 * state; they rewrite the nodal velocities the next step recomputes. */
int amsc_sim_time_p2g(amsc_sim *s, int reps, double *avg_ms);

/* ------------------------------------------------------------------ 4. */
#define AMSC_PREC_FP64 0  /* every particle/nodal value and operation FP64 */
#define AMSC_PREC_MIXED 1 /* FP32 arithmetic; FP64 positions, depth, H, the
                             free-surface difference, P2G mass/momentum sums
                             and conservation reductions */

/* An ensemble is a simulation object holding nscen scenarios; every
 * amsc_sim_* function accepts it and then refers to scenario 0. */
typedef struct amsc_sim amsc_ens;

/* Create nscen scenarios: params[i], np[i] particles; the particle arrays
 * are the scenarios' DATA arrays concatenated in scenario order (sum of np
 * entries each), Z the concatenated nodal Z arrays.  capacity bounds the
 * total particle count merge/split may reach (0 = twice the initial). */
int amsc_ens_create(int nscen, const amsc_params *params, const int *np,
                    const double *x, const double *y, const double *Mp,
                    const double *Ap, const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp, const double *Z,
                    int precision, long long capacity, amsc_ens **out);
/* Per-scenario steps taken, t, dt of the last step, 1 while the run goes
 * on, particle count; arrays of nscen entries, any may be NULL.
 * Synchronises.  amsc_sim_step advances every scenario that is still
 * running; *done is the largest number of steps any scenario took. */
int amsc_ens_status(amsc_ens *e, int *it, double *t, double *dt, int *active,
                    int *np);
/* 1 for every scenario whose run was stopped because a particle speed
 * exceeded 1e4 m/s or became NaN (optim_par would step on with a vanishing
 * dt); its state is that of the step that crossed the limit.  nscen
 * entries; synchronises. */
int amsc_ens_diverged(amsc_ens *e, int *out);
int amsc_ens_get_field(amsc_ens *e, int scen, const char *name, double *out,
                       int n);
int amsc_ens_get_field_int(amsc_ens *e, int scen, const char *name, int *out,
                           int n);
int amsc_ens_set_field(amsc_ens *e, int scen, const char *name,
                       const double *in, int n);
int amsc_ens_conservation(amsc_ens *e, int scen, double out[3]);
int amsc_ens_ms_counts(amsc_ens *e, int scen, int out[2]);
/* Scenarios, precision, device bytes held, particle capacity. */
int amsc_ens_info(amsc_ens *e, int *nscen, int *precision, long long *bytes,
                  int *capacity);
/* Served form: a new batch holding the scenarios keep[0..nkeep) of e (their
 * device state copied device-to-device, in that order) followed by nadd new
 * scenarios given as for amsc_ens_create (their momentum must be vpx*Mp).
 * On success e is destroyed and *out replaces it; on failure e is
 * untouched.  Old and new batch are both allocated during the call. */
int amsc_ens_repack(amsc_ens *e, int nkeep, const int *keep, int nadd,
                    const amsc_params *params, const int *np, const double *x,
                    const double *y, const double *Mp, const double *Ap,
                    const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp, const double *Z,
                    long long capacity, amsc_ens **out);
/* Device memory of destroyed simulations is cached for the next ones of
 * the process (at most 32 GiB); this returns it to the device. */
int amsc_release_cache(void);
/* Device bytes amsc_ens_create would allocate (admission control); -1 on
 * bad arguments. */
long long amsc_ens_bytes_estimate(int nscen, const amsc_params *params,
                                  const int *np, int precision,
                                  long long capacity);

#ifdef __cplusplus
}
#endif
#endif
