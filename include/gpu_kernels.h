/*
 * gpu_kernels.h -- drop-in replacement for the reference's main/gpu_kernels.h
 * (namespace gpu_kernels, the same 18 functions with the same parameter
 * lists, gpu_kernels.h:7-196), implemented by the gfx950 library through the
 * C ABI of amsc_mpm.h.
 *
 * Like the reference's functions these return void and complete before
 * returning (the reference's OpenMP target regions are synchronous); an
 * error prints amsc_last_error() and aborts, where the reference would abort
 * inside the offload runtime.  Pointers are device addresses (in an OpenMP
 * offload program, the ones `#pragma omp target data use_device_addr(...)`
 * hands out for the mapped arrays); `coloff` stays a host array and the
 * *_out diagnostics are ignored, as in the reference.
 */
#ifndef GPU_KERNELS_H
#define GPU_KERNELS_H

#include <cstdio>
#include <cstdlib>

#include "amsc_mpm.h"

namespace gpu_kernels {

namespace detail {
// Whether each call waits for its kernels before returning (the default:
// the reference's target regions are synchronous) or only enqueues them on
// the library's stream (set_synchronous(false)).
inline bool &synchronous() {
  static bool on = true;
  return on;
}

inline void check(int rc, const char *what) {
  if (rc == AMSC_OK && synchronous()) rc = amsc_synchronize();
  if (rc != AMSC_OK) {
    std::fprintf(stderr, "gpu_kernels::%s: %s\n", what, amsc_last_error());
    std::abort();
  }
}
}  // namespace detail

// No counterpart in the reference.  set_synchronous(false) lets the calls
// return once their kernels are enqueued, in order, on the library's stream:
// for callers that pass the results of one call to the next on the device
// only and read them on the host after amsc_synchronize().  A kernel that
// fails while running is then reported by that later synchronisation (or
// by the next call that has to wait, such as amsc_bin_particles or
// amsc_cfl_dt), not by the call that launched it.
inline void set_synchronous(bool on) { detail::synchronous() = on; }

inline void reset_nodal_arrays(
    int nn, double *d_Mv, double *d_mom_vx, double *d_mom_vy,
    double *d_F_ext_vx, double *d_F_ext_vy, double *d_F_int_vx,
    double *d_F_int_vy, double *d_Ftot_vx, double *d_Ftot_vy, double *d_avx,
    double *d_avy, double *d_vvx, double *d_vvy, double *d_vvxL,
    double *d_vvyL, double *d_Fric_x, double *d_Fric_y, double *d_FPxv,
    double *d_FPyv, double *d_HV, double *d_pv_rho, double *d_pv_vvx,
    double *d_pv_vvy, double *d_pv_avx, double *d_pv_avy) {
  detail::check(amsc_reset_nodal_arrays(
                    nn, d_Mv, d_mom_vx, d_mom_vy, d_F_ext_vx, d_F_ext_vy,
                    d_F_int_vx, d_F_int_vy, d_Ftot_vx, d_Ftot_vy, d_avx, d_avy,
                    d_vvx, d_vvy, d_vvxL, d_vvyL, d_Fric_x, d_Fric_y, d_FPxv,
                    d_FPyv, d_HV, d_pv_rho, d_pv_vvx, d_pv_vvy, d_pv_avx,
                    d_pv_avy),
                "reset_nodal_arrays");
}

inline void scatter_mass_momentum(int np, int nrows, double hx, double hy,
                                  const int *cstart, const int *cptcls,
                                  const int *ccidx, const int *coloff,
                                  const double *d_x, const double *d_y,
                                  const double *d_Mp, const double *d_mom_px,
                                  const double *d_mom_py, double *d_Mv,
                                  double *d_mom_vx, double *d_mom_vy) {
  detail::check(amsc_scatter_mass_momentum(np, nrows, hx, hy, cstart, cptcls,
                                           ccidx, coloff, d_x, d_y, d_Mp,
                                           d_mom_px, d_mom_py, d_Mv, d_mom_vx,
                                           d_mom_vy),
                "scatter_mass_momentum");
}

inline void gradZ_to_particles(int np, int nrows, double hx, double hy,
                               const double *d_x, const double *d_y,
                               const int *d_p2g, const double *d_Z,
                               double *d_dZxp, double *d_dZyp) {
  detail::check(amsc_gradZ_to_particles(np, nrows, hx, hy, d_x, d_y, d_p2g,
                                        d_Z, d_dZxp, d_dZyp),
                "gradZ_to_particles");
}

inline void gravity_force(int np, double g_c, const double *d_hp,
                          const double *d_H, const double *d_Mp,
                          const double *d_dZxp, const double *d_dZyp,
                          double *d_Fpx, double *d_Fpy) {
  detail::check(amsc_gravity_force(np, g_c, d_hp, d_H, d_Mp, d_dZxp, d_dZyp,
                                   d_Fpx, d_Fpy),
                "gravity_force");
}

inline void friction_particle_force(int np, const double *d_Ap,
                                    const double *d_Fb_x, const double *d_Fb_y,
                                    double *d_Fric_px, double *d_Fric_py) {
  detail::check(amsc_friction_particle_force(np, d_Ap, d_Fb_x, d_Fb_y,
                                             d_Fric_px, d_Fric_py),
                "friction_particle_force");
}

inline void scatter_gravity_force(int np, int nrows, double hx, double hy,
                                  const int *cstart, const int *cptcls,
                                  const int *ccidx, const int *coloff,
                                  const double *d_x, const double *d_y,
                                  const double *d_Fpx, const double *d_Fpy,
                                  double *d_FPxv, double *d_FPyv) {
  detail::check(amsc_scatter_gravity_force(np, nrows, hx, hy, cstart, cptcls,
                                           ccidx, coloff, d_x, d_y, d_Fpx,
                                           d_Fpy, d_FPxv, d_FPyv),
                "scatter_gravity_force");
}

inline void scatter_friction_force(int np, int nrows, double hx, double hy,
                                   const int *cstart, const int *cptcls,
                                   const int *ccidx, const int *coloff,
                                   const double *d_x, const double *d_y,
                                   const double *d_Fric_px,
                                   const double *d_Fric_py, double *d_Fric_x,
                                   double *d_Fric_y) {
  detail::check(amsc_scatter_friction_force(np, nrows, hx, hy, cstart, cptcls,
                                            ccidx, coloff, d_x, d_y, d_Fric_px,
                                            d_Fric_py, d_Fric_x, d_Fric_y),
                "scatter_friction_force");
}

inline void compute_F_ext(int nn, const double *d_FPxv, const double *d_FPyv,
                          const double *d_Fric_x, const double *d_Fric_y,
                          double *d_F_ext_vx, double *d_F_ext_vy) {
  detail::check(amsc_compute_F_ext(nn, d_FPxv, d_FPyv, d_Fric_x, d_Fric_y,
                                   d_F_ext_vx, d_F_ext_vy),
                "compute_F_ext");
}

inline void scatter_internal_forces(int np, int nrows, double hx, double hy,
                                    const int *cstart, const int *cptcls,
                                    const int *ccidx, const int *coloff,
                                    const double *d_x, const double *d_y,
                                    const double *d_Vp, const double *d_F11,
                                    const double *d_F12, const double *d_F21,
                                    const double *d_F22, double *d_F_int_vx,
                                    double *d_F_int_vy) {
  detail::check(amsc_scatter_internal_forces(
                    np, nrows, hx, hy, cstart, cptcls, ccidx, coloff, d_x, d_y,
                    d_Vp, d_F11, d_F12, d_F21, d_F22, d_F_int_vx, d_F_int_vy),
                "scatter_internal_forces");
}

inline void nodal_force_and_momentum(int nn, double dt,
                                     const double *d_F_ext_vx,
                                     const double *d_F_ext_vy,
                                     const double *d_F_int_vx,
                                     const double *d_F_int_vy,
                                     double *d_Ftot_vx, double *d_Ftot_vy,
                                     double *d_mom_vx, double *d_mom_vy) {
  detail::check(amsc_nodal_force_and_momentum(nn, dt, d_F_ext_vx, d_F_ext_vy,
                                              d_F_int_vx, d_F_int_vy,
                                              d_Ftot_vx, d_Ftot_vy, d_mom_vx,
                                              d_mom_vy),
                "nodal_force_and_momentum");
}

inline void nodal_velocity_and_bc(int nn, int nrows, int ncols, double dt,
                                  bool bc_flag, const double *d_Mv,
                                  const double *d_Ftot_vx,
                                  const double *d_Ftot_vy, double *d_mom_vx,
                                  double *d_mom_vy, double *d_avx,
                                  double *d_avy, double *d_vvx, double *d_vvy,
                                  double *d_vvxL, double *d_vvyL) {
  detail::check(amsc_nodal_velocity_and_bc(nn, nrows, ncols, dt,
                                           bc_flag ? 1 : 0, d_Mv, d_Ftot_vx,
                                           d_Ftot_vy, d_mom_vx, d_mom_vy,
                                           d_avx, d_avy, d_vvx, d_vvy, d_vvxL,
                                           d_vvyL),
                "nodal_velocity_and_bc");
}

inline void g2p_velocity_and_advect(
    int np, int nrows, int ncols, double hx, double hy, double dt,
    const int *d_p2g, const double *d_vvxL, const double *d_vvyL,
    const double *d_avx, const double *d_avy, double *d_x, double *d_y,
    double *d_vpx, double *d_vpy, double *d_apx, double *d_apy,
    double *d_mom_px, double *d_mom_py, double *d_vpxL, double *d_vpyL,
    const double *d_Mp) {
  detail::check(amsc_g2p_velocity_and_advect(
                    np, nrows, ncols, hx, hy, dt, d_p2g, d_vvxL, d_vvyL, d_avx,
                    d_avy, d_x, d_y, d_vpx, d_vpy, d_apx, d_apy, d_mom_px,
                    d_mom_py, d_vpxL, d_vpyL, d_Mp),
                "g2p_velocity_and_advect");
}

inline void g2pd_gradients_and_height_update(
    int np, int nrows, double hx, double hy, double dt, const double *d_x,
    const double *d_y, const int *d_p2g, const double *d_vvxL,
    const double *d_vvyL, double *d_vpx_dx, double *d_vpx_dy,
    double *d_vpy_dx, double *d_vpy_dy, double *d_hp, double *d_Vp,
    double *d_Ap, double *min_sc_out, double *min_hp_out) {
  detail::check(amsc_g2pd_gradients_and_height_update(
                    np, nrows, hx, hy, dt, d_x, d_y, d_p2g, d_vvxL, d_vvyL,
                    d_vpx_dx, d_vpx_dy, d_vpy_dx, d_vpy_dy, d_hp, d_Vp, d_Ap,
                    min_sc_out, min_hp_out),
                "g2pd_gradients_and_height_update");
}

inline void voellmy_friction(int np, double rho, double g_c, double tan_fa,
                             double xi, double fric_on, const double *d_vpx,
                             const double *d_vpy, const double *d_hp,
                             double *d_Fb_x, double *d_Fb_y) {
  detail::check(amsc_voellmy_friction(np, rho, g_c, tan_fa, xi, fric_on, d_vpx,
                                      d_vpy, d_hp, d_Fb_x, d_Fb_y),
                "voellmy_friction");
}

inline void bingham_stress_update(
    int np, double rho, double g_c, double mu, double tau_Y, double cc,
    double A_coeff, double C_coeff, const double *d_vpx, const double *d_vpy,
    const double *d_hp, const double *d_H, const double *d_vpx_dx,
    const double *d_vpx_dy, const double *d_vpy_dx, const double *d_vpy_dy,
    double *d_F11, double *d_F12, double *d_F21, double *d_F22,
    double *max_coeff_out, double *min_hp_out, double *max_nv_out,
    double *max_F11_out, double *max_dev_out, double *max_pres_out,
    double *max_Dxx_out) {
  detail::check(amsc_bingham_stress_update(
                    np, rho, g_c, mu, tau_Y, cc, A_coeff, C_coeff, d_vpx,
                    d_vpy, d_hp, d_H, d_vpx_dx, d_vpx_dy, d_vpy_dx, d_vpy_dy,
                    d_F11, d_F12, d_F21, d_F22, max_coeff_out, min_hp_out,
                    max_nv_out, max_F11_out, max_dev_out, max_pres_out,
                    max_Dxx_out),
                "bingham_stress_update");
}

inline void g2p_height(int np, int nrows, double hx, double hy,
                       const double *d_x, const double *d_y, const int *d_p2g,
                       const double *d_Z, const double *d_hp, double *d_Zp,
                       double *d_hpZ) {
  detail::check(amsc_g2p_height(np, nrows, hx, hy, d_x, d_y, d_p2g, d_Z, d_hp,
                                d_Zp, d_hpZ),
                "g2p_height");
}

inline void scatter_plotvars(int np, int nrows, double hx, double hy,
                             const int *cstart, const int *cptcls,
                             const int *ccidx, const int *coloff,
                             const double *d_x, const double *d_y,
                             const double *d_Mp, const double *d_vpx,
                             const double *d_vpy, const double *d_apx,
                             const double *d_apy, const double *d_hp,
                             double *d_pv_rho, double *d_pv_vvx,
                             double *d_pv_vvy, double *d_pv_avx,
                             double *d_pv_avy, double *d_HV) {
  detail::check(amsc_scatter_plotvars(np, nrows, hx, hy, cstart, cptcls, ccidx,
                                      coloff, d_x, d_y, d_Mp, d_vpx, d_vpy,
                                      d_apx, d_apy, d_hp, d_pv_rho, d_pv_vvx,
                                      d_pv_vvy, d_pv_avx, d_pv_avy, d_HV),
                "scatter_plotvars");
}

inline void normalize_plotvars(int nn, double *d_pv_rho, double *d_pv_vvx,
                               double *d_pv_vvy, double *d_pv_avx,
                               double *d_pv_avy) {
  detail::check(amsc_normalize_plotvars(nn, d_pv_rho, d_pv_vvx, d_pv_vvy,
                                        d_pv_avx, d_pv_avy),
                "normalize_plotvars");
}

}  // namespace gpu_kernels

#endif /* GPU_KERNELS_H */
