/*
 * mpm_oracle.c -- FP64 CPU oracle for the depth-averaged MPM time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): never linked into the product.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -march, no FMA).
 *
 * Expression order follows the reference sources cited per function, so
 * that a GPU implementation compiled without FMA contraction can be
 * compared bit for bit.
 */
#include "mpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Per-particle and per-node loops of the time step run in parallel when the
 * library is built with OpenMP and the arrays are large (each iteration
 * writes only its own entry, so results do not depend on the thread count);
 * scatters, sums and binning stay serial, in the reference's order. */
#ifdef _OPENMP
#include <omp.h>
#define OR_PAR_NP _Pragma("omp parallel for schedule(static) if (np > 16384)")
#define OR_PAR_NN _Pragma("omp parallel for schedule(static) if (nn > 16384)")
#else
#define OR_PAR_NP
#define OR_PAR_NN
#endif

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int or_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ===================================================================== */
/* quadgrid layout helpers: main/gpu_kernels.cpp:7-59                     */
/* node id r + c*(nrows+1); cell id r + c*nrows; local nodes               */
/* 0:(c,r) 1:(c,r+1) 2:(c+1,r) 3:(c+1,r+1)                                 */
/* ===================================================================== */
static inline int gt(int inode, int c, int r, int nrows) {
  int bl = r + c * (nrows + 1);
  switch (inode) {
  case 0: return bl;
  case 1: return bl + 1;
  case 2: return bl + (nrows + 1);
  default: return bl + (nrows + 2);
  }
}

/* pnode(0,0,...) = c*hx and pnode(1,0,...) = r*hy (gpu_kernels.cpp:21-31) */
static inline double shp(double x, double y, int inode, int c, int r,
                         double hx, double hy) {
  /* gpu_kernels.cpp:33-41 */
  switch (inode) {
  case 3: return ((x - c * hx) / hx) * ((y - r * hy) / hy);
  case 2: return ((x - c * hx) / hx) * (1.0 - (y - r * hy) / hy);
  case 1: return (1.0 - (x - c * hx) / hx) * ((y - r * hy) / hy);
  default: return (1.0 - (x - c * hx) / hx) * (1.0 - (y - r * hy) / hy);
  }
}

static inline double shg(double x, double y, int idir, int inode, int c,
                         int r, double hx, double hy) {
  /* gpu_kernels.cpp:43-59 */
  switch (inode) {
  case 3:
    if (idir == 0) return (1.0 / hx) * ((y - r * hy) / hy);
    else return ((x - c * hx) / hx) * (1.0 / hy);
  case 2:
    if (idir == 0) return (1.0 / hx) * (1.0 - (y - r * hy) / hy);
    else return ((x - c * hx) / hx) * (-1.0 / hy);
  case 1:
    if (idir == 0) return (-1.0 / hx) * ((y - r * hy) / hy);
    else return (1.0 - (x - c * hx) / hx) * (1.0 / hy);
  default:
    if (idir == 0) return (-1.0 / hx) * (1.0 - (y - r * hy) / hy);
    else return (1.0 - (x - c * hx) / hx) * (-1.0 / hy);
  }
}

/* ===================================================================== */
/* 1. per-kernel restatements, main/gpu_kernels.cpp                       */
/* ===================================================================== */

/* gpu_kernels.cpp:66-94 */
void or_reset_nodal_arrays(int nn, double *Mv, double *mom_vx, double *mom_vy,
                           double *F_ext_vx, double *F_ext_vy,
                           double *F_int_vx, double *F_int_vy,
                           double *Ftot_vx, double *Ftot_vy, double *avx,
                           double *avy, double *vvx, double *vvy, double *vvxL,
                           double *vvyL, double *Fric_x, double *Fric_y,
                           double *FPxv, double *FPyv, double *HV,
                           double *pv_rho, double *pv_vvx, double *pv_vvy,
                           double *pv_avx, double *pv_avy) {
  double *a[25] = {Mv,     mom_vx, mom_vy, F_ext_vx, F_ext_vy, F_int_vx,
                   F_int_vy, Ftot_vx, Ftot_vy, avx, avy, vvx, vvy, vvxL,
                   vvyL,   Fric_x, Fric_y, FPxv, FPyv, HV, pv_rho, pv_vvx,
                   pv_vvy, pv_avx, pv_avy};
  for (int k = 0; k < 25; ++k)
    for (int i = 0; i < nn; ++i) a[k][i] = 0;
}

/* Colour-ordered scatter driver shared by the P2G restatements: colours
 * 0..3 in turn (gpu_kernels.cpp:104-107), cells of a colour in ccidx order,
 * the cell's particles in cptcls order, nodes 0..3 (gpu_kernels.cpp:110-126).
 */
#define COLOUR_LOOP_BEGIN                                                     \
  for (int color = 0; color < 4; ++color) {                                   \
    const int cbegin = coloff[color], cend = coloff[color + 1];               \
    for (int ic = cbegin; ic < cend; ++ic) {                                  \
      const int cell = ccidx[ic];                                             \
      const int r = cell % nrows, c = cell / nrows;                           \
      for (int jp = cstart[cell]; jp < cstart[cell + 1]; ++jp) {              \
        const int ip = cptcls[jp];                                            \
        const double xx = x[ip], yy = y[ip];
#define COLOUR_LOOP_END                                                       \
  }                                                                           \
  }                                                                           \
  }

/* gpu_kernels.cpp:97-128 */
void or_scatter_mass_momentum(int np, int nrows, double hx, double hy,
                              const int *cstart, const int *cptcls,
                              const int *ccidx, const int *coloff,
                              const double *x, const double *y,
                              const double *Mp, const double *mom_px,
                              const double *mom_py, double *Mv, double *mom_vx,
                              double *mom_vy) {
  (void)np;
  COLOUR_LOOP_BEGIN
  const double mp = Mp[ip], mpx = mom_px[ip], mpy = mom_py[ip];
  for (int inode = 0; inode < 4; ++inode) {
    const double N = shp(xx, yy, inode, c, r, hx, hy);
    const int n = gt(inode, c, r, nrows);
    Mv[n] += N * mp;
    mom_vx[n] += N * mpx;
    mom_vy[n] += N * mpy;
  }
  COLOUR_LOOP_END
}

/* gpu_kernels.cpp:131-152 */
void or_gradZ_to_particles(int np, int nrows, double hx, double hy,
                           const double *x, const double *y, const int *p2g,
                           const double *Z, double *dZxp, double *dZyp) {
  for (int ip = 0; ip < np; ++ip) {
    const double xx = x[ip], yy = y[ip];
    const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
    double dzx = 0.0, dzy = 0.0;
    for (int inode = 0; inode < 4; ++inode) {
      const int n = gt(inode, c, r, nrows);
      dzx += shg(xx, yy, 0, inode, c, r, hx, hy) * Z[n];
      dzy += shg(xx, yy, 1, inode, c, r, hx, hy) * Z[n];
    }
    dZxp[ip] = dzx;
    dZyp[ip] = dzy;
  }
}

/* gpu_kernels.cpp:155-167 */
void or_gravity_force(int np, double g, const double *hp, const double *H,
                      const double *Mp, const double *dZxp, const double *dZyp,
                      double *Fpx, double *Fpy) {
  for (int ip = 0; ip < np; ++ip) {
    const double ratio = hp[ip] > 1e-10 ? H[ip] / hp[ip] : 0.0;
    Fpx[ip] = -g * Mp[ip] * (1.0 - ratio) * dZxp[ip];
    Fpy[ip] = -g * Mp[ip] * (1.0 - ratio) * dZyp[ip];
  }
}

/* gpu_kernels.cpp:170-180 */
void or_friction_particle_force(int np, const double *Ap, const double *Fb_x,
                                const double *Fb_y, double *Fric_px,
                                double *Fric_py) {
  for (int ip = 0; ip < np; ++ip) {
    Fric_px[ip] = Ap[ip] * Fb_x[ip];
    Fric_py[ip] = Ap[ip] * Fb_y[ip];
  }
}

/* gpu_kernels.cpp:183-213 */
void or_scatter_gravity_force(int np, int nrows, double hx, double hy,
                              const int *cstart, const int *cptcls,
                              const int *ccidx, const int *coloff,
                              const double *x, const double *y,
                              const double *Fpx, const double *Fpy,
                              double *FPxv, double *FPyv) {
  (void)np;
  COLOUR_LOOP_BEGIN
  const double fx = Fpx[ip], fy = Fpy[ip];
  for (int inode = 0; inode < 4; ++inode) {
    const double N = shp(xx, yy, inode, c, r, hx, hy);
    const int n = gt(inode, c, r, nrows);
    FPxv[n] += N * fx;
    FPyv[n] += N * fy;
  }
  COLOUR_LOOP_END
}

/* gpu_kernels.cpp:216-246 */
void or_scatter_friction_force(int np, int nrows, double hx, double hy,
                               const int *cstart, const int *cptcls,
                               const int *ccidx, const int *coloff,
                               const double *x, const double *y,
                               const double *Fric_px, const double *Fric_py,
                               double *Fric_x, double *Fric_y) {
  (void)np;
  COLOUR_LOOP_BEGIN
  const double fx = Fric_px[ip], fy = Fric_py[ip];
  for (int inode = 0; inode < 4; ++inode) {
    const double N = shp(xx, yy, inode, c, r, hx, hy);
    const int n = gt(inode, c, r, nrows);
    Fric_x[n] += N * fx;
    Fric_y[n] += N * fy;
  }
  COLOUR_LOOP_END
}

/* gpu_kernels.cpp:249-260 */
void or_compute_F_ext(int nn, const double *FPxv, const double *FPyv,
                      const double *Fric_x, const double *Fric_y,
                      double *F_ext_vx, double *F_ext_vy) {
  for (int i = 0; i < nn; ++i) {
    F_ext_vx[i] = FPxv[i] + Fric_x[i];
    F_ext_vy[i] = FPyv[i] + Fric_y[i];
  }
}

/* gpu_kernels.cpp:263-296 */
void or_scatter_internal_forces(int np, int nrows, double hx, double hy,
                                const int *cstart, const int *cptcls,
                                const int *ccidx, const int *coloff,
                                const double *x, const double *y,
                                const double *Vp, const double *F11,
                                const double *F12, const double *F21,
                                const double *F22, double *F_int_vx,
                                double *F_int_vy) {
  (void)np;
  COLOUR_LOOP_BEGIN
  const double vp = Vp[ip];
  const double f11 = F11[ip], f12 = F12[ip], f21 = F21[ip], f22 = F22[ip];
  for (int inode = 0; inode < 4; ++inode) {
    const double Nx = shg(xx, yy, 0, inode, c, r, hx, hy);
    const double Ny = shg(xx, yy, 1, inode, c, r, hx, hy);
    const int n = gt(inode, c, r, nrows);
    F_int_vx[n] += (Nx * f11 + Ny * f21) * vp;
    F_int_vy[n] += (Nx * f12 + Ny * f22) * vp;
  }
  COLOUR_LOOP_END
}

/* gpu_kernels.cpp:299-315 (the momentum update is commented out there) */
void or_nodal_force_and_momentum(int nn, double dt, const double *F_ext_vx,
                                 const double *F_ext_vy,
                                 const double *F_int_vx,
                                 const double *F_int_vy, double *Ftot_vx,
                                 double *Ftot_vy, double *mom_vx,
                                 double *mom_vy) {
  (void)dt; (void)mom_vx; (void)mom_vy;
  for (int i = 0; i < nn; ++i) {
    Ftot_vx[i] = F_ext_vx[i] + F_int_vx[i];
    Ftot_vy[i] = F_ext_vy[i] + F_int_vy[i];
  }
}

/* gpu_kernels.cpp:318-353 */
void or_nodal_velocity_and_bc(int nn, int nrows, int ncols, double dt,
                              int bc_flag, const double *Mv,
                              const double *Ftot_vx, const double *Ftot_vy,
                              double *mom_vx, double *mom_vy, double *avx,
                              double *avy, double *vvx, double *vvy,
                              double *vvxL, double *vvyL) {
  for (int iv = 0; iv < nn; ++iv) {
    const int active = Mv[iv] > 1e-2;
    double ax = active ? Ftot_vx[iv] / Mv[iv] : 0.0;
    double ay = active ? Ftot_vy[iv] / Mv[iv] : 0.0;
    double vx = active ? mom_vx[iv] / Mv[iv] : 0.0;
    double vy = active ? mom_vy[iv] / Mv[iv] : 0.0;
    double vxL = active ? dt * ax + vx : 0.0;
    double vyL = active ? dt * ay + vy : 0.0;
    if (bc_flag) {
      const int r = iv % (nrows + 1), c = iv / (nrows + 1);
      if (c == 0 || c == 1 || c == ncols - 1 || c == ncols) {
        mom_vx[iv] = 0.0; vx = 0.0; vxL = 0.0; ax = 0.0;
      }
      if (r == 0 || r == 1 || r == nrows - 1 || r == nrows) {
        mom_vy[iv] = 0.0; vy = 0.0; vyL = 0.0; ay = 0.0;
      }
    }
    avx[iv] = ax; avy[iv] = ay;
    vvx[iv] = vx; vvy[iv] = vy;
    vvxL[iv] = vxL; vvyL[iv] = vyL;
  }
}

/* gpu_kernels.cpp:356-407 */
void or_g2p_velocity_and_advect(int np, int nrows, int ncols, double hx,
                                double hy, double dt, const int *p2g,
                                const double *vvxL, const double *vvyL,
                                const double *avx, const double *avy,
                                double *x, double *y, double *vpx, double *vpy,
                                double *apx, double *apy, double *mom_px,
                                double *mom_py, double *vpxL, double *vpyL,
                                const double *Mp) {
  for (int ip = 0; ip < np; ++ip) {
    const double xx = x[ip], yy = y[ip];
    const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
    double ax = 0, ay = 0, vxL = 0, vyL = 0;
    for (int inode = 0; inode < 4; ++inode) {
      const double N = shp(xx, yy, inode, c, r, hx, hy);
      const int n = gt(inode, c, r, nrows);
      ax += N * avx[n];
      ay += N * avy[n];
      vxL += N * vvxL[n];
      vyL += N * vvyL[n];
    }
    vpx[ip] = vxL; vpy[ip] = vyL;
    apx[ip] = ax; apy[ip] = ay;
    x[ip] += dt * vxL;
    y[ip] += dt * vyL;
    const double eps = 1e-10, Lx = ncols * hx, Ly = nrows * hy;
    if (x[ip] < eps) x[ip] = eps;
    if (x[ip] > Lx - eps) x[ip] = Lx - eps;
    if (y[ip] < eps) y[ip] = eps;
    if (y[ip] > Ly - eps) y[ip] = Ly - eps;
    vpxL[ip] = vxL; vpyL[ip] = vyL;
    mom_px[ip] = vxL * Mp[ip];
    mom_py[ip] = vyL * Mp[ip];
  }
}

/* gpu_kernels.cpp:410-454 (the min reductions are never written out) */
void or_g2pd_gradients_and_height_update(
    int np, int nrows, double hx, double hy, double dt, const double *x,
    const double *y, const int *p2g, const double *vvxL, const double *vvyL,
    double *vpx_dx, double *vpx_dy, double *vpy_dx, double *vpy_dy,
    double *hp, double *Vp, double *Ap) {
  for (int ip = 0; ip < np; ++ip) {
    const double xx = x[ip], yy = y[ip];
    const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
    double vxdx = 0, vxdy = 0, vydx = 0, vydy = 0;
    for (int inode = 0; inode < 4; ++inode) {
      const int n = gt(inode, c, r, nrows);
      const double Nx = shg(xx, yy, 0, inode, c, r, hx, hy);
      const double Ny = shg(xx, yy, 1, inode, c, r, hx, hy);
      vxdx += Nx * vvxL[n];
      vxdy += Ny * vvxL[n];
      vydx += Nx * vvyL[n];
      vydy += Ny * vvyL[n];
    }
    vpx_dx[ip] = vxdx; vpx_dy[ip] = vxdy;
    vpy_dx[ip] = vydx; vpy_dy[ip] = vydy;
    const double sc = 1.0 + dt * (vxdx + vydy);
    hp[ip] /= sc;
    Ap[ip] = Vp[ip] / hp[ip];
  }
}

/* gpu_kernels.cpp:457-475 */
void or_voellmy_friction(int np, double rho, double g, double tan_fa,
                         double xi, double fric_on, const double *vpx,
                         const double *vpy, const double *hp, double *Fb_x,
                         double *Fb_y) {
  for (int ip = 0; ip < np; ++ip) {
    const double vx = vpx[ip], vy = vpy[ip], h = hp[ip];
    const double nv = sqrt(vx * vx + vy * vy);
    if (fric_on > 0 && nv > 1e-10 && xi > 0) {
      const double v2 = vx * vx + vy * vy;
      Fb_x[ip] = -fric_on * (rho * g * h * tan_fa + rho * g * v2 / xi) * vx / nv;
      Fb_y[ip] = -fric_on * (rho * g * h * tan_fa + rho * g * v2 / xi) * vy / nv;
    } else {
      Fb_x[ip] = 0.0;
      Fb_y[ip] = 0.0;
    }
  }
}

/* gpu_kernels.cpp:478-546 (diagnostic maxima are never written out) */
void or_bingham_stress_update(int np, double rho, double g, double mu,
                              double tau_Y, double cc, double A_coeff,
                              double C_coeff, const double *vpx,
                              const double *vpy, const double *hp,
                              const double *H, const double *vpx_dx,
                              const double *vpx_dy, const double *vpy_dx,
                              const double *vpy_dy, double *F11, double *F12,
                              double *F21, double *F22) {
  for (int ip = 0; ip < np; ++ip) {
    const double vx = vpx[ip], vy = vpy[ip], h = hp[ip];
    const double nv = sqrt(vx * vx + vy * vy);
    const double alf =
        (h > 1e-3 && tau_Y > 1e-12) ? (6.0 * mu * nv) / ((h + 0.001) * tau_Y)
                                    : 0.0;
    const double bv = -114.0 / 32.0 - alf;
    const double sq = sqrt(bv * bv - 4.0 * A_coeff * C_coeff);
    const double z1 = (-bv + sq) / (2.0 * A_coeff);
    const double z2 = (-bv - sq) / (2.0 * A_coeff);
    const double zz = (fabs(z1 - 0.5) <= 0.5) ? z1 : z2;
    const double Dxx = vpx_dx[ip];
    const double Dxy = 0.5 * (vpx_dy[ip] + vpy_dx[ip]);
    const double Dyy = vpy_dy[ip];
    const double Dzz = -(Dxx + Dyy);
    const double Dzx =
        h > 1e-3 ? 0.5 * (3.0 / (2.0 + zz)) * (vx / (h + 0.001)) : 0.0;
    const double Dzy =
        h > 1e-3 ? 0.5 * (3.0 / (2.0 + zz)) * (vy / (h + 0.001)) : 0.0;
    const double inv2 =
        0.5 * (Dxx * Dxx + Dyy * Dyy + Dzz * Dzz) + Dzx * Dzx + Dzy * Dzy +
        Dxy * Dxy;
    const double coeff_max = 5.0e6;
    double inv2_floor = (tau_Y / coeff_max) * (tau_Y / coeff_max);
    if (inv2_floor < 1e-9) inv2_floor = 1e-9;
    const double coeff = tau_Y / sqrt(inv2 + inv2_floor) + 2.0 * mu;
    const double h_corr = h > 1e-10 ? h - H[ip] * H[ip] / h : 0.0;
    F11[ip] = -cc * coeff * Dxx + 0.5 * rho * g * h_corr;
    F22[ip] = -cc * coeff * Dyy + 0.5 * rho * g * h_corr;
    F12[ip] = -cc * coeff * Dxy;
    F21[ip] = -cc * coeff * Dxy;
  }
}

/* gpu_kernels.cpp:549-569 */
void or_g2p_height(int np, int nrows, double hx, double hy, const double *x,
                   const double *y, const int *p2g, const double *Z,
                   const double *hp, double *Zp, double *hpZ) {
  for (int ip = 0; ip < np; ++ip) {
    const double xx = x[ip], yy = y[ip];
    const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
    double zp = 0.0;
    for (int inode = 0; inode < 4; ++inode)
      zp += shp(xx, yy, inode, c, r, hx, hy) * Z[gt(inode, c, r, nrows)];
    Zp[ip] = zp;
    hpZ[ip] = hp[ip] + zp;
  }
}

/* gpu_kernels.cpp:572-609 */
void or_scatter_plotvars(int np, int nrows, double hx, double hy,
                         const int *cstart, const int *cptcls,
                         const int *ccidx, const int *coloff, const double *x,
                         const double *y, const double *Mp, const double *vpx,
                         const double *vpy, const double *apx,
                         const double *apy, const double *hp, double *pv_rho,
                         double *pv_vvx, double *pv_vvy, double *pv_avx,
                         double *pv_avy, double *HV) {
  (void)np;
  COLOUR_LOOP_BEGIN
  const double mp = Mp[ip], vx = vpx[ip], vy = vpy[ip];
  const double ax = apx[ip], ay = apy[ip], hh = hp[ip];
  for (int inode = 0; inode < 4; ++inode) {
    const double N = shp(xx, yy, inode, c, r, hx, hy);
    const int n = gt(inode, c, r, nrows);
    pv_rho[n] += N * mp;
    pv_vvx[n] += N * vx;
    pv_vvy[n] += N * vy;
    pv_avx[n] += N * ax;
    pv_avy[n] += N * ay;
    HV[n] += N * hh;
  }
  COLOUR_LOOP_END
}

/* gpu_kernels.cpp:612-627 */
void or_normalize_plotvars(int nn, double *pv_rho, double *pv_vvx,
                           double *pv_vvy, double *pv_avx, double *pv_avy) {
  for (int i = 0; i < nn; ++i) {
    const double mv = pv_rho[i];
    if (mv > 1e-8) {
      pv_vvx[i] /= mv;
      pv_vvy[i] /= mv;
      pv_avx[i] /= mv;
      pv_avy[i] /= mv;
    }
  }
}

/* ===================================================================== */
/* 2. binning: quadgrid particles_t::init_particle_mesh (un-vendored;     */
/*    called at optim_par.cpp:341 and gpu_offload.cpp:243,380,549).        */
/*    Cell = (floor(y/hy), floor(x/hx)) clamped to the grid (the clamp is  */
/*    an assumption: advection already keeps x in [1e-10, L-1e-10],        */
/*    gpu_kernels.cpp:393-399).  CSR in ascending cell, particles in index */
/*    order within a cell (grd_to_ptcl is an ordered map filled in index   */
/*    order).  Colouring: colour = (r&1) + 2*(c&1); cells of a colour in   */
/*    ascending order.  Adjacent cells never share a colour, which is the  */
/*    property the colour-parallel scatter needs (gpu_kernels.cpp:104-126). */
/* ===================================================================== */
int or_cell_of(double x, double y, int nrows, int ncols, double hx,
               double hy) {
  int c = (int)floor(x / hx), r = (int)floor(y / hy);
  if (c < 0) c = 0;
  if (c > ncols - 1) c = ncols - 1;
  if (r < 0) r = 0;
  if (r > nrows - 1) r = nrows - 1;
  return r + c * nrows;
}

void or_bin_particles(int np, const double *x, const double *y, int nrows,
                      int ncols, double hx, double hy, int *p2g, int *cstart,
                      int *cptcls, int *ccidx, int *coloff) {
  const int nc = nrows * ncols;
  for (int i = 0; i <= nc; ++i) cstart[i] = 0;
  for (int ip = 0; ip < np; ++ip) {
    p2g[ip] = or_cell_of(x[ip], y[ip], nrows, ncols, hx, hy);
    cstart[p2g[ip] + 1]++;
  }
  for (int i = 0; i < nc; ++i) cstart[i + 1] += cstart[i];
  int *fill = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  for (int i = 0; i < nc; ++i) fill[i] = cstart[i];
  for (int ip = 0; ip < np; ++ip) cptcls[fill[p2g[ip]]++] = ip;
  free(fill);
  int k = 0;
  for (int col = 0; col < 4; ++col) {
    coloff[col] = k;
    for (int cell = 0; cell < nc; ++cell) {
      const int r = cell % nrows, c = cell / nrows;
      if (((r & 1) + 2 * (c & 1)) == col) ccidx[k++] = cell;
    }
  }
  coloff[4] = k;
}

/* ===================================================================== */
/* 3. the optim_par step                                                  */
/* ===================================================================== */
static const char *FIELD_NAMES[OR_NFIELDS] = {
    "x",      "y",      "Mp",     "Ap",      "vpx",     "vpy",  "mom_px",
    "mom_py", "hp",     "Vp",     "H",       "F_11",    "F_12", "F_21",
    "F_22",   "Fb_x",   "Fb_y",   "vpx_dx",  "vpx_dy",  "vpy_dx", "vpy_dy",
    "apx",    "apy",    "dZxp",   "dZyp",    "Zp",      "hpZ",  "Fpx",
    "Fpy",    "Fric_px", "Fric_py", "vpxL",  "vpyL",    "xp",   "yp"};

/* extensive props of merge_split_ops_cmes.h:49-50 */
static int is_extensive(int f) {
  return f == OR_MP || f == OR_VP || f == OR_AP || f == OR_MOM_PX ||
         f == OR_MOM_PY;
}

const char *or_field_name(int f) {
  return (f >= 0 && f < OR_NFIELDS) ? FIELD_NAMES[f] : "";
}

enum {
  NV_MV, NV_MOMX, NV_MOMY, NV_FEXTX, NV_FEXTY, NV_FINTX, NV_FINTY, NV_FRX,
  NV_FRY, NV_AVX, NV_AVY, NV_VVX, NV_VVY, NV_FTOTX, NV_FTOTY, NV_HV, NV_FPX,
  NV_FPY, NV_VVXL, NV_VVYL, NV_COUNT
};
static const char *NODAL_NAMES[NV_COUNT] = {
    "Mv",     "mom_vx", "mom_vy", "F_ext_vx", "F_ext_vy", "F_int_vx",
    "F_int_vy", "Fric_x", "Fric_y", "avx",    "avy",      "vvx",
    "vvy",    "Ftot_vx", "Ftot_vy", "HV",     "FPxv",     "FPyv",
    "vvxL",   "vvyL"};

struct or_sim {
  or_params p;
  int np, cap, nn, nc;
  double *f[OR_NFIELDS];
  int *label, *level;
  double *Z;
  double *nv[NV_COUNT];
  int *p2g, *cstart, *cptcls, *ccidx;
  int coloff[5];
  int it;
  double t, dt;
  double cons[3];
  int ms_counts[2];
};

static void sim_reserve(or_sim *s, int cap) {
  if (cap <= s->cap) return;
  for (int k = 0; k < OR_NFIELDS; ++k)
    s->f[k] = (double *)realloc(s->f[k], sizeof(double) * cap);
  s->label = (int *)realloc(s->label, sizeof(int) * cap);
  s->level = (int *)realloc(s->level, sizeof(int) * cap);
  s->p2g = (int *)realloc(s->p2g, sizeof(int) * cap);
  s->cptcls = (int *)realloc(s->cptcls, sizeof(int) * cap);
  s->cap = cap;
}

/* p2g restatement, optim_par order: ascending cell, index order within the
 * cell, nodes 0..3; accumulates (+=) like quadgrid's p2g (D19). */
static void p2g_field(const or_sim *s, const double *pv, double *gv) {
  const int nrows = s->p.nrows;
  const double hx = s->p.hx, hy = s->p.hy;
  const double *x = s->f[OR_X], *y = s->f[OR_Y];
  for (int cell = 0; cell < s->nc; ++cell) {
    const int r = cell % nrows, c = cell / nrows;
    for (int jp = s->cstart[cell]; jp < s->cstart[cell + 1]; ++jp) {
      const int ip = s->cptcls[jp];
      for (int inode = 0; inode < 4; ++inode)
        gv[gt(inode, c, r, nrows)] +=
            shp(x[ip], y[ip], inode, c, r, hx, hy) * pv[ip];
    }
  }
}

/* p2gd restatement (optim_par.cpp:476): g += (Nx*a + Ny*b)*w */
static void p2gd_field(const or_sim *s, const double *a, const double *b,
                       const double *w, double *gv) {
  const int nrows = s->p.nrows;
  const double hx = s->p.hx, hy = s->p.hy;
  const double *x = s->f[OR_X], *y = s->f[OR_Y];
  for (int cell = 0; cell < s->nc; ++cell) {
    const int r = cell % nrows, c = cell / nrows;
    for (int jp = s->cstart[cell]; jp < s->cstart[cell + 1]; ++jp) {
      const int ip = s->cptcls[jp];
      for (int inode = 0; inode < 4; ++inode) {
        const double Nx = shg(x[ip], y[ip], 0, inode, c, r, hx, hy);
        const double Ny = shg(x[ip], y[ip], 1, inode, c, r, hx, hy);
        gv[gt(inode, c, r, nrows)] += (Nx * a[ip] + Ny * b[ip]) * w[ip];
      }
    }
  }
}

/* g2p restatement: p = 0 + sum_inode N*g (the destinations are zeroed by
 * the caller first, optim_par.cpp:534-537). */
static double g2p_one(const or_sim *s, int ip, const double *gv) {
  const int nrows = s->p.nrows, ci = s->p2g[ip];
  const int r = ci % nrows, c = ci / nrows;
  double acc = 0.0;
  for (int inode = 0; inode < 4; ++inode)
    acc += shp(s->f[OR_X][ip], s->f[OR_Y][ip], inode, c, r, s->p.hx,
               s->p.hy) *
           gv[gt(inode, c, r, nrows)];
  return acc;
}

static void g2pd_one(const or_sim *s, int ip, const double *gv, double *dx,
                     double *dy) {
  const int nrows = s->p.nrows, ci = s->p2g[ip];
  const int r = ci % nrows, c = ci / nrows;
  double ax = 0.0, ay = 0.0;
  for (int inode = 0; inode < 4; ++inode) {
    const int n = gt(inode, c, r, nrows);
    ax += shg(s->f[OR_X][ip], s->f[OR_Y][ip], 0, inode, c, r, s->p.hx,
              s->p.hy) * gv[n];
    ay += shg(s->f[OR_X][ip], s->f[OR_Y][ip], 1, inode, c, r, s->p.hx,
              s->p.hy) * gv[n];
  }
  *dx = ax;
  *dy = ay;
}

static void sim_bin(or_sim *s) {
  or_bin_particles(s->np, s->f[OR_X], s->f[OR_Y], s->p.nrows, s->p.ncols,
                   s->p.hx, s->p.hy, s->p2g, s->cstart, s->cptcls, s->ccidx,
                   s->coloff);
}

or_sim *or_sim_create(const or_params *p, int np, const double *x,
                      const double *y, const double *Mp, const double *Ap,
                      const double *vpx, const double *vpy,
                      const double *mom_px, const double *mom_py,
                      const double *hp, const double *Vp, const double *Z) {
  or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
  s->p = *p;
  s->np = np;
  s->nn = (p->nrows + 1) * (p->ncols + 1);
  s->nc = p->nrows * p->ncols;
  sim_reserve(s, np > 0 ? np : 1);
  for (int k = 0; k < OR_NFIELDS; ++k) memset(s->f[k], 0, sizeof(double) * np);
  memcpy(s->f[OR_X], x, sizeof(double) * np);
  memcpy(s->f[OR_Y], y, sizeof(double) * np);
  memcpy(s->f[OR_MP], Mp, sizeof(double) * np);
  memcpy(s->f[OR_AP], Ap, sizeof(double) * np);
  memcpy(s->f[OR_VPX], vpx, sizeof(double) * np);
  memcpy(s->f[OR_VPY], vpy, sizeof(double) * np);
  memcpy(s->f[OR_MOM_PX], mom_px, sizeof(double) * np);
  memcpy(s->f[OR_MOM_PY], mom_py, sizeof(double) * np);
  memcpy(s->f[OR_HP], hp, sizeof(double) * np);
  memcpy(s->f[OR_VP], Vp, sizeof(double) * np);
  memcpy(s->f[OR_XP], x, sizeof(double) * np);
  memcpy(s->f[OR_YP], y, sizeof(double) * np);
  for (int i = 0; i < np; ++i) { s->label[i] = i; s->level[i] = 0; }
  s->Z = (double *)malloc(sizeof(double) * s->nn);
  memcpy(s->Z, Z, sizeof(double) * s->nn);
  for (int k = 0; k < NV_COUNT; ++k)
    s->nv[k] = (double *)calloc(s->nn, sizeof(double));
  s->cstart = (int *)malloc(sizeof(int) * (s->nc + 1));
  s->ccidx = (int *)malloc(sizeof(int) * (s->nc > 0 ? s->nc : 1));

  /* step-0 initialisation, optim_par.cpp:235-264: Zp = g2p(Z) on the
   * initial mesh; H = eq_level - Zp where positive (only if eq_level > 0);
   * hydrostatic F11 = F22 = -1/2 rho g (hp - H^2/hp). */
  sim_bin(s);
  for (int ip = 0; ip < np; ++ip) {
    const double zp = g2p_one(s, ip, s->Z);
    s->f[OR_ZP][ip] = zp;
    s->f[OR_HPZ][ip] = s->f[OR_HP][ip] + zp;
    const double diff = p->eq_level - zp;
    s->f[OR_H][ip] = (p->eq_level > 0 && diff > 0) ? diff : 0.0;
  }
  for (int ip = 0; ip < np; ++ip) {
    const double h = s->f[OR_HP][ip], H = s->f[OR_H][ip];
    const double h_corr = h > 1e-10 ? h - (H * H) / h : 0.0;
    s->f[OR_F11][ip] = -.5 * p->rho * p->g * h_corr;
    s->f[OR_F12][ip] = 0.0;
    s->f[OR_F21][ip] = 0.0;
    s->f[OR_F22][ip] = -.5 * p->rho * p->g * h_corr;
  }
  s->it = 0;
  s->t = 0.0;
  s->dt = 1.0e-3; /* gpu_offload.cpp:255 (D6) */
  return s;
}

void or_sim_destroy(or_sim *s) {
  if (!s) return;
  for (int k = 0; k < OR_NFIELDS; ++k) free(s->f[k]);
  for (int k = 0; k < NV_COUNT; ++k) free(s->nv[k]);
  free(s->label); free(s->level); free(s->Z); free(s->p2g); free(s->cstart);
  free(s->cptcls); free(s->ccidx);
  free(s);
}

int or_sim_np(const or_sim *s) { return s->np; }
int or_sim_it(const or_sim *s) { return s->it; }
double or_sim_t(const or_sim *s) { return s->t; }
double or_sim_dt(const or_sim *s) { return s->dt; }
void or_sim_conservation(const or_sim *s, double out[3]) {
  out[0] = s->cons[0]; out[1] = s->cons[1]; out[2] = s->cons[2];
}
void or_sim_ms_counts(const or_sim *s, int out[2]) {
  out[0] = s->ms_counts[0]; out[1] = s->ms_counts[1];
}

int or_sim_get(const or_sim *s, const char *name, double *out) {
  for (int k = 0; k < OR_NFIELDS; ++k)
    if (!strcmp(name, FIELD_NAMES[k])) {
      memcpy(out, s->f[k], sizeof(double) * s->np);
      return s->np;
    }
  for (int k = 0; k < NV_COUNT; ++k)
    if (!strcmp(name, NODAL_NAMES[k])) {
      memcpy(out, s->nv[k], sizeof(double) * s->nn);
      return s->nn;
    }
  if (!strcmp(name, "Z")) {
    memcpy(out, s->Z, sizeof(double) * s->nn);
    return s->nn;
  }
  return -1;
}

int or_sim_get_int(const or_sim *s, const char *name, int *out) {
  const int *src = NULL;
  if (!strcmp(name, "label")) src = s->label;
  else if (!strcmp(name, "level")) src = s->level;
  else if (!strcmp(name, "p2g")) src = s->p2g;
  else if (!strcmp(name, "cptcls")) src = s->cptcls;
  else if (!strcmp(name, "cstart")) {
    memcpy(out, s->cstart, sizeof(int) * (s->nc + 1));
    return s->nc + 1;
  }
  if (!src) return -1;
  memcpy(out, src, sizeof(int) * s->np);
  return s->np;
}

int or_sim_set(or_sim *s, const char *name, const double *in) {
  for (int k = 0; k < OR_NFIELDS; ++k)
    if (!strcmp(name, FIELD_NAMES[k])) {
      memcpy(s->f[k], in, sizeof(double) * s->np);
      return s->np;
    }
  return -1;
}

/* ---- merge/split (merge_split_ops_cmes.h) ---- */

/* principal_stretch_dir, merge_split_ops_cmes.h:93-110 */
static void principal_stretch_dir(double exx, double eyy, double exy,
                                  double *ux, double *uy) {
  const double tr = exx + eyy;
  const double diff = 0.5 * (exx - eyy);
  const double disc = sqrt(diff * diff + exy * exy);
  const double l1 = 0.5 * tr + disc;
  const double vx = l1 - eyy, vy = exy;
  const double n = sqrt(vx * vx + vy * vy);
  if (n < 1e-12) { *ux = 1.0; *uy = 0.0; }
  else { *ux = vx / n; *uy = vy / n; }
}

/* decide_actions, merge_split_ops_cmes.h:116-253, with the cells of the
 * given CSR (ascending cell = grd_to_ptcl's map order). */
static void ms_flags_cmes(const or_params *p, int np, int nc, double dt,
                          const double *x, const double *y, const double *Ap,
                          const double *hp, const double *exx_,
                          const double *exy1, const double *exy2,
                          const double *eyy_, const double *vpx,
                          const double *vpy, const int *lvl, int *act,
                          int *partner) {
  const double min_h = p->hx < p->hy ? p->hx : p->hy;
  const double r_split = p->ms_alpha * min_h;
  const double r_merge = p->ms_beta * min_h;
  for (int ip = 0; ip < np; ++ip) {
    act[ip] = 0;
    partner[ip] = -1;
    const double r_i = sqrt(Ap[ip]);
    const double exx = exx_[ip], eyy = eyy_[ip];
    const double exy = 0.5 * (exy1[ip] + exy2[ip]);
    const double gdot = sqrt(2.0 * (exx * exx + eyy * eyy + 2.0 * exy * exy));
    const int stretched = r_i > r_split;
    const int sheared = gdot * dt > p->ms_shear_split;
    const int thick = r_i < r_merge;
    const int thin_skip = hp[ip] < p->ms_split_hp_min;
    if ((stretched || sheared) && !thin_skip && lvl[ip] > p->ms_min_level)
      act[ip] = 1;
    else if (thick && hp[ip] > p->ms_hp_min && lvl[ip] < p->ms_max_level)
      act[ip] = 2;
  }
}

/* The per-cell part of decide_actions (merge_split_ops_cmes.h:176-249 and,
 * with split_w = 4 and min_ppc = 2, merge_split_ops.h:405-497): over the
 * cells in ascending order, each cell's particles in ascending id order
 * (the quadgrid order assumed throughout): the global split cap, greedy
 * nearest-partner pairing of the merge candidates under the global merge
 * cap, candidates left unpaired kept, then merges undone (first primary in
 * id order) until the cell keeps min_ppc particles.  act holds the flags on
 * entry (1 split, 2 merge candidate) and the actions on exit. */
static void ms_pair_csr(const or_params *p, int np, int nc, const double *x,
                        const double *y, const double *vpx, const double *vpy,
                        const int *cstart, const int *cptcls, int *act,
                        int *partner, int counts[2], int split_w,
                        int min_ppc) {
  const int max_ops = (int)p->ms_max_ops; /* gpu_offload.cpp:317 */
  for (int ip = 0; ip < np; ++ip) partner[ip] = -1;
  const double max_dv2 = p->ms_max_dv * p->ms_max_dv;
  int n_split = 0, n_merge = 0;
  int *cand = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  char *paired = (char *)malloc(np > 0 ? np : 1);
  for (int cell = 0; cell < nc; ++cell) {
    const int b = cstart[cell], e = cstart[cell + 1];
    if (b == e) continue;
    for (int j = b; j < e; ++j) {
      const int q = cptcls[j];
      if (act[q] == 1) {
        if (n_split < max_ops) ++n_split;
        else act[q] = 0;
      }
    }
    int nc_ = 0;
    for (int j = b; j < e; ++j)
      if (act[cptcls[j]] == 2) cand[nc_++] = cptcls[j];
    for (int i = 0; i < nc_; ++i) paired[i] = 0;
    for (int i = 0; i < nc_ && n_merge < max_ops; ++i) {
      if (paired[i]) continue;
      const int i1 = cand[i];
      double best = 1.7976931348623157e308; /* DBL_MAX */
      int bj = nc_;
      for (int j = i + 1; j < nc_; ++j) {
        if (paired[j]) continue;
        const int i2 = cand[j];
        const double dx = x[i1] - x[i2], dy = y[i1] - y[i2];
        const double d2 = dx * dx + dy * dy;
        const double dvx = vpx[i1] - vpx[i2], dvy = vpy[i1] - vpy[i2];
        if (d2 < best && (dvx * dvx + dvy * dvy) < max_dv2) {
          best = d2;
          bj = j;
        }
      }
      if (bj < nc_) {
        act[i1] = 2;
        act[cand[bj]] = 3;
        partner[i1] = cand[bj];
        paired[i] = 1;
        paired[bj] = 1;
        ++n_merge;
      } else {
        act[i1] = 0;
      }
    }
    for (int i = 0; i < nc_; ++i)
      if (!paired[i]) act[cand[i]] = 0;
    int projected = 0;
    for (int j = b; j < e; ++j) {
      const int a = act[cptcls[j]];
      projected += a == 0 ? 1 : a == 1 ? split_w : a == 2 ? 1 : 0;
    }
    while (projected < min_ppc) {
      int undone = 0;
      for (int j = b; j < e; ++j) {
        const int q = cptcls[j];
        if (act[q] == 2) {
          const int qq = partner[q];
          if (qq >= 0) act[qq] = 0;
          partner[q] = -1;
          act[q] = 0;
          --n_merge;
          projected += 1;
          undone = 1;
          break;
        }
      }
      if (!undone) break;
    }
  }
  free(cand);
  free(paired);
  counts[0] = n_split;
  counts[1] = n_merge;
}

void or_ms_decide(const or_params *p, int np, double dt, const double *x,
                  const double *y, const double *Ap, const double *hp,
                  const double *vpx_dx, const double *vpx_dy,
                  const double *vpy_dx, const double *vpy_dy,
                  const double *vpx, const double *vpy, const int *level,
                  int *action, int *partner, int counts[2]) {
  const int nc = p->nrows * p->ncols;
  int *p2g = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  int *cptcls = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  int *cstart = (int *)malloc(sizeof(int) * (nc + 1));
  int *ccidx = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  int coloff[5];
  or_bin_particles(np, x, y, p->nrows, p->ncols, p->hx, p->hy, p2g, cstart,
                   cptcls, ccidx, coloff);
  ms_flags_cmes(p, np, nc, dt, x, y, Ap, hp, vpx_dx, vpx_dy, vpy_dx, vpy_dy,
                vpx, vpy, level, action, partner);
  ms_pair_csr(p, np, nc, x, y, vpx, vpy, cstart, cptcls, action, partner,
              counts, 2, p->ms_min_ppc);
  free(p2g); free(cptcls); free(cstart); free(ccidx);
}

/* execute_merge_split, merge_split_ops_cmes.h:258-427: exclusive scan of
 * weights {1,2,1,0}, then new SoA in input order. */
/* execute_merge_split (merge_split_ops_cmes.h:258-427) from arrays: f and
 * nf hold the OR_NFIELDS fields in FIELD_NAMES order, nf/nlabel/nlevel room
 * for the new count (np + splits); returns the new count. */
static int ms_execute_core(const or_params *p, int N, const int *act,
                           const int *partner, double *const *f,
                           const int *label, const int *level, double **nf,
                           int *nlabel, int *nlevel) {
  static const int weights[4] = {1, 2, 1, 0};
  int *offset = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
  int Nnew = 0;
  for (int i = 0; i < N; ++i) {
    offset[i] = Nnew;
    Nnew += weights[act[i]];
  }
  const double hx = p->hx, hy = p->hy;
  const double Lx = p->ncols * hx, Ly = p->nrows * hy, eps = 1e-10;
#define CLAMPX(v) ((v) < eps ? eps : ((v) > Lx - eps ? Lx - eps : (v)))
#define CLAMPY(v) ((v) < eps ? eps : ((v) > Ly - eps ? Ly - eps : (v)))
  for (int i = 0; i < N; ++i) {
    const int a = act[i];
    if (a == 0) {
      const int o = offset[i];
      for (int k = 0; k < OR_NFIELDS; ++k) nf[k][o] = f[k][i];
      nlabel[o] = label[i];
      nlevel[o] = level[i];
    } else if (a == 1) {
      const int o1 = offset[i], o2 = offset[i] + 1;
      double ux, uy;
      const double exx = f[OR_VPX_DX][i], eyy = f[OR_VPY_DY][i];
      const double exy = 0.5 * (f[OR_VPX_DY][i] + f[OR_VPY_DX][i]);
      principal_stretch_dir(exx, eyy, exy, &ux, &uy);
      const double r_i = sqrt(f[OR_AP][i]);
      const double off = 0.25 * r_i;
      const double cx = f[OR_X][i], cy = f[OR_Y][i];
      const double nx1 = CLAMPX(cx + off * ux), ny1 = CLAMPY(cy + off * uy);
      const double nx2 = CLAMPX(cx - off * ux), ny2 = CLAMPY(cy - off * uy);
      for (int k = 0; k < OR_NFIELDS; ++k) {
        if (k == OR_X || k == OR_Y) continue;
        const double val = f[k][i] / (is_extensive(k) ? 2.0 : 1.0);
        nf[k][o1] = val;
        nf[k][o2] = val;
      }
      const double h = nf[OR_HP][o1];
      nf[OR_AP][o1] = nf[OR_VP][o1] / h;
      nf[OR_AP][o2] = nf[OR_VP][o2] / h;
      nf[OR_X][o1] = nx1; nf[OR_Y][o1] = ny1;
      nf[OR_X][o2] = nx2; nf[OR_Y][o2] = ny2;
      nf[OR_XP][o1] = nx1; nf[OR_XP][o2] = nx2;
      nf[OR_YP][o1] = ny1; nf[OR_YP][o2] = ny2;
      nf[OR_MOM_PX][o1] = nf[OR_MP][o1] * nf[OR_VPX][o1];
      nf[OR_MOM_PX][o2] = nf[OR_MP][o2] * nf[OR_VPX][o2];
      nf[OR_MOM_PY][o1] = nf[OR_MP][o1] * nf[OR_VPY][o1];
      nf[OR_MOM_PY][o2] = nf[OR_MP][o2] * nf[OR_VPY][o2];
      nlevel[o1] = nlevel[o2] = level[i] - 1;
      nlabel[o1] = nlabel[o2] = -1;
    } else if (a == 2) {
      const int j = partner[i], o = offset[i];
      const double M1 = f[OR_MP][i], M2 = f[OR_MP][j], Mt = M1 + M2;
      const double nx = (M1 * f[OR_X][i] + M2 * f[OR_X][j]) / Mt;
      const double ny = (M1 * f[OR_Y][i] + M2 * f[OR_Y][j]) / Mt;
      for (int k = 0; k < OR_NFIELDS; ++k) {
        if (k == OR_X || k == OR_Y) continue;
        nf[k][o] = is_extensive(k) ? (f[k][i] + f[k][j])
                                   : (M1 * f[k][i] + M2 * f[k][j]) / Mt;
      }
      nf[OR_HP][o] = nf[OR_VP][o] / nf[OR_AP][o];
      nf[OR_X][o] = nx; nf[OR_Y][o] = ny;
      nf[OR_XP][o] = nx; nf[OR_YP][o] = ny;
      nf[OR_MOM_PX][o] = nf[OR_MP][o] * nf[OR_VPX][o];
      nf[OR_MOM_PY][o] = nf[OR_MP][o] * nf[OR_VPY][o];
      nlevel[o] = level[i] + 1;
      nlabel[o] = -1;
    }
  }
#undef CLAMPX
#undef CLAMPY
  free(offset);
  return Nnew;
}

static int ms_new_count(int N, const int *act) {
  static const int weights[4] = {1, 2, 1, 0};
  int Nnew = 0;
  for (int i = 0; i < N; ++i) Nnew += weights[act[i]];
  return Nnew;
}

static void ms_execute(or_sim *s, const int *act, const int *partner) {
  const int N = s->np, Nnew = ms_new_count(N, act);
  double *nf[OR_NFIELDS];
  for (int k = 0; k < OR_NFIELDS; ++k)
    nf[k] = (double *)malloc(sizeof(double) * (Nnew > 0 ? Nnew : 1));
  int *nlabel = (int *)malloc(sizeof(int) * (Nnew > 0 ? Nnew : 1));
  int *nlevel = (int *)malloc(sizeof(int) * (Nnew > 0 ? Nnew : 1));
  ms_execute_core(&s->p, N, act, partner, s->f, s->label, s->level, nf,
                  nlabel, nlevel);
  sim_reserve(s, Nnew > 0 ? Nnew : 1);
  for (int k = 0; k < OR_NFIELDS; ++k) {
    memcpy(s->f[k], nf[k], sizeof(double) * Nnew);
    free(nf[k]);
  }
  memcpy(s->label, nlabel, sizeof(int) * Nnew);
  memcpy(s->level, nlevel, sizeof(int) * Nnew);
  free(nlabel); free(nlevel);
  s->np = Nnew;
}

int or_ms_execute(const or_params *p, int np, const int *act,
                  const int *partner, double *const *in, const int *label,
                  const int *level, double **out, int *nlabel, int *nlevel) {
  return ms_execute_core(p, np, act, partner, in, label, level, out, nlabel,
                         nlevel);
}

int or_field_count(void) { return OR_NFIELDS; }

/* ---- legacy ELFS merge/split (merge_split_ops.h:86-755), ms_mode 1 ---- */

/* mark_exterior_cells, merge_split_ops.h:86-165: the empty cells reachable
 * from the grid edge's empty cells through empty cells (4-neighbours). */
static void elfs_exterior(int nrows, int ncols, const int *cstart, char *ext,
                          int *queue) {
  const int nc = nrows * ncols;
  int head = 0, tail = 0;
  for (int i = 0; i < nc; ++i) ext[i] = 0;
#define WET(i) (cstart[(i) + 1] > cstart[i])
  for (int c = 0; c < ncols; ++c) {
    const int rs[2] = {0, nrows - 1};
    for (int k = 0; k < 2; ++k) {
      const int idx = rs[k] + nrows * c;
      if (!WET(idx) && !ext[idx]) { ext[idx] = 1; queue[tail++] = idx; }
    }
  }
  for (int r = 0; r < nrows; ++r) {
    const int cs[2] = {0, ncols - 1};
    for (int k = 0; k < 2; ++k) {
      const int idx = r + nrows * cs[k];
      if (!WET(idx) && !ext[idx]) { ext[idx] = 1; queue[tail++] = idx; }
    }
  }
  static const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
  while (head < tail) {
    const int cidx = queue[head++];
    const int r = cidx % nrows, c = cidx / nrows;
    for (int k = 0; k < 4; ++k) {
      const int nr = r + dr[k], ncl = c + dc[k];
      if (nr < 0 || nr >= nrows || ncl < 0 || ncl >= ncols) continue;
      const int n = nr + nrows * ncl;
      if (!WET(n) && !ext[n]) { ext[n] = 1; queue[tail++] = n; }
    }
  }
}

/* compute_boundary_distance, merge_split_ops.h:170-262: hop count (times
 * min(hx, hy)) from the wet cells off the domain edge with an exterior
 * cell to their left or right, through wet cells; DBL_MAX elsewhere. */
static void elfs_distance(int nrows, int ncols, double h, const int *cstart,
                          const char *ext, double *dist, int *qc, int *qd) {
  const int nc = nrows * ncols;
  int head = 0, tail = 0;
  for (int i = 0; i < nc; ++i) dist[i] = 1.7976931348623157e308; /* DBL_MAX */
  for (int c = 0; c < ncols; ++c)
    for (int r = 0; r < nrows; ++r) {
      const int idx = r + nrows * c;
      if (!WET(idx)) continue;
      const int edge = r == 0 || r == nrows - 1 || c == 0 || c == ncols - 1;
      int on = 0;
      if (!edge) {
        if (c > 0 && ext[idx - nrows]) on = 1;
        if (c < ncols - 1 && ext[idx + nrows]) on = 1;
      }
      if (on) { dist[idx] = 0.0; qc[tail] = idx; qd[tail] = 0; ++tail; }
    }
  static const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
  while (head < tail) {
    const int cidx = qc[head], d = qd[head];
    ++head;
    const int r = cidx % nrows, c = cidx / nrows;
    for (int k = 0; k < 4; ++k) {
      const int nr = r + dr[k], ncl = c + dc[k];
      if (nr < 0 || nr >= nrows || ncl < 0 || ncl >= ncols) continue;
      const int n = nr + nrows * ncl;
      if (!WET(n)) continue;
      const double nd = (d + 1) * h;
      if (nd < dist[n]) { dist[n] = nd; qc[tail] = n; qd[tail] = d + 1; ++tail; }
    }
  }
#undef WET
}

/* compute_elfs, merge_split_ops.h:266-342: the cell distance, refined near
 * the boundary by the distance from the particle to the exterior cells of
 * its row within two columns. */
static double elfs_one(const or_params *p, int cell, double px, double py,
                       const double *dist, const char *ext) {
  const int nrows = p->nrows, ncols = p->ncols;
  const double hx = p->hx, hy = p->hy;
  const double thr = hx > hy ? hx : hy;
  const double base = dist[cell];
  if (base > thr) return base;
  double mind = base;
  const int r0 = cell % nrows, c0 = cell / nrows;
  for (int dcl = -2; dcl <= 2; ++dcl) {
    const int r = r0, c = c0 + dcl;
    if (r < 0 || r >= nrows || c < 0 || c >= ncols) continue;
    const int n = r + nrows * c;
    if (!ext[n]) continue;
    const double cx_lo = c * hx, cx_hi = (c + 1) * hx;
    const double cy_lo = r * hy, cy_hi = (r + 1) * hy;
    const double nx = px < cx_lo ? cx_lo : (px > cx_hi ? cx_hi : px);
    const double ny = py < cy_lo ? cy_lo : (py > cy_hi ? cy_hi : py);
    const double dx = px - nx, dy = py - ny;
    const double d = sqrt(dx * dx + dy * dy);
    if (d < mind) mind = d;
  }
  return mind;
}

/* The two cell passes for a given occupancy (cstart), exported for tests. */
void or_elfs_cells(int nrows, int ncols, double h, const int *cstart,
                   char *ext, double *dist) {
  const int nc = nrows * ncols;
  int *q1 = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  int *q2 = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  elfs_exterior(nrows, ncols, cstart, ext, q1);
  elfs_distance(nrows, ncols, h, cstart, ext, dist, q1, q2);
  free(q1); free(q2);
}

/* decide_actions, merge_split_ops.h:349-510 (flags :388-402, then the
 * per-cell pass with a 1 -> 4 split and at least 2 particles per cell), on
 * arrays binned by p2g / cstart / cptcls. */
static void elfs_decide_core(const or_params *p, int np, const double *x,
                             const double *y, const double *Ap,
                             const double *hp, const double *vpx_dx,
                             const double *vpy_dy, const double *vpx,
                             const double *vpy, const int *level,
                             const int *p2g, const int *cstart,
                             const int *cptcls, int *act, int *partner,
                             int counts[2]) {
  const int nc = p->nrows * p->ncols;
  char *ext = (char *)malloc(nc > 0 ? nc : 1);
  double *dist = (double *)malloc(sizeof(double) * (nc > 0 ? nc : 1));
  int *q1 = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  int *q2 = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  elfs_exterior(p->nrows, p->ncols, cstart, ext, q1);
  elfs_distance(p->nrows, p->ncols, p->hx < p->hy ? p->hx : p->hy, cstart,
                ext, dist, q1, q2);
  for (int ip = 0; ip < np; ++ip) {
    const double e = elfs_one(p, p2g[ip], x[ip], y[ip], dist, ext);
    const double r_i = sqrt(Ap[ip]);
    const double divergence = vpx_dx[ip] + vpy_dy[ip];
    act[ip] = 0;
    if ((e < p->ms_alpha * r_i) || (divergence > p->ms_stretch)) {
      if (level[ip] > p->ms_min_level) act[ip] = 1;
    } else if (e > p->ms_beta * r_i && hp[ip] > p->ms_hp_min) {
      if (level[ip] < p->ms_max_level) act[ip] = 2;
    }
  }
  ms_pair_csr(p, np, nc, x, y, vpx, vpy, cstart, cptcls, act, partner, counts,
              4, 2);
  free(ext); free(dist); free(q1); free(q2);
}

static void elfs_decide(or_sim *s, int *act, int *partner, int counts[2]) {
  double **f = s->f;
  elfs_decide_core(&s->p, s->np, f[OR_X], f[OR_Y], f[OR_AP], f[OR_HP],
                   f[OR_VPX_DX], f[OR_VPY_DY], f[OR_VPX], f[OR_VPY], s->level,
                   s->p2g, s->cstart, s->cptcls, act, partner, counts);
}

void or_elfs_decide(const or_params *p, int np, const double *x,
                    const double *y, const double *Ap, const double *hp,
                    const double *vpx_dx, const double *vpy_dy,
                    const double *vpx, const double *vpy, const int *level,
                    int *action, int *partner, int counts[2]) {
  const int nc = p->nrows * p->ncols;
  int *p2g = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  int *cptcls = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  int *cstart = (int *)malloc(sizeof(int) * (nc + 1));
  int *ccidx = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
  int coloff[5];
  or_bin_particles(np, x, y, p->nrows, p->ncols, p->hx, p->hy, p2g, cstart,
                   cptcls, ccidx, coloff);
  elfs_decide_core(p, np, x, y, Ap, hp, vpx_dx, vpy_dy, vpx, vpy, level, p2g,
                   cstart, cptcls, action, partner, counts);
  free(p2g); free(cptcls); free(cstart); free(ccidx);
}

/* execute_merge_split, merge_split_ops.h:513-716: weights {1,4,1,0}; a
 * split leaves four children at (x +- r/2, y +- r/2), r = sqrt(Ap), clamped
 * to the domain, extensive fields quartered, level - 1, label -1; a merge
 * averages the intensive fields by mass, sums the extensive ones, keeps
 * the primary's label, level + 1; momentum = Mp * v for both.  Arrays as
 * ms_execute_core; returns the new count. */
static int elfs_execute_core(const or_params *p, int N, const int *act,
                             const int *partner, double *const *f,
                             const int *label, const int *level, double **nf,
                             int *nlabel, int *nlevel) {
  static const int weights[4] = {1, 4, 1, 0};
  int *offset = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
  int Nnew = 0;
  for (int i = 0; i < N; ++i) {
    offset[i] = Nnew;
    Nnew += weights[act[i]];
  }
  const double Lx = p->ncols * p->hx, Ly = p->nrows * p->hy;
  for (int i = 0; i < N; ++i) {
    const int a = act[i];
    if (a == 0) {
      const int o = offset[i];
      for (int k = 0; k < OR_NFIELDS; ++k) nf[k][o] = f[k][i];
      nlabel[o] = label[i];
      nlevel[o] = level[i];
    } else if (a == 1) {
      const int o = offset[i];
      const double px = f[OR_X][i], py = f[OR_Y][i];
      const double off = sqrt(f[OR_AP][i]) / 2.0;
      double cx[4] = {px + off, px + off, px - off, px - off};
      double cy[4] = {py + off, py - off, py + off, py - off};
      for (int k = 0; k < 4; ++k) {
        if (cx[k] < 1e-10 || cx[k] > Lx - 1e-10) {
          const double m = Lx - 1e-10 < cx[k] ? Lx - 1e-10 : cx[k];
          cx[k] = 1e-10 > m ? 1e-10 : m;
        }
        if (cy[k] < 1e-10 || cy[k] > Ly - 1e-10) {
          const double m = Ly - 1e-10 < cy[k] ? Ly - 1e-10 : cy[k];
          cy[k] = 1e-10 > m ? 1e-10 : m;
        }
      }
      for (int c = 0; c < 4; ++c) {
        for (int k = 0; k < OR_NFIELDS; ++k) {
          if (k == OR_X || k == OR_Y) continue;
          nf[k][o + c] = f[k][i] / (is_extensive(k) ? 4.0 : 1.0);
        }
        nf[OR_X][o + c] = cx[c];
        nf[OR_Y][o + c] = cy[c];
        nf[OR_XP][o + c] = cx[c];
        nf[OR_YP][o + c] = cy[c];
        nf[OR_MOM_PX][o + c] = nf[OR_MP][o + c] * nf[OR_VPX][o + c];
        nf[OR_MOM_PY][o + c] = nf[OR_MP][o + c] * nf[OR_VPY][o + c];
        nlevel[o + c] = level[i] - 1;
        nlabel[o + c] = -1;
      }
    } else if (a == 2) {
      const int j = partner[i], o = offset[i];
      const double M1 = f[OR_MP][i], M2 = f[OR_MP][j], Mt = M1 + M2;
      const double nx = (M1 * f[OR_X][i] + M2 * f[OR_X][j]) / Mt;
      const double ny = (M1 * f[OR_Y][i] + M2 * f[OR_Y][j]) / Mt;
      for (int k = 0; k < OR_NFIELDS; ++k) {
        if (k == OR_X || k == OR_Y) continue;
        nf[k][o] = is_extensive(k) ? (f[k][i] + f[k][j])
                                   : (M1 * f[k][i] + M2 * f[k][j]) / Mt;
      }
      nf[OR_X][o] = nx; nf[OR_Y][o] = ny;
      nf[OR_XP][o] = nx; nf[OR_YP][o] = ny;
      nf[OR_MOM_PX][o] = nf[OR_MP][o] * nf[OR_VPX][o];
      nf[OR_MOM_PY][o] = nf[OR_MP][o] * nf[OR_VPY][o];
      nlabel[o] = label[i];
      nlevel[o] = level[i] + 1;
    }
  }
  free(offset);
  return Nnew;
}

static void elfs_execute(or_sim *s, const int *act, const int *partner) {
  static const int weights[4] = {1, 4, 1, 0};
  const int N = s->np;
  int Nnew = 0;
  for (int i = 0; i < N; ++i) Nnew += weights[act[i]];
  double *nf[OR_NFIELDS];
  for (int k = 0; k < OR_NFIELDS; ++k)
    nf[k] = (double *)malloc(sizeof(double) * (Nnew > 0 ? Nnew : 1));
  int *nlabel = (int *)malloc(sizeof(int) * (Nnew > 0 ? Nnew : 1));
  int *nlevel = (int *)malloc(sizeof(int) * (Nnew > 0 ? Nnew : 1));
  elfs_execute_core(&s->p, N, act, partner, s->f, s->label, s->level, nf,
                    nlabel, nlevel);
  sim_reserve(s, Nnew > 0 ? Nnew : 1);
  for (int k = 0; k < OR_NFIELDS; ++k) {
    memcpy(s->f[k], nf[k], sizeof(double) * Nnew);
    free(nf[k]);
  }
  memcpy(s->label, nlabel, sizeof(int) * Nnew);
  memcpy(s->level, nlevel, sizeof(int) * Nnew);
  free(nlabel); free(nlevel);
  s->np = Nnew;
}

int or_elfs_execute(const or_params *p, int np, const int *act,
                    const int *partner, double *const *in, const int *label,
                    const int *level, double **out, int *nlabel, int *nlevel) {
  return elfs_execute_core(p, np, act, partner, in, label, level, out, nlabel,
                           nlevel);
}

/* stress_tensor_t::operator(), optim_par.cpp:68-120 (Bingham, with the
 * hydrostatic well-balanced term) */
void or_stress_optim_par(int np, const or_params *p, const double *vpx,
                         const double *vpy, const double *hp, const double *Hh,
                         const double *vpx_dx, const double *vpx_dy,
                         const double *vpy_dx, const double *vpy_dy,
                         double *F11, double *F12, double *F21, double *F22) {
  OR_PAR_NP
  for (int ip = 0; ip < np; ++ip) {
    const double A = 3. / 2., C = 65. / 32.;
    const double BING = p->bingham_on, mu = p->mu, tau_Y = p->tauy;
    const double vx = vpx[ip], vy = vpy[ip], h = hp[ip];
    const double nrm = sqrt(vx * vx + vy * vy);
    const double ALF = h > 1.e-3 ? (6. * mu * nrm) / ((h + 0.001) * tau_Y) : 0.0;
    const double B = -114. / 32. - ALF;
    const double Z1 = (-B + sqrt(B * B - 4. * A * C)) / (2. * A);
    const double Z2 = (-B - sqrt(B * B - 4. * A * C)) / (2. * A);
    const double ZZ = fabs(Z1 - .5) <= .5 ? Z1 : Z2;
    const double s_xx = vpx_dx[ip];
    const double s_xy = 0.5 * (vpx_dy[ip] + vpy_dx[ip]);
    const double s_yy = vpy_dy[ip];
    const double D_xx = s_xx, D_xy = s_xy, D_yy = s_yy;
    const double D_zx =
        h > 1.e-3 ? 0.5 * (3. / (2. + ZZ)) * (vx / (h + 0.001)) : 0.0;
    const double D_zy =
        h > 1.e-3 ? 0.5 * (3. / (2. + ZZ)) * (vy / (h + 0.001)) : 0.0;
    const double D_zz = -(vpx_dx[ip] + vpy_dy[ip]);
    const double invII = 0.5 * (D_xx * D_xx + D_yy * D_yy + D_zz * D_zz) +
                         D_zx * D_zx + D_zy * D_zy + D_xy * D_xy;
    const double sig_xx =
        invII != 0 ? (tau_Y / sqrt(invII) + 2. * mu) * D_xx : 0.0;
    const double sig_xy =
        invII != 0 ? (tau_Y / sqrt(invII) + 2. * mu) * D_xy : 0.0;
    const double sig_yy =
        invII != 0 ? (tau_Y / sqrt(invII) + 2. * mu) * D_yy : 0.0;
    const double H = Hh[ip];
    const double h_corr = h > 1e-10 ? h - (H * H) / h : 0.0;
    F11[ip] = BING * sig_xx - .5 * p->rho * p->g * h_corr;
    F12[ip] = BING * sig_xy;
    F21[ip] = BING * sig_xy;
    F22[ip] = BING * sig_yy - .5 * p->rho * p->g * h_corr;
  }
}

int or_sim_step(or_sim *s) {
  const or_params *p = &s->p;
  const int fixed = p->dt_fixed > 0.0;
  const int np0 = s->np;
  /* loop condition, optim_par.cpp:281 (+ optional step cap, D15) */
  if (fixed ? !(s->it < p->nsteps) : !(s->t < p->T)) return 0;
  if (!fixed && p->max_steps > 0 && s->it >= p->max_steps) return 0;
  double dt = s->dt;
  if (fixed) {
    dt = p->dt_fixed;
  } else {
    /* optim_par.cpp:288-304 */
    double maxvx = 0, minvx = 0, maxvy = 0, minvy = 0, hmax = 0;
    if (np0 > 0) {
      maxvx = minvx = s->f[OR_VPX][0];
      maxvy = minvy = s->f[OR_VPY][0];
      hmax = s->f[OR_HP][0];
    }
    for (int i = 1; i < np0; ++i) {
      const double vx = s->f[OR_VPX][i], vy = s->f[OR_VPY][i];
      if (vx > maxvx) maxvx = vx;
      if (vx < minvx) minvx = vx;
      if (vy > maxvy) maxvy = vy;
      if (vy < minvy) minvy = vy;
      if (s->f[OR_HP][i] > hmax) hmax = s->f[OR_HP][i];
    }
    const double mvx = fabs(maxvx) > fabs(minvx) ? fabs(maxvx) : fabs(minvx);
    const double mvy = fabs(maxvy) > fabs(minvy) ? fabs(maxvy) : fabs(minvy);
    const double a = sqrt(p->g * hmax) + mvx, b = sqrt(p->g * hmax) + mvy;
    const double cel = fabs(a > b ? a : b);
    const double minh = p->hx < p->hy ? p->hx : p->hy;
    if (s->it > 0) dt = p->cfl * minh / (1e-2 + cel);
  }
  s->dt = dt;
  s->it++;
  s->ms_counts[0] = s->ms_counts[1] = 0;

  /* (0) connectivity, optim_par.cpp:341 */
  sim_bin(s);

  /* merge/split at gpu_offload's position (gpu_offload.cpp:321-401), on
   * the current state (D7) */
  if (p->ms_on && s->it % p->ms_interval == 0 && s->it > 0) {
    const int N = s->np;
    int *act = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    int *partner = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    if (p->ms_mode == 1) {
      elfs_decide(s, act, partner, s->ms_counts);
    } else {
      ms_flags_cmes(p, N, s->nc, dt, s->f[OR_X], s->f[OR_Y], s->f[OR_AP],
                    s->f[OR_HP], s->f[OR_VPX_DX], s->f[OR_VPX_DY],
                    s->f[OR_VPY_DX], s->f[OR_VPY_DY], s->f[OR_VPX],
                    s->f[OR_VPY], s->level, act, partner);
      ms_pair_csr(p, N, s->nc, s->f[OR_X], s->f[OR_Y], s->f[OR_VPX],
                  s->f[OR_VPY], s->cstart, s->cptcls, act, partner,
                  s->ms_counts, 2, p->ms_min_ppc);
    }
    int any = 0;
    for (int i = 0; i < N; ++i)
      if (act[i] != 0) { any = 1; break; }
    if (any) {
      if (p->ms_mode == 1) elfs_execute(s, act, partner);
      else ms_execute(s, act, partner);
      sim_bin(s);
    }
    free(act);
    free(partner);
  }

  const int np = s->np, nn = s->nn;
  double **f = s->f;
  double **v = s->nv;
  for (int k = 0; k < NV_COUNT; ++k) memset(v[k], 0, sizeof(double) * nn);

  /* (1) P2G mass/momentum, optim_par.cpp:383 */
  p2g_field(s, f[OR_MP], v[NV_MV]);
  p2g_field(s, f[OR_MOM_PX], v[NV_MOMX]);
  p2g_field(s, f[OR_MOM_PY], v[NV_MOMY]);
  /* g2pd(Z), optim_par.cpp:373-374,387 */
  OR_PAR_NP
  for (int ip = 0; ip < np; ++ip)
    g2pd_one(s, ip, s->Z, &f[OR_DZXP][ip], &f[OR_DZYP][ip]);

  /* conservation, optim_par.cpp:395-428 */
  {
    double tm = 0, tx = 0, ty = 0, nm = 0, nx = 0, ny = 0;
    for (int ip = 0; ip < np; ++ip) {
      tm += f[OR_MP][ip];
      tx += f[OR_MOM_PX][ip];
      ty += f[OR_MOM_PY][ip];
    }
    for (int i = 0; i < nn; ++i) {
      nm += v[NV_MV][i];
      nx += v[NV_MOMX][i];
      ny += v[NV_MOMY][i];
    }
    const double err_mass = fabs(tm - nm) / tm;
    const double a = fabs(tx) > fabs(ty) ? fabs(tx) : fabs(ty);
    const double norm = 1.0 > a ? 1.0 : a;
    const double ex = fabs(tx - nx) / norm, ey = fabs(ty - ny) / norm;
    s->cons[0] = err_mass;
    s->cons[1] = ex > ey ? ex : ey;
    s->cons[2] = s->t;
  }

  /* (2) external forces, optim_par.cpp:447-470 */
  OR_PAR_NP
  for (int ip = 0; ip < np; ++ip) {
    const double wb =
        f[OR_HP][ip] > 1e-10 ? 1.0 - (f[OR_H][ip] / f[OR_HP][ip]) : 1.0;
    f[OR_FPX][ip] = -p->g * f[OR_VP][ip] * p->rho * f[OR_DZXP][ip] * wb;
    f[OR_FPY][ip] = -p->g * f[OR_VP][ip] * p->rho * f[OR_DZYP][ip] * wb;
    f[OR_FRIC_PX][ip] = f[OR_AP][ip] * f[OR_FB_X][ip];
    f[OR_FRIC_PY][ip] = f[OR_AP][ip] * f[OR_FB_Y][ip];
  }
  p2g_field(s, f[OR_FPX], v[NV_FPX]);
  p2g_field(s, f[OR_FPY], v[NV_FPY]);
  p2g_field(s, f[OR_FRIC_PX], v[NV_FRX]);
  p2g_field(s, f[OR_FRIC_PY], v[NV_FRY]);
  OR_PAR_NN
  for (int i = 0; i < nn; ++i) {
    v[NV_FEXTX][i] = v[NV_FPX][i] + v[NV_FRX][i];
    v[NV_FEXTY][i] = v[NV_FPY][i] + v[NV_FRY][i];
  }
  /* (3) internal forces and momentum balance, optim_par.cpp:476-485 */
  p2gd_field(s, f[OR_F11], f[OR_F12], f[OR_VP], v[NV_FINTX]);
  p2gd_field(s, f[OR_F21], f[OR_F22], f[OR_VP], v[NV_FINTY]);
  OR_PAR_NN
  for (int i = 0; i < nn; ++i) {
    v[NV_FTOTX][i] = v[NV_FEXTX][i] - v[NV_FINTX][i];
    v[NV_FTOTY][i] = v[NV_FEXTY][i] - v[NV_FINTY][i];
    v[NV_MOMX][i] = dt * v[NV_FTOTX][i] + v[NV_MOMX][i];
    v[NV_MOMY][i] = dt * v[NV_FTOTY][i] + v[NV_MOMY][i];
  }
  /* (4) nodal accelerations and velocities, optim_par.cpp:492-497 (D1) */
  OR_PAR_NN
  for (int i = 0; i < nn; ++i) {
    const double m = v[NV_MV][i];
    v[NV_AVX][i] = m > 1.e-2 ? v[NV_FTOTX][i] / m : 0.0;
    v[NV_AVY][i] = m > 1.e-2 ? v[NV_FTOTY][i] / m : 0.0;
    v[NV_VVX][i] = m > 1.e-2 ? v[NV_MOMX][i] / m : 0.0;
    v[NV_VVY][i] = m > 1.e-2 ? v[NV_MOMY][i] / m : 0.0;
    v[NV_VVXL][i] = dt * v[NV_AVX][i] + v[NV_VVX][i];
    v[NV_VVYL][i] = dt * v[NV_AVY][i] + v[NV_VVY][i];
  }
  /* (5) boundary conditions, optim_par.cpp:503-529: the nodes of the
   * first/last cell column lose their x components, those of the first/last
   * cell row their y components (cell edges e(2),e(3) / e(0),e(1)). */
  if (p->bc_flag) {
    const int nrows = p->nrows, ncols = p->ncols;
    for (int cell = 0; cell < s->nc; ++cell) {
      const int r = cell % nrows, c = cell / nrows;
      if (c == 0 || c == ncols - 1)
        for (int inode = 0; inode < 4; ++inode) {
          const int n = gt(inode, c, r, nrows);
          v[NV_AVX][n] = 0.0; v[NV_VVX][n] = 0.0;
          v[NV_VVXL][n] = 0.0; v[NV_MOMX][n] = 0.0;
        }
      if (r == 0 || r == nrows - 1)
        for (int inode = 0; inode < 4; ++inode) {
          const int n = gt(inode, c, r, nrows);
          v[NV_AVY][n] = 0.0; v[NV_VVY][n] = 0.0;
          v[NV_VVYL][n] = 0.0; v[NV_MOMY][n] = 0.0;
        }
    }
  }
  /* (6) G2P, PIC velocity, advection, clamp: optim_par.cpp:534-564 */
  OR_PAR_NP
  for (int ip = 0; ip < np; ++ip) {
    f[OR_VPXL][ip] = g2p_one(s, ip, v[NV_VVXL]);
    f[OR_VPYL][ip] = g2p_one(s, ip, v[NV_VVYL]);
    f[OR_APX][ip] = g2p_one(s, ip, v[NV_AVX]);
    f[OR_APY][ip] = g2p_one(s, ip, v[NV_AVY]);
  }
  {
    const double eps = 1e-10, Lx = p->ncols * p->hx, Ly = p->nrows * p->hy;
    OR_PAR_NP
    for (int ip = 0; ip < np; ++ip) {
      f[OR_VPX][ip] = f[OR_VPXL][ip];
      f[OR_VPY][ip] = f[OR_VPYL][ip];
      f[OR_X][ip] = f[OR_X][ip] + dt * f[OR_VPX][ip];
      f[OR_Y][ip] = f[OR_Y][ip] + dt * f[OR_VPY][ip];
      if (f[OR_X][ip] < eps) f[OR_X][ip] = eps;
      if (f[OR_X][ip] > Lx - eps) f[OR_X][ip] = Lx - eps;
      if (f[OR_Y][ip] < eps) f[OR_Y][ip] = eps;
      if (f[OR_Y][ip] > Ly - eps) f[OR_Y][ip] = Ly - eps;
    }
  }
  /* (7) g2pd of the updated nodal velocities at the new positions with the
   * step-start cells (D18), height update, floor, momentum, area:
   * optim_par.cpp:569-607 */
  OR_PAR_NP
  for (int ip = 0; ip < np; ++ip) {
    g2pd_one(s, ip, v[NV_VVXL], &f[OR_VPX_DX][ip], &f[OR_VPX_DY][ip]);
    g2pd_one(s, ip, v[NV_VVYL], &f[OR_VPY_DX][ip], &f[OR_VPY_DY][ip]);
    const double div = f[OR_VPX_DX][ip] + f[OR_VPY_DY][ip];
    double h = f[OR_HP][ip] / (1 + dt * div);
    h = h < 1e-2 ? 1e-2 : h;
    f[OR_HP][ip] = h;
    f[OR_MOM_PX][ip] = f[OR_VPX][ip] * f[OR_MP][ip];
    f[OR_MOM_PY][ip] = f[OR_VPY][ip] * f[OR_MP][ip];
    f[OR_AP][ip] = h > 1.e-2 ? f[OR_VP][ip] / h : 0.0;
  }
  /* (7b) Voellmy friction, optim_par.cpp:613-626 */
  {
    const double fric_ang = p->phi * M_PI / 180.;
    const double tanf = tan(fric_ang);
    OR_PAR_NP
    for (int ip = 0; ip < np; ++ip) {
      const double vx = f[OR_VPX][ip], vy = f[OR_VPY][ip], h = f[OR_HP][ip];
      const double nv = sqrt(vx * vx + vy * vy);
      if (p->friction_on > 0 && nv > 1e-10 && p->xi > 0) {
        const double v2 = vx * vx + vy * vy;
        f[OR_FB_X][ip] = -p->friction_on *
                         (p->rho * p->g * h * tanf + p->rho * p->g * v2 / p->xi) *
                         vx / nv;
        f[OR_FB_Y][ip] = -p->friction_on *
                         (p->rho * p->g * h * tanf + p->rho * p->g * v2 / p->xi) *
                         vy / nv;
      } else {
        f[OR_FB_X][ip] = 0.0;
        f[OR_FB_Y][ip] = 0.0;
      }
    }
  }
  /* (8) Bingham stress, optim_par.cpp:633-635 */
  or_stress_optim_par(np, p, f[OR_VPX], f[OR_VPY], f[OR_HP], f[OR_H],
                      f[OR_VPX_DX], f[OR_VPX_DY], f[OR_VPY_DX], f[OR_VPY_DY],
                      f[OR_F11], f[OR_F12], f[OR_F21], f[OR_F22]);
  /* g2p(Z) at the new positions, hpZ (output fields), optim_par.cpp:637-646 */
  OR_PAR_NP
  for (int ip = 0; ip < np; ++ip) {
    f[OR_ZP][ip] = g2p_one(s, ip, s->Z);
    f[OR_HPZ][ip] = f[OR_HP][ip] + f[OR_ZP][ip];
  }
  /* p2g(hp -> HV), optim_par.cpp:650 (output field) */
  p2g_field(s, f[OR_HP], v[NV_HV]);
  s->t += dt;
  return 1;
}
