// offload_loop.cpp -- the time loop of the reference's gpu_offload driver
// (main/gpu_offload.cpp:276-575, merge/split off) written against
// the kernel-level drop-in include/gpu_kernels.h: the same eighteen
// gpu_kernels:: calls in the same order (K1-K10), with the host CSR rebuild
// after each step (init_particle_mesh, gpu_offload.cpp:551-562) replaced by
// amsc_bin_particles and the adaptive dt (host max_element over copies of
// the particle arrays, :282-301) by amsc_cfl_dt.  It is the worked example of keeping gpu_offload's
// loop and swapping its kernels, and a test driver: state in and out as
// binary files.
//
//   amsc_offload_loop <in.bin> <out.bin> [--async]
//
// --async: gpu_kernels::set_synchronous(false), the calls only enqueue their
// kernels (the default waits after each call, as the reference's target
// regions do); the results are the same bits.
//
// in.bin:  int32  np nrows ncols nsteps bc_flag
//          float64 hx hy dt g rho mu tau_Y cc A_coeff C_coeff tan_fa xi fric_on
//                  cfl  (dt > 0: fixed dt; dt <= 0: adaptive, CFL = cfl)
//          float64 particle arrays (np each, kParticle order), Z (nn)
// out.bin: float64 particle arrays (kParticle order), nodal arrays
//          (kNodal order), then int32 p2g (np), cstart (nc+1), cptcls (np)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpu_kernels.h"

namespace {

enum P {
  X, Y, MP, VP, AP, HP, H, VPX, VPY, MOMX, MOMY, FBX, FBY, F11, F12, F21, F22,
  FRPX, FRPY, DZX, DZY, FPX, FPY, APX, APY, VPXL, VPYL, VXDX, VXDY, VYDX, VYDY,
  ZP, HPZ, NP_FIELDS
};
enum N {
  MV, MOMVX, MOMVY, FEXTX, FEXTY, FINTX, FINTY, FTOTX, FTOTY, AVX, AVY, VVX,
  VVY, VVXL, VVYL, FRX, FRY, FPXV, FPYV, HV, PVRHO, PVVX, PVVY, PVAX, PVAY,
  NN_FIELDS
};

[[noreturn]] void die(const char *what) {
  std::fprintf(stderr, "amsc_offload_loop: %s (%s)\n", what, amsc_last_error());
  std::exit(1);
}

void ok(int rc, const char *what) {
  if (rc != AMSC_OK) die(what);
}

template <typename T> T *dev(size_t n) {
  void *p = nullptr;
  ok(amsc_device_alloc(sizeof(T) * (n > 0 ? n : 1), &p), "device allocation");
  return static_cast<T *>(p);
}

template <typename T> void rd(FILE *f, T *v, size_t n) {
  if (n && std::fread(v, sizeof(T), n, f) != n) die("short input file");
}

}  // namespace

int main(int argc, char **argv) {
  const bool async = argc == 4 && std::string(argv[3]) == "--async";
  if (argc != 3 && !async) {
    std::fprintf(stderr,
                 "usage: amsc_offload_loop <in.bin> <out.bin> [--async]\n");
    return 2;
  }
  gpu_kernels::set_synchronous(!async);
  FILE *in = std::fopen(argv[1], "rb");
  if (!in) die("cannot open the input file");
  int hdr[5];
  rd(in, hdr, 5);
  const int np = hdr[0], nrows = hdr[1], ncols = hdr[2], nsteps = hdr[3];
  const bool bc_flag = hdr[4] != 0;
  double s[14];
  rd(in, s, 14);
  const double hx = s[0], hy = s[1], dt_fixed = s[2], g = s[3], rho = s[4],
               mu = s[5], tau_Y = s[6], cc = s[7], A_coeff = s[8],
               C_coeff = s[9], tan_fa = s[10], xi = s[11], fric_on = s[12],
               cfl = s[13];
  const bool fixed = dt_fixed > 0.0;
  double dt = fixed ? dt_fixed : 1.0e-3;  // gpu_offload.cpp:255
  const int nn = (nrows + 1) * (ncols + 1), nc = nrows * ncols;
  std::vector<double> host((size_t)np);
  double *p[NP_FIELDS], *v[NN_FIELDS];
  for (int k = 0; k < NP_FIELDS; ++k) {
    p[k] = dev<double>(np);
    rd(in, host.data(), np);
    ok(amsc_copy_to_device(p[k], host.data(), sizeof(double) * np), "upload");
  }
  double *Z = dev<double>(nn);
  std::vector<double> hz((size_t)nn);
  rd(in, hz.data(), nn);
  std::fclose(in);
  ok(amsc_copy_to_device(Z, hz.data(), sizeof(double) * nn), "upload Z");
  for (int k = 0; k < NN_FIELDS; ++k) v[k] = dev<double>(nn);
  int *p2g = dev<int>(np), *cstart = dev<int>(nc + 1), *cptcls = dev<int>(np),
      *ccidx = dev<int>(nc);
  int coloff[5];
  ok(amsc_bin_particles(np, nrows, ncols, hx, hy, p[X], p[Y], p2g, cstart,
                        cptcls, ccidx, coloff),
     "binning");

  namespace gk = gpu_kernels;
  ok(amsc_synchronize(), "synchronize");
  // the reference's timing: the clock starts at step WARMUP
  // (gpu_offload.cpp:252,278-279) and counts the steps after it (:582)
  const int WARMUP = 5;
  auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < nsteps; ++it) {
    if (it == WARMUP) {
      ok(amsc_synchronize(), "synchronize");
      t0 = std::chrono::steady_clock::now();
    }
    // dt, gpu_offload.cpp:282-301 (it = steps taken so far)
    if (!fixed)
      ok(amsc_cfl_dt(np, p[VPX], p[VPY], p[HP], g, cfl, hx, hy, it, &dt,
                     nullptr),
         "cfl");
    // K1, gpu_offload.cpp:410-422
    gk::reset_nodal_arrays(nn, v[MV], v[MOMVX], v[MOMVY], v[FEXTX], v[FEXTY],
                           v[FINTX], v[FINTY], v[FTOTX], v[FTOTY], v[AVX],
                           v[AVY], v[VVX], v[VVY], v[VVXL], v[VVYL], v[FRX],
                           v[FRY], v[FPXV], v[FPYV], v[HV], v[PVRHO], v[PVVX],
                           v[PVVY], v[PVAX], v[PVAY]);
    gk::scatter_mass_momentum(np, nrows, hx, hy, cstart, cptcls, ccidx, coloff,
                              p[X], p[Y], p[MP], p[MOMX], p[MOMY], v[MV],
                              v[MOMVX], v[MOMVY]);
    // K2-K5, :427-462
    gk::gradZ_to_particles(np, nrows, hx, hy, p[X], p[Y], p2g, Z, p[DZX],
                           p[DZY]);
    gk::gravity_force(np, g, p[HP], p[H], p[MP], p[DZX], p[DZY], p[FPX],
                      p[FPY]);
    gk::friction_particle_force(np, p[AP], p[FBX], p[FBY], p[FRPX], p[FRPY]);
    gk::scatter_gravity_force(np, nrows, hx, hy, cstart, cptcls, ccidx, coloff,
                              p[X], p[Y], p[FPX], p[FPY], v[FPXV], v[FPYV]);
    gk::scatter_friction_force(np, nrows, hx, hy, cstart, cptcls, ccidx,
                               coloff, p[X], p[Y], p[FRPX], p[FRPY], v[FRX],
                               v[FRY]);
    gk::compute_F_ext(nn, v[FPXV], v[FPYV], v[FRX], v[FRY], v[FEXTX],
                      v[FEXTY]);
    gk::scatter_internal_forces(np, nrows, hx, hy, cstart, cptcls, ccidx,
                                coloff, p[X], p[Y], p[VP], p[F11], p[F12],
                                p[F21], p[F22], v[FINTX], v[FINTY]);
    // K6-K9, :476-522
    gk::nodal_force_and_momentum(nn, dt, v[FEXTX], v[FEXTY], v[FINTX],
                                 v[FINTY], v[FTOTX], v[FTOTY], v[MOMVX],
                                 v[MOMVY]);
    gk::nodal_velocity_and_bc(nn, nrows, ncols, dt, bc_flag, v[MV], v[FTOTX],
                              v[FTOTY], v[MOMVX], v[MOMVY], v[AVX], v[AVY],
                              v[VVX], v[VVY], v[VVXL], v[VVYL]);
    gk::g2p_velocity_and_advect(np, nrows, ncols, hx, hy, dt, p2g, v[VVXL],
                                v[VVYL], v[AVX], v[AVY], p[X], p[Y], p[VPX],
                                p[VPY], p[APX], p[APY], p[MOMX], p[MOMY],
                                p[VPXL], p[VPYL], p[MP]);
    double min_sc = 1e30, min_hp_g8 = 1e30;
    gk::g2pd_gradients_and_height_update(
        np, nrows, hx, hy, dt, p[X], p[Y], p2g, v[VVXL], v[VVYL], p[VXDX],
        p[VXDY], p[VYDX], p[VYDY], p[HP], p[VP], p[AP], &min_sc, &min_hp_g8);
    gk::voellmy_friction(np, rho, g, tan_fa, xi, fric_on, p[VPX], p[VPY],
                         p[HP], p[FBX], p[FBY]);
    double d[7] = {0, 0, 0, 0, 0, 0, 0};
    gk::bingham_stress_update(np, rho, g, mu, tau_Y, cc, A_coeff, C_coeff,
                              p[VPX], p[VPY], p[HP], p[H], p[VXDX], p[VXDY],
                              p[VYDX], p[VYDY], p[F11], p[F12], p[F21], p[F22],
                              &d[0], &d[1], &d[2], &d[3], &d[4], &d[5], &d[6]);
    // K10, :530-541
    gk::g2p_height(np, nrows, hx, hy, p[X], p[Y], p2g, Z, p[HP], p[ZP],
                   p[HPZ]);
    gk::scatter_plotvars(np, nrows, hx, hy, cstart, cptcls, ccidx, coloff,
                         p[X], p[Y], p[MP], p[VPX], p[VPY], p[APX], p[APY],
                         p[HP], v[PVRHO], v[PVVX], v[PVVY], v[PVAX], v[PVAY],
                         v[HV]);
    gk::normalize_plotvars(nn, v[PVRHO], v[PVVX], v[PVVY], v[PVAX], v[PVAY]);
    // the CSR of the new positions, :551-562
    ok(amsc_bin_particles(np, nrows, ncols, hx, hy, p[X], p[Y], p2g, cstart,
                          cptcls, ccidx, coloff),
       "binning");
  }

  ok(amsc_synchronize(), "synchronize");
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  FILE *out = std::fopen(argv[2], "wb");
  if (!out) die("cannot open the output file");
  for (int k = 0; k < NP_FIELDS; ++k) {
    ok(amsc_copy_to_host(host.data(), p[k], sizeof(double) * np), "download");
    std::fwrite(host.data(), sizeof(double), np, out);
  }
  for (int k = 0; k < NN_FIELDS; ++k) {
    ok(amsc_copy_to_host(hz.data(), v[k], sizeof(double) * nn), "download");
    std::fwrite(hz.data(), sizeof(double), nn, out);
  }
  std::vector<int> hi((size_t)(np > nc + 1 ? np : nc + 1));
  ok(amsc_copy_to_host(hi.data(), p2g, sizeof(int) * np), "download");
  std::fwrite(hi.data(), sizeof(int), np, out);
  ok(amsc_copy_to_host(hi.data(), cstart, sizeof(int) * (nc + 1)), "download");
  std::fwrite(hi.data(), sizeof(int), nc + 1, out);
  ok(amsc_copy_to_host(hi.data(), cptcls, sizeof(int) * np), "download");
  std::fwrite(hi.data(), sizeof(int), np, out);
  const bool bad = std::ferror(out);
  std::fclose(out);
  for (int k = 0; k < NP_FIELDS; ++k) amsc_device_free(p[k]);
  for (int k = 0; k < NN_FIELDS; ++k) amsc_device_free(v[k]);
  amsc_device_free(Z);
  amsc_device_free(p2g);
  amsc_device_free(cstart);
  amsc_device_free(cptcls);
  amsc_device_free(ccidx);
  if (bad) die("write failed");
  // the reference driver's PERF line (gpu_offload.cpp:580-586): the steps
  // after the warm-up (all of them, clock from the start, for runs of at
  // most WARMUP steps)
  const int counted = nsteps > WARMUP ? nsteps - WARMUP : nsteps;
  std::printf("PERF np=%d counted_steps=%d tot_s=%.6f per_step_ms=%.4f\n", np,
              counted, secs, counted > 0 ? 1e3 * secs / counted : 0.0);
  return 0;
}
