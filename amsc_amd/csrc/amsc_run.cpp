// amsc_run.cpp -- command-line driver: the time loop of main/gpu_offload.cpp
// (reads ./DATA.json or the given path, runs to the end condition, prints
// the reference's PERF line, gpu_offload.cpp:580-586) on the device-resident
// path of libamsc_mpm.
//
//   amsc_run [DATA.json] [--max-steps N] [--capacity N] [--output]
//
// --output writes the reference solver's files into the working directory
// (optim_par.cpp:269-275,311-337,431,655-665): GRID_forZ.vts before the
// loop, conservation_errors.csv (time,err_mass,err_mom every step, time =
// start of the step), nc_particles_<it>.csv at the start of every step
// it % 10 == 0 and nc_grid_<it>.vts after it.  The timed loop then steps
// one step at a time and includes the writes, as the reference's does.
#include "amsc_mpm.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

int fail_out(const char *what) {
  std::fprintf(stderr, "amsc_run: %s: %s\n", what, amsc_last_error());
  return 1;
}

// One step with the reference's outputs around it; *done = steps taken (0
// once the run has ended).
int output_step(amsc_sim *sim, int it, FILE *cons, int *done) {
  if (it % 10 == 0) {
    const std::string fn = "nc_particles_" + std::to_string(it) + ".csv";
    if (amsc_sim_write_particles_csv(sim, fn.c_str())) return fail_out("csv");
  }
  if (amsc_sim_step(sim, 1, done)) return fail_out("step");
  if (*done == 0) return 0;
  double c[3];
  if (amsc_sim_conservation(sim, c)) return fail_out("conservation");
  std::fprintf(cons, "%g,%g,%g\n", c[2], c[0], c[1]);
  if (it % 10 == 0) {
    const std::string fn = "nc_grid_" + std::to_string(it) + ".vts";
    if (amsc_sim_write_grid_vts(sim, fn.c_str())) return fail_out("vts");
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const char *path = "DATA.json";
  int max_steps = 0, capacity = 0;
  bool output = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--output"))
      output = true;
    else if (!std::strcmp(argv[i], "--max-steps") && i + 1 < argc)
      max_steps = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--capacity") && i + 1 < argc)
      capacity = std::atoi(argv[++i]);
    else
      path = argv[i];
  }
  amsc_sim *sim = nullptr;
  if (amsc_sim_create_from_json(path, max_steps, capacity, &sim)) {
    std::fprintf(stderr, "amsc_run: %s\n", amsc_last_error());
    return 1;
  }
  const int WARMUP = 5;  // gpu_offload.cpp:252
  FILE *cons = nullptr;
  if (output) {
    if (amsc_sim_set_aux(sim, 1)) return fail_out("aux");
    if (amsc_sim_write_grid_vts(sim, "GRID_forZ.vts")) return fail_out("vts");
    cons = std::fopen("conservation_errors.csv", "w");
    if (!cons) return fail_out("conservation_errors.csv");
    std::fprintf(cons, "time,err_mass,err_mom\n");
  }
  int done = 0, it = 0;
  if (output) {
    for (int k = 0; k < WARMUP; ++k) {
      if (output_step(sim, it + 1, cons, &done)) return 1;
      it += done;
      if (!done) break;
    }
  } else if (amsc_sim_step(sim, WARMUP, &done)) {
    return fail_out("step");
  }
  const auto t0 = std::chrono::high_resolution_clock::now();
  int counted = 0;
  while (true) {
    int n = 0;
    if (output) {
      if (output_step(sim, it + 1, cons, &n)) return 1;
      it += n;
      counted += n;
      if (n < 1) break;
      continue;
    }
    if (amsc_sim_step(sim, 100, &n)) return fail_out("step");
    counted += n;
    if (n < 100) break;
  }
  if (cons) std::fclose(cons);
  const auto t1 = std::chrono::high_resolution_clock::now();
  const double sec = std::chrono::duration<double>(t1 - t0).count();
  int np = 0, steps = 0, diverged = 0;
  double t = 0, dt = 0, cerr[3] = {0, 0, 0};
  amsc_sim_num_particles(sim, &np);
  amsc_sim_state(sim, &steps, &t, &dt);
  amsc_sim_conservation(sim, cerr);
  amsc_ens_diverged(sim, &diverged);
  std::printf("steps=%d t=%.6g dt=%.6g err_mass=%.3e err_mom=%.3e%s\n", steps,
              t, dt, cerr[0], cerr[1],
              diverged ? " (stopped: particle speed above 1e4 m/s)" : "");
  std::printf("PERF np=%d counted_steps=%d tot_s=%.6f per_step_ms=%.6f\n", np,
              counted, sec, counted > 0 ? 1e3 * sec / counted : 0.0);
  amsc_sim_destroy(sim);
  return 0;
}
