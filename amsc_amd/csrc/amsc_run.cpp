// amsc_run.cpp -- command-line driver: the time loop of main/gpu_offload.cpp
// (reads ./DATA.json or the given path, runs to the end condition, prints
// the reference's PERF line, gpu_offload.cpp:580-586) on the device-resident
// path of libamsc_mpm.
//
//   amsc_run [DATA.json] [--max-steps N] [--capacity N]
#include "amsc_mpm.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char **argv) {
  const char *path = "DATA.json";
  int max_steps = 0, capacity = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--max-steps") && i + 1 < argc)
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
  int done = 0;
  if (amsc_sim_step(sim, WARMUP, &done)) {
    std::fprintf(stderr, "amsc_run: %s\n", amsc_last_error());
    return 1;
  }
  const auto t0 = std::chrono::high_resolution_clock::now();
  int counted = 0;
  while (true) {
    int n = 0;
    if (amsc_sim_step(sim, 100, &n)) {
      std::fprintf(stderr, "amsc_run: %s\n", amsc_last_error());
      return 1;
    }
    counted += n;
    if (n < 100) break;
  }
  const auto t1 = std::chrono::high_resolution_clock::now();
  const double sec = std::chrono::duration<double>(t1 - t0).count();
  int np = 0, it = 0;
  double t = 0, dt = 0, cons[3] = {0, 0, 0};
  amsc_sim_num_particles(sim, &np);
  amsc_sim_state(sim, &it, &t, &dt);
  amsc_sim_conservation(sim, cons);
  std::printf("steps=%d t=%.6g dt=%.6g err_mass=%.3e err_mom=%.3e\n", it, t,
              dt, cons[0], cons[1]);
  std::printf("PERF np=%d counted_steps=%d tot_s=%.6f per_step_ms=%.6f\n", np,
              counted, sec, counted > 0 ? 1e3 * sec / counted : 0.0);
  amsc_sim_destroy(sim);
  return 0;
}
