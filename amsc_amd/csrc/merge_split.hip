// merge_split.hip -- adaptive merge/split decisions on the device
// (decide_actions, main/merge_split_ops_cmes.h:116-253, which marks on the
// device and then pairs on the host over a std::map, serially).
//
//   flag kernel   one thread per particle: SPLIT / merge candidate / KEEP;
//   pair kernel   one thread per cell: the cell's particles in ascending id
//                 (cptcls order = grd_to_ptcl order), greedy nearest
//                 co-moving partner, then the min-particles-per-cell undo.
// With ms_max_ops not binding (the default 1e9, mpm_data.h:112) cells are
// independent, so the parallel result equals the serial one.  When the
// totals show the cap could bind, a single-thread kernel replays the exact
// serial loop (global split/merge counters in ascending cell order).
#include "merge_split.hpp"
#include "amsc_mpm.h"

namespace amsc {

namespace {

__global__ void ms_flag_kernel(int np, MsParams m, const double *dt_dev,
                               double dt_host, MsInputs in, int *act,
                               int *partner) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const double dt = dt_dev ? *dt_dev : dt_host;
  act[ip] = ms_flag_one(m, dt, in, ip);
  partner[ip] = -1;
}

__global__ void ms_pair_kernel(int nc, MsParams m, MsInputs in,
                               const int *__restrict__ cstart,
                               const int *__restrict__ cptcls, int *act,
                               int *partner, int *cnt) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nc) return;
  const int b = cstart[cell], e = cstart[cell + 1];
  if (b == e) return;
  int nsplit = 0;
  for (int i = b; i < e; ++i) nsplit += act[cptcls[i]] == MS_SPLIT;
  const int formed =
      pair_cell(b, e, cptcls, in, m.max_dv2, act, partner, nullptr, 0);
  const int undone = undo_cell(b, e, cptcls, m.min_ppc, 2, act, partner);
  if (nsplit) atomicAdd(&cnt[0], nsplit);
  if (formed) atomicAdd(&cnt[1], formed);
  if (formed - undone) atomicAdd(&cnt[2], formed - undone);
}

// Exact serial replay with the global caps (merge_split_ops_cmes.h:176-249).
__global__ void ms_serial_kernel(int nc, MsParams m, MsInputs in,
                                 const int *__restrict__ cstart,
                                 const int *__restrict__ cptcls, int *act,
                                 int *partner, int *cnt) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  ms_serial_cells(0, nc, m, in, cstart, cptcls, act, partner, cnt);
}

}  // namespace

int ms_decide(const MsParams &m, int np, int nc, const double *dt_dev,
              double dt_host, const MsInputs &in, const int *cstart,
              const int *cptcls, int *act, int *partner, int *cnt_dev,
              int *counts, hipStream_t st) {
  counts[0] = counts[1] = 0;
  if (np <= 0) return AMSC_OK;
  AMSC_HIP(hipMemsetAsync(cnt_dev, 0, sizeof(int) * 4, st));
  ms_flag_kernel<<<grid_for(np, 256), 256, 0, st>>>(np, m, dt_dev, dt_host,
                                                    in, act, partner);
  ms_pair_kernel<<<grid_for(nc, 256), 256, 0, st>>>(nc, m, in, cstart,
                                                    cptcls, act, partner,
                                                    cnt_dev);
  AMSC_LAUNCH_CHECK();
  int h[4];
  AMSC_HIP(hipMemcpyAsync(h, cnt_dev, sizeof(h), hipMemcpyDeviceToHost, st));
  AMSC_HIP(hipStreamSynchronize(st));
  if (h[0] > m.max_ops || h[1] > m.max_ops) {
    // the cap binds somewhere: replay the serial loop from fresh flags
    ms_flag_kernel<<<grid_for(np, 256), 256, 0, st>>>(np, m, dt_dev, dt_host,
                                                      in, act, partner);
    ms_serial_kernel<<<1, 1, 0, st>>>(nc, m, in, cstart, cptcls, act,
                                      partner, cnt_dev);
    AMSC_LAUNCH_CHECK();
    AMSC_HIP(hipMemcpyAsync(h, cnt_dev, sizeof(h), hipMemcpyDeviceToHost, st));
    AMSC_HIP(hipStreamSynchronize(st));
  }
  counts[0] = h[0];
  counts[1] = h[2];
  return AMSC_OK;
}

MsParams ms_params_from(const amsc_params &p) {
  MsParams m;
  const double min_h = p.hx < p.hy ? p.hx : p.hy;
  m.r_split = p.ms_alpha * min_h;
  m.r_merge = p.ms_beta * min_h;
  m.shear_split = p.ms_shear_split;
  m.split_hp_min = p.ms_split_hp_min;
  m.hp_min = p.ms_hp_min;
  m.max_dv2 = p.ms_max_dv * p.ms_max_dv;
  m.min_level = p.ms_min_level;
  m.max_level = p.ms_max_level;
  m.min_ppc = p.ms_min_ppc;
  // (int)data.ms_max_ops, gpu_offload.cpp:317 (saturated instead of UB)
  m.max_ops = p.ms_max_ops >= 2147483647.0 ? 2147483647 : (int)p.ms_max_ops;
  m.mode = p.ms_mode;
  m.alpha = p.ms_alpha;
  m.beta = p.ms_beta;
  m.stretch = p.ms_stretch;
  return m;
}

}  // namespace amsc

using namespace amsc;

extern "C" int amsc_ms_decide(const amsc_params *p, int np, double dt,
                              const double *d_x, const double *d_y,
                              const double *d_Ap, const double *d_hp,
                              const double *d_vpx_dx, const double *d_vpx_dy,
                              const double *d_vpy_dx, const double *d_vpy_dy,
                              const double *d_vpx, const double *d_vpy,
                              const int *d_level, const int *d_cstart,
                              const int *d_cptcls, int *d_action,
                              int *d_partner, int *counts) {
  if (!p || np < 0 || p->nrows <= 0 || p->ncols <= 0 || !counts)
    return fail(AMSC_ERR_INVALID, "amsc_ms_decide: bad arguments");
  if (p->ms_mode != 0)
    return fail(AMSC_ERR_INVALID,
                "amsc_ms_decide implements decide_actions of "
                "merge_split_ops_cmes.h (ms_mode 0); the legacy ELFS rule "
                "runs in the time loop");
  const MsParams m = ms_params_from(*p);
  MsInputs in;
  in.x = d_x;
  in.y = d_y;
  in.hp = d_hp;
  in.Ap = d_Ap;
  in.vxdx = d_vpx_dx;
  in.vxdy = d_vpx_dy;
  in.vydx = d_vpy_dx;
  in.vydy = d_vpy_dy;
  in.vpx = d_vpx;
  in.vpy = d_vpy;
  in.level = d_level;
  int *cnt = nullptr;
  if (hipMalloc(&cnt, sizeof(int) * 4) != hipSuccess)
    return fail(AMSC_ERR_NOMEM, "amsc_ms_decide: counter allocation");
  const int rc = ms_decide(m, np, p->nrows * p->ncols, nullptr, dt, in,
                           d_cstart, d_cptcls, d_action, d_partner, cnt,
                           counts, current_stream());
  (void)hipFree(cnt);
  return rc;
}
