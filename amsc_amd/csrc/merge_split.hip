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
  const double r_i = sqrt(in.Ap[ip]);
  const double exx = in.vxdx[ip], eyy = in.vydy[ip];
  const double exy = 0.5 * (in.vxdy[ip] + in.vydx[ip]);
  const double gdot = sqrt(2.0 * (exx * exx + eyy * eyy + 2.0 * exy * exy));
  const bool stretched = r_i > m.r_split;
  const bool sheared = gdot * dt > m.shear_split;
  const bool thick = r_i < m.r_merge;
  const bool thin_skip = in.hp[ip] < m.split_hp_min;
  const int lvl = in.level[ip];
  int a = MS_KEEP;
  if ((stretched || sheared) && !thin_skip && lvl > m.min_level)
    a = MS_SPLIT;
  else if (thick && in.hp[ip] > m.hp_min && lvl < m.max_level)
    a = MS_CANDIDATE;
  act[ip] = a;
  partner[ip] = -1;
}

// Greedy pairing of one cell's candidates, ascending id; returns the number
// of pairs formed.  `limit` stops pairing once *n_merge reaches it (serial
// replay); the parallel kernel passes a null counter.
__device__ int pair_cell(int b, int e, const int *cptcls, const MsInputs &in,
                         double max_dv2, int *act, int *partner,
                         int *n_merge, int limit) {
  int formed = 0;
  for (int i = b; i < e; ++i) {
    if (n_merge && *n_merge >= limit) break;
    const int q = cptcls[i];
    if (act[q] != MS_CANDIDATE) continue;
    double best = 1.7976931348623157e308;  // DBL_MAX
    int bj = -1;
    const double xq = in.x[q], yq = in.y[q], vxq = in.vpx[q], vyq = in.vpy[q];
    for (int j = i + 1; j < e; ++j) {
      const int q2 = cptcls[j];
      if (act[q2] != MS_CANDIDATE) continue;
      const double dx = xq - in.x[q2], dy = yq - in.y[q2];
      const double d2 = dx * dx + dy * dy;
      const double dvx = vxq - in.vpx[q2], dvy = vyq - in.vpy[q2];
      if (d2 < best && (dvx * dvx + dvy * dvy) < max_dv2) {
        best = d2;
        bj = j;
      }
    }
    if (bj >= 0) {
      const int q2 = cptcls[bj];
      act[q] = MS_PRIMARY;
      act[q2] = MS_SECONDARY;
      partner[q] = q2;
      ++formed;
      if (n_merge) ++*n_merge;
    } else {
      act[q] = MS_KEEP;
    }
  }
  // candidates never paired (or left over when the cap stopped the loop)
  for (int i = b; i < e; ++i)
    if (act[cptcls[i]] == MS_CANDIDATE) act[cptcls[i]] = MS_KEEP;
  return formed;
}

// Undo merges (first primary in id order) until the cell keeps min_ppc
// particles; returns the number undone.
__device__ int undo_cell(int b, int e, const int *cptcls, int min_ppc,
                         int *act, int *partner) {
  int projected = 0;
  for (int i = b; i < e; ++i) {
    const int a = act[cptcls[i]];
    projected += a == MS_KEEP ? 1 : a == MS_SPLIT ? 2 : a == MS_PRIMARY ? 1 : 0;
  }
  int undone = 0;
  while (projected < min_ppc) {
    bool did = false;
    for (int i = b; i < e; ++i) {
      const int q = cptcls[i];
      if (act[q] == MS_PRIMARY) {
        const int q2 = partner[q];
        if (q2 >= 0) act[q2] = MS_KEEP;
        partner[q] = -1;
        act[q] = MS_KEEP;
        ++undone;
        projected += 1;
        did = true;
        break;
      }
    }
    if (!did) break;
  }
  return undone;
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
  const int undone = undo_cell(b, e, cptcls, m.min_ppc, act, partner);
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
  int n_split = 0, n_merge = 0;
  for (int cell = 0; cell < nc; ++cell) {
    const int b = cstart[cell], e = cstart[cell + 1];
    if (b == e) continue;
    for (int i = b; i < e; ++i) {
      const int q = cptcls[i];
      if (act[q] == MS_SPLIT) {
        if (n_split < m.max_ops) ++n_split;
        else act[q] = MS_KEEP;
      }
    }
    pair_cell(b, e, cptcls, in, m.max_dv2, act, partner, &n_merge, m.max_ops);
    n_merge -= undo_cell(b, e, cptcls, m.min_ppc, act, partner);
  }
  cnt[0] = n_split;
  cnt[1] = n_merge;
  cnt[2] = n_merge;
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
  const MsParams m = ms_params_from(*p);
  MsInputs in = {d_x,      d_y,      d_Ap,  d_hp,  d_vpx_dx, d_vpx_dy,
                 d_vpy_dx, d_vpy_dy, d_vpx, d_vpy, d_level};
  int *cnt = nullptr;
  if (hipMalloc(&cnt, sizeof(int) * 4) != hipSuccess)
    return fail(AMSC_ERR_NOMEM, "amsc_ms_decide: counter allocation");
  const int rc = ms_decide(m, np, p->nrows * p->ncols, nullptr, dt, in,
                           d_cstart, d_cptcls, d_action, d_partner, cnt,
                           counts, current_stream());
  (void)hipFree(cnt);
  return rc;
}
