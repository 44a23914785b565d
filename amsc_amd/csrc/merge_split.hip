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
#include "binning.hpp"
#include "amsc_mpm.h"

#include <cstring>
#include <string>

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

// ---- execute_merge_split (merge_split_ops_cmes.h:258-427) on the caller's
// device arrays, for callers that keep their own particle storage (the
// kernel-level drop-in; the time loop executes merge/split itself).

namespace amsc {
namespace {

constexpr int kMsMaxFields = 64;

// The caller's particle fields: in/out pointers, which are extensive
// (merge_split_ops_cmes.h:49-50), and where the fields with a role are.
struct MsExecFields {
  const double *in[kMsMaxFields];
  double *out[kMsMaxFields];
  unsigned char ext[kMsMaxFields];
  int n;
  int x, y, xp, yp, Mp, Vp, Ap, hp, vpx, vpy, momx, momy, vxdx, vxdy, vydx, vydy;
};

__global__ void ms_weight_kernel(int np, const int *__restrict__ act, int *w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int a = act[i];
  w[i] = a == MS_KEEP ? 1 : a == MS_SPLIT ? 2 : a == MS_PRIMARY ? 1 : 0;
}

__global__ void ms_exec_kernel(int np, double Lx, double Ly,
                               const int *__restrict__ act,
                               const int *__restrict__ partner,
                               const int *__restrict__ offset, MsExecFields f,
                               const int *__restrict__ label_in,
                               const int *__restrict__ level_in,
                               int *label_out, int *level_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int a = act[i], o = offset[i];
  const double eps = 1e-10;
  if (a == MS_KEEP) {
    for (int k = 0; k < f.n; ++k) f.out[k][o] = f.in[k][i];
    label_out[o] = label_in[i];
    level_out[o] = level_in[i];
  } else if (a == MS_SPLIT) {
    const int o1 = o, o2 = o + 1;
    double ux, uy;
    const double exx = f.in[f.vxdx][i], eyy = f.in[f.vydy][i];
    const double exy = 0.5 * (f.in[f.vxdy][i] + f.in[f.vydx][i]);
    principal_stretch_dir(exx, eyy, exy, &ux, &uy);
    const double off = 0.25 * sqrt(f.in[f.Ap][i]);
    const double cx = f.in[f.x][i], cy = f.in[f.y][i];
    double x1 = cx + off * ux, y1 = cy + off * uy;
    double x2 = cx - off * ux, y2 = cy - off * uy;
    x1 = x1 < eps ? eps : (x1 > Lx - eps ? Lx - eps : x1);
    y1 = y1 < eps ? eps : (y1 > Ly - eps ? Ly - eps : y1);
    x2 = x2 < eps ? eps : (x2 > Lx - eps ? Lx - eps : x2);
    y2 = y2 < eps ? eps : (y2 > Ly - eps ? Ly - eps : y2);
    for (int k = 0; k < f.n; ++k) {
      if (k == f.x || k == f.y) continue;
      const double v = f.in[k][i] / (f.ext[k] ? 2.0 : 1.0);
      f.out[k][o1] = v;
      f.out[k][o2] = v;
    }
    const double h = f.out[f.hp][o1];
    f.out[f.Ap][o1] = f.out[f.Vp][o1] / h;
    f.out[f.Ap][o2] = f.out[f.Vp][o2] / h;
    f.out[f.x][o1] = x1;
    f.out[f.y][o1] = y1;
    f.out[f.x][o2] = x2;
    f.out[f.y][o2] = y2;
    if (f.xp >= 0) {
      f.out[f.xp][o1] = x1;
      f.out[f.xp][o2] = x2;
    }
    if (f.yp >= 0) {
      f.out[f.yp][o1] = y1;
      f.out[f.yp][o2] = y2;
    }
    for (int c = 0; c < 2; ++c) {
      const int oc = o + c;
      if (f.momx >= 0) f.out[f.momx][oc] = f.out[f.Mp][oc] * f.out[f.vpx][oc];
      if (f.momy >= 0) f.out[f.momy][oc] = f.out[f.Mp][oc] * f.out[f.vpy][oc];
      label_out[oc] = -1;
      level_out[oc] = level_in[i] - 1;
    }
  } else if (a == MS_PRIMARY) {
    const int j = partner[i];
    const double M1 = f.in[f.Mp][i], M2 = f.in[f.Mp][j], Mt = M1 + M2;
    const double nx = (M1 * f.in[f.x][i] + M2 * f.in[f.x][j]) / Mt;
    const double ny = (M1 * f.in[f.y][i] + M2 * f.in[f.y][j]) / Mt;
    for (int k = 0; k < f.n; ++k) {
      if (k == f.x || k == f.y) continue;
      f.out[k][o] = f.ext[k] ? (f.in[k][i] + f.in[k][j])
                             : (M1 * f.in[k][i] + M2 * f.in[k][j]) / Mt;
    }
    f.out[f.hp][o] = f.out[f.Vp][o] / f.out[f.Ap][o];
    f.out[f.x][o] = nx;
    f.out[f.y][o] = ny;
    if (f.xp >= 0) f.out[f.xp][o] = nx;
    if (f.yp >= 0) f.out[f.yp][o] = ny;
    if (f.momx >= 0) f.out[f.momx][o] = f.out[f.Mp][o] * f.out[f.vpx][o];
    if (f.momy >= 0) f.out[f.momy][o] = f.out[f.Mp][o] * f.out[f.vpy][o];
    label_out[o] = -1;
    level_out[o] = level_in[i] + 1;
  }
}

}  // namespace
}  // namespace amsc

extern "C" int amsc_ms_execute(const amsc_params *p, int np,
                               const int *d_action, const int *d_partner,
                               int nfields, const char *const *names,
                               const double *const *d_in, double *const *d_out,
                               const int *d_label_in, const int *d_level_in,
                               int *d_label_out, int *d_level_out,
                               int capacity, int *np_out) {
  using namespace amsc;
  if (!p || np < 0 || nfields <= 0 || nfields > kMsMaxFields || !names ||
      !d_in || !d_out || !np_out || capacity < 0 ||
      (np > 0 && (!d_action || !d_partner || !d_label_in || !d_level_in ||
                  !d_label_out || !d_level_out)))
    return fail(AMSC_ERR_INVALID, "amsc_ms_execute: bad arguments");
  MsExecFields f;
  std::memset(&f, 0, sizeof(f));
  f.n = nfields;
  int *roles[] = {&f.x, &f.y, &f.xp, &f.yp, &f.Mp, &f.Vp, &f.Ap, &f.hp,
                  &f.vpx, &f.vpy, &f.momx, &f.momy, &f.vxdx, &f.vxdy,
                  &f.vydx, &f.vydy};
  const char *role_names[] = {"x", "y", "xp", "yp", "Mp", "Vp", "Ap", "hp",
                              "vpx", "vpy", "mom_px", "mom_py", "vpx_dx",
                              "vpx_dy", "vpy_dx", "vpy_dy"};
  for (int *r : roles) *r = -1;
  for (int k = 0; k < nfields; ++k) {
    if (!names[k] || !d_in[k] || !d_out[k])
      return fail(AMSC_ERR_INVALID, "amsc_ms_execute: NULL field");
    f.in[k] = d_in[k];
    f.out[k] = d_out[k];
    for (int r = 0; r < 16; ++r)
      if (!std::strcmp(names[k], role_names[r])) {
        if (*roles[r] >= 0)
          return fail(AMSC_ERR_INVALID,
                      std::string("amsc_ms_execute: field listed twice: ") +
                          names[k]);
        *roles[r] = k;
      }
    const char *e = names[k];
    f.ext[k] = !std::strcmp(e, "Mp") || !std::strcmp(e, "Vp") ||
               !std::strcmp(e, "Ap") || !std::strcmp(e, "mom_px") ||
               !std::strcmp(e, "mom_py");
  }
  for (int r = 0; r < 16; ++r)
    if (*roles[r] < 0 && r != 2 && r != 3 && r != 10 && r != 11)
      return fail(AMSC_ERR_INVALID,
                  std::string("amsc_ms_execute: required field missing: ") +
                      role_names[r]);
  if (np == 0) {
    *np_out = 0;
    return AMSC_OK;
  }
  hipStream_t st = current_stream();
  int *w = nullptr, *offs = nullptr;
  if (hipMalloc(&w, sizeof(int) * np) != hipSuccess ||
      hipMalloc(&offs, sizeof(int) * (np + 1)) != hipSuccess) {
    if (w) (void)hipFree(w);
    return fail(AMSC_ERR_NOMEM, "amsc_ms_execute: workspace allocation");
  }
  ScanWork scan;
  int nnew = 0;
  int rc = AMSC_OK;
  ms_weight_kernel<<<grid_for(np, 256), 256, 0, st>>>(np, d_action, w);
  if (hipGetLastError() != hipSuccess)
    rc = fail(AMSC_ERR_HIP, "amsc_ms_execute: launch");
  if (!rc) rc = exclusive_scan(w, offs, np, scan, st, &nnew);
  if (!rc && nnew > capacity)
    rc = fail(AMSC_ERR_CAPACITY, "amsc_ms_execute: " + std::to_string(nnew) +
                                     " particles after merge/split, capacity " +
                                     std::to_string(capacity));
  if (!rc) {
    ms_exec_kernel<<<grid_for(np, 256), 256, 0, st>>>(
        np, p->ncols * p->hx, p->nrows * p->hy, d_action, d_partner, offs, f,
        d_label_in, d_level_in, d_label_out, d_level_out);
    if (hipGetLastError() != hipSuccess)
      rc = fail(AMSC_ERR_HIP, "amsc_ms_execute: launch");
  }
  const hipError_t e = hipStreamSynchronize(st);
  if (!rc && e != hipSuccess) rc = fail(AMSC_ERR_HIP, hipGetErrorString(e));
  (void)hipFree(w);
  (void)hipFree(offs);
  if (scan.tmp) (void)hipFree(scan.tmp);
  if (!rc) *np_out = nnew;
  return rc;
}
