// merge_split.hpp -- device merge/split decisions (merge_split_ops_cmes.h).
#pragma once

#include "common.hpp"
#include "amsc_mpm.h"

namespace amsc {

enum : int {
  MS_KEEP = 0,
  MS_SPLIT = 1,
  MS_PRIMARY = 2,
  MS_SECONDARY = 3,
  MS_CANDIDATE = 4  // transient: merge candidate not yet paired
};

struct MsParams {
  double r_split, r_merge, shear_split, split_hp_min, hp_min, max_dv2;
  int min_level, max_level, min_ppc, max_ops;
};

struct MsInputs {
  const double *x, *y, *Ap, *hp, *vxdx, *vxdy, *vydx, *vydy, *vpx, *vpy;
  const int *level;
};

// ms_config from the solver parameters (gpu_offload.cpp:308-319).
MsParams ms_params_from(const amsc_params &p);

// Decide actions for np particles binned in (cstart, cptcls) over nc
// cells.  dt is read from dt_dev when non-null, else dt_host.  counts (host
// int[2]) receives splits and merges.  Synchronises the stream.
int ms_decide(const MsParams &m, int np, int nc, const double *dt_dev,
              double dt_host, const MsInputs &in, const int *cstart,
              const int *cptcls, int *act, int *partner, int *cnt_dev,
              int *counts, hipStream_t st);

__device__ __forceinline__ void principal_stretch_dir(double exx, double eyy,
                                                      double exy, double *ux,
                                                      double *uy) {
  // merge_split_ops_cmes.h:93-110
  const double tr = exx + eyy;
  const double diff = 0.5 * (exx - eyy);
  const double disc = sqrt(diff * diff + exy * exy);
  const double l1 = 0.5 * tr + disc;
  const double vx = l1 - eyy, vy = exy;
  const double n = sqrt(vx * vx + vy * vy);
  if (n < 1e-12) {
    *ux = 1.0;
    *uy = 0.0;
  } else {
    *ux = vx / n;
    *uy = vy / n;
  }
}

}  // namespace amsc
