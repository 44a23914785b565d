// merge_split.hpp -- device merge/split decisions (merge_split_ops_cmes.h,
// and the legacy ELFS rule of merge_split_ops.h in the time loop): the
// per-particle flag, the per-cell greedy pairing and the
// min-particles-per-cell undo, shared by the standalone decision entry point
// and the fused time loop.  Templated on the storage type R of the velocity,
// gradient and area fields; the comparisons are evaluated in double.
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
  // rule: 0 merge_split_ops_cmes.h, 1 the legacy ELFS rule of
  // merge_split_ops.h (split where ELFS < alpha sqrt(Ap) or the divergence
  // exceeds stretch, merge where ELFS > beta sqrt(Ap); a split makes four
  // particles; at least two particles per cell)
  int mode;
  double alpha, beta, stretch;
};

__host__ __device__ __forceinline__ int split_weight(const MsParams &m) {
  return m.mode == 1 ? 4 : 2;
}
__host__ __device__ __forceinline__ int cell_min_ppc(const MsParams &m) {
  return m.mode == 1 ? 2 : m.min_ppc;  // merge_split_ops.h:478 fixes 2
}

template <typename R> struct MsIn {
  const double *x, *y, *hp;
  const R *Ap, *vxdx, *vxdy, *vydx, *vydy, *vpx, *vpy;
  const int *level;
};
using MsInputs = MsIn<double>;

// ms_config from the solver parameters (gpu_offload.cpp:308-319).
MsParams ms_params_from(const amsc_params &p);

// Decide actions for np particles binned in (cstart, cptcls) over nc
// cells.  dt is read from dt_dev when non-null, else dt_host.  counts (host
// int[2]) receives splits and merges.  Synchronises the stream.
int ms_decide(const MsParams &m, int np, int nc, const double *dt_dev,
              double dt_host, const MsInputs &in, const int *cstart,
              const int *cptcls, int *act, int *partner, int *cnt_dev,
              int *counts, hipStream_t st);

// Action flag of one particle, merge_split_ops_cmes.h:143-172, in two
// parts: the tests on its area and velocity gradients (as stored) and dt,
// as bits (MS_BIT_STRAIN: stretched or sheared; MS_BIT_THICK: area below
// the merge radius), then the tests on its depth and level.
enum : int { MS_BIT_STRAIN = 1, MS_BIT_THICK = 2 };
__device__ __forceinline__ int ms_strain_bits(const MsParams &m, double dt,
                                              double Ap, double vxdx,
                                              double vxdy, double vydx,
                                              double vydy) {
  const double r_i = sqrt(Ap);
  const double exx = vxdx, eyy = vydy;
  const double exy = 0.5 * (vxdy + vydx);
  const double gdot = sqrt(2.0 * (exx * exx + eyy * eyy + 2.0 * exy * exy));
  const bool stretched = r_i > m.r_split;
  const bool sheared = gdot * dt > m.shear_split;
  const bool thick = r_i < m.r_merge;
  return (stretched || sheared ? MS_BIT_STRAIN : 0) | (thick ? MS_BIT_THICK : 0);
}
__device__ __forceinline__ int ms_flag_of_bits(const MsParams &m, int bits,
                                               double hp, int lvl) {
  const bool thin_skip = hp < m.split_hp_min;
  if ((bits & MS_BIT_STRAIN) && !thin_skip && lvl > m.min_level) return MS_SPLIT;
  if ((bits & MS_BIT_THICK) && hp > m.hp_min && lvl < m.max_level)
    return MS_CANDIDATE;
  return MS_KEEP;
}
template <typename R>
__device__ __forceinline__ int ms_flag_one(const MsParams &m, double dt,
                                           const MsIn<R> &in, int ip) {
  const int bits = ms_strain_bits(m, dt, (double)in.Ap[ip], (double)in.vxdx[ip],
                                  (double)in.vxdy[ip], (double)in.vydx[ip],
                                  (double)in.vydy[ip]);
  return ms_flag_of_bits(m, bits, in.hp[ip], in.level[ip]);
}

// Greedy pairing of one cell's candidates, ascending id
// (merge_split_ops_cmes.h:195-228); returns the number of pairs formed.
// `limit` stops pairing once *n_merge reaches it (serial replay); the
// parallel kernels pass a null counter.
template <typename R>
__device__ int pair_cell(int b, int e, const int *cptcls, const MsIn<R> &in,
                         double max_dv2, int *act, int *partner, int *n_merge,
                         int limit) {
  int formed = 0;
  for (int i = b; i < e; ++i) {
    if (n_merge && *n_merge >= limit) break;
    const int q = cptcls[i];
    if (act[q] != MS_CANDIDATE) continue;
    double best = 1.7976931348623157e308;  // DBL_MAX
    int bj = -1;
    const double xq = in.x[q], yq = in.y[q];
    const double vxq = in.vpx[q], vyq = in.vpy[q];
    for (int j = i + 1; j < e; ++j) {
      const int q2 = cptcls[j];
      if (act[q2] != MS_CANDIDATE) continue;
      const double dx = xq - in.x[q2], dy = yq - in.y[q2];
      const double d2 = dx * dx + dy * dy;
      const double dvx = vxq - (double)in.vpx[q2];
      const double dvy = vyq - (double)in.vpy[q2];
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
// particles (merge_split_ops_cmes.h:230-249); returns the number undone.
__device__ inline int undo_cell(int b, int e, const int *cptcls, int min_ppc,
                                int split_w, int *act, int *partner) {
  int projected = 0;
  for (int i = b; i < e; ++i) {
    const int a = act[cptcls[i]];
    projected += a == MS_KEEP ? 1 : a == MS_SPLIT ? split_w : a == MS_PRIMARY ? 1 : 0;
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

// The serial decision loop over cells [c0, c1) with the global caps
// (merge_split_ops_cmes.h:176-249, merge_split_ops.h:405-497), from flags
// already set; cnt[0..2] = splits, merges, merges.  With early_stop the
// loop ends after the first cell at which both caps are spent and cnt[3]
// receives the first cell not visited (every later split and candidate is
// then kept, which the caller applies); otherwise cnt[3] = c1.
template <typename R>
__device__ void ms_serial_cells(int c0, int c1, const MsParams &m,
                                const MsIn<R> &in, const int *cstart,
                                const int *cptcls, int *act, int *partner,
                                int *cnt, bool early_stop = false) {
  int n_split = 0, n_merge = 0, cell = c0;
  for (; cell < c1; ++cell) {
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
    n_merge -= undo_cell(b, e, cptcls, cell_min_ppc(m), split_weight(m), act,
                         partner);
    // with both caps spent no later cell can split, pair or undo
    if (early_stop && n_split >= m.max_ops && n_merge >= m.max_ops) {
      ++cell;
      break;
    }
  }
  cnt[0] = n_split;
  cnt[1] = n_merge;
  cnt[2] = n_merge;
  cnt[3] = cell;
}

// Per-cell inputs of the legacy ELFS rule: exterior cells and hop distances
// to the fluid boundary (-1: not reached), indexed by global cell.
struct ElfsCells {
  const int *ext, *hops;
};

// compute_elfs (merge_split_ops.h:266-342) of one particle in local cell
// `cell`: the cell distance hops * min(hx, hy), refined below max(hx, hy)
// by the distance to the exterior cells of its row within two columns.
__device__ __forceinline__ double elfs_value(const ElfsCells &E, int coff,
                                             int nrows, int ncols, double hx,
                                             double hy, int cell, double px,
                                             double py) {
  const double h = hx < hy ? hx : hy, thr = hx > hy ? hx : hy;
  const int hop = E.hops[coff + cell];
  const double base = hop < 0 ? 1.7976931348623157e308 : hop * h;
  if (base > thr) return base;
  double mind = base;
  const int r0 = cell % nrows, c0 = cell / nrows;
  for (int dc = -2; dc <= 2; ++dc) {
    const int c = c0 + dc;
    if (c < 0 || c >= ncols) continue;
    if (!E.ext[coff + r0 + nrows * c]) continue;
    const double cx_lo = c * hx, cx_hi = (c + 1) * hx;
    const double cy_lo = r0 * hy, cy_hi = (r0 + 1) * hy;
    const double nx = px < cx_lo ? cx_lo : (px > cx_hi ? cx_hi : px);
    const double ny = py < cy_lo ? cy_lo : (py > cy_hi ? cy_hi : py);
    const double dx = px - nx, dy = py - ny;
    const double d = sqrt(dx * dx + dy * dy);
    if (d < mind) mind = d;
  }
  return mind;
}

// Action flag of one particle under the ELFS rule, merge_split_ops.h:388-402.
template <typename R>
__device__ __forceinline__ int elfs_flag_one(const MsParams &m, double elfs,
                                             const MsIn<R> &in, int ip) {
  const double r_i = sqrt((double)in.Ap[ip]);
  const double divergence = (double)in.vxdx[ip] + (double)in.vydy[ip];
  const int lvl = in.level[ip];
  if ((elfs < m.alpha * r_i) || (divergence > m.stretch))
    return lvl > m.min_level ? MS_SPLIT : MS_KEEP;
  if (elfs > m.beta * r_i && in.hp[ip] > m.hp_min)
    return lvl < m.max_level ? MS_CANDIDATE : MS_KEEP;
  return MS_KEEP;
}

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
