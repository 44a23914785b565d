// sim_kernels.hpp -- device kernels of the fused time loop (sim.hip), for
// one scenario or a batch of scenarios, templated on the arithmetic type R
// of the particle and nodal fields.
//
// Precision.  R = double is the plain FP64 path: every expression is the
// oracle's, in the oracle's order, so results are bit-identical to the FP64
// CPU restatement of optim_par.  R = float is the mixed path: positions,
// depth hp, H, the well-balance terms (1 - H/hp and hp - H^2/hp, i.e. the
// free-surface difference hp + Zp - eq_level), the bathymetry gradient, the
// P2G mass and momentum sums and the conservation reductions are FP64;
// velocities, stresses, friction, forces, velocity gradients, volumes,
// masses and the nodal velocities are FP32, and so is their arithmetic.
//
// Batch layout.  Scenario s owns particles [poff[s], poff[s+1]) (device
// offsets, moved by merge/split), cells [coff, coff + nrows*ncols) (the
// binning key is the global cell coff + r + c*nrows), nodes
// [noff, noff + (nrows+1)*(ncols+1)) and P2G tiles [toff, toff + ntiles).
// Particle kernels find a particle's scenario from wscen (scenario of the
// first particle of every 64-particle wavefront) and poff; P2G tiles from
// tscen.
#pragma once

#include <cstddef>

#include "binning.hpp"
#include "common.hpp"
#include "merge_split.hpp"

namespace amsc {

// ------------------------------------------------------------- layout
enum PF : int {
  // persistent state
  PF_X, PF_Y, PF_VPX, PF_VPY, PF_HP, PF_VP, PF_MP, PF_H, PF_F11, PF_F12,
  PF_F22, PF_FRX, PF_FRY,
  // written on output / merge steps only
  PF_AP, PF_FBX, PF_FBY, PF_VXDX, PF_VXDY, PF_VYDX, PF_VYDY, PF_APX, PF_APY,
  PF_ZP, PF_HPZ, PF_DZX, PF_DZY, PF_FPX, PF_FPY, PF_MOMX, PF_MOMY,
  PF_COUNT
};
constexpr int PF_CORE = PF_AP;

// Fields stored in FP64 in both precisions; the others are of type R.
__host__ __device__ constexpr bool pf_wide(int k) {
  return k == PF_X || k == PF_Y || k == PF_HP || k == PF_H;
}

__host__ __device__ constexpr bool pf_extensive(int k) {
  // merge_split_ops_cmes.h:49-50
  return k == PF_MP || k == PF_VP || k == PF_AP || k == PF_MOMX ||
         k == PF_MOMY;
}

enum NF : int {
  NF_VVXL, NF_VVYL,  // always, type R
  NF_AVX, NF_AVY, NF_MV, NF_MOMX, NF_MOMY, NF_FEXTX, NF_FEXTY, NF_FINTX,
  NF_FINTY, NF_FRX, NF_FRY, NF_FPX, NF_FPY, NF_FTOTX, NF_FTOTY, NF_VVX,
  NF_VVY,  // output steps only, FP64
  NF_COUNT
};
__host__ __device__ constexpr bool nf_narrow(int k) {
  return k == NF_VVXL || k == NF_VVYL;
}

struct PPtr {
  void *f[PF_COUNT];
  int *label, *level;
};
struct NPtr {
  void *f[NF_COUNT];
};

template <typename T>
__host__ __device__ __forceinline__ T *fld(const PPtr &p, int k) {
  return static_cast<T *>(p.f[k]);
}

// Device control block, one per scenario.
// Aligned to a cache line (128 bytes), as is each CFL slot below.
struct alignas(128) Ctrl {
  double t;        // time after the last step taken
  double dt;       // dt of the last step taken
  double t_start;  // time at the start of the last step taken
  double cons[2];  // err_mass, err_mom of the last step
  int it;          // steps taken
  int active;      // 0 once the run has ended
  int diverged;    // 1: a particle left the physical range (run stopped)
  // The next P2G's friction force Fric_p (optim_par.cpp:456-459) is
  // Ap * Fb of the last particle update, a function of the stored vpx, vpy,
  // hp and Vp, which P2G recomputes instead of reading, except where the
  // stored field holds something else: at creation (Fb = 0), after
  // merge/split (mass-averaged Fb), after a field was overwritten.
  // fric_set: the host asks the next step to read the stored field;
  // fric_read: the prologue's copy of it for the step being taken.
  int fric_set, fric_read;
  // largest particle count of a cell in any binning since creation (the
  // particle update and re-binning raise it); the host flags a scenario for
  // the dense-cell P2G from it
  int max_ppc;
  int pad_[16];
  // max |vpx|, |vpy|, hp of the particles as bit patterns of non-negative
  // doubles, by step parity ([slot][parity][value]).  The particle kernel's
  // workgroups add theirs with one atomic per value into one of 64 slots,
  // each a cache line of its own: device-scope atomics on one line are
  // executed one after another (about 40 ns each here), and with the 64
  // slots packed into 12 lines they held the 4M workload's particle kernel
  // back (0.4197 -> 0.4015 ms per FP64 step with a line per slot, 0.4387 ->
  // 0.3891 with workgroups of 256, which make twice as many)
  unsigned long long cfl[64][2][8];
  // the step whose merge/split strain bits the previous particle update
  // wrote (see flags_due), or -1
  int flags_it;
};
constexpr size_t kCtrlHead = 128;  // bytes before cfl (what the host reads)
static_assert(offsetof(Ctrl, cfl) == kCtrlHead, "control block header size");

// Divergence stop (no counterpart in optim_par, whose loop would step on
// with a vanishing dt): a particle speed above this, or NaN, flags the
// scenario and the next step's prologue ends its run.  Depth-averaged
// flows move at tens of m/s.
constexpr double kSpeedDiverged = 1.0e4;

struct KParams {
  int nrows, ncols;
  double hx, hy, g, rho, mu, tauy, xi, tan_fa, bing, fric, cfl, eq_level,
      dt_fixed, T;
  int bc, nsteps, max_steps;
};

// One scenario of the batch.
struct Scen {
  KParams P;
  Recip rc;            // FP64 basis-gradient factors
  RecipT<float> rcf;   // the same, rounded to FP32
  FastDiv fd;          // division by nrows
  MsParams ms;
  int ms_on, ms_interval;
  int coff, noff;      // first global cell, node
  int toff, ntiles, ntx;  // first P2G tile, tiles, tiles along x
};

// Arithmetic type of a storage type: FP64 computes in FP64; the mixed form
// computes in FP32 (or in FP64 over FP32 storage when built with
// AMSC_MIXED_FP64_MATH).
template <typename R> struct ArithT { using type = R; };
#ifdef AMSC_MIXED_FP64_MATH
template <> struct ArithT<float> { using type = double; };
#endif
template <typename R> using Arith = typename ArithT<R>::type;

template <typename R> __device__ __forceinline__ RecipT<R> recip_of(const Scen &S);
template <>
__device__ __forceinline__ RecipT<double> recip_of<double>(const Scen &S) {
  return RecipT<double>{S.rc.ix, S.rc.nix, S.rc.iy, S.rc.niy};
}
template <>
__device__ __forceinline__ RecipT<float> recip_of<float>(const Scen &S) {
  return S.rcf;
}

struct Tables {
  const Scen *scen;   // [S]
  Ctrl *ctrl;         // [S]
  const int *poff;    // [S+1]; poff[S] = particles in the batch
  const int *wscen;   // [capacity/64]
  int S;
};

__device__ __forceinline__ int scen_of(const Tables &T, int ip) {
  int s = T.wscen[ip >> 6];
  while (s + 1 < T.S && ip >= T.poff[s + 1]) ++s;
  return s;
}

// Merge/split runs on step numbers divisible by ms_call_interval
// (gpu_offload.cpp:321: it % interval == 0 after it++).
__device__ __forceinline__ bool ms_due(const Scen &S, const Ctrl &C) {
  return S.ms_on && S.ms_interval > 0 && C.active &&
         C.it % S.ms_interval == 0;
}

// Which optional fields a step writes: none; what the next merge/split
// reads (Ap, Fb, velocity gradients); every output field and the nodal set.
enum : int { AUX_NONE = 0, AUX_MS = 1, AUX_FULL = 2 };

// Whether the particle update of step `it` evaluates the strain tests of
// the merge/split flag of step it + 1 itself (ms_strain_bits): the next step
// merges and splits under the solvers' rule with a fixed dt (the tests need
// that dt).  The update then stores the bits of every particle and writes
// area, friction and velocity gradients only where a bit is set (no other
// particle can split or merge; the friction force of every particle is still
// written, for the next P2G), and the flag kernel finishes the flag from the
// bits, the depth and the level.
__device__ __forceinline__ bool flags_due(const Scen &S, int it) {
  return S.ms.mode == 0 && S.P.dt_fixed > 0.0 && S.ms_on &&
         S.ms_interval > 0 && (it + 1) % S.ms_interval == 0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(v, d);
    v = o > v ? o : v;
  }
  return v;
}

// ------------------------------------------------------------- prologue
// One wavefront per scenario: end test, dt, t, it for the step about to run.
__global__ void __launch_bounds__(64)
    prologue_kernel(const Scen *__restrict__ scen, Ctrl *ctrls) {
  const int lane = threadIdx.x;
  const KParams &P = scen[blockIdx.x].P;
  Ctrl *ctrl = &ctrls[blockIdx.x];
  if (!ctrl->active) return;
  if (ctrl->diverged) {
    if (lane == 0) ctrl->active = 0;
    return;
  }
  const bool fixed = P.dt_fixed > 0.0;
  const int it = ctrl->it;
  // loop condition, optim_par.cpp:281; optional step cap for adaptive runs
  const bool ended = (fixed ? !(it < P.nsteps) : !(ctrl->t < P.T)) ||
                     (!fixed && P.max_steps > 0 && it >= P.max_steps);
  if (ended) {
    if (lane == 0) ctrl->active = 0;
    return;
  }
  const int s = it + 1;  // this step
  const unsigned long long m0 = wave_max_u64(ctrl->cfl[lane][(s - 1) & 1][0]);
  const unsigned long long m1 = wave_max_u64(ctrl->cfl[lane][(s - 1) & 1][1]);
  const unsigned long long m2 = wave_max_u64(ctrl->cfl[lane][(s - 1) & 1][2]);
  double dt = ctrl->dt;
  if (fixed) {
    dt = P.dt_fixed;
  } else if (it > 0) {
    // optim_par.cpp:289-304: max |v| and max hp over the particles
    const double mvx = __longlong_as_double((long long)m0);
    const double mvy = __longlong_as_double((long long)m1);
    const double hmax = __longlong_as_double((long long)m2);
    const double a = sqrt(P.g * hmax) + mvx, b = sqrt(P.g * hmax) + mvy;
    const double cel = fabs(a > b ? a : b);
    const double minh = P.hx < P.hy ? P.hx : P.hy;
    dt = P.cfl * minh / (1e-2 + cel);
  }
  ctrl->cfl[lane][s & 1][0] = 0ull;
  ctrl->cfl[lane][s & 1][1] = 0ull;
  ctrl->cfl[lane][s & 1][2] = 0ull;
  if (lane == 0) {
    ctrl->dt = dt;
    ctrl->t_start = ctrl->t;
    ctrl->t = ctrl->t + dt;
    ctrl->it = s;
    ctrl->fric_read = ctrl->fric_set;
    ctrl->fric_set = 0;
  }
}

// Voellmy basal friction Fb (optim_par.cpp:613-626) and area Ap
// (optim_par.cpp:605-607) of a particle from its updated velocity, depth
// and volume; Fric_p = Ap * Fb.
template <typename A>
__device__ __forceinline__ void voellmy(const KParams &P, A vpx, A vpy,
                                        double h, A Vp, A &Fbx, A &Fby,
                                        A &Ap) {
  const A hR = (A)h;
  Ap = h > 1.e-2 ? Vp / hR : A(0);
  const A rho = (A)P.rho, g = (A)P.g;
  Fbx = 0;
  Fby = 0;
  const A nv = sqrt(vpx * vpx + vpy * vpy);
  if (P.fric > 0 && nv > A(1e-10) && P.xi > 0) {
    const A v2 = vpx * vpx + vpy * vpy;
    const A fric = (A)P.fric, tan_fa = (A)P.tan_fa, xi = (A)P.xi;
    Fbx = -fric * (rho * g * hR * tan_fa + rho * g * v2 / xi) * vpx / nv;
    Fby = -fric * (rho * g * hR * tan_fa + rho * g * v2 / xi) * vpy / nv;
  }
}

// ------------------------------------------------------------- step 0
// optim_par.cpp:235-264 (cells from the initial binning): Zp, hpZ, H,
// hydrostatic F, Fric = Ap*Fb with Fb = 0.
// Particles [p_begin, poff[S]) (all of them at creation, the scenarios just
// added when a served batch is repacked); mom_*_in are indexed from p_begin.
template <typename R>
__global__ void init_kernel(Tables T, PPtr p, const int *key,
                            const double *__restrict__ Z,
                            const double *mom_x_in, const double *mom_y_in,
                            int p_begin) {
  const int ip = p_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= T.poff[T.S]) return;
  const int sid = scen_of(T, ip);
  const Scen &S = T.scen[sid];
  const KParams &P = S.P;
  const int cell = key[ip] - S.coff, r = cell % P.nrows, c = cell / P.nrows;
  const Local L = local_coords(fld<double>(p, PF_X)[ip],
                               fld<double>(p, PF_Y)[ip], c, r, P.hx, P.hy);
  double zp = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    zp += shape(L, k) * Z[S.noff + node_id(k, c, r, P.nrows)];
  const double hp = fld<double>(p, PF_HP)[ip];
  fld<R>(p, PF_ZP)[ip] = (R)zp;
  fld<R>(p, PF_HPZ)[ip] = (R)(hp + zp);
  const double diff = P.eq_level - zp;
  const double H = (P.eq_level > 0 && diff > 0) ? diff : 0.0;
  fld<double>(p, PF_H)[ip] = H;
  const double h_corr = hp > 1e-10 ? hp - (H * H) / hp : 0.0;
  fld<R>(p, PF_F11)[ip] = (R)(-.5 * P.rho * P.g * h_corr);
  fld<R>(p, PF_F12)[ip] = R(0);
  fld<R>(p, PF_F22)[ip] = (R)(-.5 * P.rho * P.g * h_corr);
  fld<R>(p, PF_FBX)[ip] = R(0);
  fld<R>(p, PF_FBY)[ip] = R(0);
  fld<R>(p, PF_FRX)[ip] = fld<R>(p, PF_AP)[ip] * R(0);
  fld<R>(p, PF_FRY)[ip] = fld<R>(p, PF_AP)[ip] * R(0);
  const int zero_fields[] = {PF_VXDX, PF_VXDY, PF_VYDX, PF_VYDY, PF_APX,
                             PF_APY,  PF_DZX,  PF_DZY,  PF_FPX,  PF_FPY};
#pragma unroll
  for (int k = 0; k < 10; ++k) fld<R>(p, zero_fields[k])[ip] = R(0);
  fld<R>(p, PF_MOMX)[ip] = (R)mom_x_in[ip - p_begin];
  fld<R>(p, PF_MOMY)[ip] = (R)mom_y_in[ip - p_begin];
  p.label[ip] = ip - T.poff[sid];
  p.level[ip] = 0;
}

// Stored friction force Ap * Fb of the particles of scenario sid from their
// stored vpx, vpy, hp and Vp: what the last particle update would have
// written, for a reader of the field after a step that did not write it.
template <typename R>
__global__ void fric_kernel(Tables T, PPtr p, int sid) {
  using A = Arith<R>;
  const int ip = T.poff[sid] + blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= T.poff[sid + 1]) return;
  A fbx, fby, ap;
  voellmy<A>(T.scen[sid].P, (A)fld<R>(p, PF_VPX)[ip], (A)fld<R>(p, PF_VPY)[ip],
             fld<double>(p, PF_HP)[ip], (A)fld<R>(p, PF_VP)[ip], fbx, fby, ap);
  fld<R>(p, PF_FRX)[ip] = ap * fbx;
  fld<R>(p, PF_FRY)[ip] = ap * fby;
}

// Cell key and arrival slot of every particle (re-binning after creation or
// merge/split).
__global__ void key_kernel(Tables T, const double *__restrict__ x,
                           const double *__restrict__ y, int *key, int *slot,
                           int *count) {
  const int np = T.poff[T.S];
  const int b0 = blockIdx.x * blockDim.x;
  if (b0 >= np) return;
  const int ip = b0 + threadIdx.x;
  const bool valid = ip < np;
  int k = 0, sid = 0;
  if (valid) {
    sid = scen_of(T, ip);
    const Scen &S = T.scen[sid];
    k = S.coff + cell_of(x[ip], y[ip], S.P.nrows, S.P.ncols, S.P.hx, S.P.hy,
                         S.rc);
  }
  const int sl = wave_slot(count, k, valid);
  if (valid) {
    key[ip] = k;
    slot[ip] = sl;
  }
  // the largest cell count (slot + 1), raised in the control block past the
  // value stored there: one atomic per wavefront when the wavefront's
  // particles belong to one scenario, one per lane otherwise
  const int s0 = __builtin_amdgcn_readfirstlane(sid);
  if (__ballot(valid && sid != s0) == 0ull) {
    int m = valid ? sl + 1 : 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const int o = __shfl_xor(m, d);
      m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m > T.ctrl[s0].max_ppc)
      atomicMax(&T.ctrl[s0].max_ppc, m);
  } else if (valid && sl + 1 > T.ctrl[sid].max_ppc) {
    atomicMax(&T.ctrl[sid].max_ppc, sl + 1);
  }
}

// ------------------------------------------------------------- P2G
// Output-only destinations of p2g_nodal (read from device memory only on
// steps that write them, which keeps the kernel's argument block small).
struct P2GOut {
  void *dzx, *dzy, *fpx, *fpy, *momx, *momy;  // particle fields, type R
  double *nod[NF_COUNT];                       // nodal set (FP64 ones)
};

// The staged P2G values of one particle for the dense-cell P2G: FP64 local
// coordinates and momentum, then Mp, Fpx, Fpy, frx, fry, F11, F12, F22, Vp
// of type R; one record per binned position (of cptcls), 16-byte aligned so
// that a node thread reads a record with a few 16-byte loads.
template <typename R> struct alignas(16) P2GRec {
  double xi, eta, mx, my;
  R v[9];
};
struct P2GStaged {
  void *rec;  // P2GRec<R>[capacity]
};

struct P2GArgs {
  Tables T;
  const int *tscen;
  int aux, mom_stored;
  // per scenario: 1 = P2G by the dense-cell kernels (the tiled P2G skips
  // its tiles); null = no scenario is
  const int *dense;
  const int *key;  // binned cell of each particle
  P2GStaged st;
  const int *cstart, *cptcls;
  const double *x, *y, *hp, *H;
  // type R: Mp, momentum source (vpx/vpy, or the DATA momentum on the first
  // step), Vp, friction force, stress
  const void *Mp, *m1, *m2, *Vp, *frx, *fry, *f11, *f12, *f22;
  const double *Z;
  void *vvxL, *vvyL;  // type R
  double *partials;   // [tiles][6] or null
  const P2GOut *out;
};

// A pointer read from device memory is generic to the compiler, and a store
// through it is a FLAT store, which the LDS wait counter also counts: a FLAT
// store that may still be in flight forces every later LDS read to wait for
// all LDS reads before it.  These pointers always address global memory.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *global_ptr(void *p) {
  return (__attribute__((address_space(1))) T *)p;
}

template <typename R> struct P2GRaw {
  double x, y, hp, H;
  R Mp, m1, m2, Vp, frx, fry, f11, f12, f22;
};
// Where a tile-list position lives: particle id, cell column segment and
// cell row within the segment.
struct P2GIdx {
  int ip, s, j;
};

// fip: ip where the stored friction force is read, 0 where it is
// recomputed (a load of one shared line, so that both cases issue the same
// loads).
template <typename R>
__device__ __forceinline__ void p2g_fields(const P2GArgs &a, int ip, int fip,
                                           P2GRaw<R> &W) {
  W.x = a.x[ip];
  W.y = a.y[ip];
  W.Mp = static_cast<const R *>(a.Mp)[ip];
  W.m1 = static_cast<const R *>(a.m1)[ip];
  W.m2 = static_cast<const R *>(a.m2)[ip];
  W.Vp = static_cast<const R *>(a.Vp)[ip];
  W.hp = a.hp[ip];
  W.H = a.H[ip];
  W.frx = static_cast<const R *>(a.frx)[fip];
  W.fry = static_cast<const R *>(a.fry)[fip];
  W.f11 = static_cast<const R *>(a.f11)[ip];
  W.f12 = static_cast<const R *>(a.f12)[ip];
  W.f22 = static_cast<const R *>(a.f22)[ip];
}

// Nodal update of node (c, r) from its finished P2G sums: force balance and
// nodal velocities (optim_par.cpp:467-497, D1), boundary conditions
// (optim_par.cpp:503-529: nodes of the first/last cell column lose their x
// components, of the first/last cell row their y components), the
// velocities vL for G2P and, on output steps, the whole nodal set.
template <typename R>
__device__ __forceinline__ void p2g_node_update(
    const P2GArgs &a, int n, int c, int r, int nrows, int ncols, int bc,
    double dt, double Mv, double mvx, double mvy, Arith<R> FPx, Arith<R> FPy,
    Arith<R> Frx, Arith<R> Fry, Arith<R> Fix, Arith<R> Fiy) {
  using A = Arith<R>;
  const A Fex = FPx + Frx, Fey = FPy + Fry;
  const A Ftx = Fex - Fix, Fty = Fey - Fiy;
  double mx = dt * (double)Ftx + mvx, my = dt * (double)Fty + mvy;
  double ax = Mv > 1.e-2 ? (double)Ftx / Mv : 0.0;
  double ay = Mv > 1.e-2 ? (double)Fty / Mv : 0.0;
  double vx = Mv > 1.e-2 ? mx / Mv : 0.0;
  double vy = Mv > 1.e-2 ? my / Mv : 0.0;
  double vxL = dt * ax + vx, vyL = dt * ay + vy;
  if (bc) {
    if (c == 0 || c == 1 || c == ncols - 1 || c == ncols) {
      ax = 0.0; vx = 0.0; vxL = 0.0; mx = 0.0;
    }
    if (r == 0 || r == 1 || r == nrows - 1 || r == nrows) {
      ay = 0.0; vy = 0.0; vyL = 0.0; my = 0.0;
    }
  }
  static_cast<R *>(a.vvxL)[n] = (R)vxL;
  static_cast<R *>(a.vvyL)[n] = (R)vyL;
  if (a.aux >= AUX_FULL) {
    double *const *nod = a.out->nod;
    global_ptr<double>(nod[NF_AVX])[n] = ax;
    global_ptr<double>(nod[NF_AVY])[n] = ay;
    global_ptr<double>(nod[NF_MV])[n] = Mv;
    global_ptr<double>(nod[NF_MOMX])[n] = mx;
    global_ptr<double>(nod[NF_MOMY])[n] = my;
    global_ptr<double>(nod[NF_FEXTX])[n] = Fex;
    global_ptr<double>(nod[NF_FEXTY])[n] = Fey;
    global_ptr<double>(nod[NF_FINTX])[n] = Fix;
    global_ptr<double>(nod[NF_FINTY])[n] = Fiy;
    global_ptr<double>(nod[NF_FRX])[n] = Frx;
    global_ptr<double>(nod[NF_FRY])[n] = Fry;
    global_ptr<double>(nod[NF_FPX])[n] = FPx;
    global_ptr<double>(nod[NF_FPY])[n] = FPy;
    global_ptr<double>(nod[NF_FTOTX])[n] = Ftx;
    global_ptr<double>(nod[NF_FTOTY])[n] = Fty;
    global_ptr<double>(nod[NF_VVX])[n] = vx;
    global_ptr<double>(nod[NF_VVY])[n] = vy;
  }
}

// Staged particle i lives at i + i/16 in each LDS array: the node threads of
// a wavefront walk neighbouring cells in lockstep, i.e. at a stride of about
// 4 particles, which without the pad maps 16 lanes onto one bank.
__device__ __forceinline__ int lds_slot(int i) { return i + (i >> 4); }

// P2G + nodal update for one tile of TC x TR nodes.
//
// The tile's particles are the CSR ranges of its (TC+1) x (TR+1) cells,
// taken column segment by column segment ("tile list"), and staged through
// LDS TC*TR at a time.  Each chunk goes through a two-stage pipeline: while
// chunk k is folded, the fields of chunk k+1 and the particle ids of chunk
// k+2 are in flight.  A per-chunk owner map (cell of every tile-list
// position, built cell-parallel) turns a position into its particle id with
// one LDS read.  Each node thread folds its four cells in ascending cell id,
// each cell in ascending particle id: the serial p2g order, so the sums are
// the oracle's bit for bit.  Tile order: linear (row of tiles by row of
// tiles), or with DESC column-major from the last tile column down: particles
// are stored about in column-major cell order, so the particles the
// particle kernel wrote last (still in the memory-side cache) are staged
// first, and the first ones, which the next particle kernel reads first,
// last.
template <typename R, int TC, int TR, bool DESC>
__global__ void __launch_bounds__(TC * TR) p2g_nodal_kernel(P2GArgs a) {
  constexpr int NT = TC * TR, CH = NT, CHP = CH + CH / 16, SEG = TC + 1;
  static_assert(TR + 1 < 256 && SEG < 256, "owner map packs (s, j) in 16 bits");
  const int nb = gridDim.x, b = blockIdx.x;
  const int tile = DESC ? nb - 1 - b : b;
  // workgroup-uniform, said explicitly
  const int sid = __builtin_amdgcn_readfirstlane(a.tscen[tile]);
  const Scen &sc = a.T.scen[sid];
  // the parameters the staging reads, loaded here, in uniform control flow
  // (scalar loads): read through sc inside the friction branches they were
  // per-lane loads, and each one's wait also waited for the next chunk's
  // prefetched fields
  const KParams Pv = sc.P;
  Ctrl *ctrl = &a.T.ctrl[sid];
  if (!ctrl->active || (a.dense && a.dense[sid])) return;
  // stored friction force (first step, after merge/split or a field
  // write) or Ap * Fb recomputed from the stored vpx, vpy, hp, Vp
  const bool fric_read = ctrl->fric_read || ms_due(sc, *ctrl);
  // read before the kernel's first store: a load of it after one is a
  // vector load that the node update waits for, column after column
  const double step_dt = ctrl->dt;
  const int nrows = sc.P.nrows, ncols = sc.P.ncols, bc = sc.P.bc;
  const double hx = sc.P.hx, hy = sc.P.hy, grav = sc.P.g, rho = sc.P.rho;
  const int coff = sc.coff, noff = sc.noff, toff = sc.toff, ntx = sc.ntx;
  const Recip rc = sc.rc;
  using A = Arith<R>;
  const RecipT<A> rcr = recip_of<A>(sc);
  const int lt = tile - toff;
  int tcx = lt % ntx, tcy = lt / ntx;
  if (DESC) {
    const int nty = sc.ntiles / ntx;
    tcx = lt / nty;
    tcy = lt % nty;
  }
  const int c0 = tcx * TC, r0 = tcy * TR;
  const int t = threadIdx.x;
  const int lr = t % TR, lc = t / TR;
  const int c = c0 + lc, r = r0 + lr;
  const bool node_ok = c <= ncols && r <= nrows;
  const int *cstart = a.cstart + coff;

  __shared__ int s_cs[SEG][TR + 2];  // cstart, cell rows rl..rh+1
  __shared__ int s_sego[SEG + 1];    // tile-list offset per column
  __shared__ double s_Z[(TC + 2) * (TR + 2)];
  __shared__ double s_w[4][CHP];     // xi, eta, mom x, mom y (FP64)
  __shared__ R s_n[9][CHP];          // Mp, Fpx, Fpy, frx, fry, F11, F12, F22, Vp
  __shared__ unsigned short s_own[CH];
  __shared__ double s_red[6][NT / 64];

  const int ccl = c0 - 1 < 0 ? 0 : c0 - 1;
  const int cch = c0 + TC - 1 > ncols - 1 ? ncols - 1 : c0 + TC - 1;
  const int rl = r0 - 1 < 0 ? 0 : r0 - 1;
  const int rh = r0 + TR - 1 > nrows - 1 ? nrows - 1 : r0 + TR - 1;
  const int nseg = cch - ccl + 1, nr = rh - rl + 1;

  for (int i = t; i < nseg * (nr + 1); i += NT) {
    const int s = i / (nr + 1), j = i % (nr + 1);
    s_cs[s][j] = cstart[(ccl + s) * nrows + rl + j];
  }
  // A tile whose cells hold no particle has nothing to stage or fold, and
  // its nodes are written only on output steps: every wavefront tests the
  // window's column ranges itself (the loads above have fetched the same
  // words) and the whole workgroup leaves together, with zero partials.
  if (a.aux < AUX_FULL) {
    const int lane = t & 63;
    int n = 0;
    if (lane < nseg) {
      const int base = (ccl + lane) * nrows;
      n = cstart[base + rh + 1] - cstart[base + rl];
    }
    if (__ballot(n != 0) == 0ull) {
      if (a.partials && t < 6) a.partials[tile * 6 + t] = 0.0;
      return;
    }
  }
  __syncthreads();
  if (t == 0) {
    int o = 0;
    for (int s = 0; s < nseg; ++s) {
      s_sego[s] = o;
      o += s_cs[s][nr] - s_cs[s][0];
    }
    s_sego[nseg] = o;
  }
  __syncthreads();
  const int L = s_sego[nseg];

  // owner map of the chunk starting at base (positions base .. base+CH)
  auto build_own = [&](int base) {
    if (base >= L) return;
    for (int ci = t; ci < nseg * nr; ci += NT) {
      const int s = ci / nr, j = ci % nr;
      const int lo0 = s_sego[s] + (s_cs[s][j] - s_cs[s][0]);
      const int hi0 = lo0 + (s_cs[s][j + 1] - s_cs[s][j]);
      const int lo = lo0 > base ? lo0 : base;
      const int hi = hi0 < base + CH ? hi0 : base + CH;
      for (int q = lo; q < hi; ++q)
        s_own[q - base] = (unsigned short)((s << 8) | j);
    }
  };
  // Positions past the list are clamped to its last entry, so that every
  // thread issues the same loads whether or not its position is in the
  // list: the compiler can then count the loads in flight exactly and wait
  // for just the ones a stage needs (a load behind a branch may or may not
  // have been issued, which forces a wait for all of them).
  auto fetch_idx = [&](int q, P2GIdx &I) {
    q = q < L ? q : L - 1;
    const int v = s_own[q % CH];
    I.s = v >> 8;
    I.j = v & 0xFF;
    I.ip = a.cptcls[s_cs[I.s][0] + (q - s_sego[I.s])];
  };

  double pm = 0, ppx = 0, ppy = 0;  // particle sums (owned cells)
  double Mv = 0, mvx = 0, mvy = 0;
  A FPx = 0, FPy = 0, Frx = 0, Fry = 0, Fix = 0, Fiy = 0;

  if (L > 0) {
    P2GIdx I1, I2, I3;
    P2GRaw<R> W;
    build_own(0);
    // Z of nodes (c0-1 .. c0+TC, r0-1 .. r0+TR)
    for (int i = t; i < (TC + 2) * (TR + 2); i += NT) {
      const int zc = c0 - 1 + i / (TR + 2), zr = r0 - 1 + i % (TR + 2);
      s_Z[i] = (zc >= 0 && zc <= ncols && zr >= 0 && zr <= nrows)
                   ? a.Z[noff + zr + zc * (nrows + 1)]
                   : 0.0;
    }
    // this node's four cells as tile-list ranges, ascending cell id:
    // (c-1,r-1) corner 3, (c-1,r) corner 2, (c,r-1) corner 1, (c,r) corner 0
    int tb[4], te[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int cc = c - 1 + (k >> 1), cr = r - 1 + (k & 1);
      tb[k] = te[k] = 0;
      if (node_ok && cc >= 0 && cc < ncols && cr >= 0 && cr < nrows) {
        const int s = cc - ccl, j = cr - rl;
        tb[k] = s_sego[s] + (s_cs[s][j] - s_cs[s][0]);
        te[k] = s_sego[s] + (s_cs[s][j + 1] - s_cs[s][0]);
      }
    }
    __syncthreads();
    fetch_idx(t, I1);
    __syncthreads();
    build_own(CH);
    __syncthreads();
    fetch_idx(CH + t, I2);
    p2g_fields<R>(a, I1.ip, fric_read ? I1.ip : 0, W);
    __syncthreads();  // s_own is read above and rebuilt in the loop
    // Let the loads above land before the loop: otherwise the loop header
    // merges a path on which they are still in flight, and the compiler
    // then waits for every load in flight at the first reuse of one of
    // their registers, which falls inside the fold, i.e. it would wait for
    // the next chunk's prefetch before folding this one.
    __builtin_amdgcn_s_waitcnt(0);

    for (int base = 0; base < L; base += CH) {
      if (base + t < L) {
        // the particle's cell is the CSR cell it was found in
        const int pc = ccl + I1.s, pr = rl + I1.j;
        const Local Lc = local_coords(W.x, W.y, pc, pr, hx, hy, rc);
        // grad Z at the particle (g2pd(Z), optim_par.cpp:387)
        const int zb = (pc - (c0 - 1)) * (TR + 2) + (pr - (r0 - 1));
        const double z0 = s_Z[zb], z1 = s_Z[zb + 1];
        const double z2 = s_Z[zb + (TR + 2)], z3 = s_Z[zb + (TR + 3)];
        double dzx = 0.0, dzy = 0.0;
        dzx += shape_dx(Lc, 0, rc) * z0;
        dzy += shape_dy(Lc, 0, rc) * z0;
        dzx += shape_dx(Lc, 1, rc) * z1;
        dzy += shape_dy(Lc, 1, rc) * z1;
        dzx += shape_dx(Lc, 2, rc) * z2;
        dzy += shape_dy(Lc, 2, rc) * z2;
        dzx += shape_dx(Lc, 3, rc) * z3;
        dzy += shape_dy(Lc, 3, rc) * z3;
        // gravity along the slope with the well-balance factor,
        // optim_par.cpp:450-453
        const double wb = W.hp > 1e-10 ? 1.0 - (W.H / W.hp) : 1.0;
        const double Vpd = W.Vp;
        const double Fpx = -grav * Vpd * rho * dzx * wb;
        const double Fpy = -grav * Vpd * rho * dzy * wb;
        // mom_p = vpx * Mp (optim_par.cpp:600-601), DATA momentum at step 1
        const double Mpd = W.Mp;
        const double mx = a.mom_stored ? (double)W.m1 : (double)W.m1 * Mpd;
        const double my = a.mom_stored ? (double)W.m2 : (double)W.m2 * Mpd;
        const int ts = lds_slot(t);
        s_w[0][ts] = Lc.xi;
        s_w[1][ts] = Lc.eta;
        s_w[2][ts] = mx;
        s_w[3][ts] = my;
        s_n[0][ts] = W.Mp;
        s_n[1][ts] = (R)Fpx;
        s_n[2][ts] = (R)Fpy;
        R frx = W.frx, fry = W.fry;
        if (!fric_read) {
          A fbx, fby, ap;
          voellmy<A>(Pv, (A)W.m1, (A)W.m2, W.hp, (A)W.Vp, fbx, fby, ap);
          frx = ap * fbx;
          fry = ap * fby;
        }
        s_n[3][ts] = frx;
        s_n[4][ts] = fry;
        s_n[5][ts] = W.f11;
        s_n[6][ts] = W.f12;
        s_n[7][ts] = W.f22;
        s_n[8][ts] = W.Vp;
        const bool owned = pc >= c0 && pc < c0 + TC && pr >= r0 && pr < r0 + TR;
        if (owned) {
          pm += Mpd;
          ppx += mx;
          ppy += my;
          if (a.aux >= AUX_FULL) {
            const P2GOut *o = a.out;
            global_ptr<R>(o->dzx)[I1.ip] = (R)dzx;
            global_ptr<R>(o->dzy)[I1.ip] = (R)dzy;
            global_ptr<R>(o->fpx)[I1.ip] = (R)Fpx;
            global_ptr<R>(o->fpy)[I1.ip] = (R)Fpy;
            global_ptr<R>(o->momx)[I1.ip] = (R)mx;
            global_ptr<R>(o->momy)[I1.ip] = (R)my;
          }
        }
      }
      build_own(base + 2 * CH);
      __syncthreads();
      // in flight during the fold: ids of chunk k+2, fields of chunk k+1
      // issued at a raised wave priority, as in the swept kernel
      __builtin_amdgcn_s_setprio(2);
      fetch_idx(base + 2 * CH + t, I3);
      p2g_fields<R>(a, I2.ip, fric_read ? I2.ip : 0, W);
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int lo = tb[k] > base ? tb[k] : base;
        const int hi = te[k] < base + CH ? te[k] : base + CH;
        const int corner = 3 - k;
        for (int qq = lo; qq < hi; ++qq) {
          const int i = lds_slot(qq - base);
          const double xi = s_w[0][i], eta = s_w[1][i];
          const double N = shape_t<double>(xi, eta, corner);
          Mv += N * (double)s_n[0][i];
          mvx += N * s_w[2][i];
          mvy += N * s_w[3][i];
          const A Nr = (A)N;
          FPx += Nr * (A)s_n[1][i];
          FPy += Nr * (A)s_n[2][i];
          Frx += Nr * (A)s_n[3][i];
          Fry += Nr * (A)s_n[4][i];
          const A xr = (A)xi, er = (A)eta;
          const A Nx = shape_dx_t<A>(xr, er, corner, rcr);
          const A Ny = shape_dy_t<A>(xr, er, corner, rcr);
          const A vp = s_n[8][i];
          Fix += (Nx * (A)s_n[5][i] + Ny * (A)s_n[6][i]) * vp;
          Fiy += (Nx * (A)s_n[6][i] + Ny * (A)s_n[7][i]) * vp;
        }
      }
      __syncthreads();
      I1 = I2;
      I2 = I3;
    }
  }

  if (node_ok && (L > 0 || a.aux >= AUX_FULL))
    p2g_node_update<R>(a, noff + r + c * (nrows + 1), c, r, nrows, ncols, bc,
                       step_dt, Mv, mvx, mvy, FPx, FPy, Frx, Fry, Fix, Fiy);

  if (a.partials) {
    // conservation sums (optim_par.cpp:391-428), FP64, fixed order: tile
    // partials here, the scenario's reduction in conservation_kernel
    const int lane = t & 63, w = t >> 6;
    const double v[6] = {pm, ppx, ppy, node_ok ? Mv : 0.0, node_ok ? mvx : 0.0,
                         node_ok ? mvy : 0.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double sm = wave_sum(v[k]);
      if (lane == 0) s_red[k][w] = sm;
    }
    __syncthreads();
    if (t < 6) {
      double sm = 0.0;
      for (int ww = 0; ww < NT / 64; ++ww) sm += s_red[t][ww];
      a.partials[tile * 6 + t] = sm;
    }
  }
}

// ------------------------------------------------------------- swept P2G
// P2G + nodal update for one tile of W node columns x SR node rows, by one
// wavefront that sweeps the tile's cell columns from left to right.
//
// Node (c, r) adds the particles of its cells (c-1, r-1), (c-1, r), (c, r-1),
// (c, r) in that order: cell column c-1 rows r-1..r, then cell column c rows
// r-1..r, each a contiguous CSR range.  Two lanes per node row hold the two
// node columns beside the current cell column j: one folds node (r, j), which
// then has all four of its cells and is finished, the other starts node
// (r, j+1); the roles swap from column to column.  A lane walks the CSR range
// of cells (j, r-1) and (j, r) in order, so every node adds its particles in
// ascending cell, ascending id order: the serial sums of the tiled kernel,
// bit for bit.  Each column's particles are staged through LDS once (up to
// CAP positions a round); at four particles per cell every lane then folds
// eight particles per column, so the whole wavefront is busy in every column
// (a chunk of the tiled kernel keeps only the nodes beside its one or two
// cell columns busy).  The wavefront is the workgroup, so no barrier waits
// for another wavefront.  Pipeline: the fields of column j+1 load while
// column j is folded, its particle ids (and the cell row of each binned
// position) one column earlier, its CSR starts one column before that.
struct SweepIds {
  int ip[2], row[2];
};

// The cell row (0 .. nr-1) of binned position q in a column whose CSR starts
// lane l holds in cs (l <= nr): the last row whose start is <= q.  A binary
// search whose every round runs on every lane: the rounds read other lanes'
// registers (ds_bpermute), which a lane that has left the loop would no
// longer provide.  Six rounds cover nr <= 64.
__device__ __forceinline__ int sweep_row_of(int cs, int q, int nr) {
  int lo = 0, hi = nr;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int mid = (lo + hi) >> 1;
    const int v = __shfl(cs, mid);
    if (hi - lo > 1) {
      if (v <= q) lo = mid;
      else hi = mid;
    }
  }
  return lo;
}

template <typename R, int W, int SR, bool DESC>
__device__ __forceinline__ void p2g_sweep_body(P2GArgs a) {
  constexpr int CAP = 128, CAPP = CAP + CAP / 4, ZR = SR + 2;
  static_assert(2 * SR <= 64 && SR + 2 <= 64 && W + 1 <= 64, "tile too large");
  const int nb = gridDim.x, b = blockIdx.x;
  const int tile = DESC ? nb - 1 - b : b;
  const int sid = __builtin_amdgcn_readfirstlane(a.tscen[tile]);
  const Scen &sc = a.T.scen[sid];
  const KParams Pv = sc.P;
  Ctrl *ctrl = &a.T.ctrl[sid];
  if (!ctrl->active || (a.dense && a.dense[sid])) return;
  const bool fric_read = ctrl->fric_read || ms_due(sc, *ctrl);
  // read before the kernel's first store: a load of it after one is a
  // vector load that the node update waits for, column after column
  const double step_dt = ctrl->dt;
  const int nrows = sc.P.nrows, ncols = sc.P.ncols, bc = sc.P.bc;
  const double hx = sc.P.hx, hy = sc.P.hy, grav = sc.P.g, rho = sc.P.rho;
  const int coff = sc.coff, noff = sc.noff, toff = sc.toff, ntx = sc.ntx;
  const Recip rc = sc.rc;
  using A = Arith<R>;
  const RecipT<A> rcr = recip_of<A>(sc);
  const int lt = tile - toff;
  int tcx = lt % ntx, tcy = lt / ntx;
  if (DESC) {
    const int nty = sc.ntiles / ntx;
    tcx = lt / nty;
    tcy = lt % nty;
  }
  const int c0 = tcx * W, r0 = tcy * SR;
  const int lane = threadIdx.x, row = lane >> 1, side = lane & 1;
  const int r = r0 + row;  // this lane's node row
  const bool row_ok = row < SR && r <= nrows;
  const int *cstart = a.cstart + coff;
  // cell rows rl..rh of every column of the window, cell columns ccl..cch
  const int rl = r0 - 1 < 0 ? 0 : r0 - 1;
  const int rh = r0 + SR - 1 > nrows - 1 ? nrows - 1 : r0 + SR - 1;
  const int nr = rh - rl + 1;
  const int ccl = c0 - 1 < 0 ? 0 : c0 - 1;
  const int cch = c0 + W - 1 > ncols - 1 ? ncols - 1 : c0 + W - 1;

  __shared__ double s_w[4][CAPP];  // xi, eta, mom x, mom y (FP64)
  __shared__ R s_n[9][CAPP];       // Mp, Fpx, Fpy, frx, fry, F11, F12, F22, Vp
  __shared__ unsigned char s_own[CAP];  // cell row of each staged position
  __shared__ double s_Z[(W + 2) * ZR];

  // An empty window: nothing to stage or fold (its nodes are written on
  // output steps only).  qsafe: a binned position of this tile, where loads
  // for positions past a column's end are pointed (every load is issued
  // whether or not its position exists, see the tiled kernel).
  int qsafe = -1;
  {
    int n = 0, q0 = 0;
    if (lane < cch - ccl + 1) {
      const int base = (ccl + lane) * nrows;
      q0 = cstart[base + rl];
      n = cstart[base + rh + 1] - q0;
    }
    const unsigned long long m = __ballot(n != 0);
    if (m == 0ull) {
      if (a.aux < AUX_FULL) {
        if (a.partials && lane < 6) a.partials[tile * 6 + lane] = 0.0;
        return;
      }
    } else {
      qsafe = __shfl(q0, __builtin_ctzll(m));
    }
  }
  const bool any = qsafe >= 0;

  // CSR starts of cell column j, rows rl..rh+1, one per lane (0 past them)
  auto load_cs = [&](int j) -> int {
    return (j >= 0 && j < ncols && lane <= nr) ? cstart[j * nrows + rl + lane] : 0;
  };
  // particle ids and cell rows of the first CAP positions of a column.  The
  // cell row of a position: in the mixed form the lane of each row marks its
  // positions in an owner map (LDS) that the lanes then read; in FP64 a
  // binary search over the CSR starts (each measured the faster of the two
  // in its precision on the 4M workload, 121 against 133 us and 170 against
  // 189 us)
  auto fetch_ids = [&](int cs, SweepIds &I) {
    const int B = __shfl(cs, 0), E = __shfl(cs, nr);
    if constexpr (sizeof(R) == 4) {
      const int cs1 = __shfl_down(cs, 1);  // start of the next row
      if (lane < nr) {
        const int hi = cs1 < B + CAP ? cs1 : B + CAP;
        for (int q = cs; q < hi; ++q) s_own[q - B] = (unsigned char)lane;
      }
      __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = B + h * 64 + lane;
      const int qc = q < E ? q : (E > B ? E - 1 : qsafe);
      I.ip[h] = a.cptcls[qc];
      if constexpr (sizeof(R) == 4)
        I.row[h] = q < E ? s_own[h * 64 + lane] : 0;
      else
        I.row[h] = sweep_row_of(cs, q, nr);
    }
    if constexpr (sizeof(R) == 4) __syncthreads();  // the map is rewritten next
  };
  auto fetch_fields = [&](const SweepIds &I, P2GRaw<R> *Wf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) p2g_fields<R>(a, I.ip[h], fric_read ? I.ip[h] : 0, Wf[h]);
  };

  double pm = 0, ppx = 0, ppy = 0;  // particle sums (owned cells)
  double nm = 0, nmx = 0, nmy = 0;  // sums of the finished nodes
  double Mv = 0, mvx = 0, mvy = 0;
  A FPx = 0, FPy = 0, Frx = 0, Fry = 0, Fix = 0, Fiy = 0;

  // stage one particle (position q, binned cell (j, pr)) into LDS slot ls
  auto stage_one = [&](const P2GRaw<R> &Wp, int ip, int j, int pr, int ls) {
    const Local Lc = local_coords(Wp.x, Wp.y, j, pr, hx, hy, rc);
    const int zb = (j - (c0 - 1)) * ZR + (pr - (r0 - 1));
    const double z0 = s_Z[zb], z1 = s_Z[zb + 1];
    const double z2 = s_Z[zb + ZR], z3 = s_Z[zb + ZR + 1];
    // grad Z at the particle (g2pd(Z), optim_par.cpp:387)
    double dzx = 0.0, dzy = 0.0;
    dzx += shape_dx(Lc, 0, rc) * z0;
    dzy += shape_dy(Lc, 0, rc) * z0;
    dzx += shape_dx(Lc, 1, rc) * z1;
    dzy += shape_dy(Lc, 1, rc) * z1;
    dzx += shape_dx(Lc, 2, rc) * z2;
    dzy += shape_dy(Lc, 2, rc) * z2;
    dzx += shape_dx(Lc, 3, rc) * z3;
    dzy += shape_dy(Lc, 3, rc) * z3;
    // gravity along the slope with the well-balance factor,
    // optim_par.cpp:450-453
    const double wb = Wp.hp > 1e-10 ? 1.0 - (Wp.H / Wp.hp) : 1.0;
    const double Vpd = Wp.Vp;
    const double Fpx = -grav * Vpd * rho * dzx * wb;
    const double Fpy = -grav * Vpd * rho * dzy * wb;
    // mom_p = vpx * Mp (optim_par.cpp:600-601), DATA momentum at step 1
    const double Mpd = Wp.Mp;
    const double mx = a.mom_stored ? (double)Wp.m1 : (double)Wp.m1 * Mpd;
    const double my = a.mom_stored ? (double)Wp.m2 : (double)Wp.m2 * Mpd;
    s_w[0][ls] = Lc.xi;
    s_w[1][ls] = Lc.eta;
    s_w[2][ls] = mx;
    s_w[3][ls] = my;
    s_n[0][ls] = Wp.Mp;
    s_n[1][ls] = (R)Fpx;
    s_n[2][ls] = (R)Fpy;
    R frx = Wp.frx, fry = Wp.fry;
    if (!fric_read) {
      A fbx, fby, ap;
      voellmy<A>(Pv, (A)Wp.m1, (A)Wp.m2, Wp.hp, (A)Wp.Vp, fbx, fby, ap);
      frx = ap * fbx;
      fry = ap * fby;
    }
    s_n[3][ls] = frx;
    s_n[4][ls] = fry;
    s_n[5][ls] = Wp.f11;
    s_n[6][ls] = Wp.f12;
    s_n[7][ls] = Wp.f22;
    s_n[8][ls] = Wp.Vp;
    const bool owned = j >= c0 && j < c0 + W && pr >= r0 && pr < r0 + SR;
    if (owned) {
      pm += Mpd;
      ppx += mx;
      ppy += my;
      if (a.aux >= AUX_FULL) {
        const P2GOut *o = a.out;
        global_ptr<R>(o->dzx)[ip] = (R)dzx;
        global_ptr<R>(o->dzy)[ip] = (R)dzy;
        global_ptr<R>(o->fpx)[ip] = (R)Fpx;
        global_ptr<R>(o->fpy)[ip] = (R)Fpy;
        global_ptr<R>(o->momx)[ip] = (R)mx;
        global_ptr<R>(o->momy)[ip] = (R)my;
      }
    }
  };
  // fold the staged positions [q0, q1) (LDS slots from position base) into
  // this lane's node; positions below mid are in the upper cell (j, r-1)
  auto fold = [&](int q0, int q1, int base, int mid, bool right) {
    for (int q = q0; q < q1; ++q) {
      const int i = q - base + ((q - base) >> 2);
      const bool upper = q < mid;
      const double xi = s_w[0][i], eta = s_w[1][i];
      // the basis of the node at corner (right ? 2 : 0) + (upper ? 1 : 0)
      // of the particle's cell, as shape_t / shape_dx_t / shape_dy_t
      const double fa = right ? xi : (1.0 - xi);
      const double fb = upper ? eta : (1.0 - eta);
      const double N = fa * fb;
      Mv += N * (double)s_n[0][i];
      mvx += N * s_w[2][i];
      mvy += N * s_w[3][i];
      const A Nr = (A)N;
      FPx += Nr * (A)s_n[1][i];
      FPy += Nr * (A)s_n[2][i];
      Frx += Nr * (A)s_n[3][i];
      Fry += Nr * (A)s_n[4][i];
      const A xr = (A)xi, er = (A)eta;
      const A Nx = (right ? rcr.ix : rcr.nix) * (upper ? er : (A(1) - er));
      const A Ny = (right ? xr : (A(1) - xr)) * (upper ? rcr.iy : rcr.niy);
      const A vp = s_n[8][i];
      Fix += (Nx * (A)s_n[5][i] + Ny * (A)s_n[6][i]) * vp;
      Fiy += (Nx * (A)s_n[6][i] + Ny * (A)s_n[7][i]) * vp;
    }
  };

  if (any)
    for (int i = lane; i < (W + 2) * ZR; i += 64) {
      const int zc = c0 - 1 + i / ZR, zr = r0 - 1 + i % ZR;
      s_Z[i] = (zc >= 0 && zc <= ncols && zr >= 0 && zr <= nrows)
                   ? a.Z[noff + zr + zc * (nrows + 1)]
                   : 0.0;
    }

  // FP64 (whose two wavefronts per SIMD leave registers to spare): the
  // positions CAP .. CAP+63 of a column that holds more than CAP are loaded
  // ahead too, their ids one column ahead and their fields during the
  // column's first fold, instead of at once after it.  Columns that long
  // appear as a flow spreads (on the 4M workload 12 % of them at step 120,
  // 22 % at step 220).
  constexpr bool OVF = sizeof(R) == 8;
  int ioA = 0, roA = 0;  // overflow id and cell row of the step's column
  auto fetch_ovf = [&](int cs, int &io, int &ro) {
    const int B = __shfl(cs, 0), E = __shfl(cs, nr);
    if (E - B > CAP) {
      const int q = B + CAP + lane;
      io = a.cptcls[q < E ? q : E - 1];
      ro = sweep_row_of(cs, q, nr);
    }
  };

  // pipeline: column of step k in csA/IA/WA, step k+1 in csB/IB, k+2 in csC
  int csA = 0, csB = 0, csC = 0;
  SweepIds IA, IB;
  P2GRaw<R> WA[2];
  if (any) {
    csA = load_cs(c0 - 1);
    csB = load_cs(c0);
    fetch_ids(csA, IA);
    fetch_ids(csB, IB);
    if (OVF) fetch_ovf(csA, ioA, roA);
    fetch_fields(IA, WA);
    csC = load_cs(c0 + 1);
  }
  __syncthreads();

  for (int k = 0; k <= W; ++k) {
    const int j = c0 - 1 + k;  // cell column of this step
    // this lane's node: column j+1 (started here, corners 2/3) or column j
    // (finished here, corners 0/1)
    const bool right = side == (k & 1);
    const int c = right ? j + 1 : j;
    const bool node_ok = row_ok && c >= c0 && c < c0 + W && c <= ncols;
    if (any) {
      const int B = __shfl(csA, 0), E = __shfl(csA, nr);
      const int n1 = E - B < CAP ? E - B : CAP;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (h * 64 + lane < n1) {
          const int ls = h * 64 + lane + ((h * 64 + lane) >> 2);
          stage_one(WA[h], IA.ip[h], j, rl + IA.row[h], ls);
        }
      // in flight during the fold: fields of step k+1, ids of step k+2,
      // CSR starts of step k+3
      P2GRaw<R> WB[2];
      SweepIds IC;
      // FP64: the loads are issued at a raised wave priority, so that a
      // wavefront that reaches them is not held behind the other
      // wavefront's fold on the SIMD (step 0.3582 -> 0.3555 ms on the 4M
      // workload over 200 steps; the mixed kernel, three wavefronts per
      // SIMD, was slower with it: 137 -> 149 us inside the steps)
      if constexpr (sizeof(R) == 8) __builtin_amdgcn_s_setprio(2);
      fetch_fields(IB, WB);
      fetch_ids(csC, IC);
      const int csD = load_cs(j + 3);
      P2GRaw<R> WO;  // this column's overflow fields, in flight during the fold
      int ioN = 0, roN = 0;  // the next column's overflow id and row
      if (OVF) {
        if (E - B > CAP) p2g_fields<R>(a, ioA, fric_read ? ioA : 0, WO);
        fetch_ovf(csB, ioN, roN);
      }
      if constexpr (sizeof(R) == 8) __builtin_amdgcn_s_setprio(0);
      __syncthreads();
      // this lane's range: cells (j, r-1) and (j, r) where they exist
      int lo = 0, mid = 0, hi = 0;
      {
        const int iu = r - 1 - rl < 0 ? 0 : r - 1 - rl;
        const int im = r - rl < 0 ? 0 : (r - rl > nr ? nr : r - rl);
        const int il = r + 1 - rl > nr ? nr : r + 1 - rl;
        const int su = __shfl(csA, iu), sm = __shfl(csA, im), sl = __shfl(csA, il);
        if (node_ok && j >= 0 && j < ncols) {
          mid = sm;
          lo = r >= 1 ? su : sm;
          hi = r <= nrows - 1 ? sl : sm;
        }
      }
      fold(lo > B ? lo : B, hi < B + n1 ? hi : B + n1, B, mid, right);
      // columns of more than CAP positions: the prefetched 64 (FP64), then
      // the rest, CAP at a time
      int sub = B + CAP;
      if (OVF && sub < E) {
        __syncthreads();
        const int ns = E - sub < 64 ? E - sub : 64;
        if (lane < ns) stage_one(WO, ioA, j, rl + roA, lane + (lane >> 2));
        __syncthreads();
        fold(lo > sub ? lo : sub, hi < sub + ns ? hi : sub + ns, sub, mid, right);
        sub += 64;
      }
      for (; sub < E; sub += CAP) {
        __syncthreads();
        const int ns = E - sub < CAP ? E - sub : CAP;
        SweepIds IS;
        P2GRaw<R> WS[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int q = sub + h * 64 + lane;
          const int qc = q < E ? q : E - 1;
          IS.ip[h] = a.cptcls[qc];
          IS.row[h] = sweep_row_of(csA, q, nr);
        }
        fetch_fields(IS, WS);
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (h * 64 + lane < ns) {
            const int ls = h * 64 + lane + ((h * 64 + lane) >> 2);
            stage_one(WS[h], IS.ip[h], j, rl + IS.row[h], ls);
          }
        __syncthreads();
        fold(lo > sub ? lo : sub, hi < sub + ns ? hi : sub + ns, sub, mid, right);
      }
      csA = csB;
      csB = csC;
      csC = csD;
      IA = IB;
      IB = IC;
      WA[0] = WB[0];
      WA[1] = WB[1];
      ioA = ioN;
      roA = roN;
    }
    if (!right && node_ok) {
      if (any || a.aux >= AUX_FULL)
        p2g_node_update<R>(a, noff + r + c * (nrows + 1), c, r, nrows, ncols, bc,
                           step_dt, Mv, mvx, mvy, FPx, FPy, Frx, Fry, Fix, Fiy);
      nm += Mv;
      nmx += mvx;
      nmy += mvy;
    }
    if (!right) {
      Mv = mvx = mvy = 0;
      FPx = FPy = Frx = Fry = Fix = Fiy = 0;
    }
    __syncthreads();  // the next column's staging overwrites the LDS
  }

  if (a.partials) {
    // conservation sums (optim_par.cpp:391-428), FP64: tile partials here,
    // the scenario's reduction in conservation_kernel
    const double v[6] = {pm, ppx, ppy, nm, nmx, nmy};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double sm = wave_sum(v[k]);
      if (lane == k) a.partials[tile * 6 + k] = sm;
    }
  }
}

// The FP64 kernel as the compiler allocates it (about 204 registers, two
// wavefronts per SIMD); the mixed one held to three wavefronts per SIMD
// (168 registers, a few spilled): the faster of each on the 4M workload.
template <typename R, int W, int SR, bool DESC>
__global__ void __launch_bounds__(64) p2g_sweep_kernel(P2GArgs a) {
  p2g_sweep_body<R, W, SR, DESC>(a);
}
template <int W, int SR, bool DESC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
    p2g_sweep_mixed_kernel(P2GArgs a) {
  p2g_sweep_body<float, W, SR, DESC>(a);
}

// ------------------------------------------------------------- dense P2G
// P2G for scenarios whose cells hold many particles (a deposit piling up at
// the foot of a slope).  The tiled P2G stages its tile's particles through
// LDS a chunk at a time and only the nodes beside a chunk's few cells have
// work, so a tile of crowded cells runs for many chunks with most threads
// idle.  Here the staging is done once per particle into a buffer in
// binned order (p2g_stage_kernel), and then one thread per node walks its
// four cells in that buffer (p2g_dense_node_kernel): all nodes fold at
// once.  Same staging and folding expressions, same order per node (cells
// ascending, particles ascending), so the sums are the tiled P2G's bit for
// bit.

// One thread per binned position q: the staged values of particle
// cptcls[q] of a dense scenario, as the tiled P2G stages them, and on
// output steps the particle's output fields.
template <typename R>
__global__ void __launch_bounds__(256) p2g_stage_kernel(P2GArgs a) {
  using A = Arith<R>;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.T.poff[a.T.S]) return;
  const int ip = a.cptcls[q];
  const int sid = scen_of(a.T, ip);
  if (!a.dense[sid]) return;
  const Scen &sc = a.T.scen[sid];
  const Ctrl &C = a.T.ctrl[sid];
  if (!C.active) return;
  const bool fric_read = C.fric_read || ms_due(sc, C);
  const KParams &P = sc.P;
  const int nrows = P.nrows;
  const double hx = P.hx, hy = P.hy, grav = P.g, rho = P.rho;
  const Recip rc = sc.rc;
  const int cell = a.key[ip] - sc.coff, c = fdiv(cell, sc.fd), r = cell - c * nrows;
  P2GRaw<R> W;
  p2g_fields<R>(a, ip, fric_read ? ip : 0, W);
  double z[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) z[k] = a.Z[sc.noff + node_id(k, c, r, nrows)];
  const Local Lc = local_coords(W.x, W.y, c, r, hx, hy, rc);
  // grad Z at the particle (g2pd(Z), optim_par.cpp:387)
  double dzx = 0.0, dzy = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    dzx += shape_dx(Lc, k, rc) * z[k];
    dzy += shape_dy(Lc, k, rc) * z[k];
  }
  // gravity with the well-balance factor, optim_par.cpp:450-453
  const double wb = W.hp > 1e-10 ? 1.0 - (W.H / W.hp) : 1.0;
  const double Vpd = W.Vp;
  const double Fpx = -grav * Vpd * rho * dzx * wb;
  const double Fpy = -grav * Vpd * rho * dzy * wb;
  // mom_p = vpx * Mp (optim_par.cpp:600-601), DATA momentum at step 1
  const double Mpd = W.Mp;
  const double mx = a.mom_stored ? (double)W.m1 : (double)W.m1 * Mpd;
  const double my = a.mom_stored ? (double)W.m2 : (double)W.m2 * Mpd;
  R frx = W.frx, fry = W.fry;
  if (!fric_read) {
    A fbx, fby, ap;
    voellmy<A>(P, (A)W.m1, (A)W.m2, W.hp, (A)W.Vp, fbx, fby, ap);
    frx = ap * fbx;
    fry = ap * fby;
  }
  P2GRec<R> rec;
  rec.xi = Lc.xi;
  rec.eta = Lc.eta;
  rec.mx = mx;
  rec.my = my;
  const R v[9] = {W.Mp, (R)Fpx, (R)Fpy, frx, fry, W.f11, W.f12, W.f22, W.Vp};
#pragma unroll
  for (int k = 0; k < 9; ++k) rec.v[k] = v[k];
  static_cast<P2GRec<R> *>(a.st.rec)[q] = rec;
  if (a.aux >= AUX_FULL) {
    const P2GOut *o = a.out;
    global_ptr<R>(o->dzx)[ip] = (R)dzx;
    global_ptr<R>(o->dzy)[ip] = (R)dzy;
    global_ptr<R>(o->fpx)[ip] = (R)Fpx;
    global_ptr<R>(o->fpy)[ip] = (R)Fpy;
    global_ptr<R>(o->momx)[ip] = (R)mx;
    global_ptr<R>(o->momy)[ip] = (R)my;
  }
}

// One thread per node of a dense scenario's tiles (the tile layout of the
// tiled P2G, whose conservation partials it writes in the same places):
// the node's four cells in ascending cell id, each cell's staged particles
// in ascending id, then the nodal update.
template <typename R, int TC, int TR>
__global__ void __launch_bounds__((TC * TR + 63) / 64 * 64)
    __attribute__((amdgpu_waves_per_eu(1, 2))) p2g_dense_node_kernel(P2GArgs a) {
  // NT nodes, launched as NB threads (whole wavefronts)
  constexpr int NT = TC * TR, NB = (NT + 63) / 64 * 64, D = 8;
  using A = Arith<R>;
  const int tile = blockIdx.x;
  const int sid = __builtin_amdgcn_readfirstlane(a.tscen[tile]);
  if (!a.dense[sid]) return;
  const Scen &sc = a.T.scen[sid];
  const Ctrl *ctrl = &a.T.ctrl[sid];
  if (!ctrl->active) return;
  const double step_dt = ctrl->dt;  // read before the first store
  const int nrows = sc.P.nrows, ncols = sc.P.ncols;
  const RecipT<A> rcr = recip_of<A>(sc);
  const int lt = tile - sc.toff;
  const int c0 = (lt % sc.ntx) * TC, r0 = (lt / sc.ntx) * TR;
  const int t = threadIdx.x;
  const int c = c0 + t / TR, r = r0 + t % TR;
  const bool node_ok = t < NT && c <= ncols && r <= nrows;
  const int *cstart = a.cstart + sc.coff;
  const P2GRec<R> *recs = static_cast<const P2GRec<R> *>(a.st.rec);
  double pm = 0, ppx = 0, ppy = 0;  // particle sums (the node's own cell)
  double Mv = 0, mvx = 0, mvy = 0;
  A FPx = 0, FPy = 0, Frx = 0, Fry = 0, Fix = 0, Fiy = 0;
  if (node_ok) {
    // cells (c-1,r-1) corner 3, (c-1,r) corner 2, (c,r-1) corner 1, (c,r)
    // corner 0: ascending cell id
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int cc = c - 1 + (k >> 1), cr = r - 1 + (k & 1);
      if (cc < 0 || cc >= ncols || cr < 0 || cr >= nrows) continue;
      const int corner = 3 - k;
      const int b = cstart[cc * nrows + cr], e = cstart[cc * nrows + cr + 1];
      // D particles' values are loaded together (every load issued before
      // the first is used: a chain of one memory round trip per particle
      // would be latency-bound), then folded in order
      for (int q0 = b; q0 < e; q0 += D) {
        const int n = e - q0 < D ? e - q0 : D;
        P2GRec<R> rr[D];
#pragma unroll
        for (int u = 0; u < D; ++u) rr[u] = recs[q0 + (u < n ? u : 0)];
        // an (empty) use of every loaded value here, so that the compiler
        // keeps the D records' loads together instead of sinking each into
        // the iteration that needs it
#pragma unroll
        for (int u = 0; u < D; ++u) {
          asm volatile("" ::"v"(rr[u].xi), "v"(rr[u].eta), "v"(rr[u].mx), "v"(rr[u].my));
#pragma unroll
          for (int f = 0; f < 9; ++f) asm volatile("" ::"v"(rr[u].v[f]));
        }
#pragma unroll
        for (int u = 0; u < D; ++u) {
          if (u >= n) break;
          const R *fv = rr[u].v;
          const double xi = rr[u].xi, eta = rr[u].eta;
          const double Mp = (double)fv[0], mx = rr[u].mx, my = rr[u].my;
          const double N = shape_t<double>(xi, eta, corner);
          Mv += N * Mp;
          mvx += N * mx;
          mvy += N * my;
          const A Nr = (A)N;
          FPx += Nr * (A)fv[1];
          FPy += Nr * (A)fv[2];
          Frx += Nr * (A)fv[3];
          Fry += Nr * (A)fv[4];
          const A xr = (A)xi, er = (A)eta;
          const A Nx = shape_dx_t<A>(xr, er, corner, rcr);
          const A Ny = shape_dy_t<A>(xr, er, corner, rcr);
          const A f11 = (A)fv[5], f12 = (A)fv[6], f22 = (A)fv[7];
          const A vp = fv[8];
          Fix += (Nx * f11 + Ny * f12) * vp;
          Fiy += (Nx * f12 + Ny * f22) * vp;
          if (k == 3) {
            pm += Mp;
            ppx += mx;
            ppy += my;
          }
        }
      }
    }
    p2g_node_update<R>(a, sc.noff + r + c * (nrows + 1), c, r, nrows, ncols,
                       sc.P.bc, step_dt, Mv, mvx, mvy, FPx, FPy, Frx, Fry,
                       Fix, Fiy);
  }
  if (a.partials) {
    // conservation sums, FP64: tile partials as the tiled P2G writes them
    __shared__ double s_red[6][NB / 64];
    const int lane = t & 63, w = t >> 6;
    const double v[6] = {pm, ppx, ppy, Mv, mvx, mvy};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double sm = wave_sum(v[k]);
      if (lane == 0) s_red[k][w] = sm;
    }
    __syncthreads();
    if (t < 6) {
      double sm = 0.0;
      for (int ww = 0; ww < NB / 64; ++ww) sm += s_red[t][ww];
      a.partials[tile * 6 + t] = sm;
    }
  }
}

// err_mass, err_mom of every scenario from its tiles' partials of the last
// P2G it ran (tiles of an ended scenario no longer write them), one
// workgroup per scenario, optim_par.cpp:416-425.
__global__ void __launch_bounds__(256)
    conservation_kernel(const Scen *__restrict__ scen,
                        const double *__restrict__ partials, Ctrl *ctrls) {
  const Scen &sc = scen[blockIdx.x];
  Ctrl *ctrl = &ctrls[blockIdx.x];
  if (ctrl->it == 0) return;  // no step taken yet
  __shared__ double s[6][4];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < sc.ntiles; i += 256)
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] += partials[(sc.toff + i) * 6 + k];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double sm = wave_sum(v[k]);
    if (lane == 0) s[k][w] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[6];
    for (int k = 0; k < 6; ++k) tot[k] = ((s[k][0] + s[k][1]) + s[k][2]) + s[k][3];
    const double tm = tot[0], tx = tot[1], ty = tot[2];
    const double nm = tot[3], nx = tot[4], ny = tot[5];
    const double am = fabs(tx) > fabs(ty) ? fabs(tx) : fabs(ty);
    const double norm = 1.0 > am ? 1.0 : am;
    const double ex = fabs(tx - nx) / norm, ey = fabs(ty - ny) / norm;
    ctrl->cons[0] = fabs(tm - nm) / tm;
    ctrl->cons[1] = ex > ey ? ex : ey;
  }
}

// ------------------------------------------------------------- particles
struct PartArgs {
  Tables T;
  PPtr p;
  const int *key_in;  // step-start cell of each particle
  int *key_out, *slot, *count;
  int *ms_bits;  // merge/split strain bits for the next step (flags_due)
  const void *vvxL, *vvyL;  // type R
  const double *avx, *avy;
  const double *Z;
};

// What the particle update reads of one particle: its step-start cell, the
// grid velocities at the cell's nodes, position, depth, volume and H.
template <typename R> struct PartIn {
  using A = Arith<R>;
  int cell;
  A gx[4], gy[4];
  double x, y, hp0, H;
  A Vp;
};

// Every load of the update is issued here, together: a load first written
// at its use sits behind the branches before it and costs a memory round
// trip of its own.
template <typename R>
__device__ __forceinline__ void particle_load(const PartArgs &a, const Scen &sc,
                                              int ip, PartIn<R> &in) {
  const PPtr &p = a.p;
  const int nrows = sc.P.nrows;
  in.cell = a.key_in[ip] - sc.coff;
  const int c = fdiv(in.cell, sc.fd), r = in.cell - c * nrows;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int nid = sc.noff + node_id(k, c, r, nrows);
    in.gx[k] = static_cast<const R *>(a.vvxL)[nid];
    in.gy[k] = static_cast<const R *>(a.vvyL)[nid];
  }
  in.x = fld<double>(p, PF_X)[ip];
  in.y = fld<double>(p, PF_Y)[ip];
  in.hp0 = fld<double>(p, PF_HP)[ip];
  in.Vp = fld<R>(p, PF_VP)[ip];
  in.H = fld<double>(p, PF_H)[ip];
}

// G2P, PIC velocity, advection, G2PD at the new position with the
// step-start cell, height update, Voellmy, Bingham, friction force and the
// new cell key of one particle of an active scenario, from its loaded inputs.
template <typename R, int AUX>
__device__ __forceinline__ void particle_update(const PartArgs &a, const Scen &sc,
                                                double dt, int it, int ip,
                                                const PartIn<R> &in,
                                                int &newkey, double &vmx,
                                                double &vmy, double &hmx) {
  using A = Arith<R>;
  const KParams &P = sc.P;
  const PPtr &p = a.p;
  const int nrows = P.nrows;
  const double hx = P.hx, hy = P.hy;
  // the merge/split parameters read before the first store (after one they
  // are vector loads, each waited for in the middle of the update)
  const bool bits_due = AUX == AUX_MS && flags_due(sc, it);
  const int coff = sc.coff;
  MsParams msp{};
  double dt_fixed = 0.0;
  if (bits_due) {
    msp = sc.ms;
    dt_fixed = P.dt_fixed;
  }
  const int cell = in.cell, c = fdiv(cell, sc.fd), r = cell - c * nrows;
  const A *gx = in.gx, *gy = in.gy;
  const double x = in.x, y = in.y, hp0 = in.hp0, H = in.H;
  const A Vp = in.Vp;
  // the grid reciprocals copied (read through the reference, they were
  // loaded again in each branch that uses them)
  const Recip rc = sc.rc;
  const Local L0 = local_coords(x, y, c, r, hx, hy, rc);
  // g2p(vvxL, vvyL), optim_par.cpp:542
  const A xi0 = (A)L0.xi, eta0 = (A)L0.eta;
  A vxL = 0, vyL = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const A N = shape_t<A>(xi0, eta0, k);
    vxL += N * gx[k];
    vyL += N * gy[k];
  }
  // PIC velocity, advection and clamp, optim_par.cpp:548-564
  const A vpx = vxL, vpy = vyL;
  double xn = x + dt * (double)vpx, yn = y + dt * (double)vpy;
  const double eps = 1e-10, Lx = P.ncols * P.hx, Ly = P.nrows * P.hy;
  if (xn < eps) xn = eps;
  if (xn > Lx - eps) xn = Lx - eps;
  if (yn < eps) yn = eps;
  if (yn > Ly - eps) yn = Ly - eps;
  // g2pd(vvxL, vvyL) at the new position, step-start cell (D18),
  // optim_par.cpp:578
  const Local L1 = local_coords(xn, yn, c, r, hx, hy, rc);
  const A xi1 = (A)L1.xi, eta1 = (A)L1.eta;
  const RecipT<A> rcr = recip_of<A>(sc);
  A vxdx = 0, vxdy = 0, vydx = 0, vydy = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const A Nx = shape_dx_t<A>(xi1, eta1, k, rcr);
    const A Ny = shape_dy_t<A>(xi1, eta1, k, rcr);
    vxdx += Nx * gx[k];
    vxdy += Ny * gx[k];
    vydx += Nx * gy[k];
    vydy += Ny * gy[k];
  }
  // height update with floor, area, optim_par.cpp:595-607 (FP64)
  const A div = vxdx + vydy;
  double h = hp0 / (1 + dt * (double)div);
  h = h < 1e-2 ? 1e-2 : h;
  const A hR = (A)h;
  A Fbx, Fby, Ap;
  // from the velocity as stored, which is what a P2G that recomputes the
  // friction force sees (the same value unless A is wider than R)
  voellmy<A>(P, (A)(R)vpx, (A)(R)vpy, h, Vp, Fbx, Fby, Ap);
  // Bingham stress, stress_tensor_t::operator(), optim_par.cpp:68-120
  A F11, F12, F22;
  {
    const A qa = A(3) / A(2), qc = A(65) / A(32);
    const A mu = (A)P.mu, tau_Y = (A)P.tauy;
    const A nrm = sqrt(vpx * vpx + vpy * vpy);
    const A ALF = hR > A(1.e-3) ? (A(6) * mu * nrm) / ((hR + A(0.001)) * tau_Y) : A(0);
    const A B = A(-114) / A(32) - ALF;
    // Z1 = (-B + sq)/(2A) is selected only if it lies in [0, 1]; when
    // -B + sq > 3.0000001 it cannot (2A = 3), so Z2 is taken without
    // evaluating Z1.  Same values as computing both.
    const A sq = sqrt(B * B - A(4) * qa * qc);
    const A num1 = -B + sq;
    A ZZ;
    if (num1 > A(3.0000001)) {
      ZZ = (-B - sq) / (A(2) * qa);
    } else {
      const A Z1 = num1 / (A(2) * qa);
      ZZ = fabs(Z1 - A(.5)) <= A(.5) ? Z1 : (-B - sq) / (A(2) * qa);
    }
    const A D_xx = vxdx;
    const A D_xy = A(0.5) * (vxdy + vydx);
    const A D_yy = vydy;
    const A D_zx = hR > A(1.e-3) ? A(0.5) * (A(3) / (A(2) + ZZ)) * (vpx / (hR + A(0.001))) : A(0);
    const A D_zy = hR > A(1.e-3) ? A(0.5) * (A(3) / (A(2) + ZZ)) * (vpy / (hR + A(0.001))) : A(0);
    const A D_zz = -(vxdx + vydy);
    const A invII = A(0.5) * (D_xx * D_xx + D_yy * D_yy + D_zz * D_zz) +
                    D_zx * D_zx + D_zy * D_zy + D_xy * D_xy;
    const A sig_xx = invII != A(0) ? (tau_Y / sqrt(invII) + A(2) * mu) * D_xx : A(0);
    const A sig_xy = invII != A(0) ? (tau_Y / sqrt(invII) + A(2) * mu) * D_xy : A(0);
    const A sig_yy = invII != A(0) ? (tau_Y / sqrt(invII) + A(2) * mu) * D_yy : A(0);
    // hydrostatic part with the well-balance correction, FP64
    const double h_corr = h > 1e-10 ? h - (H * H) / h : 0.0;
    if (sizeof(A) == sizeof(double)) {
      F11 = P.bing * sig_xx - .5 * P.rho * P.g * h_corr;
      F12 = P.bing * sig_xy;
      F22 = P.bing * sig_yy - .5 * P.rho * P.g * h_corr;
    } else {
      const A pres = (A)(.5 * P.rho * P.g * h_corr);
      const A bing = (A)P.bing;
      F11 = bing * sig_xx - pres;
      F12 = bing * sig_xy;
      F22 = bing * sig_yy - pres;
    }
  }
  fld<double>(p, PF_X)[ip] = xn;
  fld<double>(p, PF_Y)[ip] = yn;
  fld<R>(p, PF_VPX)[ip] = vpx;
  fld<R>(p, PF_VPY)[ip] = vpy;
  fld<double>(p, PF_HP)[ip] = h;
  fld<R>(p, PF_F11)[ip] = F11;
  fld<R>(p, PF_F12)[ip] = F12;
  fld<R>(p, PF_F22)[ip] = F22;
  if (AUX >= AUX_MS) {
    // friction force for a P2G that reads it (after merge/split),
    // optim_par.cpp:456-459; other P2Gs recompute it
    fld<R>(p, PF_FRX)[ip] = Ap * Fbx;
    fld<R>(p, PF_FRY)[ip] = Ap * Fby;
  }
  bool ms_fields = AUX >= AUX_MS;
  if (bits_due) {
    // the strain tests from the values as stored
    const int bits = ms_strain_bits(msp, dt_fixed, (double)(R)Ap,
                                    (double)(R)vxdx, (double)(R)vxdy,
                                    (double)(R)vydx, (double)(R)vydy);
    a.ms_bits[ip] = bits;
    ms_fields = bits != 0;
  }
  if (ms_fields) {
    fld<R>(p, PF_AP)[ip] = Ap;
    fld<R>(p, PF_FBX)[ip] = Fbx;
    fld<R>(p, PF_FBY)[ip] = Fby;
    fld<R>(p, PF_VXDX)[ip] = vxdx;
    fld<R>(p, PF_VXDY)[ip] = vxdy;
    fld<R>(p, PF_VYDX)[ip] = vydx;
    fld<R>(p, PF_VYDY)[ip] = vydy;
  }
  if (AUX >= AUX_FULL) {
    double ax = 0.0, ay = 0.0, zp = 0.0;
    int nid[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) nid[k] = sc.noff + node_id(k, c, r, nrows);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double N = shape(L0, k);
      ax += N * a.avx[nid[k]];
      ay += N * a.avy[nid[k]];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) zp += shape(L1, k) * a.Z[nid[k]];
    fld<R>(p, PF_APX)[ip] = (A)ax;
    fld<R>(p, PF_APY)[ip] = (A)ay;
    fld<R>(p, PF_ZP)[ip] = (A)zp;
    fld<R>(p, PF_HPZ)[ip] = (A)(h + zp);
    const A Mp = fld<R>(p, PF_MP)[ip];
    fld<R>(p, PF_MOMX)[ip] = vpx * Mp;
    fld<R>(p, PF_MOMY)[ip] = vpy * Mp;
  }
  newkey = coff + cell_of(xn, yn, nrows, P.ncols, hx, hy, rc);
  vmx = fabs((double)vpx);
  vmy = fabs((double)vpy);
  hmx = h > 0.0 ? h : 0.0;
}

template <typename R, int AUX>
__device__ __forceinline__ void particle_body(const PartArgs &a, const Scen &sc,
                                              double dt, int it, int ip,
                                              int &newkey, double &vmx,
                                              double &vmy, double &hmx) {
  PartIn<R> in;
  particle_load<R>(a, sc, ip, in);
  particle_update<R, AUX>(a, sc, dt, it, ip, in, newkey, vmx, vmy, hmx);
}

// Particles per thread of the particle kernel (a thread loads the inputs
// of all of them before it updates the first, so that their memory round
// trips overlap).  One: two measured slower in both precisions on the 4M
// workload, 0.3910 against 0.3760 ms per FP64 step and 0.3202 against
// 0.3106 mixed (the second particle's registers cost more waves than the
// overlap gains).
constexpr int kPartIlp = 1;

// kPartIlp particles per thread (thread t of a workgroup of BS takes
// particles t, t + BS, ... of the workgroup's kPartIlp * BS).  Particles of
// a scenario whose run has ended keep their cell (and are binned again).
// The main launch (EDGE = false) takes the workgroups whose particles all
// belong to one scenario: the scenario's parameters are then
// workgroup-uniform, and the CFL maxima are reduced in the workgroup and
// added with one atomic per value into one of 64 slots.  The few
// workgroups that straddle a scenario boundary are left to a second launch
// (EDGE = true, one workgroup per boundary) that looks the scenario up per
// particle and adds the maxima per particle.
template <typename R, int AUX, int BS, bool EDGE>
__global__ void __launch_bounds__(BS)
    __attribute__((amdgpu_waves_per_eu(EDGE ? 1 : AUX == AUX_FULL ? 4 : AUX == AUX_MS ? 5 : 6)))
    particle_kernel(PartArgs a) {
  constexpr int ILP = kPartIlp, PB = ILP * BS;
  const Tables &T = a.T;
  const int np = T.poff[T.S];
  int blk = blockIdx.x;
  if (EDGE) {
    // boundary between scenarios j and j+1; the first boundary inside a
    // workgroup's range owns that workgroup
    const int j = blockIdx.x;
    if (j + 1 >= T.S) return;
    const int pb = T.poff[j + 1];
    if (pb >= np) return;
    blk = pb / PB;
    if (j > 0 && T.poff[j] / PB == blk) return;
  }
  const int b0 = blk * PB;
  if (b0 >= np) return;
  int sb = T.wscen[b0 >> 6];
  while (sb + 1 < T.S && b0 >= T.poff[sb + 1]) ++sb;
  const int last = (b0 + PB < np ? b0 + PB : np) - 1;
  const bool uniform = last < T.poff[sb + 1];
  if (uniform == EDGE) return;
  if (EDGE) {
    for (int j = 0; j < ILP; ++j) {
      const int ip = b0 + j * BS + threadIdx.x;
      const bool valid = ip < np;
      int sid = sb;
      if (valid)
        while (sid + 1 < T.S && ip >= T.poff[sid + 1]) ++sid;
      const Ctrl &C = T.ctrl[sid];
      const bool act = C.active != 0;
      int newkey = 0;
      double vmx = 0.0, vmy = 0.0, hmx = 0.0;
      if (valid) {
        if (act) {
          particle_body<R, AUX>(a, T.scen[sid], C.dt, C.it, ip, newkey, vmx, vmy, hmx);
          if (AUX == AUX_MS && flags_due(T.scen[sid], C.it))
            T.ctrl[sid].flags_it = C.it + 1;
          if (!(vmx <= kSpeedDiverged && vmy <= kSpeedDiverged))
            atomicOr(&T.ctrl[sid].diverged, 1);
        } else {
          newkey = a.key_in[ip];
        }
      }
      const int sl = wave_slot(a.count, newkey, valid);
      if (valid) {
        a.key_out[ip] = newkey;
        a.slot[ip] = sl;
        // the largest cell count of this binning (a slot + 1 is a count
        // its cell reaches), raised only past the value stored
        if (sl + 1 > T.ctrl[sid].max_ppc) atomicMax(&T.ctrl[sid].max_ppc, sl + 1);
        if (act) {
          unsigned long long *m = T.ctrl[sid].cfl[ip & 63][C.it & 1];
          atomicMax(&m[0], (unsigned long long)__double_as_longlong(vmx));
          atomicMax(&m[1], (unsigned long long)__double_as_longlong(vmy));
          atomicMax(&m[2], (unsigned long long)__double_as_longlong(hmx));
        }
      }
    }
    return;
  }
  // one scenario per workgroup: say so to the compiler, so that the
  // scenario's parameters are read with scalar loads
  const int sid = __builtin_amdgcn_readfirstlane(sb);
  const Ctrl &C = T.ctrl[sid];
  // the control block's fields read before the first store (after one they
  // would be read again with vector loads, each waited for)
  const bool act = C.active != 0;
  const double cdt = C.dt;
  const int cit = C.it, cmax = C.max_ppc;
  const bool fdue = AUX == AUX_MS && flags_due(T.scen[sid], cit);
  int newkey[ILP];
  double vmx = 0.0, vmy = 0.0, hmx = 0.0;
  if (act) {
    const Scen &sc = T.scen[sid];
    PartIn<R> in[ILP];
#pragma unroll
    for (int j = 0; j < ILP; ++j) {
      const int ip = b0 + j * BS + threadIdx.x;
      if (ip < np) particle_load<R>(a, sc, ip, in[j]);
    }
#pragma unroll
    for (int j = 0; j < ILP; ++j) {
      const int ip = b0 + j * BS + threadIdx.x;
      newkey[j] = 0;
      if (ip < np) {
        double vx, vy, hm;
        particle_update<R, AUX>(a, sc, cdt, cit, ip, in[j], newkey[j], vx, vy, hm);
        if (!(vx <= kSpeedDiverged && vy <= kSpeedDiverged))
          atomicOr(&T.ctrl[sid].diverged, 1);
        vmx = fmax(vmx, vx);
        vmy = fmax(vmy, vy);
        hmx = fmax(hmx, hm);
      }
    }
    if (fdue && threadIdx.x == 0) T.ctrl[sid].flags_it = cit + 1;
  } else {
#pragma unroll
    for (int j = 0; j < ILP; ++j) {
      const int ip = b0 + j * BS + threadIdx.x;
      newkey[j] = ip < np ? a.key_in[ip] : 0;
    }
  }
  // the largest cell count of this binning (a slot + 1 is a count its cell
  // reaches), raised in the control block only past the value stored there
  int m = 0;
#pragma unroll
  for (int j = 0; j < ILP; ++j) {
    const int ip = b0 + j * BS + threadIdx.x;
    const bool valid = ip < np;
    const int sl = wave_slot(a.count, newkey[j], valid);
    if (valid) {
      a.key_out[ip] = newkey[j];
      a.slot[ip] = sl;
      m = sl + 1 > m ? sl + 1 : m;
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const int o = __shfl_xor(m, d);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m > cmax) atomicMax(&T.ctrl[sid].max_ppc, m);
  if (!act) return;
  __shared__ double s_cfl[3][BS / 64];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    vmx = fmax(vmx, __shfl_xor(vmx, d));
    vmy = fmax(vmy, __shfl_xor(vmy, d));
    hmx = fmax(hmx, __shfl_xor(hmx, d));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_cfl[0][w] = vmx;
    s_cfl[1][w] = vmy;
    s_cfl[2][w] = hmx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BS / 64; ++k) {
      vmx = fmax(vmx, s_cfl[0][k]);
      vmy = fmax(vmy, s_cfl[1][k]);
      hmx = fmax(hmx, s_cfl[2][k]);
    }
    unsigned long long *c = T.ctrl[sb].cfl[blk & 63][cit & 1];
    atomicMax(&c[0], (unsigned long long)__double_as_longlong(vmx));
    atomicMax(&c[1], (unsigned long long)__double_as_longlong(vmy));
    atomicMax(&c[2], (unsigned long long)__double_as_longlong(hmx));
  }
}

// ------------------------------------------------------------- merge/split
// Fresh action flag of particle ip under its scenario's rule; from the
// strain bits of the last particle update when it stored them for this
// step (flags_due; the dt is then the fixed one the bits were taken with).
template <typename R>
__device__ __forceinline__ int fresh_flag(const Scen &sc, const Ctrl &C,
                                          const MsIn<R> &in, const ElfsCells &E,
                                          const int *key, const int *bits,
                                          int ip) {
  if (sc.ms.mode == 1) {
    const double e = elfs_value(E, sc.coff, sc.P.nrows, sc.P.ncols, sc.P.hx,
                                sc.P.hy, key[ip] - sc.coff, in.x[ip], in.y[ip]);
    return elfs_flag_one(sc.ms, e, in, ip);
  }
  if (C.flags_it == C.it) return ms_flag_of_bits(sc.ms, bits[ip], in.hp[ip], in.level[ip]);
  return ms_flag_one(sc.ms, C.dt, in, ip);
}

// Fresh flags of every particle; splits counted per scenario (cnt[4*s]),
// and the cells holding a merge candidate marked in cand (the pair kernel
// visits only those and clears the mark).
template <typename R>
__global__ void ens_flag_kernel(Tables T, MsIn<R> in, ElfsCells E,
                                const int *__restrict__ key,
                                const int *__restrict__ bits, int *act,
                                int *partner, int *cnt, int *cand) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= T.poff[T.S]) return;
  const int sid = scen_of(T, ip);
  // a wavefront whose particles share a scenario reads its parameters with
  // scalar loads
  const int s0 = __builtin_amdgcn_readfirstlane(sid);
  const bool uniform = __ballot(sid != s0) == 0ull;
  int f;
  if (uniform) {
    const Scen &sc = T.scen[s0];
    const Ctrl &C = T.ctrl[s0];
    f = ms_due(sc, C) ? fresh_flag(sc, C, in, E, key, bits, ip) : MS_KEEP;
  } else {
    const Scen &sc = T.scen[sid];
    const Ctrl &C = T.ctrl[sid];
    f = ms_due(sc, C) ? fresh_flag(sc, C, in, E, key, bits, ip) : MS_KEEP;
  }
  act[ip] = f;
  partner[ip] = -1;
  if (f == MS_CANDIDATE) cand[key[ip]] = 1;
  // one atomic per wavefront when its particles share a scenario
  const unsigned long long sp = __ballot(f == MS_SPLIT);
  if (uniform) {
    if (sp && (int)__lane_id() == __ffsll((long long)__ballot(true)) - 1)
      atomicAdd(&cnt[4 * s0], __popcll(sp));
  } else if (f == MS_SPLIT) {
    atomicAdd(&cnt[4 * sid], 1);
  }
}

// The two cell passes of the legacy ELFS rule (merge_split_ops.h:86-262) for
// each scenario using it, one workgroup per scenario, level by level:
//   ext[c] = 1 for the empty cells connected (4-neighbours, through empty
//     cells) to an empty cell of the grid edge;
//   hops[c] = fewest 4-neighbour steps through wet cells from a seed (a wet
//     cell off the domain edge with an exterior cell left or right of it),
//     -1 where no seed is reached.
// Both are sets defined by reachability, so the order in which a level's
// cells are claimed (atomics) does not change them.  qa, qb: frontier
// queues; a cell enters each pass's queues once.
__global__ void __launch_bounds__(1024)
    elfs_cells_kernel(Tables T, const int *__restrict__ cstart, int *ext,
                      int *hops, int *qa, int *qb) {
  const int sid = blockIdx.x;
  const Scen &sc = T.scen[sid];
  if (sc.ms.mode != 1 || !ms_due(sc, T.ctrl[sid])) return;
  const int nrows = sc.P.nrows, ncols = sc.P.ncols, nc = nrows * ncols;
  const int *cs = cstart + sc.coff;
  int *ex = ext + sc.coff, *hp = hops + sc.coff;
  int *A = qa + sc.coff, *B = qb + sc.coff;
  const int t = threadIdx.x, NT = blockDim.x;
  __shared__ int s_n[2];
  auto wet = [&](int c) { return cs[c + 1] > cs[c]; };
  for (int i = t; i < nc; i += NT) {
    ex[i] = 0;
    hp[i] = -1;
  }
  if (t == 0) s_n[0] = 0;
  __syncthreads();
  // exterior: seeds are the empty cells of rows 0, nrows-1 and columns 0,
  // ncols-1
  const int perim = 2 * (nrows + ncols);
  for (int i = t; i < perim; i += NT) {
    int r, c;
    if (i < 2 * ncols) {
      c = i >> 1;
      r = (i & 1) ? nrows - 1 : 0;
    } else {
      const int k = i - 2 * ncols;
      r = k >> 1;
      c = (k & 1) ? ncols - 1 : 0;
    }
    const int idx = r + nrows * c;
    if (!wet(idx) && atomicCAS(&ex[idx], 0, 1) == 0) A[atomicAdd(&s_n[0], 1)] = idx;
  }
  __syncthreads();
  int n = s_n[0];
  int *cur = A, *nxt = B;
  while (n > 0) {
    __syncthreads();
    if (t == 0) s_n[1] = 0;
    __syncthreads();
    for (int k = t; k < n; k += NT) {
      const int cidx = cur[k], r = cidx % nrows, c = cidx / nrows;
      const int nb[4] = {r > 0 ? cidx - 1 : -1, r < nrows - 1 ? cidx + 1 : -1,
                         c > 0 ? cidx - nrows : -1, c < ncols - 1 ? cidx + nrows : -1};
      for (int j = 0; j < 4; ++j)
        if (nb[j] >= 0 && !wet(nb[j]) && atomicCAS(&ex[nb[j]], 0, 1) == 0)
          nxt[atomicAdd(&s_n[1], 1)] = nb[j];
    }
    __syncthreads();
    n = s_n[1];
    int *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  // distance: seeds are wet cells off the domain edge with an exterior cell
  // left or right of them
  __syncthreads();
  if (t == 0) s_n[0] = 0;
  __syncthreads();
  for (int i = t; i < nc; i += NT) {
    const int r = i % nrows, c = i / nrows;
    if (!wet(i) || r == 0 || r == nrows - 1 || c == 0 || c == ncols - 1) continue;
    if (ex[i - nrows] || ex[i + nrows]) {
      hp[i] = 0;
      A[atomicAdd(&s_n[0], 1)] = i;
    }
  }
  __syncthreads();
  n = s_n[0];
  cur = A;
  nxt = B;
  for (int level = 0; n > 0; ++level) {
    __syncthreads();
    if (t == 0) s_n[1] = 0;
    __syncthreads();
    for (int k = t; k < n; k += NT) {
      const int cidx = cur[k], r = cidx % nrows, c = cidx / nrows;
      const int nb[4] = {r > 0 ? cidx - 1 : -1, r < nrows - 1 ? cidx + 1 : -1,
                         c > 0 ? cidx - nrows : -1, c < ncols - 1 ? cidx + nrows : -1};
      for (int j = 0; j < 4; ++j)
        if (nb[j] >= 0 && wet(nb[j]) && atomicCAS(&hp[nb[j]], -1, level + 1) == -1)
          nxt[atomicAdd(&s_n[1], 1)] = nb[j];
    }
    __syncthreads();
    n = s_n[1];
    int *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
}

// One thread per cell holding a merge candidate (marked by the flag
// kernel; the mark is cleared here): greedy pairing and the min-particles
// undo (which only a pair can make necessary); counts per scenario in
// cnt[4*s + 1..2] (pairs formed, net merges).
template <typename R>
__global__ void ens_pair_kernel(Tables T, const int *__restrict__ cell_off,
                                int nc, MsIn<R> in,
                                const int *__restrict__ cstart,
                                const int *__restrict__ cptcls, int *act,
                                int *partner, int *cnt, int *cand) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nc || !cand[cell]) return;
  cand[cell] = 0;
  const int b = cstart[cell], e = cstart[cell + 1];
  int lo = 0, hi = T.S - 1;  // last scenario with cell_off <= cell
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cell_off[mid] <= cell) lo = mid;
    else hi = mid - 1;
  }
  const Scen &sc = T.scen[lo];
  const int formed =
      pair_cell(b, e, cptcls, in, sc.ms.max_dv2, act, partner, nullptr, 0);
  const int undone = formed ? undo_cell(b, e, cptcls, cell_min_ppc(sc.ms),
                                        split_weight(sc.ms), act, partner)
                            : 0;
  int *c = cnt + 4 * lo;
  if (formed) atomicAdd(&c[1], formed);
  if (formed - undone) atomicAdd(&c[2], formed - undone);
}

// Where an operation cap could bind (more splits or pairs than ms_max_ops
// in the parallel pass), the exact serial loop replaces the parallel
// result: fresh flags for the scenario's particles (in parallel), the loop
// over its cells in order until both caps are spent (one thread), and every
// later split or candidate kept (in parallel).  cnt[4*s + 3] = the first
// global cell the loop did not visit (0: no replay).
__device__ __forceinline__ bool caps_bind(const Scen &sc, const Ctrl &C,
                                          const int *c) {
  return ms_due(sc, C) && (c[0] > sc.ms.max_ops || c[1] > sc.ms.max_ops);
}

template <typename R>
__global__ void ens_refresh_kernel(Tables T, MsIn<R> in, ElfsCells E,
                                   const int *__restrict__ key,
                                   const int *__restrict__ bits, int *act,
                                   int *partner, const int *cnt) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= T.poff[T.S]) return;
  const int sid = scen_of(T, ip);
  // most wavefronts share one scenario whose caps do not bind: they leave
  // after scalar reads
  const int s0 = __builtin_amdgcn_readfirstlane(sid);
  if (__ballot(sid != s0) == 0ull && !caps_bind(T.scen[s0], T.ctrl[s0], cnt + 4 * s0))
    return;
  const Scen &sc = T.scen[sid];
  const Ctrl &C = T.ctrl[sid];
  if (!caps_bind(sc, C, cnt + 4 * sid)) return;
  act[ip] = fresh_flag(sc, C, in, E, key, bits, ip);
  partner[ip] = -1;
}

template <typename R>
__global__ void ens_serial_kernel(Tables T, MsIn<R> in,
                                  const int *__restrict__ cstart,
                                  const int *__restrict__ cptcls, int *act,
                                  int *partner, int *cnt) {
  if (threadIdx.x != 0) return;
  const int sid = blockIdx.x;
  const Scen &sc = T.scen[sid];
  int *c = cnt + 4 * sid;
  if (!caps_bind(sc, T.ctrl[sid], c)) return;
  ms_serial_cells(sc.coff, sc.coff + sc.P.nrows * sc.P.ncols, sc.ms, in,
                  cstart, cptcls, act, partner, c, true);
}

__global__ void ens_demote_kernel(Tables T, const int *__restrict__ key,
                                  int *act, const int *cnt) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= T.poff[T.S]) return;
  const int stop = cnt[4 * scen_of(T, ip) + 3];
  if (stop == 0 || key[ip] < stop) return;
  const int a = act[ip];
  if (a == MS_SPLIT || a == MS_CANDIDATE) act[ip] = MS_KEEP;
}

__global__ void weights_kernel(Tables T, int n, const int *__restrict__ act,
                               int *w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i >= T.poff[T.S]) {
    w[i] = 0;
    return;
  }
  const int a = act[i];
  w[i] = a == MS_KEEP      ? 1
         : a == MS_SPLIT   ? split_weight(T.scen[scen_of(T, i)].ms)
         : a == MS_PRIMARY ? 1
                           : 0;
}

// New particle offsets of every scenario (order is preserved, so scenario
// s starts at the scan value of its first particle); flags overflow; copies
// each scenario's active flag next to them (read back with them).  box !=
// null (host memory): the whole block cnt[0 .. 6S+2) is also written to
// box[1 ..], then box[0] = seq, released to the host.
__global__ void ms_offsets_kernel(const int *__restrict__ poff, int S,
                                  const int *__restrict__ offs, int cap,
                                  int *poff_next, int *err,
                                  const Ctrl *__restrict__ ctrl, int *active,
                                  int *box, const int *cnt, int seq) {
  for (int s = threadIdx.x; s <= S; s += blockDim.x) {
    poff_next[s] = offs[poff[s]];
    if (s < S) active[s] = ctrl[s].active;
  }
  if (threadIdx.x == 0 && offs[poff[S]] > cap) *err = 1;
  if (!box) return;
  __syncthreads();
  for (int k = threadIdx.x; k < 6 * S + 2; k += blockDim.x) box[1 + k] = cnt[k];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(&box[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void ms_commit_kernel(int *poff, const int *__restrict__ poff_next,
                                 int S, const int *__restrict__ err) {
  if (*err) return;
  for (int s = threadIdx.x; s <= S; s += blockDim.x) poff[s] = poff_next[s];
}

// Scenario of the first particle of every 64-particle wavefront.
__global__ void wscen_kernel(const int *__restrict__ poff, int S, int nw,
                             int *wscen) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const int i = w * 64;
  int lo = 0, hi = S - 1;  // first s with i < poff[s+1]
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (i < poff[mid + 1]) hi = mid;
    else lo = mid + 1;
  }
  wscen[w] = lo;
}

// Merge/split execution, merge_split_ops_cmes.h:258-427.
struct ExecArgs {
  Tables T;
  PPtr in, out;
  const int *act, *partner, *offset;
  const int *err;
  int ncopy;  // fields copied for kept particles (PF_CORE or PF_COUNT)
  // the binning of the new particle set: cell keys, slots, cell counts
  // (zero on entry)
  int *key, *slot, *count;
};

template <typename R>
__device__ __forceinline__ double ldv(const PPtr &p, int k, int i) {
  return pf_wide(k) ? fld<double>(p, k)[i] : (double)fld<R>(p, k)[i];
}
template <typename R>
__device__ __forceinline__ void stv(const PPtr &p, int k, int i, double v) {
  if (pf_wide(k)) fld<double>(p, k)[i] = v;
  else fld<R>(p, k)[i] = (R)v;
}

// execute_merge_split of the legacy ELFS rule (merge_split_ops.h:603-710):
// a split leaves four children at (x +- r/2, y +- r/2), r = sqrt(Ap),
// clamped to the domain, extensive fields quartered, level - 1, label -1; a
// merge averages the intensive fields by mass (depth included), sums the
// extensive ones, keeps the primary's label, level + 1.  Momentum = Mp * v
// and the stored friction force Ap * Fb follow the new fields.
template <typename R>
__device__ void elfs_execute_one(const ExecArgs &a, int i, int o, int act,
                                 double Lx, double Ly, double *ox, double *oy) {
  const PPtr &in = a.in, &out = a.out;
  if (act == MS_SPLIT) {
    const double px = ldv<R>(in, PF_X, i), py = ldv<R>(in, PF_Y, i);
    const double off = sqrt(ldv<R>(in, PF_AP, i)) / 2.0;
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
    // field by field (the four children share every value but x, y)
    for (int k = PF_VPX; k < PF_COUNT; ++k) {
      const double v = ldv<R>(in, k, i) / (pf_extensive(k) ? 4.0 : 1.0);
#pragma unroll
      for (int c = 0; c < 4; ++c) stv<R>(out, k, o + c, v);
    }
    const double Mc = ldv<R>(out, PF_MP, o), Ac = ldv<R>(out, PF_AP, o);
    const double mx = Mc * ldv<R>(out, PF_VPX, o), my = Mc * ldv<R>(out, PF_VPY, o);
    const double fx = Ac * ldv<R>(out, PF_FBX, o), fy = Ac * ldv<R>(out, PF_FBY, o);
    const int lvl = in.level[i] - 1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int oc = o + c;
      ox[c] = cx[c];
      oy[c] = cy[c];
      stv<R>(out, PF_X, oc, cx[c]);
      stv<R>(out, PF_Y, oc, cy[c]);
      stv<R>(out, PF_MOMX, oc, mx);
      stv<R>(out, PF_MOMY, oc, my);
      stv<R>(out, PF_FRX, oc, fx);
      stv<R>(out, PF_FRY, oc, fy);
      out.label[oc] = -1;
      out.level[oc] = lvl;
    }
  } else {
    const int j = a.partner[i];
    const double M1 = ldv<R>(in, PF_MP, i), M2 = ldv<R>(in, PF_MP, j);
    const double Mt = M1 + M2;
    const double nx = (M1 * ldv<R>(in, PF_X, i) + M2 * ldv<R>(in, PF_X, j)) / Mt;
    const double ny = (M1 * ldv<R>(in, PF_Y, i) + M2 * ldv<R>(in, PF_Y, j)) / Mt;
    for (int k = PF_VPX; k < PF_COUNT; ++k)
      stv<R>(out, k, o,
             pf_extensive(k)
                 ? (ldv<R>(in, k, i) + ldv<R>(in, k, j))
                 : (M1 * ldv<R>(in, k, i) + M2 * ldv<R>(in, k, j)) / Mt);
    stv<R>(out, PF_X, o, nx);
    stv<R>(out, PF_Y, o, ny);
    ox[0] = nx;
    oy[0] = ny;
    const double Mp = ldv<R>(out, PF_MP, o);
    stv<R>(out, PF_MOMX, o, Mp * ldv<R>(out, PF_VPX, o));
    stv<R>(out, PF_MOMY, o, Mp * ldv<R>(out, PF_VPY, o));
    const double Ap = ldv<R>(out, PF_AP, o);
    stv<R>(out, PF_FRX, o, Ap * ldv<R>(out, PF_FBX, o));
    stv<R>(out, PF_FRY, o, Ap * ldv<R>(out, PF_FBY, o));
    out.label[o] = in.label[i];
    out.level[o] = in.level[i] + 1;
  }
}

// The new particles made from particle i (at o = offset[i]): a copy, the
// two (four: ELFS rule) children of a split, the merge of a primary with
// its partner, nothing for a secondary.  Returns how many, with their
// positions in ox, oy.
// A kept particle's fields [0, NC) copied from i to o: every load issued
// before the first store, so that the round trips overlap.  *px, *py: its
// position.
template <typename R, int NC>
__device__ __forceinline__ void copy_fields(const PPtr &in, const PPtr &out,
                                            int i, int o, double *px,
                                            double *py) {
  double vw[NC];
  R vn[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    if (pf_wide(k)) vw[k] = fld<double>(in, k)[i];
    else vn[k] = fld<R>(in, k)[i];
  }
  const int lb = in.label[i], lv = in.level[i];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    if (pf_wide(k)) fld<double>(out, k)[o] = vw[k];
    else fld<R>(out, k)[o] = vn[k];
  }
  out.label[o] = lb;
  out.level[o] = lv;
  *px = vw[PF_X];
  *py = vw[PF_Y];
}

template <typename R, bool ELFS>
__device__ __forceinline__ int execute_one(const ExecArgs &a, const Scen &sc,
                                           int i, int o, int act, double *ox,
                                           double *oy) {
  if (act == MS_KEEP) {
    if (a.ncopy == PF_COUNT) copy_fields<R, PF_COUNT>(a.in, a.out, i, o, ox, oy);
    else copy_fields<R, PF_CORE>(a.in, a.out, i, o, ox, oy);
    return 1;
  }
  if (act != MS_SPLIT && act != MS_PRIMARY) return 0;
  const double Lx = sc.P.ncols * sc.P.hx, Ly = sc.P.nrows * sc.P.hy, eps = 1e-10;
  const PPtr &in = a.in, &out = a.out;
  if (ELFS && sc.ms.mode == 1) {
    elfs_execute_one<R>(a, i, o, act, Lx, Ly, ox, oy);
    return act == MS_SPLIT ? 4 : 1;
  }
  if (act == MS_SPLIT) {
    const int o1 = o, o2 = o + 1;
    double ux, uy;
    const double exx = ldv<R>(in, PF_VXDX, i), eyy = ldv<R>(in, PF_VYDY, i);
    const double exy = 0.5 * (ldv<R>(in, PF_VXDY, i) + ldv<R>(in, PF_VYDX, i));
    principal_stretch_dir(exx, eyy, exy, &ux, &uy);
    const double r_i = sqrt(ldv<R>(in, PF_AP, i));
    const double off = 0.25 * r_i;
    const double cx = ldv<R>(in, PF_X, i), cy = ldv<R>(in, PF_Y, i);
    double x1 = cx + off * ux, y1 = cy + off * uy;
    double x2 = cx - off * ux, y2 = cy - off * uy;
    x1 = x1 < eps ? eps : (x1 > Lx - eps ? Lx - eps : x1);
    y1 = y1 < eps ? eps : (y1 > Ly - eps ? Ly - eps : y1);
    x2 = x2 < eps ? eps : (x2 > Lx - eps ? Lx - eps : x2);
    y2 = y2 < eps ? eps : (y2 > Ly - eps ? Ly - eps : y2);
    for (int k = PF_VPX; k < PF_COUNT; ++k) {
      const double val = ldv<R>(in, k, i) / (pf_extensive(k) ? 2.0 : 1.0);
      stv<R>(out, k, o1, val);
      stv<R>(out, k, o2, val);
    }
    const double h = ldv<R>(out, PF_HP, o1);
    const double Ap1 = ldv<R>(out, PF_VP, o1) / h;
    stv<R>(out, PF_AP, o1, Ap1);
    stv<R>(out, PF_AP, o2, ldv<R>(out, PF_VP, o2) / h);
    stv<R>(out, PF_X, o1, x1);
    stv<R>(out, PF_Y, o1, y1);
    stv<R>(out, PF_X, o2, x2);
    stv<R>(out, PF_Y, o2, y2);
    ox[0] = x1;
    oy[0] = y1;
    ox[1] = x2;
    oy[1] = y2;
    const double Mc = ldv<R>(out, PF_MP, o1);
    const double mx = Mc * ldv<R>(out, PF_VPX, o1);
    const double my = Mc * ldv<R>(out, PF_VPY, o1);
    stv<R>(out, PF_MOMX, o1, mx);
    stv<R>(out, PF_MOMX, o2, mx);
    stv<R>(out, PF_MOMY, o1, my);
    stv<R>(out, PF_MOMY, o2, my);
    const double Fbx = ldv<R>(out, PF_FBX, o1), Fby = ldv<R>(out, PF_FBY, o1);
    stv<R>(out, PF_FRX, o1, Ap1 * Fbx);
    stv<R>(out, PF_FRX, o2, Ap1 * Fbx);
    stv<R>(out, PF_FRY, o1, Ap1 * Fby);
    stv<R>(out, PF_FRY, o2, Ap1 * Fby);
    out.label[o1] = out.label[o2] = -1;
    out.level[o1] = out.level[o2] = in.level[i] - 1;
    return 2;
  } else {
    const int j = a.partner[i];
    const double M1 = ldv<R>(in, PF_MP, i), M2 = ldv<R>(in, PF_MP, j);
    const double Mt = M1 + M2;
    const double nx = (M1 * ldv<R>(in, PF_X, i) + M2 * ldv<R>(in, PF_X, j)) / Mt;
    const double ny = (M1 * ldv<R>(in, PF_Y, i) + M2 * ldv<R>(in, PF_Y, j)) / Mt;
    for (int k = PF_VPX; k < PF_COUNT; ++k)
      stv<R>(out, k, o,
             pf_extensive(k)
                 ? (ldv<R>(in, k, i) + ldv<R>(in, k, j))
                 : (M1 * ldv<R>(in, k, i) + M2 * ldv<R>(in, k, j)) / Mt);
    stv<R>(out, PF_HP, o, ldv<R>(out, PF_VP, o) / ldv<R>(out, PF_AP, o));
    stv<R>(out, PF_X, o, nx);
    stv<R>(out, PF_Y, o, ny);
    ox[0] = nx;
    oy[0] = ny;
    const double Mp = ldv<R>(out, PF_MP, o);
    stv<R>(out, PF_MOMX, o, Mp * ldv<R>(out, PF_VPX, o));
    stv<R>(out, PF_MOMY, o, Mp * ldv<R>(out, PF_VPY, o));
    const double Ap = ldv<R>(out, PF_AP, o);
    stv<R>(out, PF_FRX, o, Ap * ldv<R>(out, PF_FBX, o));
    stv<R>(out, PF_FRY, o, Ap * ldv<R>(out, PF_FBY, o));
    out.label[o] = -1;
    out.level[o] = in.level[i] + 1;
    return 1;
  }
}

// ELFS: some scenario of the batch uses the legacy ELFS rule (its code is
// compiled only into that variant).  One thread per particle; the new
// particles are binned as they are written (cell key and slot, the
// re-binning's first pass).
template <typename R, bool ELFS>
__global__ void ms_execute_kernel(ExecArgs a) {
  if (*a.err) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < a.T.poff[a.T.S];
  int nout = 0, o = 0, sid = 0;
  double ox[4], oy[4];
  if (live) {
    sid = scen_of(a.T, i);
    o = a.offset[i];
    nout = execute_one<R, ELFS>(a, a.T.scen[sid], i, o, a.act[i], ox, oy);
  }
  const Scen &sc = a.T.scen[sid];
  int m = 0;
#pragma unroll
  for (int c = 0; c < (ELFS ? 4 : 2); ++c) {
    const bool v = c < nout;
    int k = 0;
    if (v)
      k = sc.coff + cell_of(ox[c], oy[c], sc.P.nrows, sc.P.ncols, sc.P.hx,
                            sc.P.hy, sc.rc);
    const int sl = wave_slot(a.count, k, v);
    if (v) {
      a.key[o + c] = k;
      a.slot[o + c] = sl;
      m = sl + 1 > m ? sl + 1 : m;
    }
  }
  // the largest cell count, raised in the control block past the value
  // stored there
  if (m > a.T.ctrl[sid].max_ppc) atomicMax(&a.T.ctrl[sid].max_ppc, m);
}

// ------------------------------------------------------------- repack
// A served batch rebuilt with some scenarios kept: entry i of kept scenario
// j (found by binary search over the new offsets) comes from old entry
// src_off[j] + (i - dst_off[j]).
struct RepackMap {
  const int *dst_off;  // [nkeep+1] new offsets of the kept scenarios
  const int *src_off;  // [nkeep] their old offsets
  int nkeep;
};

__device__ __forceinline__ int repack_src(const RepackMap &m, int i) {
  int lo = 0, hi = m.nkeep - 1;  // last j with dst_off[j] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.dst_off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return m.src_off[lo] + (i - m.dst_off[lo]);
}

template <typename R>
__global__ void repack_particles_kernel(PPtr in, PPtr out, RepackMap m,
                                        const int *bits_in, int *bits_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.dst_off[m.nkeep]) return;
  const int j = repack_src(m, i);
#pragma unroll 4
  for (int k = 0; k < PF_COUNT; ++k) {
    if (pf_wide(k)) fld<double>(out, k)[i] = fld<double>(in, k)[j];
    else fld<R>(out, k)[i] = fld<R>(in, k)[j];
  }
  out.label[i] = in.label[j];
  out.level[i] = in.level[j];
  bits_out[i] = bits_in[j];  // merge/split strain bits a kept run may await
}

template <typename R>
__global__ void repack_nodes_kernel(NPtr in, NPtr out, const double *zin,
                                    double *zout, RepackMap m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.dst_off[m.nkeep]) return;
  const int j = repack_src(m, i);
#pragma unroll 4
  for (int k = 0; k < NF_COUNT; ++k) {
    if (nf_narrow(k)) static_cast<R *>(out.f[k])[i] = static_cast<const R *>(in.f[k])[j];
    else static_cast<double *>(out.f[k])[i] = static_cast<const double *>(in.f[k])[j];
  }
  zout[i] = zin[j];
}

// Control blocks of the kept scenarios, one workgroup per scenario.
__global__ void repack_ctrl_kernel(const Ctrl *in, Ctrl *out, const int *keep) {
  constexpr int W = sizeof(Ctrl) / sizeof(unsigned long long);
  const unsigned long long *src =
      reinterpret_cast<const unsigned long long *>(in + keep[blockIdx.x]);
  unsigned long long *dst = reinterpret_cast<unsigned long long *>(out + blockIdx.x);
  for (int w = threadIdx.x; w < W; w += blockDim.x) dst[w] = src[w];
}

// ------------------------------------------------------------- fills
__global__ void fill_f64_kernel(double *p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_f32_kernel(float *p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_int_kernel(int *p, long long n, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// Narrow a host-uploaded FP64 particle field to FP32.
__global__ void narrow_kernel(const double *src, float *dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}
__global__ void widen_kernel(const float *src, double *dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (double)src[i];
}

}  // namespace amsc
