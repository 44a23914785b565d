// sim.hip -- device-resident depth-averaged MPM time loop (FP64, optim_par
// semantics) on gfx950.
//
// One step (main/optim_par.cpp:281-668, merge/split at the position of
// main/gpu_offload.cpp:321-401):
//   prologue         1 thread: end test, dt (fixed or CFL from the device
//                    maxima of the previous step), t, it.
//   [merge/split]    flag / pair (merge_split.hip), weight scan, execute
//                    into the other SoA buffer, re-bin.
//   p2g_nodal        tiles of 16x16 nodes: the particles of the tile's
//                    17x17 cells (CSR ranges) are staged through LDS in
//                    chunks of 256; each thread then folds its node's
//                    contributions in ascending cell, ascending id order
//                    (the serial p2g order, hence bit-identical sums),
//                    and applies the nodal update + boundary conditions in
//                    registers.  The 9 P2G accumulators never touch HBM;
//                    only vL (and a, on output steps) are written.
//   particle         1 thread per particle: G2P, PIC velocity, advection,
//                    G2PD at the new position with the step-start cell,
//                    height update, Voellmy, Bingham, friction force, the
//                    new cell key + wave-aggregated count, CFL maxima.
//   binning          scan of the counts -> cstart, placement, in-cell sort.
// FMA contraction is off for this file: every expression is evaluated in
// the oracle's order, so results match the FP64 CPU path bit for bit.
#include "binning.hpp"
#include "common.hpp"
#include "merge_split.hpp"
#include "amsc_mpm.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

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

static const char *kPFName[PF_COUNT] = {
    "x",      "y",      "vpx",    "vpy",     "hp",      "Vp",     "Mp",
    "H",      "F_11",   "F_12",   "F_22",    "Fric_px", "Fric_py", "Ap",
    "Fb_x",   "Fb_y",   "vpx_dx", "vpx_dy",  "vpy_dx",  "vpy_dy", "apx",
    "apy",    "Zp",     "hpZ",    "dZxp",    "dZyp",    "Fpx",    "Fpy",
    "mom_px", "mom_py"};

__host__ __device__ constexpr bool pf_extensive(int k) {
  // merge_split_ops_cmes.h:49-50
  return k == PF_MP || k == PF_VP || k == PF_AP || k == PF_MOMX ||
         k == PF_MOMY;
}

enum NF : int {
  NF_VVXL, NF_VVYL,  // always
  NF_AVX, NF_AVY, NF_MV, NF_MOMX, NF_MOMY, NF_FEXTX, NF_FEXTY, NF_FINTX,
  NF_FINTY, NF_FRX, NF_FRY, NF_FPX, NF_FPY, NF_FTOTX, NF_FTOTY, NF_VVX,
  NF_VVY,
  NF_COUNT
};
static const char *kNFName[NF_COUNT] = {
    "vvxL",   "vvyL",     "avx",      "avy",      "Mv",      "mom_vx",
    "mom_vy", "F_ext_vx", "F_ext_vy", "F_int_vx", "F_int_vy", "Fric_x",
    "Fric_y", "FPxv",     "FPyv",     "Ftot_vx",  "Ftot_vy",  "vvx",
    "vvy"};

struct PPtr {
  double *f[PF_COUNT];
  int *label, *level;
};
struct NPtr {
  double *f[NF_COUNT];
};

// Device control block.
struct Ctrl {
  double t;        // time after the last step taken
  double dt;       // dt of the last step taken
  double t_start;  // time at the start of the last step taken
  double cons[2];  // err_mass, err_mom of the last step
  // max |vpx|, |vpy|, hp of the particles as bit patterns of non-negative
  // doubles, by step parity, spread over kCflSlots addresses so that the
  // per-block atomics of the particle kernel do not all hit one line
  unsigned long long cfl[2][64][3];
  int it;
  int active;
};

struct KParams {
  int nrows, ncols;
  double hx, hy, g, rho, mu, tauy, xi, tan_fa, bing, fric, cfl, eq_level,
      dt_fixed, T;
  int bc, nsteps, max_steps;
};

// ------------------------------------------------------------- kernels
namespace {

constexpr int kTC = 16;                // node columns per P2G tile
constexpr int kTR = 32;                // node rows per P2G tile
constexpr int kSeg = kTC + 1;          // cell columns (incl. halo)
constexpr int kP2GThreads = kTC * kTR; // one node per thread
constexpr int kChunk = kP2GThreads;    // particles staged per round
constexpr int kNS = 13;                // staged values per particle
// Staged particle i lives at i + i/16 in each LDS array: the node threads of
// a wavefront walk neighbouring cells in lockstep, i.e. at a stride of about
// 4 particles, which without the pad maps 16 lanes onto one bank.
constexpr int kChunkPad = kChunk + kChunk / 16;
__device__ __forceinline__ int lds_slot(int i) { return i + (i >> 4); }

// Which optional fields a step writes: none; what the next merge/split
// reads (Ap, Fb, velocity gradients); every output field and the nodal set.
enum : int { AUX_NONE = 0, AUX_MS = 1, AUX_FULL = 2 };

constexpr int kCflSlots = 64;

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(v, d);
    v = o > v ? o : v;
  }
  return v;
}

// One wavefront: end test, dt, t, it for the step about to run.
__global__ void __launch_bounds__(64) prologue_kernel(KParams P, Ctrl *ctrl) {
  const int lane = threadIdx.x;
  if (!ctrl->active) return;
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
  const unsigned long long m0 = wave_max_u64(ctrl->cfl[(s - 1) & 1][lane][0]);
  const unsigned long long m1 = wave_max_u64(ctrl->cfl[(s - 1) & 1][lane][1]);
  const unsigned long long m2 = wave_max_u64(ctrl->cfl[(s - 1) & 1][lane][2]);
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
  ctrl->cfl[s & 1][lane][0] = 0ull;
  ctrl->cfl[s & 1][lane][1] = 0ull;
  ctrl->cfl[s & 1][lane][2] = 0ull;
  if (lane == 0) {
    ctrl->dt = dt;
    ctrl->t_start = ctrl->t;
    ctrl->t = ctrl->t + dt;
    ctrl->it = s;
  }
}

// Step-0 initialisation, optim_par.cpp:235-264 (cells from the initial
// binning): Zp, hpZ, H, hydrostatic F, Fric = Ap*Fb with Fb = 0.
__global__ void init_kernel(int np, KParams P, PPtr p, const int *key,
                            const double *__restrict__ Z,
                            const double *mom_x_in, const double *mom_y_in) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const int cell = key[ip], r = cell % P.nrows, c = cell / P.nrows;
  const Local L = local_coords(p.f[PF_X][ip], p.f[PF_Y][ip], c, r, P.hx, P.hy);
  double zp = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) zp += shape(L, k) * Z[node_id(k, c, r, P.nrows)];
  const double hp = p.f[PF_HP][ip];
  p.f[PF_ZP][ip] = zp;
  p.f[PF_HPZ][ip] = hp + zp;
  const double diff = P.eq_level - zp;
  const double H = (P.eq_level > 0 && diff > 0) ? diff : 0.0;
  p.f[PF_H][ip] = H;
  const double h_corr = hp > 1e-10 ? hp - (H * H) / hp : 0.0;
  p.f[PF_F11][ip] = -.5 * P.rho * P.g * h_corr;
  p.f[PF_F12][ip] = 0.0;
  p.f[PF_F22][ip] = -.5 * P.rho * P.g * h_corr;
  p.f[PF_FBX][ip] = 0.0;
  p.f[PF_FBY][ip] = 0.0;
  p.f[PF_FRX][ip] = p.f[PF_AP][ip] * 0.0;
  p.f[PF_FRY][ip] = p.f[PF_AP][ip] * 0.0;
  const int zero_fields[] = {PF_VXDX, PF_VXDY, PF_VYDX, PF_VYDY, PF_APX,
                             PF_APY,  PF_DZX,  PF_DZY,  PF_FPX,  PF_FPY};
#pragma unroll
  for (int k = 0; k < 10; ++k) p.f[zero_fields[k]][ip] = 0.0;
  p.f[PF_MOMX][ip] = mom_x_in[ip];
  p.f[PF_MOMY][ip] = mom_y_in[ip];
  p.label[ip] = ip;
  p.level[ip] = 0;
}

// Output-only destinations of p2g_nodal (read from device memory only on
// steps that write them, which keeps the kernel's argument block small).
struct P2GOut {
  double *dzx, *dzy, *fpx, *fpy, *momx, *momy;  // particle fields
  double *nod[NF_COUNT];                         // nodal set
};

struct P2GArgs {
  int nrows, ncols, bc, aux, mom_stored;
  double hx, hy, g, rho;
  Recip rc;
  const int *cstart, *cptcls;
  // x, y, Mp, momentum source (vpx/vpy, or the DATA momentum on the first
  // step), Vp, hp, H, friction force, stress
  const double *x, *y, *Mp, *m1, *m2, *Vp, *hp, *H, *frx, *fry, *f11, *f12,
      *f22;
  const double *Z;
  const Ctrl *ctrl;
  double *vvxL, *vvyL;
  double *partials;  // [nblocks][6] or null
  const P2GOut *out;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// One particle's P2G inputs, gathered through cptcls.
struct P2GRaw {
  double x, y, Mp, m1, m2, Vp, hp, H, frx, fry, f11, f12, f22;
};
// Where a tile-list position lives: particle id, cell column segment and
// cell row within the segment.
struct P2GIdx {
  int ip, s, j;
};

__device__ __forceinline__ void p2g_fields(const P2GArgs &a, int ip,
                                           P2GRaw &R) {
  R.x = a.x[ip];
  R.y = a.y[ip];
  R.Mp = a.Mp[ip];
  R.m1 = a.m1[ip];
  R.m2 = a.m2[ip];
  R.Vp = a.Vp[ip];
  R.hp = a.hp[ip];
  R.H = a.H[ip];
  R.frx = a.frx[ip];
  R.fry = a.fry[ip];
  R.f11 = a.f11[ip];
  R.f12 = a.f12[ip];
  R.f22 = a.f22[ip];
}

// P2G + nodal update for one tile of kTC x kTR nodes.
//
// The tile's particles are the CSR ranges of its (kTC+1) x (kTR+1) cells,
// taken column segment by column segment ("tile list"), and staged through
// LDS kChunk at a time.  Each chunk goes through a two-stage pipeline: while
// chunk k is folded, the fields of chunk k+1 and the particle ids of chunk
// k+2 are in flight.  A per-chunk owner map (cell of every tile-list
// position, built cell-parallel) turns a position into its particle id with
// one LDS read.
__global__ void __launch_bounds__(kP2GThreads)
    p2g_nodal_kernel(P2GArgs a) {
  const Ctrl *ctrl = a.ctrl;
  if (!ctrl->active) return;
  const int nrows = a.nrows, ncols = a.ncols;
  const double hx = a.hx, hy = a.hy;
  const int t = threadIdx.x;
  const int c0 = blockIdx.x * kTC, r0 = blockIdx.y * kTR;
  const int lr = t % kTR, lc = t / kTR;
  const int c = c0 + lc, r = r0 + lr;
  const bool node_ok = c <= ncols && r <= nrows;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;

  __shared__ int s_cs[kSeg][kTR + 2];  // cstart, cell rows rl..rh+1
  __shared__ int s_sego[kSeg + 1];     // tile-list offset per column
  __shared__ double s_Z[(kTC + 2) * (kTR + 2)];
  __shared__ double s_p[kNS][kChunkPad];
  __shared__ unsigned short s_own[kChunk];
  __shared__ double s_red[6][kP2GThreads / 64];

  const int ccl = c0 - 1 < 0 ? 0 : c0 - 1;
  const int cch = c0 + kTC - 1 > ncols - 1 ? ncols - 1 : c0 + kTC - 1;
  const int rl = r0 - 1 < 0 ? 0 : r0 - 1;
  const int rh = r0 + kTR - 1 > nrows - 1 ? nrows - 1 : r0 + kTR - 1;
  const int nseg = cch - ccl + 1, nr = rh - rl + 1;

  for (int i = t; i < nseg * (nr + 1); i += kP2GThreads) {
    const int s = i / (nr + 1), j = i % (nr + 1);
    s_cs[s][j] = a.cstart[(ccl + s) * nrows + rl + j];
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

  // owner map of the chunk starting at base (positions base .. base+kChunk)
  auto build_own = [&](int base) {
    if (base >= L) return;
    for (int ci = t; ci < nseg * nr; ci += kP2GThreads) {
      const int s = ci / nr, j = ci % nr;
      const int b = s_sego[s] + (s_cs[s][j] - s_cs[s][0]);
      const int e = b + (s_cs[s][j + 1] - s_cs[s][j]);
      const int lo = b > base ? b : base;
      const int hi = e < base + kChunk ? e : base + kChunk;
      for (int q = lo; q < hi; ++q)
        s_own[q - base] = (unsigned short)((s << 8) | j);
    }
  };
  auto fetch_idx = [&](int q, P2GIdx &I) {
    const int v = s_own[q & (kChunk - 1)];
    I.s = v >> 8;
    I.j = v & 0xFF;
    I.ip = a.cptcls[s_cs[I.s][0] + (q - s_sego[I.s])];
  };

  double pm = 0, ppx = 0, ppy = 0;  // particle sums (owned cells)
  double Mv = 0, mvx = 0, mvy = 0, FPx = 0, FPy = 0, Frx = 0, Fry = 0,
         Fix = 0, Fiy = 0;

  if (L > 0) {
    P2GIdx I1, I2, I3;
    P2GRaw R;
    build_own(0);
    // Z of nodes (c0-1 .. c0+kTC, r0-1 .. r0+kTR)
    for (int i = t; i < (kTC + 2) * (kTR + 2); i += kP2GThreads) {
      const int zc = c0 - 1 + i / (kTR + 2), zr = r0 - 1 + i % (kTR + 2);
      s_Z[i] = (zc >= 0 && zc <= ncols && zr >= 0 && zr <= nrows)
                   ? a.Z[zr + zc * (nrows + 1)]
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
    if (t < L) fetch_idx(t, I1);
    __syncthreads();
    build_own(kChunk);
    __syncthreads();
    if (kChunk + t < L) fetch_idx(kChunk + t, I2);
    if (t < L) p2g_fields(a, I1.ip, R);
    __syncthreads();  // s_own is read above and rebuilt in the loop

    for (int base = 0; base < L; base += kChunk) {
      if (base + t < L) {
        // the particle's cell is the CSR cell it was found in
        const int pc = ccl + I1.s, pr = rl + I1.j;
        const Local Lc = local_coords(R.x, R.y, pc, pr, hx, hy);
        // grad Z at the particle (g2pd(Z), optim_par.cpp:387)
        const int zb = (pc - (c0 - 1)) * (kTR + 2) + (pr - (r0 - 1));
        const double z0 = s_Z[zb], z1 = s_Z[zb + 1];
        const double z2 = s_Z[zb + (kTR + 2)], z3 = s_Z[zb + (kTR + 3)];
        double dzx = 0.0, dzy = 0.0;
        dzx += shape_dx(Lc, 0, a.rc) * z0;
        dzy += shape_dy(Lc, 0, a.rc) * z0;
        dzx += shape_dx(Lc, 1, a.rc) * z1;
        dzy += shape_dy(Lc, 1, a.rc) * z1;
        dzx += shape_dx(Lc, 2, a.rc) * z2;
        dzy += shape_dy(Lc, 2, a.rc) * z2;
        dzx += shape_dx(Lc, 3, a.rc) * z3;
        dzy += shape_dy(Lc, 3, a.rc) * z3;
        // gravity along the slope with the well-balance factor,
        // optim_par.cpp:450-453
        const double wb = R.hp > 1e-10 ? 1.0 - (R.H / R.hp) : 1.0;
        const double Fpx = -a.g * R.Vp * a.rho * dzx * wb;
        const double Fpy = -a.g * R.Vp * a.rho * dzy * wb;
        // mom_p = vpx * Mp (optim_par.cpp:600-601), DATA momentum at step 1
        const double mx = a.mom_stored ? R.m1 : R.m1 * R.Mp;
        const double my = a.mom_stored ? R.m2 : R.m2 * R.Mp;
        const int ts = lds_slot(t);
        s_p[0][ts] = Lc.xi;
        s_p[1][ts] = Lc.eta;
        s_p[2][ts] = R.Mp;
        s_p[3][ts] = mx;
        s_p[4][ts] = my;
        s_p[5][ts] = Fpx;
        s_p[6][ts] = Fpy;
        s_p[7][ts] = R.frx;
        s_p[8][ts] = R.fry;
        s_p[9][ts] = R.f11;
        s_p[10][ts] = R.f12;
        s_p[11][ts] = R.f22;
        s_p[12][ts] = R.Vp;
        const bool owned = pc >= c0 && pc < c0 + kTC && pr >= r0 && pr < r0 + kTR;
        if (owned) {
          pm += R.Mp;
          ppx += mx;
          ppy += my;
          if (a.aux >= AUX_FULL) {
            const P2GOut *o = a.out;
            o->dzx[I1.ip] = dzx;
            o->dzy[I1.ip] = dzy;
            o->fpx[I1.ip] = Fpx;
            o->fpy[I1.ip] = Fpy;
            o->momx[I1.ip] = mx;
            o->momy[I1.ip] = my;
          }
        }
      }
      build_own(base + 2 * kChunk);
      __syncthreads();
      // in flight during the fold: ids of chunk k+2, fields of chunk k+1
      if (base + 2 * kChunk + t < L) fetch_idx(base + 2 * kChunk + t, I3);
      if (base + kChunk + t < L) p2g_fields(a, I2.ip, R);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int lo = tb[k] > base ? tb[k] : base;
        const int hi = te[k] < base + kChunk ? te[k] : base + kChunk;
        const int corner = 3 - k;
        for (int qq = lo; qq < hi; ++qq) {
          const int i = lds_slot(qq - base);
          Local Lq;
          Lq.xi = s_p[0][i];
          Lq.eta = s_p[1][i];
          const double N = shape(Lq, corner);
          Mv += N * s_p[2][i];
          mvx += N * s_p[3][i];
          mvy += N * s_p[4][i];
          FPx += N * s_p[5][i];
          FPy += N * s_p[6][i];
          Frx += N * s_p[7][i];
          Fry += N * s_p[8][i];
          const double Nx = shape_dx(Lq, corner, a.rc);
          const double Ny = shape_dy(Lq, corner, a.rc);
          const double vp = s_p[12][i];
          Fix += (Nx * s_p[9][i] + Ny * s_p[10][i]) * vp;
          Fiy += (Nx * s_p[10][i] + Ny * s_p[11][i]) * vp;
        }
      }
      __syncthreads();
      I1 = I2;
      I2 = I3;
    }
  }

  if (node_ok && (L > 0 || a.aux >= AUX_FULL)) {
    const int n = r + c * (nrows + 1);
    const double dt = ctrl->dt;
    // force balance and nodal velocities, optim_par.cpp:467-497 (D1)
    const double Fex = FPx + Frx, Fey = FPy + Fry;
    const double Ftx = Fex - Fix, Fty = Fey - Fiy;
    double mx = dt * Ftx + mvx, my = dt * Fty + mvy;
    double ax = Mv > 1.e-2 ? Ftx / Mv : 0.0;
    double ay = Mv > 1.e-2 ? Fty / Mv : 0.0;
    double vx = Mv > 1.e-2 ? mx / Mv : 0.0;
    double vy = Mv > 1.e-2 ? my / Mv : 0.0;
    double vxL = dt * ax + vx, vyL = dt * ay + vy;
    // boundary conditions, optim_par.cpp:503-529 (nodes of the first/last
    // cell column: x components; first/last cell row: y components)
    if (a.bc) {
      if (c == 0 || c == 1 || c == ncols - 1 || c == ncols) {
        ax = 0.0; vx = 0.0; vxL = 0.0; mx = 0.0;
      }
      if (r == 0 || r == 1 || r == nrows - 1 || r == nrows) {
        ay = 0.0; vy = 0.0; vyL = 0.0; my = 0.0;
      }
    }
    a.vvxL[n] = vxL;
    a.vvyL[n] = vyL;
    if (a.aux >= AUX_FULL) {
      double *const *nod = a.out->nod;
      nod[NF_AVX][n] = ax;
      nod[NF_AVY][n] = ay;
      nod[NF_MV][n] = Mv;
      nod[NF_MOMX][n] = mx;
      nod[NF_MOMY][n] = my;
      nod[NF_FEXTX][n] = Fex;
      nod[NF_FEXTY][n] = Fey;
      nod[NF_FINTX][n] = Fix;
      nod[NF_FINTY][n] = Fiy;
      nod[NF_FRX][n] = Frx;
      nod[NF_FRY][n] = Fry;
      nod[NF_FPX][n] = FPx;
      nod[NF_FPY][n] = FPy;
      nod[NF_FTOTX][n] = Ftx;
      nod[NF_FTOTY][n] = Fty;
      nod[NF_VVX][n] = vx;
      nod[NF_VVY][n] = vy;
    }
  }

  if (a.partials) {
    // deterministic block sums for the conservation check
    double v[6] = {pm, ppx, ppy, node_ok ? Mv : 0.0, node_ok ? mvx : 0.0,
                   node_ok ? mvy : 0.0};
    const int lane = t & 63, w = t >> 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double sm = wave_sum(v[k]);
      if (lane == 0) s_red[k][w] = sm;
    }
    __syncthreads();
    if (t < 6) {
      double sm = 0.0;
      for (int ww = 0; ww < kP2GThreads / 64; ++ww) sm += s_red[t][ww];
      a.partials[bid * 6 + t] = sm;
    }
  }
}

// err_mass, err_mom from the block partials, optim_par.cpp:416-425.
__global__ void conservation_kernel(const double *partials, int nb,
                                    Ctrl *ctrl) {
  __shared__ double s[6][256];
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += 256)
    for (int k = 0; k < 6; ++k) v[k] += partials[b * 6 + k];
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 6; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && ctrl->active) {
    const double tm = s[0][0], tx = s[1][0], ty = s[2][0];
    const double nm = s[3][0], nx = s[4][0], ny = s[5][0];
    const double am = fabs(tx) > fabs(ty) ? fabs(tx) : fabs(ty);
    const double norm = 1.0 > am ? 1.0 : am;
    const double ex = fabs(tx - nx) / norm, ey = fabs(ty - ny) / norm;
    ctrl->cons[0] = fabs(tm - nm) / tm;
    ctrl->cons[1] = ex > ey ? ex : ey;
  }
}

struct PartArgs {
  KParams P;
  Recip rc;
  FastDiv fd;  // division by nrows
  int np;
  PPtr p;
  const int *key_in;  // step-start cell of each particle
  int *key_out, *slot, *count;
  const double *vvxL, *vvyL, *avx, *avy;
  const double *Z;
  Ctrl *ctrl;
  int aux;
};

template <int AUX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
particle_kernel(PartArgs a) {
  Ctrl *ctrl = a.ctrl;
  const KParams &P = a.P;
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctrl->active) {
    // step after the end of the run: carry the keys over unchanged
    if (ip < a.np) a.key_out[ip] = a.key_in[ip];
    return;
  }
  const bool valid = ip < a.np;
  const double dt = ctrl->dt;
  int newkey = 0;
  double vmx = 0.0, vmy = 0.0, hmx = 0.0;
  if (valid) {
    const int nrows = P.nrows;
    const double hx = P.hx, hy = P.hy;
    const int cell = a.key_in[ip], c = fdiv(cell, a.fd), r = cell - c * nrows;
    int nid[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) nid[k] = node_id(k, c, r, nrows);
    double gx[4], gy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gx[k] = a.vvxL[nid[k]];
      gy[k] = a.vvyL[nid[k]];
    }
    const double x = a.p.f[PF_X][ip], y = a.p.f[PF_Y][ip];
    const Local L0 = local_coords(x, y, c, r, hx, hy);
    // g2p(vvxL, vvyL), optim_par.cpp:542
    double vxL = 0.0, vyL = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double N = shape(L0, k);
      vxL += N * gx[k];
      vyL += N * gy[k];
    }
    // PIC velocity, advection and clamp, optim_par.cpp:548-564
    const double vpx = vxL, vpy = vyL;
    double xn = x + dt * vpx, yn = y + dt * vpy;
    const double eps = 1e-10, Lx = P.ncols * P.hx, Ly = P.nrows * P.hy;
    if (xn < eps) xn = eps;
    if (xn > Lx - eps) xn = Lx - eps;
    if (yn < eps) yn = eps;
    if (yn > Ly - eps) yn = Ly - eps;
    // g2pd(vvxL, vvyL) at the new position, step-start cell (D18),
    // optim_par.cpp:578
    const Local L1 = local_coords(xn, yn, c, r, hx, hy);
    double vxdx = 0.0, vxdy = 0.0, vydx = 0.0, vydy = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double Nx = shape_dx(L1, k, a.rc), Ny = shape_dy(L1, k, a.rc);
      vxdx += Nx * gx[k];
      vxdy += Ny * gx[k];
      vydx += Nx * gy[k];
      vydy += Ny * gy[k];
    }
    // height update with floor, area, optim_par.cpp:595-607
    const double div = vxdx + vydy;
    double h = a.p.f[PF_HP][ip] / (1 + dt * div);
    h = h < 1e-2 ? 1e-2 : h;
    const double Vp = a.p.f[PF_VP][ip];
    const double Ap = h > 1.e-2 ? Vp / h : 0.0;
    // Voellmy basal friction, optim_par.cpp:613-626
    double Fbx = 0.0, Fby = 0.0;
    {
      const double nv = sqrt(vpx * vpx + vpy * vpy);
      if (P.fric > 0 && nv > 1e-10 && P.xi > 0) {
        const double v2 = vpx * vpx + vpy * vpy;
        Fbx = -P.fric * (P.rho * P.g * h * P.tan_fa + P.rho * P.g * v2 / P.xi) * vpx / nv;
        Fby = -P.fric * (P.rho * P.g * h * P.tan_fa + P.rho * P.g * v2 / P.xi) * vpy / nv;
      }
    }
    // Bingham stress, stress_tensor_t::operator(), optim_par.cpp:68-120
    double F11, F12, F22;
    {
      const double A = 3. / 2., C = 65. / 32.;
      const double mu = P.mu, tau_Y = P.tauy;
      const double nrm = sqrt(vpx * vpx + vpy * vpy);
      const double ALF = h > 1.e-3 ? (6. * mu * nrm) / ((h + 0.001) * tau_Y) : 0.0;
      const double B = -114. / 32. - ALF;
      // Z1 = (-B + sq)/(2A) is selected only if it lies in [0, 1]; when
      // -B + sq > 3.0000001 it cannot (2A = 3), so Z2 is taken without
      // evaluating Z1.  Same values as computing both.
      const double sq = sqrt(B * B - 4. * A * C);
      const double num1 = -B + sq;
      double ZZ;
      if (num1 > 3.0000001) {
        ZZ = (-B - sq) / (2. * A);
      } else {
        const double Z1 = num1 / (2. * A);
        ZZ = fabs(Z1 - .5) <= .5 ? Z1 : (-B - sq) / (2. * A);
      }
      const double D_xx = vxdx;
      const double D_xy = 0.5 * (vxdy + vydx);
      const double D_yy = vydy;
      const double D_zx = h > 1.e-3 ? 0.5 * (3. / (2. + ZZ)) * (vpx / (h + 0.001)) : 0.0;
      const double D_zy = h > 1.e-3 ? 0.5 * (3. / (2. + ZZ)) * (vpy / (h + 0.001)) : 0.0;
      const double D_zz = -(vxdx + vydy);
      const double invII = 0.5 * (D_xx * D_xx + D_yy * D_yy + D_zz * D_zz) +
                           D_zx * D_zx + D_zy * D_zy + D_xy * D_xy;
      const double sig_xx = invII != 0 ? (tau_Y / sqrt(invII) + 2. * mu) * D_xx : 0.0;
      const double sig_xy = invII != 0 ? (tau_Y / sqrt(invII) + 2. * mu) * D_xy : 0.0;
      const double sig_yy = invII != 0 ? (tau_Y / sqrt(invII) + 2. * mu) * D_yy : 0.0;
      const double H = a.p.f[PF_H][ip];
      const double h_corr = h > 1e-10 ? h - (H * H) / h : 0.0;
      F11 = P.bing * sig_xx - .5 * P.rho * P.g * h_corr;
      F12 = P.bing * sig_xy;
      F22 = P.bing * sig_yy - .5 * P.rho * P.g * h_corr;
    }
    a.p.f[PF_X][ip] = xn;
    a.p.f[PF_Y][ip] = yn;
    a.p.f[PF_VPX][ip] = vpx;
    a.p.f[PF_VPY][ip] = vpy;
    a.p.f[PF_HP][ip] = h;
    a.p.f[PF_F11][ip] = F11;
    a.p.f[PF_F12][ip] = F12;
    a.p.f[PF_F22][ip] = F22;
    // friction force for the next P2G, optim_par.cpp:456-459
    a.p.f[PF_FRX][ip] = Ap * Fbx;
    a.p.f[PF_FRY][ip] = Ap * Fby;
    if (AUX >= AUX_MS) {
      a.p.f[PF_AP][ip] = Ap;
      a.p.f[PF_FBX][ip] = Fbx;
      a.p.f[PF_FBY][ip] = Fby;
      a.p.f[PF_VXDX][ip] = vxdx;
      a.p.f[PF_VXDY][ip] = vxdy;
      a.p.f[PF_VYDX][ip] = vydx;
      a.p.f[PF_VYDY][ip] = vydy;
    }
    if (AUX >= AUX_FULL) {
      double ax = 0.0, ay = 0.0, zp = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double N = shape(L0, k);
        ax += N * a.avx[nid[k]];
        ay += N * a.avy[nid[k]];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) zp += shape(L1, k) * a.Z[nid[k]];
      a.p.f[PF_APX][ip] = ax;
      a.p.f[PF_APY][ip] = ay;
      a.p.f[PF_ZP][ip] = zp;
      a.p.f[PF_HPZ][ip] = h + zp;
      const double Mp = a.p.f[PF_MP][ip];
      a.p.f[PF_MOMX][ip] = vpx * Mp;
      a.p.f[PF_MOMY][ip] = vpy * Mp;
    }
    newkey = cell_of(xn, yn, P.nrows, P.ncols, hx, hy);
    vmx = fabs(vpx);
    vmy = fabs(vpy);
    hmx = h > 0.0 ? h : 0.0;
  }
  const int sl = wave_slot(a.count, newkey, valid);
  if (valid) {
    a.key_out[ip] = newkey;
    a.slot[ip] = sl;
  }
  // CFL maxima: wave, then block, then one atomic per block into one of
  // kCflSlots slots (non-negative doubles order like their bit patterns)
  __shared__ double s_cfl[3][4];
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
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      vmx = fmax(vmx, s_cfl[0][k]);
      vmy = fmax(vmy, s_cfl[1][k]);
      hmx = fmax(hmx, s_cfl[2][k]);
    }
    unsigned long long *m = ctrl->cfl[ctrl->it & 1][blockIdx.x & (kCflSlots - 1)];
    atomicMax(&m[0], (unsigned long long)__double_as_longlong(vmx));
    atomicMax(&m[1], (unsigned long long)__double_as_longlong(vmy));
    atomicMax(&m[2], (unsigned long long)__double_as_longlong(hmx));
  }
}

// Merge/split execution, merge_split_ops_cmes.h:258-427.
struct ExecArgs {
  int np, nrows, ncols;
  double hx, hy;
  PPtr in, out;
  const int *act, *partner, *offset;
};

__global__ void ms_execute_kernel(ExecArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const int act = a.act[i];
  const int o = a.offset[i];
  const double Lx = a.ncols * a.hx, Ly = a.nrows * a.hy, eps = 1e-10;
  if (act == MS_KEEP) {
#pragma unroll 4
    for (int k = 0; k < PF_COUNT; ++k) a.out.f[k][o] = a.in.f[k][i];
    a.out.label[o] = a.in.label[i];
    a.out.level[o] = a.in.level[i];
  } else if (act == MS_SPLIT) {
    const int o1 = o, o2 = o + 1;
    double ux, uy;
    const double exx = a.in.f[PF_VXDX][i], eyy = a.in.f[PF_VYDY][i];
    const double exy = 0.5 * (a.in.f[PF_VXDY][i] + a.in.f[PF_VYDX][i]);
    principal_stretch_dir(exx, eyy, exy, &ux, &uy);
    const double r_i = sqrt(a.in.f[PF_AP][i]);
    const double off = 0.25 * r_i;
    const double cx = a.in.f[PF_X][i], cy = a.in.f[PF_Y][i];
    double x1 = cx + off * ux, y1 = cy + off * uy;
    double x2 = cx - off * ux, y2 = cy - off * uy;
    x1 = x1 < eps ? eps : (x1 > Lx - eps ? Lx - eps : x1);
    y1 = y1 < eps ? eps : (y1 > Ly - eps ? Ly - eps : y1);
    x2 = x2 < eps ? eps : (x2 > Lx - eps ? Lx - eps : x2);
    y2 = y2 < eps ? eps : (y2 > Ly - eps ? Ly - eps : y2);
    for (int k = PF_VPX; k < PF_COUNT; ++k) {
      const double val = a.in.f[k][i] / (pf_extensive(k) ? 2.0 : 1.0);
      a.out.f[k][o1] = val;
      a.out.f[k][o2] = val;
    }
    const double h = a.out.f[PF_HP][o1];
    const double Ap1 = a.out.f[PF_VP][o1] / h;
    a.out.f[PF_AP][o1] = Ap1;
    a.out.f[PF_AP][o2] = a.out.f[PF_VP][o2] / h;
    a.out.f[PF_X][o1] = x1; a.out.f[PF_Y][o1] = y1;
    a.out.f[PF_X][o2] = x2; a.out.f[PF_Y][o2] = y2;
    const double Mc = a.out.f[PF_MP][o1];
    a.out.f[PF_MOMX][o1] = a.out.f[PF_MOMX][o2] = Mc * a.out.f[PF_VPX][o1];
    a.out.f[PF_MOMY][o1] = a.out.f[PF_MOMY][o2] = Mc * a.out.f[PF_VPY][o1];
    const double Fbx = a.out.f[PF_FBX][o1], Fby = a.out.f[PF_FBY][o1];
    a.out.f[PF_FRX][o1] = a.out.f[PF_FRX][o2] = Ap1 * Fbx;
    a.out.f[PF_FRY][o1] = a.out.f[PF_FRY][o2] = Ap1 * Fby;
    a.out.label[o1] = a.out.label[o2] = -1;
    a.out.level[o1] = a.out.level[o2] = a.in.level[i] - 1;
  } else if (act == MS_PRIMARY) {
    const int j = a.partner[i];
    const double M1 = a.in.f[PF_MP][i], M2 = a.in.f[PF_MP][j], Mt = M1 + M2;
    const double nx = (M1 * a.in.f[PF_X][i] + M2 * a.in.f[PF_X][j]) / Mt;
    const double ny = (M1 * a.in.f[PF_Y][i] + M2 * a.in.f[PF_Y][j]) / Mt;
    for (int k = PF_VPX; k < PF_COUNT; ++k)
      a.out.f[k][o] = pf_extensive(k)
                          ? (a.in.f[k][i] + a.in.f[k][j])
                          : (M1 * a.in.f[k][i] + M2 * a.in.f[k][j]) / Mt;
    a.out.f[PF_HP][o] = a.out.f[PF_VP][o] / a.out.f[PF_AP][o];
    a.out.f[PF_X][o] = nx;
    a.out.f[PF_Y][o] = ny;
    const double Mp = a.out.f[PF_MP][o];
    a.out.f[PF_MOMX][o] = Mp * a.out.f[PF_VPX][o];
    a.out.f[PF_MOMY][o] = Mp * a.out.f[PF_VPY][o];
    const double Ap = a.out.f[PF_AP][o];
    a.out.f[PF_FRX][o] = Ap * a.out.f[PF_FBX][o];
    a.out.f[PF_FRY][o] = Ap * a.out.f[PF_FBY][o];
    a.out.label[o] = -1;
    a.out.level[o] = a.in.level[i] + 1;
  }
}

__global__ void weights_kernel(int np, const int *act, int *w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int a = act[i];
  w[i] = a == MS_KEEP ? 1 : a == MS_SPLIT ? 2 : a == MS_PRIMARY ? 1 : 0;
}

__global__ void fill_f64_kernel(double *p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

// ------------------------------------------------------------- host side
struct TimedLaunch {
  const char *name;
  hipEvent_t a, b;
};

struct Sim {
  amsc_params p{};
  KParams kp{};
  int np = 0, cap = 0, nn = 0, nc = 0;
  hipStream_t st = nullptr;
  PPtr buf[2]{};
  int cur = 0;
  int *key[2] = {nullptr, nullptr};
  int kcur = 0;
  int *slot = nullptr, *cptcls = nullptr;
  Binning bin;
  NPtr nod{};
  double *Z = nullptr;
  Ctrl *ctrl = nullptr;
  Ctrl *ctrl_host = nullptr;  // pinned
  double *partials = nullptr;
  P2GOut *p2g_out[2] = {nullptr, nullptr};  // device copies, one per buffer
  int nblk_x = 0, nblk_y = 0;
  // merge/split workspace
  int *act = nullptr, *partner = nullptr, *wts = nullptr, *offs = nullptr;
  int *ms_cnt = nullptr;
  MsParams msp{};
  int ms_counts[2] = {0, 0};
  // host mirrors
  int it_host = 0;
  long long particle_steps = 0;  // sum of np over launched active steps
  bool finished = false;
  bool aux_on = false;
  bool diag = true;
  bool mom_stored = false;
  bool aux_valid = true;   // aux fields current for the present state
  bool nodal_valid = false;
  bool timing = false;
  std::vector<TimedLaunch> launches;
  std::vector<hipEvent_t> event_pool;
  size_t event_used = 0;
};

namespace {

int ev_get(Sim &s, hipEvent_t *e) {
  if (s.event_used < s.event_pool.size()) {
    *e = s.event_pool[s.event_used++];
    return AMSC_OK;
  }
  AMSC_HIP(hipEventCreate(e));
  s.event_pool.push_back(*e);
  s.event_used++;
  return AMSC_OK;
}

// Bracket a launch with events when timing is on.
struct Timer {
  Sim &s;
  TimedLaunch tl{};
  bool on;
  Timer(Sim &sim, const char *name) : s(sim), on(sim.timing) {
    if (on) {
      tl.name = name;
      if (ev_get(s, &tl.a) || ev_get(s, &tl.b)) {
        on = false;
        return;
      }
      (void)hipEventRecord(tl.a, s.st);
    }
  }
  ~Timer() {
    if (on) {
      (void)hipEventRecord(tl.b, s.st);
      s.launches.push_back(tl);
    }
  }
};

PPtr &cur_buf(Sim &s) { return s.buf[s.cur]; }

int alloc_ppt(PPtr &p, int cap) {
  for (int k = 0; k < PF_COUNT; ++k)
    if (hipMalloc(&p.f[k], sizeof(double) * cap) != hipSuccess)
      return fail(AMSC_ERR_NOMEM, "particle array allocation failed");
  if (hipMalloc(&p.label, sizeof(int) * cap) != hipSuccess ||
      hipMalloc(&p.level, sizeof(int) * cap) != hipSuccess)
    return fail(AMSC_ERR_NOMEM, "particle array allocation failed");
  return AMSC_OK;
}

void free_ppt(PPtr &p) {
  for (int k = 0; k < PF_COUNT; ++k)
    if (p.f[k]) (void)hipFree(p.f[k]);
  if (p.label) (void)hipFree(p.label);
  if (p.level) (void)hipFree(p.level);
  p = PPtr();
}

void sim_free(Sim *s) {
  if (!s) return;
  if (s->st) (void)hipStreamSynchronize(s->st);
  free_ppt(s->buf[0]);
  free_ppt(s->buf[1]);
  for (int i = 0; i < 2; ++i)
    if (s->key[i]) (void)hipFree(s->key[i]);
  int *ip[] = {s->slot, s->cptcls, s->act, s->partner, s->wts, s->offs, s->ms_cnt};
  for (int *q : ip)
    if (q) (void)hipFree(q);
  binning_free(s->bin);
  for (int k = 0; k < NF_COUNT; ++k)
    if (s->nod.f[k]) (void)hipFree(s->nod.f[k]);
  if (s->Z) (void)hipFree(s->Z);
  if (s->ctrl) (void)hipFree(s->ctrl);
  if (s->ctrl_host) (void)hipHostFree(s->ctrl_host);
  if (s->partials) (void)hipFree(s->partials);
  for (int b = 0; b < 2; ++b)
    if (s->p2g_out[b]) (void)hipFree(s->p2g_out[b]);
  for (auto e : s->event_pool) (void)hipEventDestroy(e);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

KParams kparams_from(const amsc_params &p) {
  KParams k;
  k.nrows = p.nrows;
  k.ncols = p.ncols;
  k.hx = p.hx;
  k.hy = p.hy;
  k.g = p.g;
  k.rho = p.rho;
  k.mu = p.mu;
  k.tauy = p.tauy;
  k.xi = p.xi;
  // fric_ang = phi * M_PI / 180 (optim_par.cpp:193), tan on the host
  k.tan_fa = std::tan(p.phi * M_PI / 180.);
  k.bing = p.bingham_on;
  k.fric = p.friction_on;
  k.cfl = p.cfl;
  k.eq_level = p.eq_level;
  k.dt_fixed = p.dt_fixed;
  k.T = p.T;
  k.bc = p.bc_flag;
  k.nsteps = p.nsteps;
  k.max_steps = p.max_steps;
  return k;
}

// Bin the current particle set into key[kcur], cptcls, cstart.
int rebin(Sim &s) {
  PPtr &p = cur_buf(s);
  {
    Timer tm(s, "key_count");
    if (int rc = launch_key_count(s.np, p.f[PF_X], p.f[PF_Y], s.p.nrows,
                                  s.p.ncols, s.p.hx, s.p.hy, s.key[s.kcur],
                                  s.slot, s.bin.count, nullptr, s.st))
      return rc;
  }
  Timer tm(s, "binning");
  return launch_finish_binning(s.bin, s.np, s.key[s.kcur], s.slot, s.cptcls,
                               nullptr, s.st);
}

int read_ctrl(Sim &s) {
  AMSC_HIP(hipMemcpyAsync(s.ctrl_host, s.ctrl, sizeof(Ctrl),
                          hipMemcpyDeviceToHost, s.st));
  AMSC_HIP(hipStreamSynchronize(s.st));
  return AMSC_OK;
}

int merge_split(Sim &s) {
  PPtr &p = cur_buf(s);
  MsInputs in = {p.f[PF_X],    p.f[PF_Y],    p.f[PF_AP],   p.f[PF_HP],
                 p.f[PF_VXDX], p.f[PF_VXDY], p.f[PF_VYDX], p.f[PF_VYDY],
                 p.f[PF_VPX],  p.f[PF_VPY],  p.level};
  int counts[2];
  {
    Timer tm(s, "ms_decide");
    if (int rc = ms_decide(s.msp, s.np, s.nc, &s.ctrl->dt, 0.0, in,
                           s.bin.cstart, s.cptcls, s.act, s.partner,
                           s.ms_cnt, counts, s.st))
      return rc;
  }
  s.ms_counts[0] = counts[0];
  s.ms_counts[1] = counts[1];
  if (counts[0] == 0 && counts[1] == 0) return AMSC_OK;  // no-op
  int nnew = 0;
  {
    Timer tm(s, "ms_execute");
    weights_kernel<<<grid_for(s.np, 256), 256, 0, s.st>>>(s.np, s.act, s.wts);
    AMSC_LAUNCH_CHECK();
    if (int rc = exclusive_scan(s.wts, s.offs, s.np, false, s.bin.scan, s.st,
                                &nnew))
      return rc;
    if (nnew > s.cap)
      return fail(AMSC_ERR_CAPACITY,
                  "merge/split needs " + std::to_string(nnew) +
                      " particles, capacity is " + std::to_string(s.cap));
    ExecArgs ea;
    ea.np = s.np;
    ea.nrows = s.p.nrows;
    ea.ncols = s.p.ncols;
    ea.hx = s.p.hx;
    ea.hy = s.p.hy;
    ea.in = s.buf[s.cur];
    ea.out = s.buf[1 - s.cur];
    ea.act = s.act;
    ea.partner = s.partner;
    ea.offset = s.offs;
    ms_execute_kernel<<<grid_for(s.np, 256), 256, 0, s.st>>>(ea);
    AMSC_LAUNCH_CHECK();
  }
  s.cur = 1 - s.cur;
  s.np = nnew;
  return rebin(s);
}

int launch_step(Sim &s) {
  const bool ms_step = s.p.ms_on && s.p.ms_interval > 0 &&
                       ((s.it_host + 1) % s.p.ms_interval == 0);
  const bool ms_next = s.p.ms_on && s.p.ms_interval > 0 &&
                       ((s.it_host + 2) % s.p.ms_interval == 0);
  const int aux = s.aux_on ? AUX_FULL : (ms_next ? AUX_MS : AUX_NONE);
  {
    Timer tm(s, "prologue");
    prologue_kernel<<<1, 64, 0, s.st>>>(s.kp, s.ctrl);
    AMSC_LAUNCH_CHECK();
  }
  if (ms_step) {
    if (int rc = read_ctrl(s)) return rc;
    if (!s.ctrl_host->active) {
      s.finished = true;
      return AMSC_OK;
    }
    if (int rc = merge_split(s)) return rc;
  }
  PPtr &p = cur_buf(s);
  {
    Timer tm(s, "p2g_nodal");
    P2GArgs a;
    a.nrows = s.p.nrows;
    a.ncols = s.p.ncols;
    a.bc = s.p.bc_flag;
    a.aux = aux;
    a.mom_stored = s.mom_stored ? 1 : 0;
    a.hx = s.p.hx;
    a.hy = s.p.hy;
    a.g = s.p.g;
    a.rho = s.p.rho;
    a.rc = make_recip(s.p.hx, s.p.hy);
    a.cstart = s.bin.cstart;
    a.cptcls = s.cptcls;
    a.x = p.f[PF_X];
    a.y = p.f[PF_Y];
    a.Mp = p.f[PF_MP];
    a.m1 = s.mom_stored ? p.f[PF_MOMX] : p.f[PF_VPX];
    a.m2 = s.mom_stored ? p.f[PF_MOMY] : p.f[PF_VPY];
    a.Vp = p.f[PF_VP];
    a.hp = p.f[PF_HP];
    a.H = p.f[PF_H];
    a.frx = p.f[PF_FRX];
    a.fry = p.f[PF_FRY];
    a.f11 = p.f[PF_F11];
    a.f12 = p.f[PF_F12];
    a.f22 = p.f[PF_F22];
    a.Z = s.Z;
    a.ctrl = s.ctrl;
    a.vvxL = s.nod.f[NF_VVXL];
    a.vvyL = s.nod.f[NF_VVYL];
    a.partials = s.diag ? s.partials : nullptr;
    a.out = s.p2g_out[s.cur];
    dim3 grid(s.nblk_x, s.nblk_y);
    p2g_nodal_kernel<<<grid, kP2GThreads, 0, s.st>>>(a);
    AMSC_LAUNCH_CHECK();
  }
  if (s.diag) {
    Timer tm(s, "conservation");
    conservation_kernel<<<1, 256, 0, s.st>>>(s.partials, s.nblk_x * s.nblk_y,
                                             s.ctrl);
    AMSC_LAUNCH_CHECK();
  }
  s.mom_stored = false;
  {
    Timer tm(s, "particle");
    PartArgs a;
    a.P = s.kp;
    a.rc = make_recip(s.p.hx, s.p.hy);
    a.fd = make_fastdiv(s.p.nrows);
    a.np = s.np;
    a.p = p;
    a.key_in = s.key[s.kcur];
    a.key_out = s.key[1 - s.kcur];
    a.slot = s.slot;
    a.count = s.bin.count;
    a.vvxL = s.nod.f[NF_VVXL];
    a.vvyL = s.nod.f[NF_VVYL];
    a.avx = s.nod.f[NF_AVX];
    a.avy = s.nod.f[NF_AVY];
    a.Z = s.Z;
    a.ctrl = s.ctrl;
    a.aux = aux;
    if (s.np > 0) {
      const unsigned g = grid_for(s.np, 256);
      if (aux == AUX_FULL) particle_kernel<AUX_FULL><<<g, 256, 0, s.st>>>(a);
      else if (aux == AUX_MS) particle_kernel<AUX_MS><<<g, 256, 0, s.st>>>(a);
      else particle_kernel<AUX_NONE><<<g, 256, 0, s.st>>>(a);
      AMSC_LAUNCH_CHECK();
    }
  }
  s.kcur = 1 - s.kcur;
  {
    Timer tm(s, "binning");
    if (int rc = launch_finish_binning(s.bin, s.np, s.key[s.kcur], s.slot,
                                       s.cptcls, &s.ctrl->active, s.st))
      return rc;
  }
  s.aux_valid = aux == AUX_FULL;
  s.nodal_valid = aux == AUX_FULL;
  s.it_host++;
  s.particle_steps += s.np;
  if (s.p.dt_fixed > 0.0 && s.it_host >= s.p.nsteps) s.finished = true;
  if (s.p.dt_fixed <= 0.0 && s.p.max_steps > 0 && s.it_host >= s.p.max_steps)
    s.finished = true;
  return AMSC_OK;
}

int find_pf(const char *name) {
  for (int k = 0; k < PF_COUNT; ++k)
    if (!std::strcmp(name, kPFName[k])) return k;
  return -1;
}
int find_nf(const char *name) {
  for (int k = 0; k < NF_COUNT; ++k)
    if (!std::strcmp(name, kNFName[k])) return k;
  return -1;
}

}  // namespace

}  // namespace amsc

using namespace amsc;

struct amsc_sim : amsc::Sim {};

extern "C" {

int amsc_sim_create(const amsc_params *p, int np, const double *x,
                    const double *y, const double *Mp, const double *Ap,
                    const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp, const double *Z,
                    int capacity, amsc_sim **out) {
  if (!p || !out || np < 0 || p->nrows <= 0 || p->ncols <= 0 ||
      !(p->hx > 0) || !(p->hy > 0))
    return fail(AMSC_ERR_INVALID, "amsc_sim_create: bad parameters");
  if ((long long)(p->nrows + 1) * (p->ncols + 1) >= (1ll << 31))
    return fail(AMSC_ERR_INVALID, "amsc_sim_create: grid too large");
  if (np > 0 && (!x || !y || !Mp || !Ap || !vpx || !vpy || !mom_px ||
                 !mom_py || !hp || !Vp))
    return fail(AMSC_ERR_INVALID, "amsc_sim_create: NULL particle array");
  if (!Z) return fail(AMSC_ERR_INVALID, "amsc_sim_create: NULL Z");
  *out = nullptr;
  amsc_sim *s = new amsc_sim();
  s->p = *p;
  s->kp = kparams_from(*p);
  s->msp = ms_params_from(*p);
  s->np = np;
  s->cap = capacity > 0 ? capacity : 2 * (np > 0 ? np : 1);
  if (s->cap < np) s->cap = np;
  s->nn = (p->nrows + 1) * (p->ncols + 1);
  s->nc = p->nrows * p->ncols;
  s->nblk_x = (p->ncols + 1 + kTC - 1) / kTC;
  s->nblk_y = (p->nrows + 1 + kTR - 1) / kTR;
  auto bail = [&](int rc) {
    sim_free(s);
    return rc;
  };
  if (hipStreamCreate(&s->st) != hipSuccess)
    return bail(fail(AMSC_ERR_HIP, "stream creation failed"));
  const int cap = s->cap;
  for (int b = 0; b < 2; ++b)
    if (int rc = alloc_ppt(s->buf[b], cap)) return bail(rc);
  bool ok = true;
  ok &= hipMalloc(&s->key[0], sizeof(int) * cap) == hipSuccess;
  ok &= hipMalloc(&s->key[1], sizeof(int) * cap) == hipSuccess;
  ok &= hipMalloc(&s->slot, sizeof(int) * cap) == hipSuccess;
  ok &= hipMalloc(&s->cptcls, sizeof(int) * cap) == hipSuccess;
  ok &= hipMalloc(&s->act, sizeof(int) * cap) == hipSuccess;
  ok &= hipMalloc(&s->partner, sizeof(int) * cap) == hipSuccess;
  ok &= hipMalloc(&s->wts, sizeof(int) * cap) == hipSuccess;
  ok &= hipMalloc(&s->offs, sizeof(int) * (cap + 1)) == hipSuccess;
  ok &= hipMalloc(&s->ms_cnt, sizeof(int) * 4) == hipSuccess;
  for (int k = 0; k < NF_COUNT; ++k)
    ok &= hipMalloc(&s->nod.f[k], sizeof(double) * s->nn) == hipSuccess;
  ok &= hipMalloc(&s->Z, sizeof(double) * s->nn) == hipSuccess;
  ok &= hipMalloc(&s->ctrl, sizeof(Ctrl)) == hipSuccess;
  ok &= hipHostMalloc(&s->ctrl_host, sizeof(Ctrl)) == hipSuccess;
  ok &= hipMalloc(&s->partials,
                  sizeof(double) * 6 * s->nblk_x * s->nblk_y) == hipSuccess;
  for (int b = 0; b < 2; ++b)
    ok &= hipMalloc(&s->p2g_out[b], sizeof(P2GOut)) == hipSuccess;
  if (!ok) return bail(fail(AMSC_ERR_NOMEM, "amsc_sim_create: allocation failed"));
  if (int rc = binning_alloc(s->bin, p->nrows, p->ncols)) return bail(rc);
  for (int b = 0; b < 2; ++b) {
    P2GOut o;
    o.dzx = s->buf[b].f[PF_DZX];
    o.dzy = s->buf[b].f[PF_DZY];
    o.fpx = s->buf[b].f[PF_FPX];
    o.fpy = s->buf[b].f[PF_FPY];
    o.momx = s->buf[b].f[PF_MOMX];
    o.momy = s->buf[b].f[PF_MOMY];
    for (int k = 0; k < NF_COUNT; ++k) o.nod[k] = s->nod.f[k];
    if (hipMemcpy(s->p2g_out[b], &o, sizeof(o), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(AMSC_ERR_HIP, "upload failed"));
  }

  hipStream_t st = s->st;
  PPtr &b = s->buf[0];
  const size_t bytes = sizeof(double) * (size_t)np;
  const double *src[] = {x, y, vpx, vpy, hp, Vp, Mp, Ap};
  const int dst[] = {PF_X, PF_Y, PF_VPX, PF_VPY, PF_HP, PF_VP, PF_MP, PF_AP};
  if (np > 0)
    for (int k = 0; k < 8; ++k)
      if (hipMemcpyAsync(b.f[dst[k]], src[k], bytes, hipMemcpyHostToDevice,
                         st) != hipSuccess)
        return bail(fail(AMSC_ERR_HIP, "upload failed"));
  if (hipMemcpyAsync(s->Z, Z, sizeof(double) * s->nn, hipMemcpyHostToDevice,
                     st) != hipSuccess)
    return bail(fail(AMSC_ERR_HIP, "upload failed"));
  for (int k = 0; k < NF_COUNT; ++k)
    fill_f64_kernel<<<grid_for(s->nn, 256), 256, 0, st>>>(s->nod.f[k], s->nn, 0.0);
  // The first step's P2G uses the DATA momentum (optim_par.cpp:151-152);
  // afterwards mom_p == vpx * Mp exactly, which P2G recomputes.
  for (int i = 0; i < np && !s->mom_stored; ++i)
    if (!(mom_px[i] == vpx[i] * Mp[i]) || !(mom_py[i] == vpy[i] * Mp[i]) ||
        std::signbit(mom_px[i]) != std::signbit(vpx[i] * Mp[i]) ||
        std::signbit(mom_py[i]) != std::signbit(vpy[i] * Mp[i]))
      s->mom_stored = true;
  double *mtmp[2] = {nullptr, nullptr};
  if (np > 0) {
    ok &= hipMalloc(&mtmp[0], bytes) == hipSuccess;
    ok &= hipMalloc(&mtmp[1], bytes) == hipSuccess;
    if (!ok) return bail(fail(AMSC_ERR_NOMEM, "allocation failed"));
    (void)hipMemcpyAsync(mtmp[0], mom_px, bytes, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(mtmp[1], mom_py, bytes, hipMemcpyHostToDevice, st);
  }
  Ctrl c0;
  std::memset(&c0, 0, sizeof(c0));
  c0.dt = 1.0e-3;  // gpu_offload.cpp:255
  c0.active = 1;
  if (hipMemcpyAsync(s->ctrl, &c0, sizeof(Ctrl), hipMemcpyHostToDevice, st) !=
      hipSuccess)
    return bail(fail(AMSC_ERR_HIP, "upload failed"));
  if (int rc = rebin(*s)) return bail(rc);
  if (np > 0) {
    init_kernel<<<grid_for(np, 256), 256, 0, st>>>(np, s->kp, b, s->key[s->kcur],
                                                   s->Z, mtmp[0], mtmp[1]);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return bail(fail(AMSC_ERR_HIP, "initialisation failed"));
  if (mtmp[0]) (void)hipFree(mtmp[0]);
  if (mtmp[1]) (void)hipFree(mtmp[1]);
  s->launches.clear();
  s->event_used = 0;
  *out = s;
  return AMSC_OK;
}

int amsc_sim_destroy(amsc_sim *s) {
  sim_free(s);
  return AMSC_OK;
}

int amsc_sim_step(amsc_sim *s, int nsteps, int *done) {
  if (!s || nsteps < 0) return fail(AMSC_ERR_INVALID, "amsc_sim_step: bad arguments");
  if (int rc = read_ctrl(*s)) return rc;
  const int it0 = s->ctrl_host->it;
  s->it_host = it0;
  if (!s->ctrl_host->active) s->finished = true;
  for (int k = 0; k < nsteps && !s->finished; ++k)
    if (int rc = launch_step(*s)) return rc;
  if (int rc = read_ctrl(*s)) return rc;
  if (!s->ctrl_host->active) s->finished = true;
  s->it_host = s->ctrl_host->it;
  if (done) *done = s->ctrl_host->it - it0;
  return AMSC_OK;
}

int amsc_sim_set_aux(amsc_sim *s, int on) {
  if (!s) return fail(AMSC_ERR_INVALID, "NULL sim");
  s->aux_on = on != 0;
  return AMSC_OK;
}

int amsc_sim_set_timing(amsc_sim *s, int on) {
  if (!s) return fail(AMSC_ERR_INVALID, "NULL sim");
  s->timing = on != 0;
  return AMSC_OK;
}

int amsc_sim_kernel_times(amsc_sim *s, const char **names, double *ms,
                          int *count, int cap, int *n) {
  if (!s || !n) return fail(AMSC_ERR_INVALID, "amsc_sim_kernel_times: bad arguments");
  AMSC_HIP(hipStreamSynchronize(s->st));
  std::vector<const char *> nm;
  std::vector<double> tot;
  std::vector<int> cnt;
  for (auto &l : s->launches) {
    float e = 0.f;
    AMSC_HIP(hipEventElapsedTime(&e, l.a, l.b));
    size_t i = 0;
    while (i < nm.size() && std::strcmp(nm[i], l.name)) ++i;
    if (i == nm.size()) {
      nm.push_back(l.name);
      tot.push_back(0.0);
      cnt.push_back(0);
    }
    tot[i] += e;
    cnt[i] += 1;
  }
  *n = (int)nm.size();
  for (int i = 0; i < (int)nm.size() && i < cap; ++i) {
    if (names) names[i] = nm[i];
    if (ms) ms[i] = tot[i];
    if (count) count[i] = cnt[i];
  }
  return AMSC_OK;
}

int amsc_sim_reset_kernel_times(amsc_sim *s) {
  if (!s) return fail(AMSC_ERR_INVALID, "NULL sim");
  AMSC_HIP(hipStreamSynchronize(s->st));
  s->launches.clear();
  s->event_used = 0;
  return AMSC_OK;
}

int amsc_sim_num_particles(amsc_sim *s, int *np) {
  if (!s || !np) return fail(AMSC_ERR_INVALID, "bad arguments");
  *np = s->np;
  return AMSC_OK;
}

int amsc_sim_state(amsc_sim *s, int *it, double *t, double *dt) {
  if (!s) return fail(AMSC_ERR_INVALID, "NULL sim");
  if (int rc = read_ctrl(*s)) return rc;
  if (it) *it = s->ctrl_host->it;
  if (t) *t = s->ctrl_host->t;
  if (dt) *dt = s->ctrl_host->dt;
  return AMSC_OK;
}

int amsc_sim_get_field(amsc_sim *s, const char *name, double *out, int n) {
  if (!s || !name || !out) return fail(AMSC_ERR_INVALID, "bad arguments");
  const double *src = nullptr;
  int len = 0;
  const int k = find_pf(name);
  if (k >= 0) {
    if (k >= PF_CORE && !s->aux_valid)
      return fail(AMSC_ERR_STALE, std::string("field '") + name +
                                      "' was not written by the last step "
                                      "(enable amsc_sim_set_aux)");
    src = s->buf[s->cur].f[k];
    len = s->np;
  } else if (!std::strcmp(name, "Z")) {
    src = s->Z;
    len = s->nn;
  } else if (!std::strcmp(name, "vpxL") || !std::strcmp(name, "vpyL")) {
    src = s->buf[s->cur].f[name[2] == 'x' ? PF_VPX : PF_VPY];
    len = s->np;
  } else if (!std::strcmp(name, "F_21")) {
    src = s->buf[s->cur].f[PF_F12];  // F_21 == F_12 (optim_par.cpp:116-117)
    len = s->np;
  } else {
    const int j = find_nf(name);
    if (j < 0) return fail(AMSC_ERR_INVALID, std::string("unknown field ") + name);
    if (j >= NF_AVX && !s->nodal_valid)
      return fail(AMSC_ERR_STALE, std::string("nodal field '") + name +
                                      "' was not written by the last step");
    src = s->nod.f[j];
    len = s->nn;
  }
  if (n < len) return fail(AMSC_ERR_INVALID, "output buffer too small");
  AMSC_HIP(hipMemcpyAsync(out, src, sizeof(double) * (size_t)len,
                          hipMemcpyDeviceToHost, s->st));
  AMSC_HIP(hipStreamSynchronize(s->st));
  return AMSC_OK;
}

int amsc_sim_get_field_int(amsc_sim *s, const char *name, int *out, int n) {
  if (!s || !name || !out) return fail(AMSC_ERR_INVALID, "bad arguments");
  const int *src = nullptr;
  int len = s->np;
  if (!std::strcmp(name, "label")) src = s->buf[s->cur].label;
  else if (!std::strcmp(name, "level")) src = s->buf[s->cur].level;
  else if (!std::strcmp(name, "p2g")) src = s->key[s->kcur];
  else if (!std::strcmp(name, "cptcls")) src = s->cptcls;
  else if (!std::strcmp(name, "cstart")) {
    src = s->bin.cstart;
    len = s->nc + 1;
  } else return fail(AMSC_ERR_INVALID, std::string("unknown field ") + name);
  if (n < len) return fail(AMSC_ERR_INVALID, "output buffer too small");
  AMSC_HIP(hipMemcpyAsync(out, src, sizeof(int) * (size_t)len,
                          hipMemcpyDeviceToHost, s->st));
  AMSC_HIP(hipStreamSynchronize(s->st));
  return AMSC_OK;
}

int amsc_sim_set_field(amsc_sim *s, const char *name, const double *in, int n) {
  if (!s || !name || !in) return fail(AMSC_ERR_INVALID, "bad arguments");
  const int k = find_pf(name);
  if (k < 0) return fail(AMSC_ERR_INVALID, std::string("unknown field ") + name);
  if (n != s->np) return fail(AMSC_ERR_INVALID, "length must equal num_particles");
  AMSC_HIP(hipMemcpyAsync(s->buf[s->cur].f[k], in, sizeof(double) * (size_t)n,
                          hipMemcpyHostToDevice, s->st));
  AMSC_HIP(hipStreamSynchronize(s->st));
  if (k == PF_X || k == PF_Y) return rebin(*s);
  return AMSC_OK;
}

int amsc_sim_conservation(amsc_sim *s, double out[3]) {
  if (!s || !out) return fail(AMSC_ERR_INVALID, "bad arguments");
  if (int rc = read_ctrl(*s)) return rc;
  out[0] = s->ctrl_host->cons[0];
  out[1] = s->ctrl_host->cons[1];
  out[2] = s->ctrl_host->t_start;
  return AMSC_OK;
}

int amsc_sim_particle_steps(amsc_sim *s, long long *out) {
  if (!s || !out) return fail(AMSC_ERR_INVALID, "bad arguments");
  *out = s->particle_steps;
  return AMSC_OK;
}

int amsc_sim_ms_counts(amsc_sim *s, int out[2]) {
  if (!s || !out) return fail(AMSC_ERR_INVALID, "bad arguments");
  out[0] = s->ms_counts[0];
  out[1] = s->ms_counts[1];
  return AMSC_OK;
}

}  // extern "C"
