// dropin.hip -- the 18 step kernels of the reference's gpu_kernels.h as
// gfx950 HIP kernels behind a C ABI (include/amsc_mpm.h, group 1).
//
// Per-particle and per-node entry points are one thread per item with the
// reference's expressions.  The four colour-scattered P2G transfers
// (gpu_kernels.cpp:97-128,183-246,263-296,572-609) become one LDS-tiled
// gather each: a node adds the particles of its (up to 4) cells in colour
// order, each cell's particles in cptcls order, to its current value.  That
// is the exact addition sequence the colour loop produces (cells of one
// colour never share a node), without atomics, without 4 launches and
// without read-modify-write traffic on the nodal arrays.
#include "common.hpp"
#include "amsc_mpm.h"

#include <climits>
#include <cmath>
#include <algorithm>
#include <mutex>

namespace amsc {

namespace {

constexpr int kBlock = 256;

// Device memory kept between calls: grown when a call needs more, and
// allocated again when the caller's current device has changed.  hipFree
// waits for the device, so a buffer is never freed under a kernel that an
// earlier, still running call launched on it.
struct DevCache {
  int dev = -1;
  void *p = nullptr;
  size_t cap = 0;
};

int dev_cache_get(DevCache &c, size_t bytes, void **out, const char *what) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return fail(AMSC_ERR_HIP, std::string(what) + ": no device");
  if (dev != c.dev || bytes > c.cap) {
    if (c.p) (void)hipFree(c.p);
    c = DevCache();
    if (hipMalloc(&c.p, bytes) != hipSuccess) {
      c.p = nullptr;
      return fail(AMSC_ERR_NOMEM, std::string(what) + ": workspace allocation");
    }
    c.dev = dev;
    c.cap = bytes;
  }
  *out = c.p;
  return AMSC_OK;
}

// Colour of each cell, rebuilt from (ccidx, coloff) on every scatter call.
// One map serves every call; a call holds g_colour_mu from the rebuild to
// the launch of the gather that reads it, and the rebuild waits (on the
// caller's stream) for the last gather that read the map, on whatever
// stream that gather was enqueued: calls that only enqueue their kernels,
// on different streams, never overwrite a map that is still being read.
DevCache g_colour;
std::mutex g_colour_mu;
hipEvent_t g_colour_read = nullptr;  // the last gather that read the map
int g_colour_read_dev = -1;

__global__ void fill_u8_kernel(uint8_t *p, int n, uint8_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void colour_map_kernel(const int *ccidx, int nc, int o1, int o2,
                                  int o3, uint8_t *colour) {
  const int ic = blockIdx.x * blockDim.x + threadIdx.x;
  if (ic >= nc) return;
  const int cell = ccidx[ic];
  if (cell < 0 || cell >= nc) return;
  colour[cell] = (uint8_t)(ic < o1 ? 0 : ic < o2 ? 1 : ic < o3 ? 2 : 3);
}

// Validate the colouring and build the colour map; returns ncols in *ncols.
// The caller holds g_colour_mu until the map's reader is enqueued.
int prepare_colouring(int nrows, const int *ccidx, const int *coloff,
                      int *ncols, const uint8_t **colour) {
  if (nrows <= 0) return fail(AMSC_ERR_INVALID, "nrows must be positive");
  if (!coloff) return fail(AMSC_ERR_INVALID, "coloff (host int[5]) is NULL");
  if (coloff[0] != 0)
    return fail(AMSC_ERR_INVALID, "coloff[0] must be 0");
  for (int k = 0; k < 4; ++k)
    if (coloff[k + 1] < coloff[k])
      return fail(AMSC_ERR_INVALID, "coloff must be non-decreasing");
  const int nc = coloff[4];
  if (nc % nrows != 0)
    return fail(AMSC_ERR_INVALID,
                "coloff[4] must equal nrows*ncols (ccidx lists every cell)");
  *ncols = nc / nrows;
  if (nc == 0) {
    *colour = nullptr;
    return AMSC_OK;
  }
  void *p = nullptr;
  if (int rc = dev_cache_get(g_colour, (size_t)nc, &p, "colour map"))
    return rc;
  uint8_t *map = static_cast<uint8_t *>(p);
  hipStream_t st = current_stream();
  if (g_colour_read && g_colour_read_dev == g_colour.dev)
    AMSC_HIP(hipStreamWaitEvent(st, g_colour_read, 0));
  fill_u8_kernel<<<grid_for(nc, kBlock), kBlock, 0, st>>>(map, nc, 0xFF);
  colour_map_kernel<<<grid_for(nc, kBlock), kBlock, 0, st>>>(
      ccidx, nc, coloff[1], coloff[2], coloff[3], map);
  AMSC_LAUNCH_CHECK();
  *colour = map;
  return AMSC_OK;
}

// ---- the colour-ordered scatters as LDS-tiled gathers.
// A workgroup owns a tile of kTC x kTR nodes (one thread each) and the
// (kTC+1) x (kTR+1) cells around it.  The tile's particles are listed cell by
// cell in colour-major order (all its colour-0 cells, then colour 1, ...;
// box order within a colour) and staged through LDS 256 at a time; each node
// adds the particles of its cells in list order, i.e. its cells in colour
// order and each cell's particles in cptcls order, starting from the node's
// current value: the addition sequence of the reference's colour loop.
constexpr int kTC = 8, kTR = 32, kTNT = kTC * kTR;
constexpr int kTCells = (kTC + 1) * (kTR + 1);
constexpr int kTCH = kTNT, kTCHP = kTCH + kTCH / 16;
static_assert(kTCells <= 2 * kTNT, "two box cells per thread");

// Staged position i at i + i/16 (node threads walk neighbouring cells at a
// stride of a few positions).
__device__ __forceinline__ int tg_slot(int i) { return i + (i >> 4); }

template <int NIN, int NOUT>
struct TGFields {
  const double *in[NIN];
  double *out[NOUT];
};

// INTERNAL: in = {Vp, F11, F12, F21, F22}, out = {F_int_vx, F_int_vy},
// out += (Nx*F11 + Ny*F21)*Vp, (Nx*F12 + Ny*F22)*Vp (gpu_kernels.cpp:286-292);
// otherwise out[k] += N * in[k].
template <int NIN, int NOUT, bool INTERNAL>
__global__ void __launch_bounds__(kTNT)
    tiled_gather_kernel(int nrows, int ncols, double hx, double hy, int ntx,
                        const int *__restrict__ cstart,
                        const int *__restrict__ cptcls,
                        const uint8_t *__restrict__ colour,
                        const double *__restrict__ x,
                        const double *__restrict__ y, TGFields<NIN, NOUT> f) {
  __shared__ int s_beg[kTCells], s_cnt[kTCells], s_off[kTCells];
  __shared__ unsigned short s_own[kTCH];
  __shared__ double s_xy[2][kTCHP];
  __shared__ double s_in[NIN][kTCHP];
  __shared__ int s_w[17];
  const int t = threadIdx.x;
  const int c0 = (blockIdx.x % ntx) * kTC, r0 = (blockIdx.x / ntx) * kTR;
  const int ccl = c0 - 1 < 0 ? 0 : c0 - 1;
  const int cch = c0 + kTC - 1 > ncols - 1 ? ncols - 1 : c0 + kTC - 1;
  const int rl = r0 - 1 < 0 ? 0 : r0 - 1;
  const int rh = r0 + kTR - 1 > nrows - 1 ? nrows - 1 : r0 + kTR - 1;
  const int nr = rh - rl + 1, ncl = (cch - ccl + 1) * nr;
  const Recip R = recip_dev(hx, hy);

  // a tile whose cells hold no particle adds nothing to its nodes: every
  // wavefront tests the box's column ranges itself and the workgroup
  // leaves together, before any barrier
  {
    const int lane = t & 63;
    int np = 0;
    if (lane <= cch - ccl) {
      const int col0 = (ccl + lane) * nrows;
      np = cstart[col0 + rh + 1] - cstart[col0 + rl];
    }
    if (__ballot(np != 0) == 0ull) return;
  }

  // cell table: thread t holds box cells 2t and 2t+1 (column-major order)
  int cnt[2] = {0, 0}, col[2] = {0xFF, 0xFF};
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int j = 2 * t + u;
    if (j < ncl) {
      const int cell = (ccl + j / nr) * nrows + (rl + j % nr);
      const int b = cstart[cell];
      col[u] = colour[cell];
      cnt[u] = col[u] == 0xFF ? 0 : cstart[cell + 1] - b;
      s_beg[j] = b;
      s_cnt[j] = cnt[u];
    }
  }
  // tile-list offsets, colour-major
  int base = 0, off[2] = {0, 0};
  for (int k = 0; k < 4; ++k) {
    const int v0 = col[0] == k ? cnt[0] : 0, v1 = col[1] == k ? cnt[1] : 0;
    int total;
    const int ex = block_excl_scan(v0 + v1, s_w, &total);
    if (col[0] == k) off[0] = base + ex;
    if (col[1] == k) off[1] = base + ex + v0;
    base += total;
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
    if (2 * t + u < ncl) s_off[2 * t + u] = off[u];
  const int L = base;
  __syncthreads();

  // this node's cells as tile-list ranges, in list order
  const int c = c0 + t / kTR, r = r0 + t % kTR;
  const bool node_ok = c <= ncols && r <= nrows;
  const int n = r + c * (nrows + 1);
  int tb[4], te[4], cn[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // (c-1,r-1) corner 3, (c-1,r) corner 2, (c,r-1) corner 1, (c,r) corner 0
    const int cc = c - 1 + (k >> 1), cr = r - 1 + (k & 1);
    const bool in = node_ok && cc >= 0 && cc < ncols && cr >= 0 && cr < nrows;
    const int j = in ? (cc - ccl) * nr + (cr - rl) : 0;
    tb[k] = in ? s_off[j] : INT_MAX;
    te[k] = in ? s_off[j] + s_cnt[j] : INT_MAX;
    cn[k] = 3 - k;
  }
  auto cx = [&](int i, int j) {
    if (tb[j] < tb[i]) {
      int q = tb[i]; tb[i] = tb[j]; tb[j] = q;
      q = te[i]; te[i] = te[j]; te[j] = q;
      q = cn[i]; cn[i] = cn[j]; cn[j] = q;
    }
  };
  cx(0, 1); cx(2, 3); cx(0, 2); cx(1, 3); cx(1, 2);
  double acc[NOUT];
#pragma unroll
  for (int o = 0; o < NOUT; ++o) acc[o] = node_ok ? f.out[o][n] : 0.0;

  for (int b0 = 0; b0 < L; b0 += kTCH) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = 2 * t + u;
      if (j < ncl) {
        const int lo0 = s_off[j] > b0 ? s_off[j] : b0;
        const int hi0 = s_off[j] + cnt[u] < b0 + kTCH ? s_off[j] + cnt[u] : b0 + kTCH;
        for (int q = lo0; q < hi0; ++q) s_own[q - b0] = (unsigned short)j;
      }
    }
    __syncthreads();
    const int q = b0 + t;
    if (q < L) {
      const int j = s_own[t];
      const int ip = cptcls[s_beg[j] + (q - s_off[j])];
      const Local Lc = local_coords(x[ip], y[ip], ccl + j / nr, rl + j % nr,
                                    hx, hy, R);
      const int i = tg_slot(t);
      s_xy[0][i] = Lc.xi;
      s_xy[1][i] = Lc.eta;
#pragma unroll
      for (int m = 0; m < NIN; ++m) s_in[m][i] = f.in[m][ip];
    }
    __syncthreads();
    if (node_ok) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int lo = tb[k] > b0 ? tb[k] : b0;
        const int hi = te[k] < b0 + kTCH ? te[k] : b0 + kTCH;
        for (int qq = lo; qq < hi; ++qq) {
          const int i = tg_slot(qq - b0);
          Local Lq;
          Lq.xi = s_xy[0][i];
          Lq.eta = s_xy[1][i];
          if (INTERNAL) {
            const double Nx = shape_dx_sel(Lq, cn[k], hx);
            const double Ny = shape_dy_sel(Lq, cn[k], hy);
            const double vp = s_in[0][i];
            acc[0] += (Nx * s_in[1][i] + Ny * s_in[3][i]) * vp;
            acc[NOUT - 1] += (Nx * s_in[2][i] + Ny * s_in[4][i]) * vp;
          } else {
            const double N = shape_sel(Lq, cn[k]);
#pragma unroll
            for (int o = 0; o < NOUT; ++o) acc[o] += N * s_in[o][i];
          }
        }
      }
    }
    __syncthreads();
  }
  if (node_ok)
#pragma unroll
    for (int o = 0; o < NOUT; ++o) f.out[o][n] = acc[o];
}

template <int NIN, int NOUT, bool INTERNAL>
int tiled_gather(int nrows, double hx, double hy, const int *cstart,
                 const int *cptcls, const int *ccidx, const int *coloff,
                 const double *x, const double *y, TGFields<NIN, NOUT> f) {
  int ncols = 0;
  const uint8_t *colour = nullptr;
  std::lock_guard<std::mutex> lk(g_colour_mu);
  if (int rc = prepare_colouring(nrows, ccidx, coloff, &ncols, &colour))
    return rc;
  if (ncols == 0) return AMSC_OK;
  const int ntx = (ncols + 1 + kTC - 1) / kTC, nty = (nrows + 1 + kTR - 1) / kTR;
  hipStream_t st = current_stream();
  tiled_gather_kernel<NIN, NOUT, INTERNAL>
      <<<ntx * nty, kTNT, 0, st>>>(nrows, ncols, hx, hy, ntx, cstart, cptcls,
                                   colour, x, y, f);
  AMSC_LAUNCH_CHECK();
  // the map's last reader, for the next rebuild to wait for
  if (g_colour_read_dev != g_colour.dev) {
    if (g_colour_read) (void)hipEventDestroy(g_colour_read);
    g_colour_read = nullptr;
    AMSC_HIP(hipEventCreateWithFlags(&g_colour_read, hipEventDisableTiming));
    g_colour_read_dev = g_colour.dev;
  }
  AMSC_HIP(hipEventRecord(g_colour_read, st));
  return AMSC_OK;
}

// ---- CFL maxima (gpu_offload.cpp:285-292): max |vpx|, max |vpy|, max hp.
// Each workgroup writes its three maxima, one workgroup reduces them (max is
// exact in any order).
constexpr int kCflBlock = 256, kCflMaxBlocks = 1024;

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const double o = __shfl_xor(v, d);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ void block_max3(double v[3], double (*s)[kCflBlock / 64]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    v[k] = wave_max_d(v[k]);
    if (lane == 0) s[k][w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double m = s[k][0];
    for (int i = 1; i < kCflBlock / 64; ++i) m = s[k][i] > m ? s[k][i] : m;
    v[k] = m;
  }
}

__global__ void __launch_bounds__(kCflBlock)
    cfl_partial_kernel(int np, const double *__restrict__ vpx,
                       const double *__restrict__ vpy,
                       const double *__restrict__ hp, double *part) {
  __shared__ double s[3][kCflBlock / 64];
  double v[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int i = blockIdx.x * kCflBlock + threadIdx.x; i < np;
       i += gridDim.x * kCflBlock) {
    const double a = fabs(vpx[i]), b = fabs(vpy[i]), h = hp[i];
    v[0] = a > v[0] ? a : v[0];
    v[1] = b > v[1] ? b : v[1];
    v[2] = h > v[2] ? h : v[2];
  }
  block_max3(v, s);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = v[k];
}

__global__ void __launch_bounds__(kCflBlock)
    cfl_final_kernel(int nb, const double *__restrict__ part, double *out) {
  __shared__ double s[3][kCflBlock / 64];
  double v[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int b = threadIdx.x; b < nb; b += kCflBlock)
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = part[3 * b + k] > v[k] ? part[3 * b + k] : v[k];
  block_max3(v, s);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) out[k] = v[k];
}

struct Ptrs25 {
  double *p[25];
};

__global__ void reset_kernel(int nn, Ptrs25 a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
#pragma unroll
  for (int k = 0; k < 25; ++k) a.p[k][i] = 0;
}

// gpu_kernels.cpp:131-152
__global__ void gradZ_kernel(int np, int nrows, double hx, double hy,
                             const double *__restrict__ x,
                             const double *__restrict__ y,
                             const int *__restrict__ p2g,
                             const double *__restrict__ Z, double *dZxp,
                             double *dZyp) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
  const Local L = local_coords(x[ip], y[ip], c, r, hx, hy, recip_dev(hx, hy));
  double dzx = 0.0, dzy = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double z = Z[node_id(k, c, r, nrows)];
    dzx += shape_dx(L, k, hx) * z;
    dzy += shape_dy(L, k, hy) * z;
  }
  dZxp[ip] = dzx;
  dZyp[ip] = dzy;
}

// gpu_kernels.cpp:155-167
__global__ void gravity_kernel(int np, double g, const double *__restrict__ hp,
                               const double *__restrict__ H,
                               const double *__restrict__ Mp,
                               const double *__restrict__ dZxp,
                               const double *__restrict__ dZyp, double *Fpx,
                               double *Fpy) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const double ratio = hp[ip] > 1e-10 ? H[ip] / hp[ip] : 0.0;
  Fpx[ip] = -g * Mp[ip] * (1.0 - ratio) * dZxp[ip];
  Fpy[ip] = -g * Mp[ip] * (1.0 - ratio) * dZyp[ip];
}

// gpu_kernels.cpp:170-180
__global__ void friction_kernel(int np, const double *__restrict__ Ap,
                                const double *__restrict__ Fbx,
                                const double *__restrict__ Fby, double *Fx,
                                double *Fy) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  Fx[ip] = Ap[ip] * Fbx[ip];
  Fy[ip] = Ap[ip] * Fby[ip];
}

// gpu_kernels.cpp:249-260
__global__ void fext_kernel(int nn, const double *__restrict__ FPx,
                            const double *__restrict__ FPy,
                            const double *__restrict__ Frx,
                            const double *__restrict__ Fry, double *Fx,
                            double *Fy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  Fx[i] = FPx[i] + Frx[i];
  Fy[i] = FPy[i] + Fry[i];
}

// gpu_kernels.cpp:299-315 (momentum update commented out in the reference)
__global__ void ftot_kernel(int nn, const double *__restrict__ Fex,
                            const double *__restrict__ Fey,
                            const double *__restrict__ Fix,
                            const double *__restrict__ Fiy, double *Ftx,
                            double *Fty) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  Ftx[i] = Fex[i] + Fix[i];
  Fty[i] = Fey[i] + Fiy[i];
}

// gpu_kernels.cpp:318-353
__global__ void nodal_vel_bc_kernel(int nn, int nrows, int ncols, double dt,
                                    int bc, const double *__restrict__ Mv,
                                    const double *__restrict__ Ftx,
                                    const double *__restrict__ Fty,
                                    double *momx, double *momy, double *avx,
                                    double *avy, double *vvx, double *vvy,
                                    double *vvxL, double *vvyL) {
  const int iv = blockIdx.x * blockDim.x + threadIdx.x;
  if (iv >= nn) return;
  const double m = Mv[iv];
  const bool active = m > 1e-2;
  double ax = active ? Ftx[iv] / m : 0.0;
  double ay = active ? Fty[iv] / m : 0.0;
  double vx = active ? momx[iv] / m : 0.0;
  double vy = active ? momy[iv] / m : 0.0;
  double vxL = active ? dt * ax + vx : 0.0;
  double vyL = active ? dt * ay + vy : 0.0;
  if (bc) {
    const int r = iv % (nrows + 1), c = iv / (nrows + 1);
    if (c == 0 || c == 1 || c == ncols - 1 || c == ncols) {
      momx[iv] = 0.0; vx = 0.0; vxL = 0.0; ax = 0.0;
    }
    if (r == 0 || r == 1 || r == nrows - 1 || r == nrows) {
      momy[iv] = 0.0; vy = 0.0; vyL = 0.0; ay = 0.0;
    }
  }
  avx[iv] = ax; avy[iv] = ay;
  vvx[iv] = vx; vvy[iv] = vy;
  vvxL[iv] = vxL; vvyL[iv] = vyL;
}

// gpu_kernels.cpp:356-407
__global__ void g2p_advect_kernel(int np, int nrows, int ncols, double hx,
                                  double hy, double dt,
                                  const int *__restrict__ p2g,
                                  const double *__restrict__ vvxL,
                                  const double *__restrict__ vvyL,
                                  const double *__restrict__ avx,
                                  const double *__restrict__ avy, double *x,
                                  double *y, double *vpx, double *vpy,
                                  double *apx, double *apy, double *mpx,
                                  double *mpy, double *vpxL, double *vpyL,
                                  const double *__restrict__ Mp) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
  const Local L = local_coords(x[ip], y[ip], c, r, hx, hy, recip_dev(hx, hy));
  double ax = 0, ay = 0, vxL = 0, vyL = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double N = shape(L, k);
    const int n = node_id(k, c, r, nrows);
    ax += N * avx[n];
    ay += N * avy[n];
    vxL += N * vvxL[n];
    vyL += N * vvyL[n];
  }
  vpx[ip] = vxL; vpy[ip] = vyL;
  apx[ip] = ax; apy[ip] = ay;
  double xn = x[ip] + dt * vxL, yn = y[ip] + dt * vyL;
  const double eps = 1e-10, Lx = ncols * hx, Ly = nrows * hy;
  if (xn < eps) xn = eps;
  if (xn > Lx - eps) xn = Lx - eps;
  if (yn < eps) yn = eps;
  if (yn > Ly - eps) yn = Ly - eps;
  x[ip] = xn; y[ip] = yn;
  vpxL[ip] = vxL; vpyL[ip] = vyL;
  mpx[ip] = vxL * Mp[ip];
  mpy[ip] = vyL * Mp[ip];
}

// gpu_kernels.cpp:410-454
__global__ void g2pd_height_kernel(int np, int nrows, double hx, double hy,
                                   double dt, const double *__restrict__ x,
                                   const double *__restrict__ y,
                                   const int *__restrict__ p2g,
                                   const double *__restrict__ vvxL,
                                   const double *__restrict__ vvyL,
                                   double *vxdx_o, double *vxdy_o,
                                   double *vydx_o, double *vydy_o, double *hp,
                                   const double *Vp, double *Ap) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
  const Local L = local_coords(x[ip], y[ip], c, r, hx, hy, recip_dev(hx, hy));
  double vxdx = 0, vxdy = 0, vydx = 0, vydy = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int n = node_id(k, c, r, nrows);
    const double Nx = shape_dx(L, k, hx), Ny = shape_dy(L, k, hy);
    vxdx += Nx * vvxL[n];
    vxdy += Ny * vvxL[n];
    vydx += Nx * vvyL[n];
    vydy += Ny * vvyL[n];
  }
  vxdx_o[ip] = vxdx; vxdy_o[ip] = vxdy;
  vydx_o[ip] = vydx; vydy_o[ip] = vydy;
  const double sc = 1.0 + dt * (vxdx + vydy);
  const double h = hp[ip] / sc;
  hp[ip] = h;
  Ap[ip] = Vp[ip] / h;
}

// gpu_kernels.cpp:457-475
__global__ void voellmy_kernel(int np, double rho, double g, double tan_fa,
                               double xi, double fric_on,
                               const double *__restrict__ vpx,
                               const double *__restrict__ vpy,
                               const double *__restrict__ hp, double *Fbx,
                               double *Fby) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const double vx = vpx[ip], vy = vpy[ip], h = hp[ip];
  const double nv = sqrt(vx * vx + vy * vy);
  if (fric_on > 0 && nv > 1e-10 && xi > 0) {
    const double v2 = vx * vx + vy * vy;
    Fbx[ip] = -fric_on * (rho * g * h * tan_fa + rho * g * v2 / xi) * vx / nv;
    Fby[ip] = -fric_on * (rho * g * h * tan_fa + rho * g * v2 / xi) * vy / nv;
  } else {
    Fbx[ip] = 0.0;
    Fby[ip] = 0.0;
  }
}

// gpu_kernels.cpp:478-546
__global__ void bingham_gpu_kernel(
    int np, double rho, double g, double mu, double tau_Y, double cc,
    double A_coeff, double C_coeff, const double *__restrict__ vpx,
    const double *__restrict__ vpy, const double *__restrict__ hp,
    const double *__restrict__ H, const double *__restrict__ vxdx,
    const double *__restrict__ vxdy, const double *__restrict__ vydx,
    const double *__restrict__ vydy, double *F11, double *F12, double *F21,
    double *F22) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const double vx = vpx[ip], vy = vpy[ip], h = hp[ip];
  const double nv = sqrt(vx * vx + vy * vy);
  const double alf = (h > 1e-3 && tau_Y > 1e-12)
                         ? (6.0 * mu * nv) / ((h + 0.001) * tau_Y)
                         : 0.0;
  const double bv = -114.0 / 32.0 - alf;
  const double sq = sqrt(bv * bv - 4.0 * A_coeff * C_coeff);
  const double z1 = (-bv + sq) / (2.0 * A_coeff);
  const double z2 = (-bv - sq) / (2.0 * A_coeff);
  const double zz = (fabs(z1 - 0.5) <= 0.5) ? z1 : z2;
  const double Dxx = vxdx[ip];
  const double Dxy = 0.5 * (vxdy[ip] + vydx[ip]);
  const double Dyy = vydy[ip];
  const double Dzz = -(Dxx + Dyy);
  const double Dzx = h > 1e-3 ? 0.5 * (3.0 / (2.0 + zz)) * (vx / (h + 0.001)) : 0.0;
  const double Dzy = h > 1e-3 ? 0.5 * (3.0 / (2.0 + zz)) * (vy / (h + 0.001)) : 0.0;
  const double inv2 = 0.5 * (Dxx * Dxx + Dyy * Dyy + Dzz * Dzz) + Dzx * Dzx +
                      Dzy * Dzy + Dxy * Dxy;
  const double coeff_max = 5.0e6;
  double inv2_floor = (tau_Y / coeff_max) * (tau_Y / coeff_max);
  if (inv2_floor < 1e-9) inv2_floor = 1e-9;
  const double coeff = tau_Y / sqrt(inv2 + inv2_floor) + 2.0 * mu;
  const double h_corr = h > 1e-10 ? h - H[ip] * H[ip] / h : 0.0;
  F11[ip] = -cc * coeff * Dxx + 0.5 * rho * g * h_corr;
  F22[ip] = -cc * coeff * Dyy + 0.5 * rho * g * h_corr;
  F12[ip] = -cc * coeff * Dxy;
  F21[ip] = -cc * coeff * Dxy;
}

// gpu_kernels.cpp:549-569
__global__ void g2p_height_kernel(int np, int nrows, double hx, double hy,
                                  const double *__restrict__ x,
                                  const double *__restrict__ y,
                                  const int *__restrict__ p2g,
                                  const double *__restrict__ Z,
                                  const double *__restrict__ hp, double *Zp,
                                  double *hpZ) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const int ci = p2g[ip], r = ci % nrows, c = ci / nrows;
  const Local L = local_coords(x[ip], y[ip], c, r, hx, hy, recip_dev(hx, hy));
  double zp = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) zp += shape(L, k) * Z[node_id(k, c, r, nrows)];
  Zp[ip] = zp;
  hpZ[ip] = hp[ip] + zp;
}

// gpu_kernels.cpp:612-627
__global__ void normalize_kernel(int nn, const double *__restrict__ rho,
                                 double *vx, double *vy, double *ax,
                                 double *ay) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double mv = rho[i];
  if (mv > 1e-8) {
    vx[i] /= mv;
    vy[i] /= mv;
    ax[i] /= mv;
    ay[i] /= mv;
  }
}

int check_n(int n, const char *what) {
  if (n < 0) return fail(AMSC_ERR_INVALID, std::string(what) + " is negative");
  return AMSC_OK;
}

}  // namespace
}  // namespace amsc

using namespace amsc;

#define LAUNCH1D(kernel, n, ...)                                             \
  do {                                                                       \
    if ((n) > 0) {                                                           \
      kernel<<<grid_for((n), kBlock), kBlock, 0, current_stream()>>>(        \
          __VA_ARGS__);                                                      \
      AMSC_LAUNCH_CHECK();                                                   \
    }                                                                        \
  } while (0)

extern "C" {

int amsc_reset_nodal_arrays(int nn, double *d_Mv, double *d_mom_vx,
                            double *d_mom_vy, double *d_F_ext_vx,
                            double *d_F_ext_vy, double *d_F_int_vx,
                            double *d_F_int_vy, double *d_Ftot_vx,
                            double *d_Ftot_vy, double *d_avx, double *d_avy,
                            double *d_vvx, double *d_vvy, double *d_vvxL,
                            double *d_vvyL, double *d_Fric_x, double *d_Fric_y,
                            double *d_FPxv, double *d_FPyv, double *d_HV,
                            double *d_pv_rho, double *d_pv_vvx,
                            double *d_pv_vvy, double *d_pv_avx,
                            double *d_pv_avy) {
  if (int rc = check_n(nn, "nn")) return rc;
  Ptrs25 a = {{d_Mv, d_mom_vx, d_mom_vy, d_F_ext_vx, d_F_ext_vy, d_F_int_vx,
               d_F_int_vy, d_Ftot_vx, d_Ftot_vy, d_avx, d_avy, d_vvx, d_vvy,
               d_vvxL, d_vvyL, d_Fric_x, d_Fric_y, d_FPxv, d_FPyv, d_HV,
               d_pv_rho, d_pv_vvx, d_pv_vvy, d_pv_avx, d_pv_avy}};
  LAUNCH1D(reset_kernel, nn, nn, a);
  return AMSC_OK;
}

int amsc_scatter_mass_momentum(int np, int nrows, double hx, double hy,
                               const int *cstart, const int *cptcls,
                               const int *ccidx, const int *coloff,
                               const double *d_x, const double *d_y,
                               const double *d_Mp, const double *d_mom_px,
                               const double *d_mom_py, double *d_Mv,
                               double *d_mom_vx, double *d_mom_vy) {
  (void)np;
  TGFields<3, 3> f = {{d_Mp, d_mom_px, d_mom_py}, {d_Mv, d_mom_vx, d_mom_vy}};
  return tiled_gather<3, 3, false>(nrows, hx, hy, cstart, cptcls, ccidx, coloff, d_x,
                         d_y, f);
}

int amsc_gradZ_to_particles(int np, int nrows, double hx, double hy,
                            const double *d_x, const double *d_y,
                            const int *d_p2g, const double *d_Z,
                            double *d_dZxp, double *d_dZyp) {
  if (int rc = check_n(np, "np")) return rc;
  LAUNCH1D(gradZ_kernel, np, np, nrows, hx, hy, d_x, d_y, d_p2g, d_Z, d_dZxp,
           d_dZyp);
  return AMSC_OK;
}

int amsc_gravity_force(int np, double g_c, const double *d_hp,
                       const double *d_H, const double *d_Mp,
                       const double *d_dZxp, const double *d_dZyp,
                       double *d_Fpx, double *d_Fpy) {
  if (int rc = check_n(np, "np")) return rc;
  LAUNCH1D(gravity_kernel, np, np, g_c, d_hp, d_H, d_Mp, d_dZxp, d_dZyp,
           d_Fpx, d_Fpy);
  return AMSC_OK;
}

int amsc_friction_particle_force(int np, const double *d_Ap,
                                 const double *d_Fb_x, const double *d_Fb_y,
                                 double *d_Fric_px, double *d_Fric_py) {
  if (int rc = check_n(np, "np")) return rc;
  LAUNCH1D(friction_kernel, np, np, d_Ap, d_Fb_x, d_Fb_y, d_Fric_px,
           d_Fric_py);
  return AMSC_OK;
}

int amsc_scatter_gravity_force(int np, int nrows, double hx, double hy,
                               const int *cstart, const int *cptcls,
                               const int *ccidx, const int *coloff,
                               const double *d_x, const double *d_y,
                               const double *d_Fpx, const double *d_Fpy,
                               double *d_FPxv, double *d_FPyv) {
  (void)np;
  TGFields<2, 2> f = {{d_Fpx, d_Fpy}, {d_FPxv, d_FPyv}};
  return tiled_gather<2, 2, false>(nrows, hx, hy, cstart, cptcls, ccidx, coloff, d_x,
                         d_y, f);
}

int amsc_scatter_friction_force(int np, int nrows, double hx, double hy,
                                const int *cstart, const int *cptcls,
                                const int *ccidx, const int *coloff,
                                const double *d_x, const double *d_y,
                                const double *d_Fric_px,
                                const double *d_Fric_py, double *d_Fric_x,
                                double *d_Fric_y) {
  (void)np;
  TGFields<2, 2> f = {{d_Fric_px, d_Fric_py}, {d_Fric_x, d_Fric_y}};
  return tiled_gather<2, 2, false>(nrows, hx, hy, cstart, cptcls, ccidx, coloff, d_x,
                         d_y, f);
}

int amsc_compute_F_ext(int nn, const double *d_FPxv, const double *d_FPyv,
                       const double *d_Fric_x, const double *d_Fric_y,
                       double *d_F_ext_vx, double *d_F_ext_vy) {
  if (int rc = check_n(nn, "nn")) return rc;
  LAUNCH1D(fext_kernel, nn, nn, d_FPxv, d_FPyv, d_Fric_x, d_Fric_y,
           d_F_ext_vx, d_F_ext_vy);
  return AMSC_OK;
}

int amsc_scatter_internal_forces(int np, int nrows, double hx, double hy,
                                 const int *cstart, const int *cptcls,
                                 const int *ccidx, const int *coloff,
                                 const double *d_x, const double *d_y,
                                 const double *d_Vp, const double *d_F11,
                                 const double *d_F12, const double *d_F21,
                                 const double *d_F22, double *d_F_int_vx,
                                 double *d_F_int_vy) {
  (void)np;
  TGFields<5, 2> f = {{d_Vp, d_F11, d_F12, d_F21, d_F22}, {d_F_int_vx, d_F_int_vy}};
  return tiled_gather<5, 2, true>(nrows, hx, hy, cstart, cptcls, ccidx, coloff,
                                  d_x, d_y, f);
}

int amsc_nodal_force_and_momentum(int nn, double dt, const double *d_F_ext_vx,
                                  const double *d_F_ext_vy,
                                  const double *d_F_int_vx,
                                  const double *d_F_int_vy, double *d_Ftot_vx,
                                  double *d_Ftot_vy, double *d_mom_vx,
                                  double *d_mom_vy) {
  (void)dt; (void)d_mom_vx; (void)d_mom_vy;
  if (int rc = check_n(nn, "nn")) return rc;
  LAUNCH1D(ftot_kernel, nn, nn, d_F_ext_vx, d_F_ext_vy, d_F_int_vx,
           d_F_int_vy, d_Ftot_vx, d_Ftot_vy);
  return AMSC_OK;
}

int amsc_nodal_velocity_and_bc(int nn, int nrows, int ncols, double dt,
                               int bc_flag, const double *d_Mv,
                               const double *d_Ftot_vx,
                               const double *d_Ftot_vy, double *d_mom_vx,
                               double *d_mom_vy, double *d_avx, double *d_avy,
                               double *d_vvx, double *d_vvy, double *d_vvxL,
                               double *d_vvyL) {
  if (int rc = check_n(nn, "nn")) return rc;
  if (nrows <= 0) return fail(AMSC_ERR_INVALID, "nrows must be positive");
  LAUNCH1D(nodal_vel_bc_kernel, nn, nn, nrows, ncols, dt, bc_flag ? 1 : 0,
           d_Mv, d_Ftot_vx, d_Ftot_vy, d_mom_vx, d_mom_vy, d_avx, d_avy,
           d_vvx, d_vvy, d_vvxL, d_vvyL);
  return AMSC_OK;
}

int amsc_g2p_velocity_and_advect(int np, int nrows, int ncols, double hx,
                                 double hy, double dt, const int *d_p2g,
                                 const double *d_vvxL, const double *d_vvyL,
                                 const double *d_avx, const double *d_avy,
                                 double *d_x, double *d_y, double *d_vpx,
                                 double *d_vpy, double *d_apx, double *d_apy,
                                 double *d_mom_px, double *d_mom_py,
                                 double *d_vpxL, double *d_vpyL,
                                 const double *d_Mp) {
  if (int rc = check_n(np, "np")) return rc;
  if (nrows <= 0) return fail(AMSC_ERR_INVALID, "nrows must be positive");
  LAUNCH1D(g2p_advect_kernel, np, np, nrows, ncols, hx, hy, dt, d_p2g, d_vvxL,
           d_vvyL, d_avx, d_avy, d_x, d_y, d_vpx, d_vpy, d_apx, d_apy,
           d_mom_px, d_mom_py, d_vpxL, d_vpyL, d_Mp);
  return AMSC_OK;
}

int amsc_g2pd_gradients_and_height_update(
    int np, int nrows, double hx, double hy, double dt, const double *d_x,
    const double *d_y, const int *d_p2g, const double *d_vvxL,
    const double *d_vvyL, double *d_vpx_dx, double *d_vpx_dy,
    double *d_vpy_dx, double *d_vpy_dy, double *d_hp, double *d_Vp,
    double *d_Ap, double *min_sc_out, double *min_hp_out) {
  (void)min_sc_out; (void)min_hp_out;
  if (int rc = check_n(np, "np")) return rc;
  if (nrows <= 0) return fail(AMSC_ERR_INVALID, "nrows must be positive");
  LAUNCH1D(g2pd_height_kernel, np, np, nrows, hx, hy, dt, d_x, d_y, d_p2g,
           d_vvxL, d_vvyL, d_vpx_dx, d_vpx_dy, d_vpy_dx, d_vpy_dy, d_hp, d_Vp,
           d_Ap);
  return AMSC_OK;
}

int amsc_voellmy_friction(int np, double rho, double g_c, double tan_fa,
                          double xi, double fric_on, const double *d_vpx,
                          const double *d_vpy, const double *d_hp,
                          double *d_Fb_x, double *d_Fb_y) {
  if (int rc = check_n(np, "np")) return rc;
  LAUNCH1D(voellmy_kernel, np, np, rho, g_c, tan_fa, xi, fric_on, d_vpx,
           d_vpy, d_hp, d_Fb_x, d_Fb_y);
  return AMSC_OK;
}

int amsc_bingham_stress_update(
    int np, double rho, double g_c, double mu, double tau_Y, double cc,
    double A_coeff, double C_coeff, const double *d_vpx, const double *d_vpy,
    const double *d_hp, const double *d_H, const double *d_vpx_dx,
    const double *d_vpx_dy, const double *d_vpy_dx, const double *d_vpy_dy,
    double *d_F11, double *d_F12, double *d_F21, double *d_F22,
    double *max_coeff_out, double *min_hp_out, double *max_nv_out,
    double *max_F11_out, double *max_dev_out, double *max_pres_out,
    double *max_Dxx_out) {
  (void)max_coeff_out; (void)min_hp_out; (void)max_nv_out; (void)max_F11_out;
  (void)max_dev_out; (void)max_pres_out; (void)max_Dxx_out;
  if (int rc = check_n(np, "np")) return rc;
  LAUNCH1D(bingham_gpu_kernel, np, np, rho, g_c, mu, tau_Y, cc, A_coeff,
           C_coeff, d_vpx, d_vpy, d_hp, d_H, d_vpx_dx, d_vpx_dy, d_vpy_dx,
           d_vpy_dy, d_F11, d_F12, d_F21, d_F22);
  return AMSC_OK;
}

int amsc_g2p_height(int np, int nrows, double hx, double hy,
                    const double *d_x, const double *d_y, const int *d_p2g,
                    const double *d_Z, const double *d_hp, double *d_Zp,
                    double *d_hpZ) {
  if (int rc = check_n(np, "np")) return rc;
  if (nrows <= 0) return fail(AMSC_ERR_INVALID, "nrows must be positive");
  LAUNCH1D(g2p_height_kernel, np, np, nrows, hx, hy, d_x, d_y, d_p2g, d_Z,
           d_hp, d_Zp, d_hpZ);
  return AMSC_OK;
}

int amsc_scatter_plotvars(int np, int nrows, double hx, double hy,
                          const int *cstart, const int *cptcls,
                          const int *ccidx, const int *coloff,
                          const double *d_x, const double *d_y,
                          const double *d_Mp, const double *d_vpx,
                          const double *d_vpy, const double *d_apx,
                          const double *d_apy, const double *d_hp,
                          double *d_pv_rho, double *d_pv_vvx,
                          double *d_pv_vvy, double *d_pv_avx,
                          double *d_pv_avy, double *d_HV) {
  (void)np;
  TGFields<6, 6> f = {{d_Mp, d_vpx, d_vpy, d_apx, d_apy, d_hp},
                 {d_pv_rho, d_pv_vvx, d_pv_vvy, d_pv_avx, d_pv_avy, d_HV}};
  return tiled_gather<6, 6, false>(nrows, hx, hy, cstart, cptcls, ccidx, coloff, d_x,
                         d_y, f);
}

int amsc_normalize_plotvars(int nn, double *d_pv_rho, double *d_pv_vvx,
                            double *d_pv_vvy, double *d_pv_avx,
                            double *d_pv_avy) {
  if (int rc = check_n(nn, "nn")) return rc;
  LAUNCH1D(normalize_kernel, nn, nn, d_pv_rho, d_pv_vvx, d_pv_vvy, d_pv_avx,
           d_pv_avy);
  return AMSC_OK;
}

int amsc_cfl_dt(int np, const double *d_vpx, const double *d_vpy,
                const double *d_hp, double g, double cfl, double hx, double hy,
                int it, double *dt, double *cel_out) {
  if (np <= 0 || !d_vpx || !d_vpy || !d_hp || !dt)
    return fail(AMSC_ERR_INVALID, "amsc_cfl_dt: bad arguments");
  const int nb = std::min((int)grid_for(np, kCflBlock), kCflMaxBlocks);
  // one workspace of the largest size, held (with the lock) until the
  // read-back below has completed
  static DevCache cache;
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  void *p = nullptr;
  if (int rc = dev_cache_get(cache, sizeof(double) * (3 * kCflMaxBlocks + 3),
                             &p, "amsc_cfl_dt"))
    return rc;
  double *part = static_cast<double *>(p);
  hipStream_t st = current_stream();
  cfl_partial_kernel<<<nb, kCflBlock, 0, st>>>(np, d_vpx, d_vpy, d_hp, part);
  cfl_final_kernel<<<1, kCflBlock, 0, st>>>(nb, part, part + 3 * nb);
  double m[3] = {0, 0, 0};
  int rc = AMSC_OK;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(m, part + 3 * nb, sizeof(m), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    rc = fail(AMSC_ERR_HIP, "amsc_cfl_dt: reduction failed");
  if (rc) return rc;
  // gpu_offload.cpp:293-298
  const double max_vx = m[0], max_vy = m[1], hmax = m[2];
  const double a = std::sqrt(g * hmax) + max_vx, b = std::sqrt(g * hmax) + max_vy;
  const double cel = std::fabs(std::max(a, b));
  if (cel_out) *cel_out = cel;
  if (it > 0) *dt = cfl * std::min(hx, hy) / (1e-2 + cel);
  return AMSC_OK;
}

}  // extern "C"
