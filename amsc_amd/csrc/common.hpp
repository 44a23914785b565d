// common.hpp -- shared device helpers and error plumbing for the gfx950
// depth-averaged MPM library.
//
// Grid layout (quadgrid, as ported in main/gpu_kernels.cpp:7-31):
//   node id  r + c*(nrows+1)   cell id  r + c*nrows
//   local nodes 0:(c,r) 1:(c,r+1) 2:(c+1,r) 3:(c+1,r+1)
// Bilinear basis with xi = (x - c*hx)/hx, eta = (y - r*hy)/hy, evaluated
// in the reference's expression order (gpu_kernels.cpp:33-59), so that with
// FMA contraction disabled every value is bit-identical to the reference's.
#pragma once

#include <cmath>
#include <hip/hip_runtime.h>

#include "amsc_mpm.h"

#include <cstdint>
#include <string>

namespace amsc {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
hipStream_t current_stream();
void set_stream(hipStream_t s);
const char *last_error_cstr();

#define AMSC_HIP(call)                                                       \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::amsc::fail(AMSC_ERR_HIP, std::string(#call) + ": " +          \
                                            hipGetErrorString(e_));          \
  } while (0)

#define AMSC_LAUNCH_CHECK()                                                  \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess)                                                    \
      return ::amsc::fail(AMSC_ERR_HIP, std::string("kernel launch: ") +     \
                                            hipGetErrorString(e_));          \
  } while (0)

inline unsigned grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

// ----------------------------------------------------------- geometry
__device__ __forceinline__ int node_id(int inode, int c, int r, int nrows) {
  const int bl = r + c * (nrows + 1);
  switch (inode) {
  case 0: return bl;
  case 1: return bl + 1;
  case 2: return bl + (nrows + 1);
  default: return bl + (nrows + 2);
  }
}

// Local coordinates of a point in cell (c, r).
struct Local {
  double xi, eta;
};

__device__ __forceinline__ Local local_coords(double x, double y, int c,
                                              int r, double hx, double hy) {
  Local L;
  L.xi = (x - c * hx) / hx;
  L.eta = (y - r * hy) / hy;
  return L;
}

// N_inode (gpu_kernels.cpp:33-41)
__device__ __forceinline__ double shape(const Local &L, int inode) {
  switch (inode) {
  case 3: return L.xi * L.eta;
  case 2: return L.xi * (1.0 - L.eta);
  case 1: return (1.0 - L.xi) * L.eta;
  default: return (1.0 - L.xi) * (1.0 - L.eta);
  }
}

// shape, shape_dx and shape_dy for a corner known only at run time (a
// node's cells in a caller's colour order), without branches: every case
// is one product of two factors, each chosen by one bit of inode, so the
// values are the switch's bit for bit.
__device__ __forceinline__ double shape_sel(const Local &L, int inode) {
  const double a = (inode & 2) ? L.xi : (1.0 - L.xi);
  const double b = (inode & 1) ? L.eta : (1.0 - L.eta);
  return a * b;
}
__device__ __forceinline__ double shape_dx_sel(const Local &L, int inode,
                                               double hx) {
  const double f = (inode & 2) ? (1.0 / hx) : (-1.0 / hx);
  const double b = (inode & 1) ? L.eta : (1.0 - L.eta);
  return f * b;
}
__device__ __forceinline__ double shape_dy_sel(const Local &L, int inode,
                                               double hy) {
  const double a = (inode & 2) ? L.xi : (1.0 - L.xi);
  const double f = (inode & 1) ? (1.0 / hy) : (-1.0 / hy);
  return a * f;
}

// dN_inode/dx (gpu_kernels.cpp:43-59, idir 0)
__device__ __forceinline__ double shape_dx(const Local &L, int inode,
                                           double hx) {
  switch (inode) {
  case 3: return (1.0 / hx) * L.eta;
  case 2: return (1.0 / hx) * (1.0 - L.eta);
  case 1: return (-1.0 / hx) * L.eta;
  default: return (-1.0 / hx) * (1.0 - L.eta);
  }
}

// dN_inode/dy (gpu_kernels.cpp:43-59, idir 1)
__device__ __forceinline__ double shape_dy(const Local &L, int inode,
                                           double hy) {
  switch (inode) {
  case 3: return L.xi * (1.0 / hy);
  case 2: return L.xi * (-1.0 / hy);
  case 1: return (1.0 - L.xi) * (1.0 / hy);
  default: return (1.0 - L.xi) * (-1.0 / hy);
  }
}

// Basis-gradient factors 1/hx, -1/hx, 1/hy, -1/hy, computed once on the
// host (IEEE division gives the same values the expressions of
// gpu_kernels.cpp:43-59 produce in every thread).
struct Recip {
  double ix, nix, iy, niy;
  // hx and hy are powers of two: u / h is then u * (1/h) exactly (the same
  // single rounding of the same real number), without the division sequence
  int pow2;
};
inline Recip make_recip(double hx, double hy) {
  int ex = 0, ey = 0;
  const bool p2 = std::fpclassify(hx) == FP_NORMAL &&
                  std::fpclassify(hy) == FP_NORMAL &&
                  std::frexp(hx, &ex) == 0.5 && std::frexp(hy, &ey) == 0.5;
  return Recip{1.0 / hx, -1.0 / hx, 1.0 / hy, -1.0 / hy, p2 ? 1 : 0};
}

// h is a positive normal power of two (its mantissa bits are zero).
__device__ __forceinline__ bool pow2_dev(double h) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(h);
  const unsigned e = (unsigned)(b >> 52) & 0x7FFu;
  return (b & 0x000FFFFFFFFFFFFFull) == 0 && e != 0 && e != 0x7FFu &&
         (b >> 63) == 0;
}

// make_recip on the device (kernels given hx, hy as arguments).
__device__ __forceinline__ Recip recip_dev(double hx, double hy) {
  Recip R;
  R.ix = 1.0 / hx;
  R.nix = -1.0 / hx;
  R.iy = 1.0 / hy;
  R.niy = -1.0 / hy;
  R.pow2 = pow2_dev(hx) && pow2_dev(hy) ? 1 : 0;
  return R;
}

// local_coords with the grid's reciprocals: the same values.
__device__ __forceinline__ Local local_coords(double x, double y, int c,
                                              int r, double hx, double hy,
                                              const Recip &R) {
  const double u = x - c * hx, v = y - r * hy;
  Local L;
  if (R.pow2) {
    L.xi = u * R.ix;
    L.eta = v * R.iy;
  } else {
    L.xi = u / hx;
    L.eta = v / hy;
  }
  return L;
}

__device__ __forceinline__ double shape_dx(const Local &L, int inode,
                                           const Recip &R) {
  switch (inode) {
  case 3: return R.ix * L.eta;
  case 2: return R.ix * (1.0 - L.eta);
  case 1: return R.nix * L.eta;
  default: return R.nix * (1.0 - L.eta);
  }
}

__device__ __forceinline__ double shape_dy(const Local &L, int inode,
                                           const Recip &R) {
  switch (inode) {
  case 3: return L.xi * R.iy;
  case 2: return L.xi * R.niy;
  case 1: return (1.0 - L.xi) * R.iy;
  default: return (1.0 - L.xi) * R.niy;
  }
}

// The basis in an arithmetic type R (the precision-generic time loop).  For
// R = double these are the expressions above, value for value.
template <typename R> struct RecipT {
  R ix, nix, iy, niy;
};
template <typename R> inline RecipT<R> make_recip_t(double hx, double hy) {
  return RecipT<R>{(R)(1.0 / hx), (R)(-1.0 / hx), (R)(1.0 / hy),
                   (R)(-1.0 / hy)};
}

template <typename R>
__device__ __forceinline__ R shape_t(R xi, R eta, int inode) {
  switch (inode) {
  case 3: return xi * eta;
  case 2: return xi * (R(1) - eta);
  case 1: return (R(1) - xi) * eta;
  default: return (R(1) - xi) * (R(1) - eta);
  }
}

template <typename R>
__device__ __forceinline__ R shape_dx_t(R xi, R eta, int inode,
                                        const RecipT<R> &rc) {
  (void)xi;
  switch (inode) {
  case 3: return rc.ix * eta;
  case 2: return rc.ix * (R(1) - eta);
  case 1: return rc.nix * eta;
  default: return rc.nix * (R(1) - eta);
  }
}

template <typename R>
__device__ __forceinline__ R shape_dy_t(R xi, R eta, int inode,
                                        const RecipT<R> &rc) {
  (void)eta;
  switch (inode) {
  case 3: return xi * rc.iy;
  case 2: return xi * rc.niy;
  case 1: return (R(1) - xi) * rc.iy;
  default: return (R(1) - xi) * rc.niy;
  }
}

// Exact division of a non-negative int (< 2^31) by a runtime constant d:
// q = (n * m) >> k with k = 31 + ceil(log2 d), m = ceil(2^k / d).  The
// error term n*(m*d - 2^k)/(d*2^k) stays below 1/d, so the floor is exact.
struct FastDiv {
  unsigned long long m;
  int k, d;
};
inline FastDiv make_fastdiv(int d) {
  int l = 0;
  while ((1ll << l) < d) ++l;
  FastDiv f;
  f.k = 31 + l;
  f.m = (unsigned long long)((((unsigned __int128)1 << f.k) + d - 1) / d);
  f.d = d;
  return f;
}
__device__ __forceinline__ int fdiv(int n, const FastDiv &f) {
  return (int)(((unsigned long long)(unsigned)n * f.m) >> f.k);
}

// Cell of a point: floor(x/hx), floor(y/hy) clamped to the grid.
__device__ __forceinline__ int cell_of(double x, double y, int nrows,
                                       int ncols, double hx, double hy) {
  int c = (int)floor(x / hx), r = (int)floor(y / hy);
  c = c < 0 ? 0 : (c > ncols - 1 ? ncols - 1 : c);
  r = r < 0 ? 0 : (r > nrows - 1 ? nrows - 1 : r);
  return r + c * nrows;
}

// cell_of with the grid's reciprocals (see Recip::pow2): the same cell.
__device__ __forceinline__ int cell_of(double x, double y, int nrows,
                                       int ncols, double hx, double hy,
                                       const Recip &R) {
  int c = (int)floor(R.pow2 ? x * R.ix : x / hx);
  int r = (int)floor(R.pow2 ? y * R.iy : y / hy);
  c = c < 0 ? 0 : (c > ncols - 1 ? ncols - 1 : c);
  r = r < 0 ? 0 : (r > nrows - 1 ? nrows - 1 : r);
  return r + c * nrows;
}

// Wave-wide helpers (64-lane wavefronts).
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread over the workgroup (blockDim.x a
// multiple of 64, at most 1024); *total = the workgroup's sum.  s_w holds 17
// ints; the caller synchronises before reusing it.
__device__ __forceinline__ int block_excl_scan(int v, int *s_w, int *total) {
  const int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  if (w == 0) {
    const int t = lane < nw ? s_w[lane] : 0;
    const int ti = wave_incl_scan(t);
    if (lane < nw) s_w[lane] = ti - t;
    if (lane == nw - 1) s_w[16] = ti;
  }
  __syncthreads();
  *total = s_w[16];
  return s_w[w] + inc - v;
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const unsigned lane = __lane_id();
  return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}
__device__ __forceinline__ unsigned long long lanemask_le() {
  return lanemask_lt() | (1ull << __lane_id());
}

}  // namespace amsc
