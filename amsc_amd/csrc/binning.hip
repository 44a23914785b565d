// binning.hip -- particles by cell on the device (the quadgrid
// init_particle_mesh that gpu_offload.cpp:546-558 rebuilt on the host every
// step, with a PCIe round trip), plus the integer exclusive scan it uses.
//
// Pipeline (all integer, so results are exact and run-to-run identical):
//   1. key = cell of the particle, slot = atomicAdd(count[key]) aggregated
//      per wavefront (fused into the particle-update kernel of a time step);
//   2. cstart = exclusive scan of count (count reset to 0 on the way);
//   3. cptcls[cstart[key] + slot] = id;
//   4. every cell's slice of cptcls sorted ascending: atomics fix which
//      slots a cell receives but not their order, so this step restores the
//      stable order of a CPU counting sort (ascending id within a cell).
#include "binning.hpp"
#include "amsc_mpm.h"

namespace amsc {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanBlock * kScanItems;  // 4096

// Exclusive scan of one value per thread over a 256-thread block.  Returns
// the exclusive prefix; *total receives the block total.
__device__ __forceinline__ int block_exclusive_scan(int v, int *total) {
  __shared__ int s_wave[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(inc, d);
    if (lane >= d) inc += t;
  }
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  int wbase = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) {
    if (w < wave) wbase += s_wave[w];
    tot += s_wave[w];
  }
  __syncthreads();
  *total = tot;
  return wbase + inc - v;
}

__global__ void __launch_bounds__(kScanBlock)
    scan_partials_kernel(const int *__restrict__ in, int n, int *bsum,
                         const int *__restrict__ active) {
  if (active && !*active) return;
  const long long base = (long long)blockIdx.x * kScanChunk;
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const long long j = base + (long long)i * kScanBlock + threadIdx.x;
    if (j < n) s += in[j];
  }
  int total;
  (void)block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Single block: exclusive scan of bsum[0..nb) in place, total in bsum[nb].
__global__ void __launch_bounds__(kScanBlock)
    scan_bsums_kernel(int *bsum, int nb, const int *__restrict__ active) {
  if (active && !*active) return;
  int carry = 0;
  for (int base = 0; base < nb; base += kScanBlock) {
    const int j = base + threadIdx.x;
    const int v = j < nb ? bsum[j] : 0;
    int total;
    const int ex = block_exclusive_scan(v, &total);
    if (j < nb) bsum[j] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ void __launch_bounds__(kScanBlock)
    scan_down_kernel(int *in, int *out, int n, const int *bsum,
                     int zero_in, const int *__restrict__ active) {
  if (active && !*active) return;
  __shared__ int s_v[kScanChunk];
  const long long base = (long long)blockIdx.x * kScanChunk;
  // coalesced load, then each thread owns kScanItems consecutive values
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int l = i * kScanBlock + threadIdx.x;
    const long long j = base + l;
    s_v[l] = j < n ? in[j] : 0;
    if (zero_in && j < n) in[j] = 0;
  }
  __syncthreads();
  int vals[kScanItems];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    vals[i] = s_v[threadIdx.x * kScanItems + i];
    s += vals[i];
  }
  int total;
  int run = block_exclusive_scan(s, &total) + bsum[blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    s_v[threadIdx.x * kScanItems + i] = run;
    run += vals[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int l = i * kScanBlock + threadIdx.x;
    const long long j = base + l;
    if (j < n) out[j] = s_v[l];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    const int nb = gridDim.x;
    out[n] = bsum[nb];
  }
}

__global__ void set_zero_kernel(int *p, const int *__restrict__ active) {
  if (active && !*active) return;
  *p = 0;
}

__global__ void key_count_kernel(int np, const double *__restrict__ x,
                                 const double *__restrict__ y, int nrows,
                                 int ncols, double hx, double hy, int *key,
                                 int *slot, int *count,
                                 const int *__restrict__ active) {
  if (active && !*active) return;
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = ip < np;
  const int k = valid ? cell_of(x[ip], y[ip], nrows, ncols, hx, hy) : 0;
  const int s = wave_slot(count, k, valid);
  if (valid) {
    key[ip] = k;
    slot[ip] = s;
  }
}

__global__ void place_kernel(int np, const int *__restrict__ np_dev,
                             const int *__restrict__ key,
                             const int *__restrict__ slot,
                             const int *__restrict__ cstart, int *cptcls,
                             const int *__restrict__ active) {
  if (active && !*active) return;
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= (np_dev ? *np_dev : np)) return;
  cptcls[cstart[key[ip]] + slot[ip]] = ip;
}

constexpr int kSortLocal = 32;

__global__ void cell_sort_kernel(int nc, const int *__restrict__ cstart,
                                 int *cptcls, const int *__restrict__ active) {
  if (active && !*active) return;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nc) return;
  const int b = cstart[cell], e = cstart[cell + 1], n = e - b;
  if (n < 2) return;
  if (n <= kSortLocal) {
    int v[kSortLocal];
    bool sorted = true;
    for (int i = 0; i < n; ++i) {
      v[i] = cptcls[b + i];
      if (i > 0 && v[i] < v[i - 1]) sorted = false;
    }
    if (sorted) return;
    for (int i = 1; i < n; ++i) {
      const int t = v[i];
      int j = i - 1;
      while (j >= 0 && v[j] > t) {
        v[j + 1] = v[j];
        --j;
      }
      v[j + 1] = t;
    }
    for (int i = 0; i < n; ++i) cptcls[b + i] = v[i];
  } else {
    for (int i = b + 1; i < e; ++i) {
      const int t = cptcls[i];
      int j = i - 1;
      while (j >= b && cptcls[j] > t) {
        cptcls[j + 1] = cptcls[j];
        --j;
      }
      cptcls[j + 1] = t;
    }
  }
}

__global__ void colouring_kernel(int nrows, int ncols, int *ccidx, int o1,
                                 int o2, int o3) {
  const int nc = nrows * ncols;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nc) return;
  const int r = cell % nrows, c = cell / nrows;
  const int k = (r & 1) + 2 * (c & 1);
  const int nr_p = (nrows - (r & 1) + 1) / 2;
  const int off = k == 0 ? 0 : k == 1 ? o1 : k == 2 ? o2 : o3;
  ccidx[off + (c >> 1) * nr_p + (r >> 1)] = cell;
}

}  // namespace

int scan_reserve(ScanWork &w, int n) {
  const int nb = (int)((n + (long long)kScanChunk - 1) / kScanChunk);
  if (w.nb_cap >= nb + 1) return AMSC_OK;
  if (w.bsum) (void)hipFree(w.bsum);
  w.bsum = nullptr;
  w.nb_cap = 0;
  if (hipMalloc(&w.bsum, sizeof(int) * (nb + 1)) != hipSuccess)
    return fail(AMSC_ERR_NOMEM, "scan workspace allocation failed");
  w.nb_cap = nb + 1;
  return AMSC_OK;
}

int exclusive_scan(const int *in, int *out, int n, bool zero_in, ScanWork &w,
                   hipStream_t st, int *total_host, const int *active) {
  if (n <= 0) {
    set_zero_kernel<<<1, 1, 0, st>>>(out, active);
    AMSC_LAUNCH_CHECK();
    if (total_host) {
      AMSC_HIP(hipStreamSynchronize(st));
      *total_host = 0;
    }
    return AMSC_OK;
  }
  const int nb = (int)((n + (long long)kScanChunk - 1) / kScanChunk);
  if (int rc = scan_reserve(w, n)) return rc;
  scan_partials_kernel<<<nb, kScanBlock, 0, st>>>(in, n, w.bsum, active);
  scan_bsums_kernel<<<1, kScanBlock, 0, st>>>(w.bsum, nb, active);
  scan_down_kernel<<<nb, kScanBlock, 0, st>>>(const_cast<int *>(in), out, n,
                                              w.bsum, zero_in ? 1 : 0, active);
  AMSC_LAUNCH_CHECK();
  if (total_host) {
    AMSC_HIP(hipMemcpyAsync(total_host, out + n, sizeof(int),
                            hipMemcpyDeviceToHost, st));
    AMSC_HIP(hipStreamSynchronize(st));
  }
  return AMSC_OK;
}

int binning_alloc(Binning &b, int nrows, int ncols) {
  b.nrows = nrows;
  b.ncols = ncols;
  const size_t nc = (size_t)nrows * ncols;
  if (hipMalloc(&b.count, sizeof(int) * (nc > 0 ? nc : 1)) != hipSuccess ||
      hipMalloc(&b.cstart, sizeof(int) * (nc + 1)) != hipSuccess)
    return fail(AMSC_ERR_NOMEM, "binning allocation failed");
  AMSC_HIP(hipMemset(b.count, 0, sizeof(int) * (nc > 0 ? nc : 1)));
  AMSC_HIP(hipMemset(b.cstart, 0, sizeof(int) * (nc + 1)));
  return AMSC_OK;
}

void binning_free(Binning &b) {
  if (b.count) (void)hipFree(b.count);
  if (b.cstart) (void)hipFree(b.cstart);
  if (b.scan.bsum) (void)hipFree(b.scan.bsum);
  b = Binning();
}

int launch_key_count(int np, const double *x, const double *y, int nrows,
                     int ncols, double hx, double hy, int *key, int *slot,
                     int *count, const int *active, hipStream_t st) {
  if (np <= 0) return AMSC_OK;
  key_count_kernel<<<grid_for(np, 256), 256, 0, st>>>(
      np, x, y, nrows, ncols, hx, hy, key, slot, count, active);
  AMSC_LAUNCH_CHECK();
  return AMSC_OK;
}

int launch_finish_binning(Binning &b, int np, const int *key,
                          const int *slot, int *cptcls, const int *active,
                          hipStream_t st, const int *np_dev) {
  const int nc = b.nrows * b.ncols;
  if (int rc = exclusive_scan(b.count, b.cstart, nc, true, b.scan, st,
                              nullptr, active))
    return rc;
  if (np > 0) {
    place_kernel<<<grid_for(np, 256), 256, 0, st>>>(np, np_dev, key, slot,
                                                    b.cstart, cptcls, active);
    cell_sort_kernel<<<grid_for(nc, 256), 256, 0, st>>>(nc, b.cstart, cptcls,
                                                        active);
    AMSC_LAUNCH_CHECK();
  }
  return AMSC_OK;
}

int launch_colouring(int nrows, int ncols, int *ccidx, int *coloff,
                     hipStream_t st) {
  long long cnt[4];
  for (int k = 0; k < 4; ++k) {
    const int rp = k & 1, cp = k >> 1;
    const long long nr = (nrows - rp + 1) / 2, ncl = (ncols - cp + 1) / 2;
    cnt[k] = (nr > 0 ? nr : 0) * (ncl > 0 ? ncl : 0);
  }
  coloff[0] = 0;
  for (int k = 0; k < 4; ++k) coloff[k + 1] = (int)(coloff[k] + cnt[k]);
  const int nc = nrows * ncols;
  if (ccidx && nc > 0) {
    colouring_kernel<<<grid_for(nc, 256), 256, 0, st>>>(
        nrows, ncols, ccidx, coloff[1], coloff[2], coloff[3]);
    AMSC_LAUNCH_CHECK();
  }
  return AMSC_OK;
}

}  // namespace amsc

using namespace amsc;

extern "C" int amsc_bin_particles(int np, int nrows, int ncols, double hx,
                                  double hy, const double *d_x,
                                  const double *d_y, int *d_p2g,
                                  int *d_cstart, int *d_cptcls, int *d_ccidx,
                                  int *coloff) {
  if (np < 0 || nrows <= 0 || ncols <= 0)
    return fail(AMSC_ERR_INVALID, "amsc_bin_particles: bad sizes");
  if ((long long)nrows * ncols >= (1ll << 31))
    return fail(AMSC_ERR_INVALID, "amsc_bin_particles: grid too large");
  if (!d_p2g || !d_cstart || (np > 0 && !d_cptcls))
    return fail(AMSC_ERR_INVALID, "amsc_bin_particles: NULL output");
  hipStream_t st = current_stream();
  Binning b;
  if (int rc = binning_alloc(b, nrows, ncols)) {
    binning_free(b);
    return rc;
  }
  int *slot = nullptr;
  if (np > 0 && hipMalloc(&slot, sizeof(int) * np) != hipSuccess) {
    binning_free(b);
    return fail(AMSC_ERR_NOMEM, "amsc_bin_particles: slot allocation");
  }
  int rc = launch_key_count(np, d_x, d_y, nrows, ncols, hx, hy, d_p2g, slot,
                            b.count, nullptr, st);
  if (!rc) rc = launch_finish_binning(b, np, d_p2g, slot, d_cptcls, nullptr, st);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(d_cstart, b.cstart,
                                  sizeof(int) * ((size_t)nrows * ncols + 1),
                                  hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) rc = fail(AMSC_ERR_HIP, hipGetErrorString(e));
  }
  if (!rc && coloff) rc = launch_colouring(nrows, ncols, d_ccidx, coloff, st);
  // the temporaries are released after the stream has used them
  hipError_t e = hipStreamSynchronize(st);
  if (!rc && e != hipSuccess) rc = fail(AMSC_ERR_HIP, hipGetErrorString(e));
  if (slot) (void)hipFree(slot);
  binning_free(b);
  return rc;
}
