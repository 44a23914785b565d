// binning.hip -- particles by cell on the device (the quadgrid
// init_particle_mesh that gpu_offload.cpp:546-558 rebuilt on the host every
// step, with a PCIe round trip), plus the integer exclusive scan it uses.
//
// Pipeline (all integer, so results are exact and run-to-run identical):
//   1. key = cell of the particle, slot = atomicAdd(count[key]) aggregated
//      per wavefront (fused into the particle-update kernel of a time step);
//   2. cstart[1..nc] = inclusive scan of count (one single-pass decoupled
//      look-back scan, rocPRIM), cstart[0] = 0 from allocation;
//   3. cptcls[cstart[key] + slot] = id;
//   4. every cell's slice of cptcls sorted ascending: atomics fix which
//      slots a cell receives but not their order, so this step restores the
//      stable order of a CPU counting sort (ascending id within a cell);
//      the cell's count is reset to 0 for the next binning.
#include "binning.hpp"
#include "amsc_mpm.h"

#include <rocprim/device/device_scan.hpp>

namespace amsc {

namespace {

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

// Heap sort of a[0..n) in place (cells above kSortLocal particles: O(n log n)
// however many particles pile into one cell).
__device__ void heap_sort(int *a, int n) {
  auto sift = [a](int root, int end) {
    while (2 * root + 1 <= end) {
      const int child = 2 * root + 1;
      int sw = root;
      if (a[sw] < a[child]) sw = child;
      if (child + 1 <= end && a[sw] < a[child + 1]) sw = child + 1;
      if (sw == root) return;
      const int t = a[root];
      a[root] = a[sw];
      a[sw] = t;
      root = sw;
    }
  };
  for (int start = (n - 2) / 2; start >= 0; --start) sift(start, n - 1);
  for (int end = n - 1; end > 0; --end) {
    const int t = a[0];
    a[0] = a[end];
    a[end] = t;
    sift(0, end - 1);
  }
}

__global__ void cell_sort_kernel(int nc, const int *__restrict__ cstart,
                                 int *cptcls, int *count,
                                 const int *__restrict__ active) {
  if (active && !*active) return;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= nc) return;
  count[cell] = 0;
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
    heap_sort(cptcls + b, n);
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

int scan_bytes(int n, size_t *bytes) {
  *bytes = 0;
  const hipError_t e = rocprim::inclusive_scan(
      nullptr, *bytes, static_cast<const int *>(nullptr),
      static_cast<int *>(nullptr), (size_t)(n > 0 ? n : 1),
      rocprim::plus<int>());
  if (e != hipSuccess)
    return fail(AMSC_ERR_HIP, std::string("scan sizing: ") + hipGetErrorString(e));
  return AMSC_OK;
}

int scan_reserve(ScanWork &w, int n) {
  size_t bytes = 0;
  if (int rc = scan_bytes(n, &bytes)) return rc;
  if (w.bytes >= bytes && w.tmp) return AMSC_OK;
  if (!w.owned)
    return fail(AMSC_ERR_INVALID, "scan workspace too small for the scan");
  if (w.tmp) (void)hipFree(w.tmp);
  w.tmp = nullptr;
  w.bytes = 0;
  if (hipMalloc(&w.tmp, bytes) != hipSuccess)
    return fail(AMSC_ERR_NOMEM, "scan workspace allocation failed");
  w.bytes = bytes;
  return AMSC_OK;
}

int exclusive_scan(const int *in, int *out, int n, ScanWork &w,
                   hipStream_t st, int *total_host) {
  if (n > 0) {
    if (int rc = scan_reserve(w, n)) return rc;
    size_t bytes = w.bytes;
    const hipError_t e = rocprim::inclusive_scan(w.tmp, bytes, in, out + 1,
                                                 (size_t)n, rocprim::plus<int>(),
                                                 st);
    if (e != hipSuccess)
      return fail(AMSC_ERR_HIP, std::string("scan: ") + hipGetErrorString(e));
  }
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
  if (b.scan.tmp) (void)hipFree(b.scan.tmp);
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
  if (active) return fail(AMSC_ERR_INVALID, "binning: unsupported gating");
  if (int rc = exclusive_scan(b.count, b.cstart, nc, b.scan, st, nullptr))
    return rc;
  if (np > 0)
    place_kernel<<<grid_for(np, 256), 256, 0, st>>>(np, np_dev, key, slot,
                                                    b.cstart, cptcls, active);
  cell_sort_kernel<<<grid_for(nc, 256), 256, 0, st>>>(nc, b.cstart, cptcls,
                                                      b.count, active);
  AMSC_LAUNCH_CHECK();
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
