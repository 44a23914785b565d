// binning.hip -- particles by cell on the device (the quadgrid
// init_particle_mesh that gpu_offload.cpp:546-558 rebuilt on the host every
// step, with a PCIe round trip), plus the integer exclusive scan it uses.
//
// Pipeline (all integer, so results are exact and run-to-run identical):
//   1. key = cell of the particle, slot = atomicAdd(count[key]) aggregated
//      per wavefront (fused into the particle-update kernel of a time step);
//   2. cstart = exclusive scan of count (reduce-then-scan: tile sums, then
//      every tile adds up the sums before it and scans its values from
//      there; no look-back state to initialise, no waits between
//      workgroups), the counts reset to 0 for the next binning as they are
//      read;
//   3. cptcls[cstart[key] + slot] = id;
//   4. every cell's slice of cptcls sorted ascending: atomics fix which
//      slots a cell receives but not their order, so this step restores the
//      stable order of a CPU counting sort (ascending id within a cell).
#include "binning.hpp"
#include "amsc_mpm.h"

#include <climits>
#include <cstdint>
#include <mutex>

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

// kPlaceItems particles per thread, a workgroup's width apart: every load
// of a thread is issued before the first cell start it needs, so the two
// dependent round trips (key and slot, then the cell start) overlap across
// its particles (4M workload: 15.9 -> 13.9 us with two).
constexpr int kPlaceItems = 2;
__global__ void place_kernel(int np, const int *__restrict__ np_dev,
                             const int *__restrict__ key,
                             const int *__restrict__ slot,
                             const int *__restrict__ cstart, int *cptcls,
                             const int *__restrict__ active) {
  if (active && !*active) return;
  const int n = np_dev ? *np_dev : np;
  const int i0 = blockIdx.x * (kPlaceItems * blockDim.x) + threadIdx.x;
  int k[kPlaceItems], sl[kPlaceItems], c[kPlaceItems];
#pragma unroll
  for (int j = 0; j < kPlaceItems; ++j) {
    const int ip = i0 + j * blockDim.x;
    k[j] = ip < n ? key[ip] : 0;
    sl[j] = ip < n ? slot[ip] : 0;
  }
#pragma unroll
  for (int j = 0; j < kPlaceItems; ++j)
    c[j] = i0 + j * (int)blockDim.x < n ? cstart[k[j]] : 0;
#pragma unroll
  for (int j = 0; j < kPlaceItems; ++j) {
    const int ip = i0 + j * blockDim.x;
    if (ip < n) cptcls[c[j] + sl[j]] = ip;
  }
}

// Heap sort of a[0..n) in place (cells above 8 particles: O(n log n)
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

// Compare-exchange of a sorting network (constant indices: registers).
__device__ __forceinline__ void cx(int &a, int &b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo;
  b = hi;
}

__global__ void cell_sort_kernel(int nc, const int *__restrict__ cstart,
                                 int *cptcls, const int *__restrict__ active,
                                 int *big, int big_cap) {
  if (active && !*active) return;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = cell < nc ? cstart[cell] : 0;
  const int n = cell < nc ? cstart[cell + 1] - b : 0;
  if (n >= 2 && n <= 8) {
    // the common case (a few particles per cell): Batcher's odd-even merge
    // network for 8 keys (19 compare-exchanges), padded with INT_MAX
    int v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = i < n ? cptcls[b + i] : INT_MAX;
    bool sorted = true;
#pragma unroll
    for (int i = 1; i < 8; ++i) sorted = sorted && v[i - 1] <= v[i];
    if (sorted) return;
    cx(v[0], v[1]); cx(v[2], v[3]); cx(v[4], v[5]); cx(v[6], v[7]);
    cx(v[0], v[2]); cx(v[1], v[3]); cx(v[4], v[6]); cx(v[5], v[7]);
    cx(v[1], v[2]); cx(v[5], v[6]);
    cx(v[0], v[4]); cx(v[1], v[5]); cx(v[2], v[6]); cx(v[3], v[7]);
    cx(v[2], v[4]); cx(v[3], v[5]);
    cx(v[1], v[2]); cx(v[3], v[4]); cx(v[5], v[6]);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (i < n) cptcls[b + i] = v[i];
    return;
  }
  if (n <= 8) return;
  // a crowded cell: queued for a wavefront of wave_sort_kernel (one atomic
  // per wavefront), or sorted here if it is too large or the queue is full
  bool queue = big && n <= kWaveSortMax;
  if (big) {
    const unsigned long long m = __ballot(queue);
    int base = 0;
    const int lead = __ffsll((long long)m) - 1;
    if (queue && (int)__lane_id() == lead)
      base = atomicAdd(big, __popcll(m));
    base = __shfl(base, lead < 0 ? 0 : lead);
    const int at = base + __popcll(m & lanemask_lt());
    if (queue && at < big_cap)
      big[1 + at] = cell;
    else
      queue = false;
  }
  if (!queue) heap_sort(cptcls + b, n);
}

// One queued cell per wavefront: its ids copied to LDS (padded to a multiple
// of four with INT_MAX), then every id is stored at its rank, the number of
// smaller ids in the cell (ids are distinct).  O(n^2 / 64) LDS reads per
// cell, all lanes reading the same four-id word at a time.
constexpr int kSortWaves = 4;
__global__ void __launch_bounds__(64 * kSortWaves)
    wave_sort_kernel(const int *__restrict__ big, int big_cap,
                     const int *__restrict__ cstart, int *cptcls) {
  __shared__ int4 s_ids[kSortWaves][kWaveSortMax / 4];
  const int lane = (int)__lane_id(), w = threadIdx.x >> 6;
  int *ids = reinterpret_cast<int *>(s_ids[w]);
  const int queued = min(big[0], big_cap);
  for (int i = blockIdx.x * kSortWaves + w; i < queued;
       i += gridDim.x * kSortWaves) {
    const int cell = big[1 + i];
    const int b = cstart[cell], n = cstart[cell + 1] - b;
    const int n4 = (n + 3) >> 2;
    for (int j = lane; j < 4 * n4; j += 64)
      ids[j] = j < n ? cptcls[b + j] : INT_MAX;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int j = lane; j < n; j += 64) {
      const int v = ids[j];
      int rank = 0;
      for (int k = 0; k < n4; ++k) {
        const int4 q = s_ids[w][k];
        rank += (q.x < v) + (q.y < v) + (q.z < v) + (q.w < v);
      }
      cptcls[b + rank] = v;
    }
    // every lane has read the cell before the next one overwrites it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ---- reduce-then-scan of int counts (exact): 4096 values per tile, four
// consecutive values per thread of a 1024-thread workgroup.
constexpr int kScanThreads = 1024, kScanItems = 4,
              kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ void load4(const int *in, long long base, int n,
                                      int v[4]) {
  if (base + 3 < n && (reinterpret_cast<uintptr_t>(in + base) & 15) == 0) {
    const int4 q = *reinterpret_cast<const int4 *>(in + base);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = base + k < n ? in[base + k] : 0;
  }
}

__global__ void __launch_bounds__(kScanThreads)
    scan_sums_kernel(const int *__restrict__ in, int n, int *tile_sum) {
  __shared__ int s_w[17];
  int v[4];
  load4(in, (long long)blockIdx.x * kScanTile + threadIdx.x * kScanItems, n, v);
  int total;
  (void)block_excl_scan(v[0] + v[1] + v[2] + v[3], s_w, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// out[i+1] = prefix through in[i]; out[0] = 0; clear != null: clear[i] = 0
// after in[i] is read (clear == in for the binning counts).  Each tile adds
// up the sums of the tiles before it itself (a few hundred cached reads),
// which spares a one-workgroup pass between the two kernels.
__global__ void __launch_bounds__(kScanThreads)
    scan_apply_kernel(const int *in, int n, const int *__restrict__ tile_sum,
                      int *out, int *clear, int *zero) {
  __shared__ int s_w[17];
  int pre = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += kScanThreads)
    pre += tile_sum[b];
  int tile_off;
  (void)block_excl_scan(pre, s_w, &tile_off);
  __syncthreads();  // s_w is reused below
  const long long base =
      (long long)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[4];
  load4(in, base, n, v);
  int total;
  const int ex = block_excl_scan(v[0] + v[1] + v[2] + v[3], s_w, &total);
  // out[i] is the exclusive prefix at i (out[n] the sum of all): this
  // thread's four values are out[base .. base+3], one 16-byte store where
  // they are aligned and inside the array
  int e[4];
  e[0] = tile_off + ex;
  e[1] = e[0] + v[0];
  e[2] = e[1] + v[1];
  e[3] = e[2] + v[2];
  if (base + 3 <= n && (reinterpret_cast<uintptr_t>(out + base) & 15) == 0) {
    *reinterpret_cast<int4 *>(out + base) = make_int4(e[0], e[1], e[2], e[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (base + k <= n) out[base + k] = e[k];
  }
  // out[n] when n is a multiple of four falls past the last thread's values
  if (base + 4 == n) out[n] = e[3] + v[3];
  // counts that are not zero are cleared (most cells of a grid hold none)
  if (clear && (v[0] | v[1] | v[2] | v[3]) != 0) {
    if (base + 3 < n && (reinterpret_cast<uintptr_t>(clear + base) & 15) == 0) {
      *reinterpret_cast<int4 *>(clear + base) = make_int4(0, 0, 0, 0);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (base + k < n) clear[base + k] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && zero) *zero = 0;
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

namespace {
int scan_tiles(int n) { return (int)(((long long)n + kScanTile - 1) / kScanTile); }

int scan_launch(const int *in, int *out, int n, ScanWork &w, hipStream_t st,
                int *clear, int *zero = nullptr) {
  if (n <= 0) return AMSC_OK;
  if (int rc = scan_reserve(w, n)) return rc;
  const int nt = scan_tiles(n);
  int *ts = static_cast<int *>(w.tmp);
  scan_sums_kernel<<<nt, kScanThreads, 0, st>>>(in, n, ts);
  scan_apply_kernel<<<nt, kScanThreads, 0, st>>>(in, n, ts, out, clear,
                                                 zero);
  AMSC_LAUNCH_CHECK();
  return AMSC_OK;
}
}  // namespace

int scan_bytes(int n, size_t *bytes) {
  const size_t nt = (size_t)(n > 0 ? scan_tiles(n) : 1);
  *bytes = (nt * sizeof(int) + 255) / 256 * 256;
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
  if (int rc = scan_launch(in, out, n, w, st, nullptr)) return rc;
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
  if (nc <= 0) return AMSC_OK;
  if (int rc = scan_launch(b.count, b.cstart, nc, b.scan, st, b.count, b.big))
    return rc;
  if (np > 0)
    place_kernel<<<grid_for(np, kPlaceItems * 256), 256, 0, st>>>(np, np_dev, key, slot,
                                                    b.cstart, cptcls, active);
  cell_sort_kernel<<<grid_for(nc, 256), 256, 0, st>>>(nc, b.cstart, cptcls,
                                                      active, b.big, b.big_cap);
  if (b.big)
    wave_sort_kernel<<<std::min<unsigned>(512, grid_for(b.big_cap,
                                                        kSortWaves)),
                       64 * kSortWaves, 0, st>>>(b.big, b.big_cap, b.cstart,
                                                 cptcls);
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

namespace {

// Workspace of amsc_bin_particles kept between calls (a time loop bins every
// step): the cell counts (zero between calls: the scan clears them), the
// scan's tile sums and the arrival slots, grown as needed, per device.  It
// is held until the process ends (never freed at exit, when the HIP runtime
// may already be gone).
struct BinCache {
  int dev = -1, nc_cap = 0, np_cap = 0;
  int *count = nullptr, *slot = nullptr;
  ScanWork scan;
};

std::mutex &bin_mutex() {
  static std::mutex m;
  return m;
}

BinCache &bin_cache() {
  static BinCache *c = new BinCache;
  return *c;
}

void bin_cache_drop(BinCache &c) {
  if (c.count) (void)hipFree(c.count);
  if (c.slot) (void)hipFree(c.slot);
  if (c.scan.tmp) (void)hipFree(c.scan.tmp);
  c = BinCache();
}

int bin_cache_reserve(BinCache &c, int nc, int np) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return fail(AMSC_ERR_HIP, "amsc_bin_particles: no device");
  if (dev != c.dev) bin_cache_drop(c);
  c.dev = dev;
  if (nc > c.nc_cap) {
    if (c.count) (void)hipFree(c.count);
    c.count = nullptr;
    c.nc_cap = 0;
    const int cap = nc + nc / 2;
    if (hipMalloc(&c.count, sizeof(int) * cap) != hipSuccess)
      return fail(AMSC_ERR_NOMEM, "amsc_bin_particles: count allocation");
    if (hipMemset(c.count, 0, sizeof(int) * cap) != hipSuccess)
      return fail(AMSC_ERR_HIP, "amsc_bin_particles: count clear");
    c.nc_cap = cap;
  }
  if (np > c.np_cap) {
    if (c.slot) (void)hipFree(c.slot);
    c.slot = nullptr;
    c.np_cap = 0;
    const int cap = np + np / 2;
    if (hipMalloc(&c.slot, sizeof(int) * cap) != hipSuccess)
      return fail(AMSC_ERR_NOMEM, "amsc_bin_particles: slot allocation");
    c.np_cap = cap;
  }
  return scan_reserve(c.scan, nc);
}

}  // namespace

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
  std::lock_guard<std::mutex> lock(bin_mutex());
  BinCache &c = bin_cache();
  if (int rc = bin_cache_reserve(c, nrows * ncols, np)) {
    bin_cache_drop(c);
    return rc;
  }
  Binning b;
  b.nrows = nrows;
  b.ncols = ncols;
  b.count = c.count;
  b.cstart = d_cstart;  // the scan writes the caller's array directly
  b.scan = c.scan;
  b.scan.owned = false;
  int rc = launch_key_count(np, d_x, d_y, nrows, ncols, hx, hy, d_p2g, c.slot,
                            c.count, nullptr, st);
  if (!rc) rc = launch_finish_binning(b, np, d_p2g, c.slot, d_cptcls, nullptr, st);
  if (!rc && coloff) rc = launch_colouring(nrows, ncols, d_ccidx, coloff, st);
  // complete before returning, like the host rebuild it replaces
  hipError_t e = hipStreamSynchronize(st);
  if (!rc && e != hipSuccess) rc = fail(AMSC_ERR_HIP, hipGetErrorString(e));
  if (rc) bin_cache_drop(c);  // the counts may not be back at zero
  return rc;
}

// ---------------------------------------------------------------------------
// amsc_bin_particles_noalloc: the same binning with its workspace (cell
// counts, arrival slots, scan tile sums) in a fixed array of the library's
// own device memory, for callers that cannot allocate (code under graph
// capture, serving loops).  Each GPU holds its own copy of the array from
// the library's load to the end of the process, at an address that
// hipGetSymbolAddress gives with that device current; the array's name in
// host code is a host-side stand-in and is never handed to a kernel.
namespace amsc {

constexpr int kNoAllocMaxCells = 1 << 22;      // a 2048 x 2048 grid
constexpr int kNoAllocMaxParticles = 1 << 22;  // 4,194,304 particles
constexpr int kNoAllocScanInts =
    ((kNoAllocMaxCells + kScanTile - 1) / kScanTile + 63) / 64 * 64;
constexpr size_t kNoAllocInts = (size_t)kNoAllocMaxCells +
                                kNoAllocMaxParticles + kNoAllocScanInts;

// counts [0, kNoAllocMaxCells), slots [.., + kNoAllocMaxParticles), scan
// tile sums after them: 32 MiB + 4 KiB per GPU
static __device__ __align__(256) int scratch[kNoAllocInts];

namespace {

constexpr int kMaxDevices = 64;

std::mutex &noalloc_mutex() {
  static std::mutex m;
  return m;
}

// Device address of this GPU's scratch, looked up once per device (the
// caller holds noalloc_mutex()).
int noalloc_scratch(int **out) {
  static int *ptr[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(AMSC_ERR_HIP, "amsc_bin_particles_noalloc: no device");
  if (!ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(AMSC_ERR_HIP, std::string("amsc_bin_particles_noalloc: "
                                            "scratch address: ") +
                                    hipGetErrorString(e));
    ptr[dev] = static_cast<int *>(p);
  }
  *out = ptr[dev];
  return AMSC_OK;
}

__global__ void clear_ints_kernel(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace
}  // namespace amsc

extern "C" int amsc_bin_particles_noalloc(int np, int nrows, int ncols,
                                          double hx, double hy,
                                          const double *d_x, const double *d_y,
                                          int *d_p2g, int *d_cstart,
                                          int *d_cptcls, int *d_ccidx,
                                          int *coloff) {
  if (np < 0 || nrows <= 0 || ncols <= 0)
    return fail(AMSC_ERR_INVALID, "amsc_bin_particles_noalloc: bad sizes");
  if ((long long)nrows * ncols > kNoAllocMaxCells)
    return fail(AMSC_ERR_INVALID,
                "amsc_bin_particles_noalloc: more than 4194304 cells (the "
                "size of its fixed workspace); use amsc_bin_particles");
  if (np > kNoAllocMaxParticles)
    return fail(AMSC_ERR_INVALID,
                "amsc_bin_particles_noalloc: more than 4194304 particles (the "
                "size of its fixed workspace); use amsc_bin_particles");
  if (!d_p2g || !d_cstart || (np > 0 && !d_cptcls))
    return fail(AMSC_ERR_INVALID, "amsc_bin_particles_noalloc: NULL output");
  const int nc = nrows * ncols;
  hipStream_t st = current_stream();
  // one call at a time per process: every call on a GPU shares its scratch
  std::lock_guard<std::mutex> lock(noalloc_mutex());
  int *ws = nullptr;
  if (int rc = noalloc_scratch(&ws)) return rc;
  int *count = ws, *slot = ws + kNoAllocMaxCells;
  // the counts start at zero (the scratch keeps whatever the last call
  // left, which after a failed call need not be zero)
  clear_ints_kernel<<<grid_for(nc, 256), 256, 0, st>>>(count, nc);
  Binning b;
  b.nrows = nrows;
  b.ncols = ncols;
  b.count = count;
  b.cstart = d_cstart;
  b.scan.tmp = ws + kNoAllocMaxCells + kNoAllocMaxParticles;
  b.scan.bytes = sizeof(int) * kNoAllocScanInts;
  b.scan.owned = false;  // never reallocated: a scan that needs more fails
  int rc = AMSC_OK;
  if (hipGetLastError() != hipSuccess)
    rc = fail(AMSC_ERR_HIP, "amsc_bin_particles_noalloc: clear failed");
  if (!rc)
    rc = launch_key_count(np, d_x, d_y, nrows, ncols, hx, hy, d_p2g, slot,
                          count, nullptr, st);
  if (!rc) rc = launch_finish_binning(b, np, d_p2g, slot, d_cptcls, nullptr, st);
  if (!rc && coloff) rc = launch_colouring(nrows, ncols, d_ccidx, coloff, st);
  // complete before returning, like amsc_bin_particles
  const hipError_t e = hipStreamSynchronize(st);
  if (!rc && e != hipSuccess) rc = fail(AMSC_ERR_HIP, hipGetErrorString(e));
  return rc;
}
