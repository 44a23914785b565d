// binning.hpp -- device binning of particles by cell and integer scans.
#pragma once

#include "common.hpp"

namespace amsc {

// Device workspace of the scan (one int per 4096-value tile).
struct ScanWork {
  void *tmp = nullptr;
  size_t bytes = 0;
  bool owned = true;  // false: tmp belongs to the caller, never reallocated
};

// Workspace bytes of a scan of n values.
int scan_bytes(int n, size_t *bytes);

// out[0..n] = exclusive prefix of in[0..n) (out[0] = 0, the total in
// out[n]; nothing is written for n = 0).  If total_host != nullptr the total
// is copied back (synchronising the stream).
int exclusive_scan(const int *in, int *out, int n, ScanWork &w,
                   hipStream_t st, int *total_host);

// Size the workspace for scans of up to n values (a scan of a size already
// reserved allocates nothing, so it may be captured into a graph).
int scan_reserve(ScanWork &w, int n);

// Device binning pipeline state for one particle set.
struct Binning {
  int nrows = 0, ncols = 0;
  int *count = nullptr;   // [nc] cell counts (kept zero between uses)
  int *cstart = nullptr;  // [nc+1], written by each binning
  // Crowded cells (9..kWaveSortMax particles) queued for a sort by a whole
  // wavefront: big[0] = how many, big[1..big_cap] the cells.  Null: every
  // cell is sorted by one thread.  Cleared by the scan of each binning.
  int *big = nullptr;
  int big_cap = 0;
  ScanWork scan;
};

// Largest cell sorted by one wavefront from LDS (larger: one thread, heap
// sort in place).
constexpr int kWaveSortMax = 1024;

int binning_alloc(Binning &b, int nrows, int ncols);
void binning_free(Binning &b);

// key[i] = cell of (x[i], y[i]); slot[i] = arrival rank in its cell.
int launch_key_count(int np, const double *x, const double *y, int nrows,
                     int ncols, double hx, double hy, int *key, int *slot,
                     int *count, const int *active, hipStream_t st);
// cstart from counts, then cptcls from (key, slot), then ascending particle
// ids within every cell, counts reset to 0.  With np_dev non-null the
// particle count is read from device memory (np then only sizes the launch
// and must bound it).  `active` must be null.
int launch_finish_binning(Binning &b, int np, const int *key,
                          const int *slot, int *cptcls, const int *active,
                          hipStream_t st, const int *np_dev = nullptr);
// ccidx for the colouring colour = (r&1) + 2*(c&1); coloff on the host.
int launch_colouring(int nrows, int ncols, int *ccidx, int *coloff,
                     hipStream_t st);

// Wave-aggregated atomic increment of count[key]: the lanes of a wavefront
// form runs of equal keys (consecutive valid lanes), and the first lane of
// each run (the leader) adds the run's length with one atomic; a lane's
// slot is the returned base plus its rank in the run (ascending lane =
// ascending particle id).  Found with one ballot whatever the keys' order;
// a key that recurs after another in the wavefront takes one atomic per
// run, and the binning's per-cell sort restores id order within the cell.
// (A ballot per distinct key instead, for wavefronts not in key order,
// measured 3 % slower per step on the 4M workload, whose wavefronts lose
// their order as the flow spreads: 23 % still ordered at step 200.)  All
// leaders issue their atomics together, so the wavefront waits for one
// atomic round trip.
__device__ __forceinline__ int wave_slot(int *count, int key, bool valid) {
  const int lane = (int)__lane_id();
  const int kv = valid ? key : 0x7fffffff;
  const int prev = __shfl_up(kv, 1);
  const bool head = valid && (lane == 0 || kv != prev);
  const unsigned long long heads = __ballot(head);
  const unsigned long long vmask = __ballot(valid);
  const unsigned long long le = heads & lanemask_le();
  const int leader = le ? 63 - __clzll((long long)le) : lane;
  int base = 0;
  if (head) {
    // the run ends at the next head or the next lane without a particle
    const unsigned long long after = (heads | ~vmask) & ~lanemask_le();
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    base = atomicAdd(&count[key], end - lane);
  }
  base = __shfl(base, leader);
  return base + (lane - leader);
}

}  // namespace amsc
