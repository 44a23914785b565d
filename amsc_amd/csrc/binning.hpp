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

// Wave-aggregated atomic increment of count[key]: lanes of a wavefront with
// equal keys form a group whose lowest lane (the leader) adds the group size
// with one atomic; a lane's slot is the returned base plus its rank in the
// group (ascending lane = ascending particle id).  When the wavefront's keys
// are non-decreasing (particles stored in cell order, the common case) the
// groups are runs found with one ballot; otherwise they are found with a
// ballot per distinct key.  Either way all leaders issue their atomics
// together, so the wavefront waits for one atomic round trip.
__device__ __forceinline__ int wave_slot(int *count, int key, bool valid) {
  const int lane = (int)__lane_id();
  const int kv = valid ? key : 0x7fffffff;
  const int next = __shfl_down(kv, 1);
  const bool ordered = lane == 63 || kv <= next;
  int my_leader = lane, my_rank = 0, my_count = 0;
  if (__ballot(!ordered) == 0ull) {
    const int prev = __shfl_up(kv, 1);
    const bool head = valid && (lane == 0 || kv != prev);
    const unsigned long long heads = __ballot(head);
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long le = heads & lanemask_le();
    my_leader = le ? 63 - __clzll((long long)le) : lane;
    my_rank = lane - my_leader;
    if (head) {
      const unsigned long long after = heads & ~lanemask_le();
      const int end = after ? __ffsll((long long)after) - 1 : __popcll(vmask);
      my_count = end - lane;
    }
  } else {
    unsigned long long todo = __ballot(valid);
    while (todo) {
      const int leader = __ffsll((long long)todo) - 1;
      const int lkey = __builtin_amdgcn_readlane(key, leader);
      const unsigned long long peers = __ballot(valid && key == lkey) & todo;
      if ((peers >> lane) & 1ull) {
        my_leader = leader;
        my_rank = __popcll(peers & lanemask_lt());
      }
      if (lane == leader) my_count = __popcll(peers);
      todo &= ~peers;
    }
  }
  int base = 0;
  if (valid && lane == my_leader) base = atomicAdd(&count[key], my_count);
  base = __shfl(base, my_leader);
  return base + my_rank;
}

}  // namespace amsc
