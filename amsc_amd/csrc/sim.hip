// sim.hip -- device-resident depth-averaged MPM time loop (optim_par
// semantics) on gfx950: one scenario (amsc_sim_*) or a batch of scenarios
// advanced together (amsc_ens_*), in FP64 or in mixed FP32/FP64.
//
// One step (main/optim_par.cpp:281-668, merge/split at the position of
// main/gpu_offload.cpp:321-401), for every scenario of the batch at once:
//   prologue      one wavefront per scenario: end test, dt (fixed, or CFL
//                 from the device maxima of the previous step), t, it.
//   [merge/split] flag, per-cell pairing, serial replay where the operation
//                 cap binds, weight scan, execution into the other SoA
//                 buffer, new particle offsets, re-binning.
//   p2g_nodal     one workgroup per tile of nodes: the tile's particles are
//                 staged through LDS and every node thread folds its cells
//                 in the serial p2g order, then applies the nodal update and
//                 boundary conditions in registers.
//   conservation  (on request, amsc_ens_conservation) one workgroup per
//                 scenario: err_mass, err_mom of the last step from the P2G
//                 tile partials, which stay on the device until then.
//   particle      one thread per particle: G2P, advection, G2PD, height,
//                 Voellmy, Bingham, friction force, new cell key with a
//                 wave-aggregated count, CFL maxima.
//   binning       scan of the counts -> cstart, placement, in-cell sort.
// Kernels are in sim_kernels.hpp.  Steps can be recorded once into HIP
// graphs (one graph per step variant) and replayed; launch extents are then
// sized by the particle capacity and the kernels read the particle count
// from device memory.
#include "sim_kernels.hpp"
#include "amsc_mpm.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <chrono>
#include <string>
#include <utility>
#include <vector>

namespace amsc {

static const char *kPFName[PF_COUNT] = {
    "x",      "y",      "vpx",    "vpy",     "hp",      "Vp",     "Mp",
    "H",      "F_11",   "F_12",   "F_22",    "Fric_px", "Fric_py", "Ap",
    "Fb_x",   "Fb_y",   "vpx_dx", "vpx_dy",  "vpy_dx",  "vpy_dy", "apx",
    "apy",    "Zp",     "hpZ",    "dZxp",    "dZyp",    "Fpx",    "Fpy",
    "mom_px", "mom_py"};

static const char *kNFName[NF_COUNT] = {
    "vvxL",   "vvyL",     "avx",      "avy",      "Mv",      "mom_vx",
    "mom_vy", "F_ext_vx", "F_ext_vy", "F_int_vx", "F_int_vy", "Fric_x",
    "Fric_y", "FPxv",     "FPyv",     "Ftot_vx",  "Ftot_vy",  "vvx",
    "vvy"};

// P2G tile shapes (node columns x node rows, one thread per node), tile
// orders (descending column-major or linear) and particle-kernel workgroup
// sizes the autotuner chooses from.  A P2G variant v is shape v % 3, order v / 3.
constexpr int kTileCfgs = 3;
constexpr int kP2GVariants = 2 * kTileCfgs;
constexpr int kTileTC[kTileCfgs] = {16, 8, 8};
constexpr int kTileTR[kTileCfgs] = {32, 31, 32};
constexpr int kPartBS[2] = {256, 512};
// A scenario whose cells have held more than this many particles runs its
// P2G on the dense-cell kernels (at 4 particles per cell the tiled P2G is
// the faster; from a few dozen per cell it serialises on the crowded cells).
constexpr int kDensePPC = 24;
constexpr int kMaxScen = 1 << 16;

struct TimedLaunch {
  const char *name;
  hipEvent_t a, b;
};

struct GraphEntry {
  hipGraphExec_t exec;
  int cur_after, kcur_after;
};

struct Sim {
  int S = 0;
  bool mixed = false;
  size_t rsz = sizeof(double);  // bytes of a narrow (type R) value
  std::vector<amsc_params> prm;
  std::vector<Scen> scen_h[kTileCfgs];
  std::vector<int> poff_h, coff_h, noff_h;
  int np = 0, cap = 0, nn = 0, nc = 0;
  hipStream_t st = nullptr;
  PPtr buf[2]{};
  int cur = 0;
  int *key[2] = {nullptr, nullptr};
  int kcur = 0;
  int *slot = nullptr, *cptcls = nullptr;
  Binning bin;
  ScanWork ms_scan;
  NPtr nod{};
  double *Z = nullptr;
  Scen *scen[kTileCfgs] = {};
  int *tscen[kTileCfgs] = {};
  int ntiles[kTileCfgs] = {};
  Ctrl *ctrl = nullptr;
  Ctrl *ctrl_host = nullptr;  // pinned, S entries (headers only are read)
  size_t ctrl_host_bytes = 0, host_ints_bytes = 0;
  int *poff = nullptr, *poff_next = nullptr, *wscen = nullptr;
  int *cell_off = nullptr, *err = nullptr;
  int *host_ints = nullptr;  // pinned staging, 4*S + S + 2 ints
  double *partials = nullptr;
  P2GOut *p2g_out[2] = {nullptr, nullptr};
  int *act = nullptr, *partner = nullptr, *wts = nullptr, *offs = nullptr;
  int *ms_cand = nullptr;  // per cell: holds a merge candidate (kept zero between merge/split steps)
  int *ms_cnt = nullptr;
  // legacy ELFS rule only (ms_mode 1): exterior flags, hop distances and
  // two frontier queues, one int per cell each
  int *elfs_ext = nullptr, *elfs_hops = nullptr, *elfs_qa = nullptr,
      *elfs_qb = nullptr;
  bool any_elfs = false;
  std::vector<int> ms_counts;  // splits, merges of the last call, per scenario
  std::vector<int> it_host;    // steps launched per scenario
  // per scenario: the stored friction force field is current (the last
  // particle update wrote it, or no step has run since it was set)
  std::vector<char> fric_cur;
  std::vector<char> ended;     // run known to have ended, per scenario
  // default tiling (8x32-node P2G tiles from the last tile column down,
  // 256-thread particle workgroups); the autotuner may pick another
  int tile_cfg = 2, tile_order = 0, part_bs = 256;
  // merge/split read-back, eager steps: ms_spec = the execute and the
  // re-binning are enqueued before the operation counts reach the host (set
  // while the last merge/split step changed particles); the offsets kernel
  // then writes the counts into ms_box (pinned, host-coherent), word 0 last
  // (= ms_seq)
  bool ms_spec = true;
  int *ms_box = nullptr;
  size_t ms_box_bytes = 0;
  int ms_seq = 0;
  // scenarios whose P2G runs on the dense-cell kernels (decided on the host
  // from the control blocks' max_ppc, once a cell has held more than
  // kDensePPC particles), their flags on the device, and the staged values
  // those kernels use (allocated when the first scenario turns dense)
  std::vector<char> dense_h;
  int *dense_d = nullptr;
  bool any_dense = false;
  P2GStaged staged{};
  // queue of crowded cells for the binning's wavefront sort, handed to the
  // binning (bin.big) once some cell has held more than 8 particles; until
  // then every cell is sorted by one thread
  int *big_q = nullptr;
  bool wave_sort = false;
  long long particle_steps = 0;
  size_t bytes = 0;  // device memory requested (cached blocks may be larger)
  std::vector<std::pair<void *, size_t>> allocs;  // every dalloc block
  bool finished = false, aux_on = false, diag = true, mom_stored = false;
  bool aux_valid = true, nodal_valid = false, timing = false;
  bool graph_on = false, broken = false;
  bool cons_stale = false;  // a P2G ran since the conservation reduction
  int cons_cfg = 0;         // tile shape of the partials' layout
  std::map<long long, GraphEntry> graphs;
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

// Launch a kernel; with timing on, the start/stop events are recorded by the
// dispatch itself (hipExtLaunchKernel), so they bracket the kernel only.
template <typename... KArgs, typename... Args>
int launch(Sim &s, const char *name, void (*kernel)(KArgs...), dim3 g,
           dim3 b, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s.timing) {
    if (int rc = ev_get(s, &e0)) return rc;
    if (int rc = ev_get(s, &e1)) return rc;
  }
  hipExtLaunchKernelGGL(kernel, g, b, 0, s.st, e0, e1, 0, args...);
  AMSC_LAUNCH_CHECK();
  if (s.timing) s.launches.push_back({name, e0, e1});
  return AMSC_OK;
}

// Events around a multi-kernel phase (binning, scans).
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

Tables tables(const Sim &s) {
  Tables T;
  T.scen = s.scen[s.tile_cfg];
  T.ctrl = s.ctrl;
  T.poff = s.poff;
  T.wscen = s.wscen;
  T.S = s.S;
  return T;
}

size_t pf_bytes(const Sim &s, int k) { return pf_wide(k) ? 8 : s.rsz; }
size_t nf_bytes(const Sim &s, int k) { return nf_narrow(k) ? s.rsz : 8; }

// Device memory of destroyed batches is kept in a cache and handed to the
// next batch of the process: a served batch is rebuilt at every join and
// leave, and hipMalloc/hipFree of its arrays would otherwise dominate the
// rebuild.  A cached block serves a request of at least half its size; the
// cache holds at most kCacheMax bytes (amsc_release_cache empties it).
constexpr size_t kCacheMax = 32ull << 30;

struct DevCache {
  std::mutex m;
  std::multimap<size_t, void *> blocks;
  size_t held = 0;
};

DevCache &dev_cache() {
  static DevCache c;
  return c;
}

bool cache_take(size_t b, void **p, size_t *got) {
  DevCache &c = dev_cache();
  std::lock_guard<std::mutex> g(c.m);
  auto it = c.blocks.lower_bound(b);
  if (it == c.blocks.end() || it->first > 2 * b) return false;
  *p = it->second;
  *got = it->first;
  c.held -= it->first;
  c.blocks.erase(it);
  return true;
}

void cache_give(void *p, size_t b) {
  DevCache &c = dev_cache();
  std::lock_guard<std::mutex> g(c.m);
  if (c.held + b > kCacheMax) {
    (void)hipFree(p);
    return;
  }
  c.blocks.emplace(b, p);
  c.held += b;
}

void cache_release() {
  DevCache &c = dev_cache();
  std::lock_guard<std::mutex> g(c.m);
  for (auto &kv : c.blocks) (void)hipFree(kv.second);
  c.blocks.clear();
  c.held = 0;
}

// Pinned host buffers, cached the same way (hipHostMalloc is slow).
struct PinCache {
  std::mutex m;
  std::multimap<size_t, void *> blocks;
};

PinCache &pin_cache() {
  static PinCache c;
  return c;
}

int pinned_take(size_t b, void **p, size_t *got) {
  {
    PinCache &c = pin_cache();
    std::lock_guard<std::mutex> g(c.m);
    auto it = c.blocks.lower_bound(b);
    if (it != c.blocks.end() && it->first <= 4 * b + 4096) {
      *p = it->second;
      *got = it->first;
      c.blocks.erase(it);
      return AMSC_OK;
    }
  }
  AMSC_HIP(hipHostMalloc(p, b));
  *got = b;
  return AMSC_OK;
}

void pinned_give(void *p, size_t b) {
  PinCache &c = pin_cache();
  std::lock_guard<std::mutex> g(c.m);
  if (c.blocks.size() >= 16) {
    (void)hipHostFree(p);
    return;
  }
  c.blocks.emplace(b, p);
}

template <typename T> int dalloc(Sim &s, T **p, size_t n) {
  const size_t b = (sizeof(T) * (n > 0 ? n : 1) + 255) & ~size_t(255);
  void *q = nullptr;
  size_t got = b;
  if (!cache_take(b, &q, &got)) {
    if (hipMalloc(&q, b) != hipSuccess) {
      // give the cache back to the device and retry once
      cache_release();
      if (hipMalloc(&q, b) != hipSuccess) {
        *p = nullptr;
        return fail(AMSC_ERR_NOMEM, "device allocation of " +
                                        std::to_string(b) + " bytes failed");
      }
    }
  }
  *p = static_cast<T *>(q);
  s.allocs.emplace_back(q, got);
  s.bytes += b;
  return AMSC_OK;
}

int alloc_ppt(Sim &s, PPtr &p) {
  for (int k = 0; k < PF_COUNT; ++k) {
    char *q = nullptr;
    if (int rc = dalloc(s, &q, pf_bytes(s, k) * (size_t)s.cap)) return rc;
    p.f[k] = q;
  }
  if (int rc = dalloc(s, &p.label, s.cap)) return rc;
  return dalloc(s, &p.level, s.cap);
}

void sim_free(Sim *s) {
  if (!s) return;
  if (s->st) (void)hipStreamSynchronize(s->st);
  for (auto &g : s->graphs) (void)hipGraphExecDestroy(g.second.exec);
  for (auto &a : s->allocs) cache_give(a.first, a.second);
  if (s->ctrl_host) pinned_give(s->ctrl_host, s->ctrl_host_bytes);
  if (s->host_ints) pinned_give(s->host_ints, s->host_ints_bytes);
  for (auto e : s->event_pool) (void)hipEventDestroy(e);
  if (s->ms_box) (void)hipHostFree(s->ms_box);
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

// Re-read the control blocks.  Between steps (mirrors = true) this also
// resets the host's per-scenario step counts and end flags; inside a step
// (after its prologue has advanced `it`) only the blocks are read.
int read_ctrl(Sim &s, bool mirrors = true) {
  AMSC_HIP(hipMemcpy2DAsync(s.ctrl_host, sizeof(Ctrl), s.ctrl, sizeof(Ctrl),
                            kCtrlHead, s.S, hipMemcpyDeviceToHost, s.st));
  AMSC_HIP(hipStreamSynchronize(s.st));
  if (!mirrors) return AMSC_OK;
  bool all = true;
  for (int i = 0; i < s.S; ++i) {
    const Ctrl &c = s.ctrl_host[i];
    const amsc_params &p = s.prm[i];
    s.it_host[i] = c.it;
    const bool fixed = p.dt_fixed > 0.0;
    s.ended[i] = !c.active || (fixed && c.it >= p.nsteps) ||
                 (!fixed && p.max_steps > 0 && c.it >= p.max_steps);
    all = all && s.ended[i];
  }
  s.finished = all;
  return AMSC_OK;
}

int read_poff(Sim &s) {
  AMSC_HIP(hipMemcpyAsync(s.host_ints, s.poff, sizeof(int) * (s.S + 1),
                          hipMemcpyDeviceToHost, s.st));
  AMSC_HIP(hipStreamSynchronize(s.st));
  s.poff_h.assign(s.host_ints, s.host_ints + s.S + 1);
  s.np = s.poff_h[s.S];
  return AMSC_OK;
}

int launch_fill_int(Sim &s, int *p, long long n, int v) {
  if (n <= 0) return AMSC_OK;
  return launch(s, "fill", fill_int_kernel, dim3(grid_for(n, 256)), dim3(256),
                p, n, v);
}

int update_wscen(Sim &s) {
  const int nw = s.cap / 64 + 1;
  return launch(s, "wscen", wscen_kernel, dim3(grid_for(nw, 256)), dim3(256),
                (const int *)s.poff, s.S, nw, s.wscen);
}

// Bin the current particle set into key[kcur], cptcls, cstart.
int rebin(Sim &s, bool capture) {
  const Tables T = tables(s);
  const int n = capture ? s.cap : s.np;
  PPtr &p = s.buf[s.cur];
  if (n > 0)
    if (int rc = launch(s, "key_count", key_kernel, dim3(grid_for(n, 256)),
                        dim3(256), T, (const double *)p.f[PF_X],
                        (const double *)p.f[PF_Y], s.key[s.kcur], s.slot,
                        s.bin.count))
      return rc;
  Timer tm(s, "binning");
  return launch_finish_binning(s.bin, n, s.key[s.kcur], s.slot, s.cptcls,
                               nullptr, s.st, capture ? s.poff + s.S : nullptr);
}

// *all_ended (eager launches only): every run had ended before this step;
// nothing was changed.
template <typename R>
int merge_split(Sim &s, bool capture, bool *all_ended) {
  *all_ended = false;
  const Tables T = tables(s);
  PPtr &p = s.buf[s.cur];
  MsIn<R> in;
  in.x = fld<double>(p, PF_X);
  in.y = fld<double>(p, PF_Y);
  in.hp = fld<double>(p, PF_HP);
  in.Ap = static_cast<const R *>(p.f[PF_AP]);
  in.vxdx = static_cast<const R *>(p.f[PF_VXDX]);
  in.vxdy = static_cast<const R *>(p.f[PF_VXDY]);
  in.vydx = static_cast<const R *>(p.f[PF_VYDX]);
  in.vydy = static_cast<const R *>(p.f[PF_VYDY]);
  in.vpx = static_cast<const R *>(p.f[PF_VPX]);
  in.vpy = static_cast<const R *>(p.f[PF_VPY]);
  in.level = p.level;
  const int n = capture ? s.cap : s.np;
  const dim3 gp(grid_for(n, 256)), b256(256);
  const int *key = s.key[s.kcur];
  ElfsCells E{s.elfs_ext, s.elfs_hops};
  if (int rc = launch_fill_int(s, s.ms_cnt, 4ll * s.S, 0)) return rc;
  if (s.any_elfs)
    if (int rc = launch(s, "ms_elfs_cells", elfs_cells_kernel, dim3(s.S),
                        dim3(1024), T, (const int *)s.bin.cstart, s.elfs_ext,
                        s.elfs_hops, s.elfs_qa, s.elfs_qb))
      return rc;
  if (n > 0)
    if (int rc = launch(s, "ms_flag", ens_flag_kernel<R>, gp, b256, T, in, E,
                        key, (const int *)s.wts, s.act, s.partner, s.ms_cnt,
                        s.ms_cand))
      return rc;
  if (int rc = launch(s, "ms_pair", ens_pair_kernel<R>, dim3(grid_for(s.nc, 256)),
                      b256, T, (const int *)s.cell_off, s.nc, in,
                      (const int *)s.bin.cstart, (const int *)s.cptcls, s.act,
                      s.partner, s.ms_cnt, s.ms_cand))
    return rc;
  if (n > 0)
    if (int rc = launch(s, "ms_refresh", ens_refresh_kernel<R>, gp, b256, T,
                        in, E, key, (const int *)s.wts, s.act, s.partner,
                        (const int *)s.ms_cnt))
      return rc;
  if (int rc = launch(s, "ms_serial", ens_serial_kernel<R>, dim3(s.S), dim3(64),
                      T, in, (const int *)s.bin.cstart, (const int *)s.cptcls,
                      s.act, s.partner, s.ms_cnt))
    return rc;
  if (n > 0)
    if (int rc = launch(s, "ms_demote", ens_demote_kernel, gp, b256, T, key,
                        s.act, (const int *)s.ms_cnt))
      return rc;
  {
    Timer tm(s, "ms_scan");
    if (n > 0)
      if (int rc = launch(s, "ms_weights", weights_kernel, gp, b256, T, n,
                          (const int *)s.act, s.wts))
        return rc;
    if (int rc = exclusive_scan(s.wts, s.offs, n, s.ms_scan, s.st, nullptr))
      return rc;
  }
  // Eager steps read back the operation counts, the overflow flag, the new
  // offsets and the scenarios' active flags (whether any run is still
  // going).  With ms_spec the execute, commit and re-binning are enqueued
  // before the counts reach the host, and the offsets kernel writes them
  // straight into host memory (word 0, the sequence number, last): the
  // device goes on with the execute while the host takes them.  (Waiting
  // for the stream left the device idle about 38 us per merge/split step
  // on the 4M workload; a copy to the host inside the stream still held
  // the next kernel back about 28 us after the copy had ended.)
  // Speculation gives the same state: with no operation the execute copies
  // every particle unchanged (so it is only used while the last merge/split
  // step changed particles); on overflow the execute and the commit do
  // nothing, the host keeps the current buffer and bins it again.
  const bool spec = !capture && s.ms_spec;
  const size_t nbox = 6 * (size_t)s.S + 3;  // sequence word + the block
  if (spec && s.ms_box_bytes < sizeof(int) * nbox) {
    if (s.ms_box) (void)hipHostFree(s.ms_box);
    s.ms_box = nullptr;
    s.ms_box_bytes = 0;
    AMSC_HIP(hipHostMalloc((void **)&s.ms_box, sizeof(int) * nbox,
                           hipHostMallocCoherent));
    s.ms_box_bytes = sizeof(int) * nbox;
    s.ms_box[0] = s.ms_seq;
  }
  if (spec) s.ms_seq = s.ms_seq == INT_MAX ? 1 : s.ms_seq + 1;
  if (int rc = launch(s, "ms_offsets", ms_offsets_kernel, dim3(1), b256,
                      (const int *)s.poff, s.S, (const int *)s.offs, s.cap,
                      s.poff_next, s.err, (const Ctrl *)s.ctrl,
                      s.ms_cnt + 5 * s.S + 2, spec ? s.ms_box : nullptr,
                      (const int *)s.ms_cnt, s.ms_seq))
    return rc;
  const int *h = spec ? s.ms_box + 1 : s.host_ints;
  // after the read-back: 0 go on, 1 nothing to do (all ended or no
  // operation), AMSC_ERR_CAPACITY on overflow
  auto take_counts = [&]() -> int {
    bool active = false;
    for (int i = 0; i < s.S; ++i) active = active || h[5 * s.S + 2 + i];
    if (!active) {
      *all_ended = true;
      return 1;
    }
    bool any = false;
    for (int i = 0; i < s.S; ++i) {
      s.ms_counts[2 * i] = h[4 * i];
      s.ms_counts[2 * i + 1] = h[4 * i + 2];
      any = any || h[4 * i] || h[4 * i + 2];
    }
    s.ms_spec = any;
    if (!any) return 1;  // no operation anywhere: nothing moves
    if (h[4 * s.S]) {
      (void)launch_fill_int(s, s.err, 1, 0);
      return fail(AMSC_ERR_CAPACITY,
                  "merge/split needs " + std::to_string(h[5 * s.S + 1]) +
                      " particles, capacity is " + std::to_string(s.cap));
    }
    return 0;
  };
  if (!capture && !spec) {
    AMSC_HIP(hipMemcpyAsync(s.host_ints, s.ms_cnt, sizeof(int) * (6 * s.S + 2),
                            hipMemcpyDeviceToHost, s.st));
    AMSC_HIP(hipStreamSynchronize(s.st));
    const int t = take_counts();
    if (t == 1) return AMSC_OK;
    if (t) return t;
  }
  ExecArgs ea;
  ea.T = T;
  ea.in = s.buf[s.cur];
  ea.out = s.buf[1 - s.cur];
  ea.act = s.act;
  ea.partner = s.partner;
  ea.offset = s.offs;
  ea.err = s.err;
  ea.ncopy = s.aux_on ? PF_COUNT : PF_CORE;
  ea.key = s.key[s.kcur];
  ea.slot = s.slot;
  ea.count = s.bin.count;
  if (n > 0)
    if (int rc = s.any_elfs
                     ? launch(s, "ms_execute", ms_execute_kernel<R, true>, gp, b256, ea)
                     : launch(s, "ms_execute", ms_execute_kernel<R, false>, gp, b256, ea))
      return rc;
  if (int rc = launch(s, "ms_commit", ms_commit_kernel, dim3(1), b256, s.poff,
                      (const int *)s.poff_next, s.S, (const int *)s.err))
    return rc;
  if (int rc = update_wscen(s)) return rc;
  s.cur = 1 - s.cur;
  if (spec) {
    // the new particles binned before their count is known: launch extents
    // from the capacity, the count read on the device
    {
      Timer tm(s, "binning");
      if (int rc = launch_finish_binning(s.bin, s.cap, s.key[s.kcur], s.slot,
                                         s.cptcls, nullptr, s.st, s.poff + s.S))
        return rc;
    }
    // wait for the sequence word (bounded: past 10 s the stream is waited
    // for, and the block copied the ordinary way if the word still lags)
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(&s.ms_box[0], __ATOMIC_ACQUIRE) != s.ms_seq) {
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10)) {
        AMSC_HIP(hipStreamSynchronize(s.st));
        if (__atomic_load_n(&s.ms_box[0], __ATOMIC_ACQUIRE) != s.ms_seq) {
          AMSC_HIP(hipMemcpy(s.ms_box + 1, s.ms_cnt, sizeof(int) * (6 * s.S + 2),
                             hipMemcpyDeviceToHost));
          s.ms_box[0] = s.ms_seq;
        }
        break;
      }
      std::this_thread::yield();
    }
    const int t = take_counts();
    if (t == 0 || (t == 1 && !*all_ended)) {
      // the committed offsets (unchanged without an operation)
      const int *o = h + 4 * s.S + 1;
      s.poff_h.assign(o, o + s.S + 1);
      s.np = s.poff_h[s.S];
      return AMSC_OK;
    }
    if (t == 1) return AMSC_OK;  // every run had ended: an unchanged copy
    // overflow: nothing was executed or committed; back to the current
    // buffer and its binning
    s.cur = 1 - s.cur;
    if (int rc = rebin(s, false)) return rc;
    return t;
  }
  if (!capture) {
    // the committed offsets, read back above
    const int *o = s.host_ints + 4 * s.S + 1;
    s.poff_h.assign(o, o + s.S + 1);
    s.np = s.poff_h[s.S];
  }
  // the execute kernel took the new particles' keys and slots
  Timer tm(s, "binning");
  return launch_finish_binning(s.bin, capture ? s.cap : s.np, s.key[s.kcur],
                               s.slot, s.cptcls, nullptr, s.st,
                               capture ? s.poff + s.S : nullptr);
}

template <typename R, bool DESC>
int launch_p2g_order(Sim &s, const P2GArgs &a) {
  const dim3 g(s.ntiles[s.tile_cfg]);
  switch (s.tile_cfg) {
  case 0:
    return launch(s, "p2g_nodal", p2g_nodal_kernel<R, 16, 32, DESC>, g, dim3(512), a);
  case 1:
    if constexpr (sizeof(R) == 4)
      return launch(s, "p2g_nodal", p2g_sweep_mixed_kernel<8, 31, DESC>, g, dim3(64), a);
    else
      return launch(s, "p2g_nodal", p2g_sweep_kernel<R, 8, 31, DESC>, g, dim3(64), a);
  default:
    return launch(s, "p2g_nodal", p2g_nodal_kernel<R, 8, 32, DESC>, g, dim3(256), a);
  }
}

template <typename R>
int launch_p2g_dense(Sim &s, const P2GArgs &a, int npos) {
  if (npos > 0)
    if (int rc = launch(s, "p2g_stage", p2g_stage_kernel<R>,
                        dim3(grid_for(npos, 256)), dim3(256), a))
      return rc;
  const dim3 g(s.ntiles[s.tile_cfg]);
  switch (s.tile_cfg) {
  case 0:
    return launch(s, "p2g_dense", p2g_dense_node_kernel<R, 16, 32>, g, dim3(512), a);
  case 1:
    return launch(s, "p2g_dense", p2g_dense_node_kernel<R, 8, 31>, g, dim3(256), a);
  default:
    return launch(s, "p2g_dense", p2g_dense_node_kernel<R, 8, 32>, g, dim3(256), a);
  }
}

// P2G of every active scenario: the dense-cell kernels for the scenarios
// flagged dense (npos bounds the binned positions), the tiled kernel for
// the others.
template <typename R>
int launch_p2g(Sim &s, const P2GArgs &a, int npos) {
  if (a.dense)
    if (int rc = launch_p2g_dense<R>(s, a, npos)) return rc;
  return s.tile_order == 0 ? launch_p2g_order<R, true>(s, a)
                           : launch_p2g_order<R, false>(s, a);
}

P2GArgs p2g_args(Sim &s, int aux, bool partials) {
  PPtr &p = s.buf[s.cur];
  P2GArgs a;
  a.T = tables(s);
  a.tscen = s.tscen[s.tile_cfg];
  a.aux = aux;
  a.mom_stored = s.mom_stored ? 1 : 0;
  a.cstart = s.bin.cstart;
  a.cptcls = s.cptcls;
  a.x = fld<double>(p, PF_X);
  a.y = fld<double>(p, PF_Y);
  a.hp = fld<double>(p, PF_HP);
  a.H = fld<double>(p, PF_H);
  a.Mp = p.f[PF_MP];
  a.m1 = s.mom_stored ? p.f[PF_MOMX] : p.f[PF_VPX];
  a.m2 = s.mom_stored ? p.f[PF_MOMY] : p.f[PF_VPY];
  a.Vp = p.f[PF_VP];
  a.frx = p.f[PF_FRX];
  a.fry = p.f[PF_FRY];
  a.f11 = p.f[PF_F11];
  a.f12 = p.f[PF_F12];
  a.f22 = p.f[PF_F22];
  a.Z = s.Z;
  a.vvxL = s.nod.f[NF_VVXL];
  a.vvyL = s.nod.f[NF_VVYL];
  a.partials = partials ? s.partials : nullptr;
  a.out = s.p2g_out[s.cur];
  a.dense = s.any_dense ? s.dense_d : nullptr;
  a.key = s.key[s.kcur];
  a.st = s.staged;
  return a;
}

template <typename R, int AUX, int BS>
int launch_particle_aux(Sim &s, const PartArgs &a, int n) {
  if (int rc = launch(s, "particle", particle_kernel<R, AUX, BS, false>,
                      dim3(grid_for(n, kPartIlp * BS)), dim3(BS), a))
    return rc;
  if (s.S == 1) return AMSC_OK;  // no scenario boundaries
  return launch(s, "particle_edge", particle_kernel<R, AUX, BS, true>,
                dim3(s.S - 1), dim3(BS), a);
}

template <typename R, int BS>
int launch_particle_bs(Sim &s, const PartArgs &a, int aux, int n) {
  if (aux == AUX_FULL) return launch_particle_aux<R, AUX_FULL, BS>(s, a, n);
  if (aux == AUX_MS) return launch_particle_aux<R, AUX_MS, BS>(s, a, n);
  return launch_particle_aux<R, AUX_NONE, BS>(s, a, n);
}

// The kernels of one step.  capture: recording into a graph (no host
// synchronisation, launch extents from the capacity).  *skipped is set when
// a merge/split step finds every run ended.
template <typename R>
int step_body(Sim &s, bool ms_step, int aux, bool capture, bool *skipped) {
  *skipped = false;
  if (int rc = launch(s, "prologue", prologue_kernel, dim3(s.S), dim3(64),
                      (const Scen *)s.scen[s.tile_cfg], s.ctrl))
    return rc;
  if (ms_step) {
    bool all_ended = false;
    if (int rc = merge_split<R>(s, capture, &all_ended)) return rc;
    if (all_ended) {
      s.finished = true;
      *skipped = true;
      return AMSC_OK;
    }
  }
  const Tables T = tables(s);
  PPtr &p = s.buf[s.cur];
  const int n = capture ? s.cap : s.np;
  if (int rc = launch_p2g<R>(s, p2g_args(s, aux, s.diag), n)) return rc;

  if (n > 0) {
    PartArgs a;
    a.T = T;
    a.p = p;
    a.key_in = s.key[s.kcur];
    a.key_out = s.key[1 - s.kcur];
    a.slot = s.slot;
    a.count = s.bin.count;
    a.vvxL = s.nod.f[NF_VVXL];
    a.vvyL = s.nod.f[NF_VVYL];
    a.avx = static_cast<const double *>(s.nod.f[NF_AVX]);
    a.avy = static_cast<const double *>(s.nod.f[NF_AVY]);
    a.Z = s.Z;
    a.ms_bits = s.wts;  // read by the next step's merge/split
    const int rc = s.part_bs == 512 ? launch_particle_bs<R, 512>(s, a, aux, n)
                                    : launch_particle_bs<R, 256>(s, a, aux, n);
    if (rc) return rc;
  }
  s.kcur = 1 - s.kcur;
  Timer tm(s, "binning");
  return launch_finish_binning(s.bin, n, s.key[s.kcur], s.slot, s.cptcls,
                               nullptr, s.st, capture ? s.poff + s.S : nullptr);
}

int run_body(Sim &s, bool ms_step, int aux, bool capture, bool *skipped) {
  return s.mixed ? step_body<float>(s, ms_step, aux, capture, skipped)
                 : step_body<double>(s, ms_step, aux, capture, skipped);
}

int launch_graph(Sim &s, bool ms_step, int aux) {
  const long long gk = (long long)ms_step | (long long)aux << 1 |
                       (long long)s.cur << 3 | (long long)s.kcur << 4 |
                       (long long)s.tile_cfg << 5 |
                       (long long)s.tile_order << 7 |
                       (long long)(s.part_bs == 512) << 12 |
                       (long long)s.diag << 9 | (long long)s.mom_stored << 10 |
                       (long long)s.aux_on << 11 | (long long)s.any_dense << 13 |
                       (long long)s.wave_sort << 14;
  auto it = s.graphs.find(gk);
  if (it == s.graphs.end()) {
    const int c0 = s.cur, k0 = s.kcur;
    AMSC_HIP(hipStreamBeginCapture(s.st, hipStreamCaptureModeThreadLocal));
    bool skipped = false;
    const int rc = run_body(s, ms_step, aux, true, &skipped);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s.st, &g);
    const int c1 = s.cur, k1 = s.kcur;
    s.cur = c0;
    s.kcur = k0;
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess)
      return fail(AMSC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess)
      return fail(AMSC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    it = s.graphs.emplace(gk, GraphEntry{ex, c1, k1}).first;
  }
  AMSC_HIP(hipGraphLaunch(it->second.exec, s.st));
  s.cur = it->second.cur_after;
  s.kcur = it->second.kcur_after;
  if (ms_step) {
    // particle counts and operation counts of the step, and overflow
    AMSC_HIP(hipMemcpyAsync(s.host_ints, s.ms_cnt, sizeof(int) * (4 * s.S + 1),
                            hipMemcpyDeviceToHost, s.st));
    AMSC_HIP(hipStreamSynchronize(s.st));
    for (int i = 0; i < s.S; ++i) {
      s.ms_counts[2 * i] = s.host_ints[4 * i];
      s.ms_counts[2 * i + 1] = s.host_ints[4 * i + 2];
    }
    if (s.host_ints[4 * s.S]) {
      s.broken = true;
      return fail(AMSC_ERR_CAPACITY,
                  "merge/split exceeded the particle capacity inside a graph "
                  "replay; the state is no longer valid");
    }
    if (int rc = read_poff(s)) return rc;
  }
  return AMSC_OK;
}

bool ms_due_host(const Sim &s, int i, int ahead) {
  const amsc_params &p = s.prm[i];
  return p.ms_on && p.ms_interval > 0 && !s.ended[i] &&
         (s.it_host[i] + ahead) % p.ms_interval == 0;
}

int launch_step(Sim &s) {
  if (s.broken) return fail(AMSC_ERR_INVALID, "the simulation state is invalid");
  bool ms_step = false, ms_next = false;
  for (int i = 0; i < s.S; ++i) {
    ms_step = ms_step || ms_due_host(s, i, 1);
    ms_next = ms_next || ms_due_host(s, i, 2);
  }
  const int aux = s.aux_on ? AUX_FULL : (ms_next ? AUX_MS : AUX_NONE);
  if (s.graph_on && !s.timing) {
    if (int rc = launch_graph(s, ms_step, aux)) return rc;
  } else {
    bool skipped = false;
    if (int rc = run_body(s, ms_step, aux, false, &skipped)) return rc;
    if (skipped) return AMSC_OK;
  }
  s.aux_valid = aux == AUX_FULL;
  s.nodal_valid = aux == AUX_FULL;
  s.mom_stored = false;
  s.cons_stale = s.diag;
  s.cons_cfg = s.tile_cfg;
  bool all = true;
  for (int i = 0; i < s.S; ++i) {
    if (!s.ended[i]) {
      s.it_host[i]++;
      s.fric_cur[i] = aux >= AUX_MS;
      s.particle_steps += s.poff_h[i + 1] - s.poff_h[i];
      const amsc_params &p = s.prm[i];
      const bool fixed = p.dt_fixed > 0.0;
      if ((fixed && s.it_host[i] >= p.nsteps) ||
          (!fixed && p.max_steps > 0 && s.it_host[i] >= p.max_steps))
        s.ended[i] = 1;
    }
    all = all && s.ended[i];
  }
  s.finished = all;
  return AMSC_OK;
}

// Flag for the dense-cell P2G the scenarios whose cells have held more than
// kDensePPC particles (from the control blocks just read); a flag, once
// set, stays.  The staged-value buffer is allocated with the first flag.
// Likewise the binning's crowded-cell queue, from the first cell above 8.
int update_dense(Sim &s) {
  for (int i = 0; i < s.S && !s.wave_sort; ++i)
    if (s.ctrl_host[i].max_ppc > 8) {
      s.wave_sort = true;
      s.bin.big = s.big_q;
    }
  bool changed = false;
  for (int i = 0; i < s.S; ++i)
    if (!s.dense_h[i] && s.ctrl_host[i].max_ppc > kDensePPC) {
      s.dense_h[i] = 1;
      changed = true;
    }
  if (!changed) return AMSC_OK;
  if (!s.staged.rec) {
    const size_t rb = s.mixed ? sizeof(P2GRec<float>) : sizeof(P2GRec<double>);
    char *q = nullptr;
    if (int rc = dalloc(s, &q, rb * (size_t)s.cap)) return rc;
    s.staged.rec = q;
  }
  std::vector<int> h(s.dense_h.begin(), s.dense_h.end());
  AMSC_HIP(hipMemcpyAsync(s.dense_d, h.data(), sizeof(int) * s.S,
                          hipMemcpyHostToDevice, s.st));
  AMSC_HIP(hipStreamSynchronize(s.st));
  s.any_dense = true;
  return AMSC_OK;
}

int step_n(Sim &s, int nsteps, int *done) {
  if (nsteps < 0) return fail(AMSC_ERR_INVALID, "negative step count");
  if (int rc = read_ctrl(s)) return rc;
  if (int rc = update_dense(s)) return rc;
  std::vector<int> it0 = s.it_host;
  for (int k = 0; k < nsteps && !s.finished; ++k)
    if (int rc = launch_step(s)) return rc;
  if (int rc = read_ctrl(s)) return rc;
  if (done) {
    int d = 0;
    for (int i = 0; i < s.S; ++i) d = std::max(d, s.it_host[i] - it0[i]);
    *done = d;
  }
  return AMSC_OK;
}

// Validate parameters of one scenario.
int check_params(const amsc_params &p, int np) {
  if (np < 0 || p.nrows <= 0 || p.ncols <= 0 || !(p.hx > 0) || !(p.hy > 0))
    return fail(AMSC_ERR_INVALID, "bad scenario parameters");
  if ((long long)(p.nrows + 1) * (p.ncols + 1) >= (1ll << 31))
    return fail(AMSC_ERR_INVALID, "grid too large");
  return AMSC_OK;
}

long long tiles_for(const amsc_params &p, int cfg) {
  const long long tx = (p.ncols + 1 + kTileTC[cfg] - 1) / kTileTC[cfg];
  const long long ty = (p.nrows + 1 + kTileTR[cfg] - 1) / kTileTR[cfg];
  return tx * ty;
}

int upload_narrow(Sim &s, void *dst, const double *src, size_t n,
                  double *stage) {
  if (n == 0) return AMSC_OK;
  if (!s.mixed) {
    AMSC_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n,
                            hipMemcpyHostToDevice, s.st));
    return AMSC_OK;
  }
  AMSC_HIP(hipMemcpyAsync(stage, src, sizeof(double) * n,
                          hipMemcpyHostToDevice, s.st));
  return launch(s, "narrow", narrow_kernel, dim3(grid_for(n, 256)), dim3(256),
                (const double *)stage, static_cast<float *>(dst), (long long)n);
}

// err_mass, err_mom of the last step from the P2G tile partials, which stay
// on the device until they are asked for.
int reduce_conservation(Sim &s) {
  if (!s.cons_stale) return AMSC_OK;
  if (int rc = launch(s, "conservation", conservation_kernel, dim3(s.S),
                      dim3(256), (const Scen *)s.scen[s.cons_cfg],
                      (const double *)s.partials, s.ctrl))
    return rc;
  s.cons_stale = false;
  return AMSC_OK;
}

// Layout, scenario tables and device allocations of a batch; zero nodal
// arrays.  Particle data, Z and the control blocks are filled by the caller.
int setup(Sim *s, int S, const amsc_params *prm, const int *np_arr,
          int precision, long long capacity, hipStream_t stream = nullptr) {
  if (S <= 0 || S > kMaxScen || !prm || !np_arr)
    return fail(AMSC_ERR_INVALID, "bad ensemble arguments");
  if (precision != AMSC_PREC_FP64 && precision != AMSC_PREC_MIXED)
    return fail(AMSC_ERR_INVALID, "unknown precision");
  s->S = S;
  s->mixed = precision == AMSC_PREC_MIXED;
  s->rsz = s->mixed ? sizeof(float) : sizeof(double);
  s->prm.assign(prm, prm + S);
  long long npt = 0, nct = 0, nnt = 0, ntl[kTileCfgs] = {0, 0, 0};
  s->poff_h.assign(S + 1, 0);
  s->coff_h.assign(S + 1, 0);
  s->noff_h.assign(S + 1, 0);
  for (int i = 0; i < S; ++i) {
    if (int rc = check_params(prm[i], np_arr[i])) return rc;
    s->poff_h[i] = (int)npt;
    s->coff_h[i] = (int)nct;
    s->noff_h[i] = (int)nnt;
    npt += np_arr[i];
    nct += (long long)prm[i].nrows * prm[i].ncols;
    nnt += (long long)(prm[i].nrows + 1) * (prm[i].ncols + 1);
    for (int c = 0; c < kTileCfgs; ++c) ntl[c] += tiles_for(prm[i], c);
  }
  if (npt >= (1ll << 30) || nct >= (1ll << 31) || nnt >= (1ll << 31))
    return fail(AMSC_ERR_INVALID, "batch too large");
  s->poff_h[S] = (int)npt;
  s->coff_h[S] = (int)nct;
  s->noff_h[S] = (int)nnt;
  s->np = (int)npt;
  long long cap = capacity > 0 ? capacity : 2 * (npt > 0 ? npt : 1);
  if (cap < npt) cap = npt;
  if (cap >= (1ll << 30)) return fail(AMSC_ERR_INVALID, "capacity too large");
  s->cap = (int)cap;
  s->nc = (int)nct;
  s->nn = (int)nnt;
  for (int c = 0; c < kTileCfgs; ++c) s->ntiles[c] = (int)ntl[c];
  // scenario tables, one per tile shape
  for (int c = 0; c < kTileCfgs; ++c) {
    s->scen_h[c].resize(S);
    std::vector<int> ts;
    ts.reserve(ntl[c]);
    for (int i = 0; i < S; ++i) {
      Scen &q = s->scen_h[c][i];
      std::memset(&q, 0, sizeof(q));
      q.P = kparams_from(prm[i]);
      q.rc = make_recip(prm[i].hx, prm[i].hy);
      q.rcf = make_recip_t<float>(prm[i].hx, prm[i].hy);
      q.fd = make_fastdiv(prm[i].nrows);
      q.ms = ms_params_from(prm[i]);
      q.ms_on = prm[i].ms_on;
      q.ms_interval = prm[i].ms_interval;
      q.coff = s->coff_h[i];
      q.noff = s->noff_h[i];
      q.toff = (int)ts.size();
      q.ntiles = (int)tiles_for(prm[i], c);
      q.ntx = (prm[i].ncols + 1 + kTileTC[c] - 1) / kTileTC[c];
      for (int t = 0; t < q.ntiles; ++t) ts.push_back(i);
    }
    if (int rc = dalloc(*s, &s->scen[c], S)) return rc;
    if (int rc = dalloc(*s, &s->tscen[c], ts.size())) return rc;
    AMSC_HIP(hipMemcpy(s->scen[c], s->scen_h[c].data(), sizeof(Scen) * S,
                       hipMemcpyHostToDevice));
    AMSC_HIP(hipMemcpy(s->tscen[c], ts.data(), sizeof(int) * ts.size(),
                       hipMemcpyHostToDevice));
  }
  s->ms_counts.assign(2 * S, 0);
  s->it_host.assign(S, 0);
  s->fric_cur.assign(S, 1);
  s->ended.assign(S, 0);

  if (stream) s->st = stream;
  else AMSC_HIP(hipStreamCreate(&s->st));
  for (int b = 0; b < 2; ++b)
    if (int rc = alloc_ppt(*s, s->buf[b])) return rc;
  int **ints[] = {&s->key[0], &s->key[1], &s->slot, &s->cptcls, &s->act,
                  &s->partner, &s->wts};
  for (int **q : ints)
    if (int rc = dalloc(*s, q, cap)) return rc;
  if (int rc = dalloc(*s, &s->offs, cap + 1)) return rc;
  // operation counts (4 per scenario), overflow flag, new offsets (S + 1)
  // and active flags (S) of a merge/split step in one block, read back with
  // one copy
  if (int rc = dalloc(*s, &s->ms_cnt, 6 * S + 2)) return rc;
  s->err = s->ms_cnt + 4 * S;
  s->poff_next = s->err + 1;
  if (int rc = dalloc(*s, &s->ms_cand, std::max(s->nc, 1))) return rc;
  if (int rc = launch_fill_int(*s, s->ms_cand, std::max(s->nc, 1), 0)) return rc;
  for (int i = 0; i < S; ++i) s->any_elfs = s->any_elfs || prm[i].ms_mode == 1;
  if (s->any_elfs)
    for (int **q : {&s->elfs_ext, &s->elfs_hops, &s->elfs_qa, &s->elfs_qb})
      if (int rc = dalloc(*s, q, std::max(s->nc, 1))) return rc;
  if (int rc = dalloc(*s, &s->poff, S + 1)) return rc;
  if (int rc = dalloc(*s, &s->wscen, cap / 64 + 1)) return rc;
  if (int rc = dalloc(*s, &s->cell_off, S + 1)) return rc;
  for (int k = 0; k < NF_COUNT; ++k) {
    char *q = nullptr;
    if (int rc = dalloc(*s, &q, nf_bytes(*s, k) * (size_t)s->nn)) return rc;
    s->nod.f[k] = q;
  }
  if (int rc = dalloc(*s, &s->Z, s->nn)) return rc;
  if (int rc = dalloc(*s, &s->ctrl, S)) return rc;
  if (int rc = dalloc(*s, &s->dense_d, S)) return rc;
  s->dense_h.assign(S, 0);
  if (int rc = launch_fill_int(*s, s->dense_d, S, 0)) return rc;
  int maxt = 0;
  for (int c = 0; c < kTileCfgs; ++c) maxt = std::max(maxt, s->ntiles[c]);
  if (int rc = dalloc(*s, &s->partials, 6 * (size_t)maxt)) return rc;
  for (int b = 0; b < 2; ++b)
    if (int rc = dalloc(*s, &s->p2g_out[b], 1)) return rc;
  if (int rc = pinned_take(sizeof(Ctrl) * S, (void **)&s->ctrl_host,
                           &s->ctrl_host_bytes))
    return rc;
  if (int rc = pinned_take(sizeof(int) * (6 * S + 8), (void **)&s->host_ints,
                           &s->host_ints_bytes))
    return rc;
  // binning arrays and scan workspaces from the device cache
  s->bin.nrows = s->nc;
  s->bin.ncols = 1;
  if (int rc = dalloc(*s, &s->bin.count, s->nc)) return rc;
  if (int rc = dalloc(*s, &s->bin.cstart, (size_t)s->nc + 1)) return rc;
  // a cell of 9 or more particles is queued for a wavefront sort: at most
  // cap / 9 of them
  s->bin.big_cap = (int)(s->cap / 9 + 1);
  if (int rc = dalloc(*s, &s->big_q, (size_t)s->bin.big_cap + 1)) return rc;
  for (ScanWork *w : {&s->bin.scan, &s->ms_scan}) {
    size_t b = 0;
    if (int rc = scan_bytes(w == &s->ms_scan ? s->cap : s->nc, &b)) return rc;
    char *q = nullptr;
    if (int rc = dalloc(*s, &q, b)) return rc;
    w->tmp = q;
    w->bytes = b;
    w->owned = false;
  }
  if (int rc = launch_fill_int(*s, s->bin.count, s->nc, 0)) return rc;
  if (int rc = launch_fill_int(*s, s->bin.cstart, 1, 0)) return rc;
  if (int rc = launch_fill_int(*s, s->big_q, 1, 0)) return rc;
  for (int b = 0; b < 2; ++b) {
    P2GOut o;
    o.dzx = s->buf[b].f[PF_DZX];
    o.dzy = s->buf[b].f[PF_DZY];
    o.fpx = s->buf[b].f[PF_FPX];
    o.fpy = s->buf[b].f[PF_FPY];
    o.momx = s->buf[b].f[PF_MOMX];
    o.momy = s->buf[b].f[PF_MOMY];
    for (int k = 0; k < NF_COUNT; ++k)
      o.nod[k] = nf_narrow(k) ? nullptr : static_cast<double *>(s->nod.f[k]);
    AMSC_HIP(hipMemcpy(s->p2g_out[b], &o, sizeof(o), hipMemcpyHostToDevice));
  }
  for (int k = 0; k < NF_COUNT; ++k) {
    int rc;
    if (nf_bytes(*s, k) == 8)
      rc = launch(*s, "fill", fill_f64_kernel, dim3(grid_for(s->nn, 256)),
                  dim3(256), static_cast<double *>(s->nod.f[k]),
                  (long long)s->nn, 0.0);
    else
      rc = launch(*s, "fill", fill_f32_kernel, dim3(grid_for(s->nn, 256)),
                  dim3(256), static_cast<float *>(s->nod.f[k]),
                  (long long)s->nn, 0.0f);
    if (rc) return rc;
  }
  if (int rc = launch_fill_int(*s, s->offs, 1, 0)) return rc;  // offs[0]
  return launch_fill_int(*s, s->err, 1, 0);
}

// DATA momentum different from vpx*Mp (the first P2G then reads it,
// optim_par.cpp:151-152; afterwards mom_p == vpx*Mp exactly).
bool momentum_stored(size_t n, const double *vpx, const double *vpy,
                     const double *Mp, const double *mom_px,
                     const double *mom_py) {
  for (size_t i = 0; i < n; ++i)
    if (!(mom_px[i] == vpx[i] * Mp[i]) || !(mom_py[i] == vpy[i] * Mp[i]) ||
        std::signbit(mom_px[i]) != std::signbit(vpx[i] * Mp[i]) ||
        std::signbit(mom_py[i]) != std::signbit(vpy[i] * Mp[i]))
      return true;
  return false;
}

// Host particle arrays of n particles into buffer 0 from particle p0, then
// the step-0 initialisation of particles [p0, p0+n) (after binning).
int upload_and_init(Sim *s, size_t p0, size_t n, const double *x,
                    const double *y, const double *Mp, const double *Ap,
                    const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp) {
  hipStream_t st = s->st;
  PPtr &b = s->buf[0];
  // staging (released with the batch, so a later rebuild reuses it)
  double *stage = nullptr, *mtmp[2] = {nullptr, nullptr};
  if (n > 0) {
    if (int rc = dalloc(*s, &stage, n)) return rc;
    if (int rc = dalloc(*s, &mtmp[0], n)) return rc;
    if (int rc = dalloc(*s, &mtmp[1], n)) return rc;
  }
  int rc = AMSC_OK;
  if (n > 0) {
    const double *wide_src[] = {x, y, hp};
    const int wide_dst[] = {PF_X, PF_Y, PF_HP};
    for (int k = 0; k < 3 && !rc; ++k)
      if (hipMemcpyAsync(static_cast<double *>(b.f[wide_dst[k]]) + p0,
                         wide_src[k], sizeof(double) * n,
                         hipMemcpyHostToDevice, st) != hipSuccess)
        rc = fail(AMSC_ERR_HIP, "upload failed");
    const double *nar_src[] = {vpx, vpy, Vp, Mp, Ap};
    const int nar_dst[] = {PF_VPX, PF_VPY, PF_VP, PF_MP, PF_AP};
    for (int k = 0; k < 5 && !rc; ++k) {
      char *dst = static_cast<char *>(b.f[nar_dst[k]]) + s->rsz * p0;
      rc = upload_narrow(*s, dst, nar_src[k], n, stage);
      // the staging buffer is reused: wait for the narrowing kernel
      if (!rc && s->mixed && hipStreamSynchronize(st) != hipSuccess)
        rc = fail(AMSC_ERR_HIP, "upload failed");
    }
    if (!rc && (hipMemcpyAsync(mtmp[0], mom_px, sizeof(double) * n,
                               hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(mtmp[1], mom_py, sizeof(double) * n,
                               hipMemcpyHostToDevice, st) != hipSuccess))
      rc = fail(AMSC_ERR_HIP, "upload failed");
  }
  if (!rc && hipMemcpyAsync(s->poff, s->poff_h.data(), sizeof(int) * (s->S + 1),
                            hipMemcpyHostToDevice, st) != hipSuccess)
    rc = fail(AMSC_ERR_HIP, "upload failed");
  if (!rc && hipMemcpyAsync(s->cell_off, s->coff_h.data(),
                            sizeof(int) * (s->S + 1), hipMemcpyHostToDevice,
                            st) != hipSuccess)
    rc = fail(AMSC_ERR_HIP, "upload failed");
  if (!rc) rc = update_wscen(*s);
  if (!rc) rc = rebin(*s, false);
  if (!rc && n > 0) {
    const Tables T = tables(*s);
    const dim3 g(grid_for(n, 256)), blk(256);
    const int pb = (int)p0;
    if (s->mixed)
      rc = launch(*s, "init", init_kernel<float>, g, blk, T, b,
                  (const int *)s->key[s->kcur], (const double *)s->Z,
                  (const double *)mtmp[0], (const double *)mtmp[1], pb);
    else
      rc = launch(*s, "init", init_kernel<double>, g, blk, T, b,
                  (const int *)s->key[s->kcur], (const double *)s->Z,
                  (const double *)mtmp[0], (const double *)mtmp[1], pb);
  }
  if (!rc && hipStreamSynchronize(st) != hipSuccess)
    rc = fail(AMSC_ERR_HIP, "initialisation failed");
  return rc;
}

int upload_ctrl(Sim *s, int first) {
  std::vector<Ctrl> c0(s->S - first);
  for (auto &c : c0) {
    std::memset(&c, 0, sizeof(c));
    c.dt = 1.0e-3;  // gpu_offload.cpp:255
    c.active = 1;
    c.fric_set = 1;  // the first P2G reads Fric_p = Ap * 0
    c.flags_it = -1;
  }
  if (c0.empty()) return AMSC_OK;
  AMSC_HIP(hipMemcpyAsync(s->ctrl + first, c0.data(), sizeof(Ctrl) * c0.size(),
                          hipMemcpyHostToDevice, s->st));
  AMSC_HIP(hipStreamSynchronize(s->st));
  return AMSC_OK;
}

int create(Sim *s, int S, const amsc_params *prm, const int *np_arr,
           const double *x, const double *y, const double *Mp,
           const double *Ap, const double *vpx, const double *vpy,
           const double *mom_px, const double *mom_py, const double *hp,
           const double *Vp, const double *Z, int precision,
           long long capacity) {
  if (!Z) return fail(AMSC_ERR_INVALID, "NULL Z");
  if (int rc = setup(s, S, prm, np_arr, precision, capacity)) return rc;
  const size_t n = (size_t)s->np;
  if (n > 0 && (!x || !y || !Mp || !Ap || !vpx || !vpy || !mom_px ||
                !mom_py || !hp || !Vp))
    return fail(AMSC_ERR_INVALID, "NULL particle array");
  AMSC_HIP(hipMemcpyAsync(s->Z, Z, sizeof(double) * s->nn,
                          hipMemcpyHostToDevice, s->st));
  if (int rc = upload_ctrl(s, 0)) return rc;
  s->mom_stored = n > 0 && momentum_stored(n, vpx, vpy, Mp, mom_px, mom_py);
  const int rc = upload_and_init(s, 0, n, x, y, Mp, Ap, vpx, vpy, mom_px,
                                 mom_py, hp, Vp);
  s->launches.clear();
  s->event_used = 0;
  return rc;
}

// New batch = scenarios keep[] of `o` (device state copied, in that order)
// followed by nadd scenarios from host data (step-0 initialised).
int repack(Sim *o, Sim *s, int nkeep, const int *keep, int nadd,
           const amsc_params *params, const int *np_new, const double *x,
           const double *y, const double *Mp, const double *Ap,
           const double *vpx, const double *vpy, const double *mom_px,
           const double *mom_py, const double *hp, const double *Vp,
           const double *Z, long long capacity) {
  if (nkeep < 0 || nadd < 0 || nkeep + nadd <= 0 || (nkeep && !keep) ||
      (nadd && (!params || !np_new || !Z)))
    return fail(AMSC_ERR_INVALID, "bad repack arguments");
  if (o->broken) return fail(AMSC_ERR_INVALID, "the simulation state is invalid");
  for (int i = 0; i < nkeep; ++i)
    if (keep[i] < 0 || keep[i] >= o->S)
      return fail(AMSC_ERR_INVALID, "bad scenario index");
  // host mirrors must be exact before copying; the kept control blocks
  // carry their last conservation errors
  if (int rc = reduce_conservation(*o)) return rc;
  AMSC_HIP(hipStreamSynchronize(o->st));
  if (int rc = read_poff(*o)) return rc;
  std::vector<amsc_params> prm;
  std::vector<int> nps;
  for (int i = 0; i < nkeep; ++i) {
    prm.push_back(o->prm[keep[i]]);
    nps.push_back(o->poff_h[keep[i] + 1] - o->poff_h[keep[i]]);
  }
  size_t nnew = 0;
  for (int i = 0; i < nadd; ++i) {
    prm.push_back(params[i]);
    nps.push_back(np_new[i]);
    if (np_new[i] < 0) return fail(AMSC_ERR_INVALID, "negative particle count");
    nnew += np_new[i];
  }
  if (nnew > 0 && (!x || !y || !Mp || !Ap || !vpx || !vpy || !mom_px ||
                   !mom_py || !hp || !Vp))
    return fail(AMSC_ERR_INVALID, "NULL particle array");
  if (nnew > 0 && momentum_stored(nnew, vpx, vpy, Mp, mom_px, mom_py))
    return fail(AMSC_ERR_INVALID,
                "scenarios added to a running batch start from momentum vpx*Mp");
  // the new batch takes over the old one's stream (ordering the copies
  // after the old batch's work)
  if (int rc = setup(s, nkeep + nadd, prm.data(), nps.data(),
                     o->mixed ? AMSC_PREC_MIXED : AMSC_PREC_FP64, capacity,
                     o->st))
    return rc;
  hipStream_t st = s->st;
  if (nkeep > 0) {
    // copy maps: new and old offsets of the kept scenarios, then keep[]
    std::vector<int> tab(5 * nkeep + 2);
    int *pd = tab.data(), *ps = pd + nkeep + 1, *nd = ps + nkeep,
        *ns = nd + nkeep + 1, *kp = ns + nkeep;
    for (int i = 0; i <= nkeep; ++i) {
      pd[i] = s->poff_h[i];
      nd[i] = s->noff_h[i];
    }
    for (int i = 0; i < nkeep; ++i) {
      ps[i] = o->poff_h[keep[i]];
      ns[i] = o->noff_h[keep[i]];
      kp[i] = keep[i];
    }
    int *dtab = nullptr;
    if (int rc = dalloc(*s, &dtab, tab.size())) return rc;
    AMSC_HIP(hipMemcpyAsync(dtab, tab.data(), sizeof(int) * tab.size(),
                            hipMemcpyHostToDevice, st));
    const RepackMap mp{dtab, dtab + nkeep + 1, nkeep};
    const RepackMap mn{dtab + 2 * nkeep + 1, dtab + 3 * nkeep + 2, nkeep};
    const int np_keep = pd[nkeep], nn_keep = nd[nkeep];
    int rc = AMSC_OK;
    if (np_keep > 0)
      rc = s->mixed ? launch(*s, "repack", repack_particles_kernel<float>,
                             dim3(grid_for(np_keep, 256)), dim3(256),
                             o->buf[o->cur], s->buf[0], mp,
                             (const int *)o->wts, s->wts)
                    : launch(*s, "repack", repack_particles_kernel<double>,
                             dim3(grid_for(np_keep, 256)), dim3(256),
                             o->buf[o->cur], s->buf[0], mp,
                             (const int *)o->wts, s->wts);
    if (!rc)
      rc = s->mixed ? launch(*s, "repack", repack_nodes_kernel<float>,
                             dim3(grid_for(nn_keep, 256)), dim3(256), o->nod,
                             s->nod, (const double *)o->Z, s->Z, mn)
                    : launch(*s, "repack", repack_nodes_kernel<double>,
                             dim3(grid_for(nn_keep, 256)), dim3(256), o->nod,
                             s->nod, (const double *)o->Z, s->Z, mn);
    if (!rc)
      rc = launch(*s, "repack", repack_ctrl_kernel, dim3(nkeep), dim3(64),
                  (const Ctrl *)o->ctrl, s->ctrl, (const int *)(dtab + 4 * nkeep + 2));
    if (rc) return rc;
    for (int i = 0; i < nkeep; ++i) {
      const int k = keep[i];
      s->it_host[i] = o->it_host[k];
      s->ended[i] = o->ended[k];
      s->ms_counts[2 * i] = o->ms_counts[2 * k];
      s->ms_counts[2 * i + 1] = o->ms_counts[2 * k + 1];
      s->fric_cur[i] = o->fric_cur[k];
    }
  }
  if (nadd > 0) {
    const size_t z0 = s->noff_h[nkeep];
    AMSC_HIP(hipMemcpyAsync(s->Z + z0, Z, sizeof(double) * (s->nn - z0),
                            hipMemcpyHostToDevice, st));
  }
  if (int rc = upload_ctrl(s, nkeep)) return rc;
  s->aux_on = o->aux_on;
  s->diag = o->diag;
  s->graph_on = o->graph_on;
  s->tile_cfg = o->tile_cfg;
  s->tile_order = o->tile_order;
  s->part_bs = o->part_bs;
  s->timing = o->timing;
  s->particle_steps = o->particle_steps;
  s->aux_valid = false;
  s->nodal_valid = false;
  const size_t p0 = s->poff_h[nkeep];
  const int rc = upload_and_init(s, p0, nnew, x, y, Mp, Ap, vpx, vpy, mom_px,
                                 mom_py, hp, Vp);
  s->launches.clear();
  s->event_used = 0;
  if (rc) return rc;
  return read_ctrl(*s);
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

int copy_out(Sim &s, const void *src, size_t elem, size_t len, double *out) {
  if (len == 0) return AMSC_OK;
  if (elem == sizeof(double)) {
    AMSC_HIP(hipMemcpyAsync(out, src, sizeof(double) * len,
                            hipMemcpyDeviceToHost, s.st));
    AMSC_HIP(hipStreamSynchronize(s.st));
    return AMSC_OK;
  }
  std::vector<float> tmp(len);
  AMSC_HIP(hipMemcpyAsync(tmp.data(), src, sizeof(float) * len,
                          hipMemcpyDeviceToHost, s.st));
  AMSC_HIP(hipStreamSynchronize(s.st));
  for (size_t i = 0; i < len; ++i) out[i] = tmp[i];
  return AMSC_OK;
}

// Make scenario sc's stored friction force current (see Ctrl::fric_set).
int fric_materialize(Sim &s, int sc) {
  if (s.fric_cur[sc]) return AMSC_OK;
  const int np = s.poff_h[sc + 1] - s.poff_h[sc];
  if (np > 0) {
    const Tables T = tables(s);
    const dim3 g(grid_for(np, 256)), b(256);
    if (int rc = s.mixed ? launch(s, "fric", fric_kernel<float>, g, b, T,
                                  s.buf[s.cur], sc)
                         : launch(s, "fric", fric_kernel<double>, g, b, T,
                                  s.buf[s.cur], sc))
      return rc;
  }
  s.fric_cur[sc] = 1;
  return AMSC_OK;
}

int get_field(Sim &s, int sc, const char *name, double *out, int n) {
  if (!name || !out || sc < 0 || sc >= s.S)
    return fail(AMSC_ERR_INVALID, "bad arguments");
  const size_t p0 = s.poff_h[sc], np = s.poff_h[sc + 1] - s.poff_h[sc];
  const size_t n0 = s.noff_h[sc], nn = s.noff_h[sc + 1] - s.noff_h[sc];
  int k = find_pf(name);
  if (!std::strcmp(name, "vpxL")) k = PF_VPX;
  if (!std::strcmp(name, "vpyL")) k = PF_VPY;
  if (!std::strcmp(name, "F_21")) k = PF_F12;  // F_21 == F_12 (optim_par.cpp:116-117)
  if (k >= 0) {
    if (k >= PF_CORE && !s.aux_valid)
      return fail(AMSC_ERR_STALE, std::string("field '") + name +
                                      "' was not written by the last step "
                                      "(enable amsc_sim_set_aux)");
    if ((size_t)n < np) return fail(AMSC_ERR_INVALID, "output buffer too small");
    if (k == PF_FRX || k == PF_FRY)
      if (int rc = fric_materialize(s, sc)) return rc;
    const size_t e = pf_bytes(s, k);
    return copy_out(s, static_cast<const char *>(s.buf[s.cur].f[k]) + e * p0,
                    e, np, out);
  }
  if (!std::strcmp(name, "Z")) {
    if ((size_t)n < nn) return fail(AMSC_ERR_INVALID, "output buffer too small");
    return copy_out(s, s.Z + n0, 8, nn, out);
  }
  const int j = find_nf(name);
  if (j < 0) return fail(AMSC_ERR_INVALID, std::string("unknown field ") + name);
  if (j >= NF_AVX && !s.nodal_valid)
    return fail(AMSC_ERR_STALE, std::string("nodal field '") + name +
                                    "' was not written by the last step");
  if ((size_t)n < nn) return fail(AMSC_ERR_INVALID, "output buffer too small");
  const size_t e = nf_bytes(s, j);
  return copy_out(s, static_cast<const char *>(s.nod.f[j]) + e * n0, e, nn, out);
}

int get_field_int(Sim &s, int sc, const char *name, int *out, int n) {
  if (!name || !out || sc < 0 || sc >= s.S)
    return fail(AMSC_ERR_INVALID, "bad arguments");
  const int p0 = s.poff_h[sc], np = s.poff_h[sc + 1] - p0;
  const int c0 = s.coff_h[sc], nc = s.coff_h[sc + 1] - c0;
  const int *src = nullptr;
  int len = np, sub = 0;
  if (!std::strcmp(name, "label")) src = s.buf[s.cur].label + p0;
  else if (!std::strcmp(name, "level")) src = s.buf[s.cur].level + p0;
  else if (!std::strcmp(name, "p2g")) {
    src = s.key[s.kcur] + p0;
    sub = c0;
  } else if (!std::strcmp(name, "cptcls")) {
    src = s.cptcls + p0;
    sub = p0;
  } else if (!std::strcmp(name, "cstart")) {
    src = s.bin.cstart + c0;
    len = nc + 1;
    sub = p0;
  } else return fail(AMSC_ERR_INVALID, std::string("unknown field ") + name);
  if (n < len) return fail(AMSC_ERR_INVALID, "output buffer too small");
  if (len > 0) {
    AMSC_HIP(hipMemcpyAsync(out, src, sizeof(int) * (size_t)len,
                            hipMemcpyDeviceToHost, s.st));
    AMSC_HIP(hipStreamSynchronize(s.st));
  }
  for (int i = 0; i < len; ++i) out[i] -= sub;
  return AMSC_OK;
}

int set_field(Sim &s, int sc, const char *name, const double *in, int n) {
  if (!name || !in || sc < 0 || sc >= s.S)
    return fail(AMSC_ERR_INVALID, "bad arguments");
  const int k = find_pf(name);
  if (k < 0) return fail(AMSC_ERR_INVALID, std::string("unknown field ") + name);
  const int p0 = s.poff_h[sc], np = s.poff_h[sc + 1] - p0;
  if (n != np) return fail(AMSC_ERR_INVALID, "length must equal num_particles");
  if (np == 0) return AMSC_OK;
  if (k == PF_AP || k == PF_VXDX || k == PF_VXDY || k == PF_VYDX || k == PF_VYDY) {
    // A step before a merge/split step of a fixed-dt run under the solvers'
    // rule keeps these only where its strain tests passed and hands the
    // tests on as bits; a new value could not reach them.
    int v[2];
    AMSC_HIP(hipMemcpyAsync(&v[0], &s.ctrl[sc].it, sizeof(int),
                            hipMemcpyDeviceToHost, s.st));
    AMSC_HIP(hipMemcpyAsync(&v[1], &s.ctrl[sc].flags_it, sizeof(int),
                            hipMemcpyDeviceToHost, s.st));
    AMSC_HIP(hipStreamSynchronize(s.st));
    if (!s.ended[sc] && v[1] == v[0] + 1)
      return fail(AMSC_ERR_STALE,
                  std::string("field '") + name +
                      "' is read by the next merge/split as the last step "
                      "left it; set it with amsc_sim_set_aux enabled before "
                      "that step");
  }
  // the next P2G reads the stored friction force as it was before the write
  if (int rc = fric_materialize(s, sc)) return rc;
  {
    const int one = 1;
    AMSC_HIP(hipMemcpyAsync(&s.ctrl[sc].fric_set, &one, sizeof(int),
                            hipMemcpyHostToDevice, s.st));
    AMSC_HIP(hipStreamSynchronize(s.st));
  }
  const size_t e = pf_bytes(s, k);
  char *dst = static_cast<char *>(s.buf[s.cur].f[k]) + e * p0;
  if (e == 8) {
    AMSC_HIP(hipMemcpyAsync(dst, in, sizeof(double) * np,
                            hipMemcpyHostToDevice, s.st));
  } else {
    std::vector<float> tmp(in, in + np);
    AMSC_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(float) * np,
                            hipMemcpyHostToDevice, s.st));
    AMSC_HIP(hipStreamSynchronize(s.st));
  }
  AMSC_HIP(hipStreamSynchronize(s.st));
  if (k == PF_X || k == PF_Y) return rebin(s, false);
  return AMSC_OK;
}

int kernel_times(Sim &s, const char **names, double *ms, int *count, int cap,
                 int *n) {
  if (!n) return fail(AMSC_ERR_INVALID, "bad arguments");
  AMSC_HIP(hipStreamSynchronize(s.st));
  std::vector<const char *> nm;
  std::vector<double> tot;
  std::vector<int> cnt;
  for (auto &l : s.launches) {
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

int reset_times(Sim &s) {
  AMSC_HIP(hipStreamSynchronize(s.st));
  s.launches.clear();
  s.event_used = 0;
  return AMSC_OK;
}

// Average time of the launches named `name` since the last reset.
int avg_time(Sim &s, const char *name, double *avg) {
  AMSC_HIP(hipStreamSynchronize(s.st));
  double tot = 0.0;
  int cnt = 0;
  for (auto &l : s.launches) {
    if (std::strcmp(l.name, name)) continue;
    float e = 0.f;
    AMSC_HIP(hipEventElapsedTime(&e, l.a, l.b));
    tot += e;
    ++cnt;
  }
  *avg = cnt ? tot / cnt : 0.0;
  return AMSC_OK;
}

// Time every P2G variant (tile shape x tile order), then every particle
// workgroup size, over `steps` live steps each (the results are the same
// whichever variant runs a step) and keep the fastest of each.
// variant_ms (may be null): kP2GVariants P2G averages, then the particle
// averages of the workgroup sizes.
int autotune(Sim &s, int steps, int *cfg_out, int *bs_out, int *used,
             double *variant_ms) {
  if (steps <= 0) return fail(AMSC_ERR_INVALID, "steps must be positive");
  const bool timing0 = s.timing;
  int total = 0;
  // one step untimed first: a run's first step stages the stored friction
  // force and the data's momentum, which later steps do not
  {
    int d = 0;
    if (int rc = step_n(s, 1, &d)) return rc;
    total += d;
  }
  s.timing = true;
  // one step of each candidate per round, `steps` rounds: every candidate
  // sees about the same stage of the flow
  auto rounds = [&](int n, const char *name, auto set, double *avg) -> int {
    std::vector<double> sum(n, 0.0);
    std::vector<int> cnt(n, 0);
    for (int r = 0; r < steps && !s.finished; ++r)
      for (int v = 0; v < n && !s.finished; ++v) {
        set(v);
        if (int rc = reset_times(s)) return rc;
        int d = 0;
        if (int rc = step_n(s, 1, &d)) return rc;
        total += d;
        double a = 0.0;
        if (int rc = avg_time(s, name, &a)) return rc;
        if (a > 0.0) {
          sum[v] += a;
          ++cnt[v];
        }
      }
    for (int v = 0; v < n; ++v) avg[v] = cnt[v] ? sum[v] / cnt[v] : 0.0;
    return AMSC_OK;
  };
  const int v0 = s.tile_cfg + kTileCfgs * s.tile_order;
  double pv[kP2GVariants], bv_ms[2];
  if (int rc = rounds(kP2GVariants, "p2g_nodal", [&](int v) {
        s.tile_cfg = v % kTileCfgs;
        s.tile_order = v / kTileCfgs;
      }, pv))
    return rc;
  // In the mixed form the swept shape (configuration 1) is taken only if it
  // beats the best tiled one by kSweepMargin: its columns of more than 128
  // positions go through an unpipelined path, and their share grows as a
  // flow spreads and cells fill unevenly (on the 4M workload: none at step
  // 60, 12 % of columns at step 120, 22 % at step 220), while the tiled
  // kernels keep their speed (over 200 steps: 0.3167 against 0.3007 ms per
  // step).  Measured at a run's start, the two are often within a few per
  // cent.  The FP64 swept kernel loads the next 64 positions of such
  // columns ahead and keeps its lead over the run (0.3619 against 0.3666
  // ms per step), so it needs no margin.
  const double kSweepMargin = s.mixed ? 0.05 : 0.0;
  int bv = v0, bt = -1;
  double best = 1e300, best_t = 1e300;
  for (int v = 0; v < kP2GVariants; ++v) {
    if (!(pv[v] > 0.0)) continue;
    if (pv[v] < best) {
      best = pv[v];
      bv = v;
    }
    if (v % kTileCfgs != 1 && pv[v] < best_t) {
      best_t = pv[v];
      bt = v;
    }
  }
  if (bv % kTileCfgs == 1 && bt >= 0 && best > (1.0 - kSweepMargin) * best_t) bv = bt;
  // Likewise the linear tile order (1) only if it beats the descending one
  // (0) of the same shape by kOrderMargin.  The descending order stages
  // first the particles that the previous update wrote last, the likeliest
  // still in the memory-side cache; at a run's start the two time within a
  // few per cent either way, over 200 steps of the 4M workload the linear
  // order was the slower in every pair (FP64 swept 0.3686 against 0.3619
  // ms per step, mixed 8x32 tiles 0.3051 against 0.3007).
  constexpr double kOrderMargin = 0.03;
  if (bv / kTileCfgs == 1) {
    const int b0 = bv % kTileCfgs;
    if (pv[b0] > 0.0 && pv[bv] > (1.0 - kOrderMargin) * pv[b0]) bv = b0;
  }
  s.tile_cfg = bv % kTileCfgs;
  s.tile_order = bv / kTileCfgs;
  const int bs0 = s.part_bs;
  if (int rc = rounds(2, "particle", [&](int i) { s.part_bs = kPartBS[i]; }, bv_ms))
    return rc;
  int bbs = bs0;
  best = 1e300;
  for (int i = 0; i < 2; ++i)
    if (bv_ms[i] > 0.0 && bv_ms[i] < best) {
      best = bv_ms[i];
      bbs = kPartBS[i];
    }
  s.part_bs = bbs;
  if (variant_ms) {
    for (int v = 0; v < kP2GVariants; ++v) variant_ms[v] = pv[v];
    for (int i = 0; i < 2; ++i) variant_ms[kP2GVariants + i] = bv_ms[i];
  }
  s.timing = timing0;
  (void)reset_times(s);
  if (cfg_out) *cfg_out = bv;
  if (bs_out) *bs_out = s.part_bs;
  if (used) *used = total;
  return AMSC_OK;
}

// Average duration of `reps` back-to-back P2G launches on the current state
// (P2G is a function of the particle state, so repeating it rewrites the
// same nodal velocities): the bracketing events are amortised over the run.
int time_p2g(Sim &s, int reps, double *avg_ms) {
  if (reps <= 0 || !avg_ms) return fail(AMSC_ERR_INVALID, "bad arguments");
  if (s.broken) return fail(AMSC_ERR_INVALID, "the simulation state is invalid");
  const bool timing0 = s.timing;
  s.timing = false;
  hipEvent_t e0, e1;
  AMSC_HIP(hipEventCreate(&e0));
  AMSC_HIP(hipEventCreate(&e1));
  const P2GArgs a = p2g_args(s, AUX_NONE, false);
  int rc = AMSC_OK;
  if (hipEventRecord(e0, s.st) != hipSuccess) rc = fail(AMSC_ERR_HIP, "event");
  for (int i = 0; i < reps && !rc; ++i)
    rc = s.mixed ? launch_p2g<float>(s, a, s.np) : launch_p2g<double>(s, a, s.np);
  if (!rc && hipEventRecord(e1, s.st) != hipSuccess) rc = fail(AMSC_ERR_HIP, "event");
  float ms = 0.f;
  if (!rc && (hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = fail(AMSC_ERR_HIP, "event timing");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s.timing = timing0;
  *avg_ms = ms / reps;
  return rc;
}

// Device bytes create() allocates for a batch (admission estimate).
long long bytes_estimate(int S, const amsc_params *prm, const int *np,
                         int precision, long long capacity) {
  long long npt = 0, nct = 0, nnt = 0, ntl = 0;
  for (int i = 0; i < S; ++i) {
    npt += np[i];
    nct += (long long)prm[i].nrows * prm[i].ncols;
    nnt += (long long)(prm[i].nrows + 1) * (prm[i].ncols + 1);
    long long m = 0;
    for (int c = 0; c < kTileCfgs; ++c) m = std::max(m, tiles_for(prm[i], c));
    ntl += m;
  }
  const long long cap = capacity > 0 ? std::max(capacity, npt) : 2 * std::max(npt, 1ll);
  const long long r = precision == AMSC_PREC_MIXED ? 4 : 8;
  // two particle buffers, the int arrays, and the dense-cell P2G's staged
  // values (allocated only if a scenario turns dense: counted anyway)
  const long long per_p = 2 * (4 * 8 + (PF_COUNT - 4) * r + 8) + 7 * 4 + 4 +
                          (r == 8 ? (long long)sizeof(P2GRec<double>)
                                  : (long long)sizeof(P2GRec<float>));
  const long long per_n = 2 * r + (NF_COUNT - 2) * 8 + 8;
  bool elfs = false;
  for (int i = 0; i < S; ++i) elfs = elfs || prm[i].ms_mode == 1;
  // + upload staging: 3 FP64 values per initial particle
  return cap * per_p + npt * 24 + (cap / 64 + 1) * 4 + (cap / 9 + 2) * 4 +
         nnt * per_n + nct * (elfs ? 28 : 12) +
         ntl * (6 * 8 + 4 * kTileCfgs) +
         S * (long long)(sizeof(Ctrl) + kTileCfgs * sizeof(Scen) + 40) +
         (1 << 16);  // allocations are rounded up to 256 bytes
}

}  // namespace

}  // namespace amsc

using namespace amsc;

struct amsc_sim : amsc::Sim {};

namespace {
int bad_handle() { return fail(AMSC_ERR_INVALID, "NULL simulation handle"); }
}  // namespace

extern "C" {

int amsc_sim_create(const amsc_params *p, int np, const double *x,
                    const double *y, const double *Mp, const double *Ap,
                    const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp, const double *Z,
                    int capacity, amsc_sim **out) {
  if (!p || !out) return fail(AMSC_ERR_INVALID, "amsc_sim_create: bad parameters");
  return amsc_ens_create(1, p, &np, x, y, Mp, Ap, vpx, vpy, mom_px, mom_py, hp,
                         Vp, Z, AMSC_PREC_FP64, capacity, out);
}

int amsc_ens_create(int nscen, const amsc_params *params, const int *np,
                    const double *x, const double *y, const double *Mp,
                    const double *Ap, const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp, const double *Z,
                    int precision, long long capacity, amsc_ens **out) {
  if (!out) return fail(AMSC_ERR_INVALID, "NULL output handle");
  *out = nullptr;
  amsc_sim *s = new amsc_sim();
  const int rc = create(s, nscen, params, np, x, y, Mp, Ap, vpx, vpy, mom_px,
                        mom_py, hp, Vp, Z, precision, capacity);
  if (rc) {
    const std::string msg = last_error_cstr();
    sim_free(s);
    return fail(rc, msg);
  }
  *out = s;
  return AMSC_OK;
}

int amsc_sim_destroy(amsc_sim *s) {
  sim_free(s);
  return AMSC_OK;
}

int amsc_sim_step(amsc_sim *s, int nsteps, int *done) {
  if (!s) return bad_handle();
  return step_n(*s, nsteps, done);
}

int amsc_sim_set_aux(amsc_sim *s, int on) {
  if (!s) return bad_handle();
  s->aux_on = on != 0;
  return AMSC_OK;
}

int amsc_sim_set_timing(amsc_sim *s, int on) {
  if (!s) return bad_handle();
  s->timing = on != 0;
  return AMSC_OK;
}

int amsc_sim_set_graph(amsc_sim *s, int on) {
  if (!s) return bad_handle();
  s->graph_on = on != 0;
  return AMSC_OK;
}

int amsc_sim_set_tiling(amsc_sim *s, int p2g_shape, int particle_block) {
  if (!s) return bad_handle();
  if (p2g_shape < 0 || p2g_shape >= kP2GVariants ||
      (particle_block != 256 && particle_block != 512))
    return fail(AMSC_ERR_INVALID, "unknown tiling");
  s->tile_cfg = p2g_shape % kTileCfgs;
  s->tile_order = p2g_shape / kTileCfgs;
  s->part_bs = particle_block;
  return AMSC_OK;
}

int amsc_sim_autotune(amsc_sim *s, int steps_per_variant, int *p2g_shape,
                      int *particle_block, int *steps_used,
                      double *variant_ms) {
  if (!s) return bad_handle();
  return autotune(*s, steps_per_variant, p2g_shape, particle_block, steps_used,
                  variant_ms);
}

int amsc_sim_time_p2g(amsc_sim *s, int reps, double *avg_ms) {
  if (!s) return bad_handle();
  return time_p2g(*s, reps, avg_ms);
}

int amsc_sim_kernel_times(amsc_sim *s, const char **names, double *ms,
                          int *count, int cap, int *n) {
  if (!s) return bad_handle();
  return kernel_times(*s, names, ms, count, cap, n);
}

int amsc_sim_reset_kernel_times(amsc_sim *s) {
  if (!s) return bad_handle();
  return reset_times(*s);
}

int amsc_sim_num_particles(amsc_sim *s, int *np) {
  if (!s || !np) return fail(AMSC_ERR_INVALID, "bad arguments");
  *np = s->poff_h[1] - s->poff_h[0];
  return AMSC_OK;
}

int amsc_sim_state(amsc_sim *s, int *it, double *t, double *dt) {
  if (!s) return bad_handle();
  if (int rc = read_ctrl(*s)) return rc;
  if (it) *it = s->ctrl_host[0].it;
  if (t) *t = s->ctrl_host[0].t;
  if (dt) *dt = s->ctrl_host[0].dt;
  return AMSC_OK;
}

int amsc_sim_get_field(amsc_sim *s, const char *name, double *out, int n) {
  if (!s) return bad_handle();
  return get_field(*s, 0, name, out, n);
}

int amsc_sim_get_field_int(amsc_sim *s, const char *name, int *out, int n) {
  if (!s) return bad_handle();
  return get_field_int(*s, 0, name, out, n);
}

int amsc_sim_set_field(amsc_sim *s, const char *name, const double *in, int n) {
  if (!s) return bad_handle();
  return set_field(*s, 0, name, in, n);
}

int amsc_sim_conservation(amsc_sim *s, double out[3]) {
  return amsc_ens_conservation(s, 0, out);
}

int amsc_sim_particle_steps(amsc_sim *s, long long *out) {
  if (!s || !out) return fail(AMSC_ERR_INVALID, "bad arguments");
  *out = s->particle_steps;
  return AMSC_OK;
}

int amsc_sim_ms_counts(amsc_sim *s, int out[2]) {
  return amsc_ens_ms_counts(s, 0, out);
}

// ---------------------------------------------------------------- batches
int amsc_ens_status(amsc_ens *e, int *it, double *t, double *dt, int *active,
                    int *np) {
  if (!e) return bad_handle();
  if (int rc = read_ctrl(*e)) return rc;
  for (int i = 0; i < e->S; ++i) {
    const Ctrl &c = e->ctrl_host[i];
    if (it) it[i] = c.it;
    if (t) t[i] = c.t;
    if (dt) dt[i] = c.dt;
    if (active) active[i] = e->ended[i] ? 0 : 1;
    if (np) np[i] = e->poff_h[i + 1] - e->poff_h[i];
  }
  return AMSC_OK;
}

int amsc_ens_diverged(amsc_ens *e, int *out) {
  if (!e) return bad_handle();
  if (!out) return fail(AMSC_ERR_INVALID, "NULL output");
  if (int rc = read_ctrl(*e)) return rc;
  for (int i = 0; i < e->S; ++i) out[i] = e->ctrl_host[i].diverged;
  return AMSC_OK;
}

int amsc_ens_get_field(amsc_ens *e, int scen, const char *name, double *out,
                       int n) {
  if (!e) return bad_handle();
  return get_field(*e, scen, name, out, n);
}

int amsc_ens_get_field_int(amsc_ens *e, int scen, const char *name, int *out,
                           int n) {
  if (!e) return bad_handle();
  return get_field_int(*e, scen, name, out, n);
}

int amsc_ens_set_field(amsc_ens *e, int scen, const char *name,
                       const double *in, int n) {
  if (!e) return bad_handle();
  return set_field(*e, scen, name, in, n);
}

int amsc_ens_conservation(amsc_ens *e, int scen, double out[3]) {
  if (!e || !out || scen < 0 || scen >= e->S)
    return fail(AMSC_ERR_INVALID, "bad arguments");
  if (int rc = reduce_conservation(*e)) return rc;
  if (int rc = read_ctrl(*e)) return rc;
  out[0] = e->ctrl_host[scen].cons[0];
  out[1] = e->ctrl_host[scen].cons[1];
  out[2] = e->ctrl_host[scen].t_start;
  return AMSC_OK;
}

int amsc_ens_ms_counts(amsc_ens *e, int scen, int out[2]) {
  if (!e || !out || scen < 0 || scen >= e->S)
    return fail(AMSC_ERR_INVALID, "bad arguments");
  out[0] = e->ms_counts[2 * scen];
  out[1] = e->ms_counts[2 * scen + 1];
  return AMSC_OK;
}

int amsc_sim_params(amsc_sim *s, amsc_params *out) {
  if (!s || !out) return fail(AMSC_ERR_INVALID, "bad arguments");
  *out = s->prm[0];
  return AMSC_OK;
}

int amsc_ens_info(amsc_ens *e, int *nscen, int *precision, long long *bytes,
                  int *capacity) {
  if (!e) return bad_handle();
  if (nscen) *nscen = e->S;
  if (precision) *precision = e->mixed ? AMSC_PREC_MIXED : AMSC_PREC_FP64;
  if (bytes) *bytes = (long long)e->bytes;
  if (capacity) *capacity = e->cap;
  return AMSC_OK;
}

int amsc_ens_repack(amsc_ens *e, int nkeep, const int *keep, int nadd,
                    const amsc_params *params, const int *np, const double *x,
                    const double *y, const double *Mp, const double *Ap,
                    const double *vpx, const double *vpy,
                    const double *mom_px, const double *mom_py,
                    const double *hp, const double *Vp, const double *Z,
                    long long capacity, amsc_ens **out) {
  if (!e || !out) return fail(AMSC_ERR_INVALID, "NULL handle");
  *out = nullptr;
  amsc_sim *s = new amsc_sim();
  const int rc = repack(e, s, nkeep, keep, nadd, params, np, x, y, Mp, Ap,
                        vpx, vpy, mom_px, mom_py, hp, Vp, Z, capacity);
  if (rc) {
    const std::string msg = last_error_cstr();
    if (s->st == e->st) s->st = nullptr;  // the stream stays with e
    sim_free(s);
    return fail(rc, msg);
  }
  e->st = nullptr;  // now owned by s
  sim_free(e);
  *out = s;
  return AMSC_OK;
}

int amsc_release_cache(void) {
  cache_release();
  return AMSC_OK;
}

long long amsc_ens_bytes_estimate(int nscen, const amsc_params *params,
                                  const int *np, int precision,
                                  long long capacity) {
  if (nscen <= 0 || !params || !np) return -1;
  return bytes_estimate(nscen, params, np, precision, capacity);
}

}  // extern "C"
