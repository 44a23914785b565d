"""Served hazard-ensemble requests on one GPU.

Requests (scenarios of uneven grid and particle counts, each run to its own
simulated end time) arrive as a Poisson process.  Two forms serve them:

* batched: every admitted request is a scenario of one device-resident
  `Ensemble`; the batch advances `chunk` steps at a time; finished scenarios
  leave (at the next join, or once they hold half the batch) and arrived
  ones join through a device-side repack (the state of a running scenario
  never leaves the GPU).  Admission keeps the device bytes of the old and
  the new batch together (both exist during a repack) within the HBM
  budget.  Steps are launched eagerly by default: a HIP graph would have to
  be recorded again after every repack;
* sequential: one `Simulation` per request, first come first served (the
  single-scenario driver loop).

Both report time-to-solution (arrival to completion, wall clock) per request,
its p50/p99, and completed scenarios per second over the run.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from .data import params_from_data, particles_from_data

PARTICLE_KEYS = ("x", "y", "Mp", "Ap", "vpx", "vpy", "mom_px", "mom_py", "hp", "Vp")


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start of the run
    prm: dict
    particles: dict
    Z: np.ndarray
    admitted: float = math.nan
    done: float = math.nan
    steps: int = 0
    result: dict | None = None
    cur_np: int = -1  # particle count while running (merge/split moves it)
    diverged: bool = False  # run stopped by the divergence limit

    @property
    def np(self) -> int:
        return len(self.particles["x"])

    @property
    def live_np(self) -> int:
        return self.cur_np if self.cur_np >= 0 else self.np


def poisson_arrivals(rate: float, n: int, seed: int) -> np.ndarray:
    """Arrival times of n requests of a Poisson process of `rate` per s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def make_requests(datas, arrivals, max_steps: int = 0):
    out = []
    for i, (d, t) in enumerate(zip(datas, arrivals)):
        prm = params_from_data(d, max_steps=max_steps)
        out.append(Request(i, float(t), prm, particles_from_data(d), np.asarray(d["Z"])))
    return out


def _pcts(tts):
    a = np.asarray(tts, np.float64)
    a = a[np.isfinite(a)]  # requests finished within the run's budget
    if a.size == 0:
        return {"p50_s": None, "p99_s": None, "mean_s": None}
    return {"p50_s": float(np.percentile(a, 50)), "p99_s": float(np.percentile(a, 99)),
            "mean_s": float(a.mean())}


INT_FIELDS = ("label", "level")


def _grab(get, get_int, fields):
    return {f: (get_int(f) if f in INT_FIELDS else get(f)) for f in fields}


def _cat(reqs, key):
    return np.ascontiguousarray(np.concatenate(
        [np.asarray(r.particles[key], np.float64) for r in reqs]))


class BatchedServer:
    """Batched form over amsc_ens_* (see module docstring)."""

    def __init__(self, precision: str = "mixed", budget_bytes: int | None = None,
                 chunk: int = 10, graph: bool = False, capacity_factor: float = 4.0,
                 collect=()):
        self.precision = precision
        self.budget = budget_bytes
        self.chunk = chunk
        self.graph = graph
        self.capacity_factor = capacity_factor
        self.collect = tuple(collect)  # fields kept from each finished request

    def _grab(self, ens, i):
        return _grab(lambda f: ens.get(i, f), lambda f: ens.get_int(i, f), self.collect)

    def _capacity(self, reqs):
        # merge/split may split particles twice (ms_min_level = -2): room for
        # 4x the present count by default
        return int(self.capacity_factor * max(1, sum(r.live_np for r in reqs)))

    def _bytes(self, reqs):
        from .simulation import Ensemble
        if not reqs:
            return 0
        return Ensemble.bytes_estimate([r.prm for r in reqs], [r.live_np for r in reqs],
                                       self.precision, self._capacity(reqs))

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep, log=None,
            budget_s=math.inf, progress_s=10.0):
        """Serve `requests`; stop after budget_s of wall clock (requests not
        finished by then keep done = nan and are counted as unfinished);
        log a progress line every progress_s."""
        import ctypes as C
        from . import native
        from .simulation import Ensemble, PRECISION
        L = native.lib()
        pending = sorted(requests, key=lambda r: r.arrival)
        active: list[Request] = []
        ens = None
        t0 = clock()
        repacks = 0
        repack_s = 0.0
        last_log = 0.0
        peak = 0  # largest device footprint admitted (old + new batch)
        while pending or active:
            now = clock() - t0
            if now > budget_s:
                if log:
                    log(f"[serve] budget of {budget_s:.0f} s reached: "
                        f"{len(active)} running, {len(pending)} waiting")
                break
            # admission, first come first served, within the budget
            joining = []
            while pending and pending[0].arrival <= now:
                cand = active + joining + [pending[0]]
                need = self._bytes(cand) + (self._bytes(active) if ens else 0)
                if self.budget is not None and need > self.budget and (active or joining):
                    break
                r = pending.pop(0)
                r.admitted = now
                joining.append(r)
                peak = max(peak, need)
            if joining:
                t_r = clock()
                if ens is not None:
                    st0 = ens.status()
                    keep = np.array([i for i in range(len(active)) if st0["active"][i]],
                                    np.int32)
                    for i in range(len(active)):
                        if not st0["active"][i] and self.collect:
                            active[i].result = self._grab(ens, i)
                else:
                    keep = np.zeros(0, np.int32)
                newreq = [active[i] for i in keep] + joining
                cap = self._capacity(newreq)
                P = (native.AmscParams * len(joining))(
                    *[native.AmscParams.from_dict(r.prm) for r in joining])
                nps = np.array([r.np for r in joining], np.int32)
                arrs = [_cat(joining, k) for k in PARTICLE_KEYS]
                Z = np.ascontiguousarray(np.concatenate(
                    [np.asarray(r.Z, np.float64).ravel() for r in joining]))
                if ens is None:
                    ens = Ensemble.__new__(Ensemble)
                    h = C.c_void_p()
                    native.check(L.amsc_ens_create(len(joining), P, nps, *arrs, Z,
                                                   PRECISION[self.precision], cap,
                                                   C.byref(h)))
                    ens._h = h
                    if self.graph:
                        ens.set_graph(True)
                else:
                    h = C.c_void_p()
                    native.check(L.amsc_ens_repack(ens._h, len(keep), keep, len(joining),
                                                   P, nps, *arrs, Z, cap, C.byref(h)))
                    ens._h = h
                    repacks += 1
                active = newreq
                ens.nscen = len(active)
                ens.prms = [r.prm for r in active]
                repack_s += clock() - t_r
            if not active:
                if pending:
                    sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.05)))
                continue
            done = C.c_int()
            native.check(L.amsc_sim_step(ens._h, self.chunk, C.byref(done)))
            st = ens.status()
            now = clock() - t0
            if log and now - last_log >= progress_s:
                last_log = now
                log(f"[serve] {now:.1f} s: {int(st['active'].sum())} running "
                    f"(steps {int(st['it'].min())}..{int(st['it'].max())}, "
                    f"sim time {float(st['t'].min()):.2f}..{float(st['t'].max()):.2f} s, "
                    f"{int(st['np'].sum())} particles), {len(pending)} waiting")
            still = []
            for i, r in enumerate(active):
                r.steps = int(st["it"][i])
                r.cur_np = int(st["np"][i])
                r.diverged = bool(st["diverged"][i])
                if st["active"][i]:
                    still.append(i)
                elif math.isnan(r.done):
                    r.done = now
                    if log:
                        log(f"[serve] request {r.rid} done: {r.steps} steps, "
                            f"{r.done - r.arrival:.3f} s")
            # finished scenarios leave at the next join, or once they hold
            # half the batch's particles (a repack rebuilds the batch)
            fin_np = sum(r.live_np for i, r in enumerate(active) if not st["active"][i])
            arrived = bool(pending) and pending[0].arrival <= now
            leave = still and (arrived or 2 * fin_np > sum(r.live_np for r in active))
            if not still or (len(still) != len(active) and leave):
                keep_set = set(still)
                if self.collect:
                    for i, r in enumerate(active):
                        if i not in keep_set:
                            r.result = self._grab(ens, i)
                if still:
                    t_r = clock()
                    keep = np.array(still, np.int32)
                    h = C.c_void_p()
                    native.check(L.amsc_ens_repack(
                        ens._h, len(keep), keep, 0, None, None, *([None] * 11),
                        self._capacity([active[i] for i in still]), C.byref(h)))
                    ens._h = h
                    repacks += 1
                    repack_s += clock() - t_r
                    active = [active[i] for i in still]
                    ens.nscen = len(active)
                    ens.prms = [r.prm for r in active]
                else:
                    ens.close()
                    ens = None
                    active = []
        wall = clock() - t0
        if ens is not None:
            ens.close()
        tts = [r.done - r.arrival for r in requests]
        fin = int(np.isfinite(tts).sum())
        return {"form": f"batched-{self.precision}", "requests": len(requests),
                "finished": fin, "diverged": int(sum(r.diverged for r in requests)),
                "wall_s": wall, "scen_per_s": fin / wall,
                "repacks": repacks, "repack_s": repack_s, "peak_bytes": int(peak),
                "budget_bytes": self.budget, **_pcts(tts)}


class SequentialServer:
    """One `Simulation` (FP64) per request, first come first served."""

    def __init__(self, chunk: int = 100, collect=(), graph: bool = False):
        self.chunk = chunk
        self.collect = tuple(collect)  # fields kept from each finished request
        self.graph = graph

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep, log=None,
            budget_s=math.inf, progress_s=10.0):
        """As BatchedServer.run."""
        from .simulation import Simulation
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        last_log = 0.0
        for r in pending:
            now = clock() - t0
            if now > budget_s:
                break
            if r.arrival > now:
                sleep(r.arrival - now)
            r.admitted = clock() - t0
            s = Simulation(r.prm, r.particles, r.Z, capacity=4 * max(r.np, 1))
            if self.graph:
                s.set_graph(True)
            while clock() - t0 <= budget_s:
                k = s.step(self.chunk)
                r.steps += k
                now = clock() - t0
                if k < self.chunk:
                    r.done = now
                    r.diverged = s.diverged
                    break
                if log and now - last_log >= progress_s:
                    last_log = now
                    it, t, dt = s.state()
                    log(f"[serve] {now:.1f} s: request {r.rid} at step {it}, "
                        f"sim time {t:.2f} s, dt {dt:.2e}, {s.num_particles} particles")
            if self.collect and not math.isnan(r.done):
                r.result = _grab(s.get, s.get_int, self.collect)
            s.close()
            if log and not math.isnan(r.done):
                log(f"[serve] request {r.rid} done: {r.steps} steps, "
                    f"{r.done - r.arrival:.3f} s")
        wall = clock() - t0
        tts = [r.done - r.arrival for r in requests]
        fin = int(np.isfinite(tts).sum())
        return {"form": "sequential-fp64", "requests": len(requests), "finished": fin,
                "diverged": int(sum(r.diverged for r in requests)),
                "wall_s": wall, "scen_per_s": fin / wall, **_pcts(tts)}
