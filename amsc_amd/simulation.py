"""Host-side driver over the C ABI: the role of main/gpu_offload.cpp (load
DATA, run the time loop, read fields back), with all state resident on the
GPU between calls; `Ensemble` advances a batch of scenarios together, in
FP64 or mixed FP32/FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .data import load_data_json, params_from_data, particles_from_data

PRECISION = {"fp64": 0, "mixed": 1}

NODAL_FIELDS = ("vvxL", "vvyL", "avx", "avy", "Mv", "mom_vx", "mom_vy",
                "F_ext_vx", "F_ext_vy", "F_int_vx", "F_int_vy", "Fric_x",
                "Fric_y", "FPxv", "FPyv", "Ftot_vx", "Ftot_vy", "vvx", "vvy",
                "Z")


class Simulation:
    """Device-resident FP64 depth-averaged MPM run (optim_par semantics)."""

    def __init__(self, prm: dict, particles: dict, Z, capacity: int = 0):
        L = native.lib()
        self.prm = dict(prm)
        self.nn = (prm["nrows"] + 1) * (prm["ncols"] + 1)
        self.nc = prm["nrows"] * prm["ncols"]
        P = native.AmscParams.from_dict(prm)
        a = [np.ascontiguousarray(particles[k], np.float64) for k in
             ("x", "y", "Mp", "Ap", "vpx", "vpy", "mom_px", "mom_py", "hp", "Vp")]
        h = C.c_void_p()
        native.check(L.amsc_sim_create(C.byref(P), len(a[0]), *a,
                                       np.ascontiguousarray(Z, np.float64),
                                       int(capacity), C.byref(h)))
        self._h = h

    @classmethod
    def from_data(cls, d: dict, capacity: int = 0, max_steps: int = 0):
        prm = params_from_data(d, max_steps=max_steps or int(d.get("MAX_STEPS", 0)))
        return cls(prm, particles_from_data(d), d["Z"], capacity)

    @classmethod
    def from_json(cls, path: str, capacity: int = 0, max_steps: int = 0):
        return cls.from_data(load_data_json(path), capacity, max_steps)

    def close(self):
        if getattr(self, "_h", None):
            native.lib().amsc_sim_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n: int = 1) -> int:
        done = C.c_int()
        native.check(native.lib().amsc_sim_step(self._h, int(n), C.byref(done)))
        return done.value

    def run(self, chunk: int = 100) -> int:
        total = 0
        while True:
            k = self.step(chunk)
            total += k
            if k < chunk:
                return total

    def set_aux(self, on: bool = True):
        native.check(native.lib().amsc_sim_set_aux(self._h, int(bool(on))))

    def set_timing(self, on: bool = True):
        native.check(native.lib().amsc_sim_set_timing(self._h, int(bool(on))))

    def reset_kernel_times(self):
        native.check(native.lib().amsc_sim_reset_kernel_times(self._h))

    def kernel_times(self) -> dict:
        cap = 32
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_int * cap)()
        n = C.c_int()
        native.check(native.lib().amsc_sim_kernel_times(self._h, names, ms, cnt,
                                                        cap, C.byref(n)))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(min(n.value, cap))}

    @property
    def num_particles(self) -> int:
        n = C.c_int()
        native.check(native.lib().amsc_sim_num_particles(self._h, C.byref(n)))
        return n.value

    def state(self):
        it, t, dt = C.c_int(), C.c_double(), C.c_double()
        native.check(native.lib().amsc_sim_state(self._h, C.byref(it), C.byref(t),
                                                 C.byref(dt)))
        return it.value, t.value, dt.value

    @property
    def diverged(self) -> bool:
        """True once the run was stopped because a particle speed exceeded
        1e4 m/s or became NaN."""
        out = np.zeros(1, np.int32)
        native.check(native.lib().amsc_ens_diverged(self._h, out))
        return bool(out[0])

    def get(self, name: str) -> np.ndarray:
        n = self.nn if name in NODAL_FIELDS else self.num_particles
        out = np.empty(max(n, 1), np.float64)
        native.check(native.lib().amsc_sim_get_field(self._h, name.encode(), out,
                                                     len(out)))
        return out[:n]

    def get_int(self, name: str) -> np.ndarray:
        n = self.nc + 1 if name == "cstart" else self.num_particles
        out = np.empty(max(n, 1), np.int32)
        native.check(native.lib().amsc_sim_get_field_int(self._h, name.encode(),
                                                         out, len(out)))
        return out[:n]

    def set(self, name: str, arr):
        a = np.ascontiguousarray(arr, np.float64)
        native.check(native.lib().amsc_sim_set_field(self._h, name.encode(), a,
                                                     len(a)))

    def conservation(self) -> np.ndarray:
        out = np.zeros(3, np.float64)
        native.check(native.lib().amsc_sim_conservation(self._h, out))
        return out

    def params(self) -> dict:
        """Solver parameters (scenario 0 of a batch) as a dict."""
        p = native.AmscParams()
        native.check(native.lib().amsc_sim_params(self._h, C.byref(p)))
        return {name: getattr(p, name) for name, _ in p._fields_}

    def write_particles_csv(self, path: str):
        """nc_particles CSV of the current state (quoted header, columns by
        name; output-only fields when the last step produced them)."""
        native.check(native.lib().amsc_sim_write_particles_csv(self._h,
                                                               str(path).encode()))

    def write_grid_vts(self, path: str):
        """VTK XML structured grid of Z and the nodal fields of the last step."""
        native.check(native.lib().amsc_sim_write_grid_vts(self._h,
                                                          str(path).encode()))

    @property
    def particle_steps(self) -> int:
        out = C.c_longlong()
        native.check(native.lib().amsc_sim_particle_steps(self._h, C.byref(out)))
        return out.value

    def active_nodes(self) -> int:
        """Nodes that are corners of occupied cells (current binning)."""
        cs = self.get_int("cstart")
        occ = np.nonzero(np.diff(cs) > 0)[0]
        nrows = self.prm["nrows"]
        r, c = occ % nrows, occ // nrows
        nodes = np.concatenate([r + c * (nrows + 1), r + 1 + c * (nrows + 1),
                                r + (c + 1) * (nrows + 1), r + 1 + (c + 1) * (nrows + 1)])
        return int(np.unique(nodes).size)

    def ms_counts(self) -> np.ndarray:
        out = np.zeros(2, np.int32)
        native.check(native.lib().amsc_sim_ms_counts(self._h, out))
        return out

    def set_graph(self, on: bool = True):
        """Replay each step variant from a HIP graph recorded once."""
        native.check(native.lib().amsc_sim_set_graph(self._h, int(bool(on))))

    def set_tiling(self, p2g_shape: int, particle_block: int):
        native.check(native.lib().amsc_sim_set_tiling(self._h, int(p2g_shape),
                                                      int(particle_block)))

    def autotune(self, steps_per_variant: int = 2, report: dict | None = None):
        """Pick the fastest P2G variant and particle workgroup size on live
        steps; returns (variant, block, steps used).  `report` (optional)
        receives the average ms of every variant."""
        a, b, u = C.c_int(), C.c_int(), C.c_int()
        ms = (C.c_double * 8)()
        native.check(native.lib().amsc_sim_autotune(self._h, int(steps_per_variant),
                                                    C.byref(a), C.byref(b),
                                                    C.byref(u), ms))
        if report is not None:
            report.update({f"p2g_{v}": round(ms[v], 5) for v in range(6)})
            report.update({"particle_256": round(ms[6], 5),
                           "particle_512": round(ms[7], 5)})
        return a.value, b.value, u.value

    def time_p2g(self, reps: int = 20) -> float:
        """Average ms of `reps` back-to-back P2G launches on the current
        state."""
        out = C.c_double()
        native.check(native.lib().amsc_sim_time_p2g(self._h, int(reps), C.byref(out)))
        return out.value


class Ensemble(Simulation):
    """A batch of independent scenarios advanced together (amsc_ens_*).

    scenarios: list of (prm, particles, Z) as for Simulation.  precision
    "fp64" (bit-identical to running each scenario alone) or "mixed".
    """

    def __init__(self, scenarios, precision: str = "fp64", capacity: int = 0):
        L = native.lib()
        self.nscen = len(scenarios)
        self.prms = [dict(p) for p, _, _ in scenarios]
        self.prm = self.prms[0]
        self.precision = precision
        P = (native.AmscParams * self.nscen)(
            *[native.AmscParams.from_dict(p) for p in self.prms])
        nps = np.array([len(part["x"]) for _, part, _ in scenarios], np.int32)
        cat = [np.ascontiguousarray(np.concatenate(
            [np.asarray(part[k], np.float64) for _, part, _ in scenarios]))
            for k in ("x", "y", "Mp", "Ap", "vpx", "vpy", "mom_px", "mom_py",
                      "hp", "Vp")]
        Z = np.ascontiguousarray(np.concatenate(
            [np.asarray(z, np.float64).ravel() for _, _, z in scenarios]))
        h = C.c_void_p()
        native.check(L.amsc_ens_create(self.nscen, P, nps, *cat, Z,
                                       PRECISION[precision], int(capacity),
                                       C.byref(h)))
        self._h = h
        self.nns = [(p["nrows"] + 1) * (p["ncols"] + 1) for p in self.prms]
        self.ncs = [p["nrows"] * p["ncols"] for p in self.prms]
        self.nn, self.nc = self.nns[0], self.ncs[0]

    @classmethod
    def from_data(cls, ds, precision: str = "fp64", capacity: int = 0,
                  max_steps: int = 0):
        sc = []
        for d in ds:
            prm = params_from_data(d, max_steps=max_steps or int(d.get("MAX_STEPS", 0)))
            sc.append((prm, particles_from_data(d), d["Z"]))
        return cls(sc, precision, capacity)

    def status(self) -> dict:
        n = self.nscen
        it, act, npart = (np.zeros(n, np.int32) for _ in range(3))
        t, dt = np.zeros(n), np.zeros(n)
        native.check(native.lib().amsc_ens_status(self._h, it, t, dt, act, npart))
        div = np.zeros(n, np.int32)
        native.check(native.lib().amsc_ens_diverged(self._h, div))
        return {"it": it, "t": t, "dt": dt, "active": act, "np": npart,
                "diverged": div}

    def num(self, scen: int) -> int:
        return int(self.status()["np"][scen])

    def get(self, scen: int, name: str) -> np.ndarray:
        n = self.nns[scen] if name in NODAL_FIELDS else self.num(scen)
        out = np.empty(max(n, 1), np.float64)
        native.check(native.lib().amsc_ens_get_field(self._h, int(scen),
                                                     name.encode(), out, len(out)))
        return out[:n]

    def get_int(self, scen: int, name: str) -> np.ndarray:
        n = self.ncs[scen] + 1 if name == "cstart" else self.num(scen)
        out = np.empty(max(n, 1), np.int32)
        native.check(native.lib().amsc_ens_get_field_int(self._h, int(scen),
                                                         name.encode(), out, len(out)))
        return out[:n]

    def set(self, scen: int, name: str, arr):
        a = np.ascontiguousarray(arr, np.float64)
        native.check(native.lib().amsc_ens_set_field(self._h, int(scen),
                                                     name.encode(), a, len(a)))

    def conservation(self, scen: int = 0) -> np.ndarray:
        out = np.zeros(3, np.float64)
        native.check(native.lib().amsc_ens_conservation(self._h, int(scen), out))
        return out

    def ms_counts(self, scen: int = 0) -> np.ndarray:
        out = np.zeros(2, np.int32)
        native.check(native.lib().amsc_ens_ms_counts(self._h, int(scen), out))
        return out

    def info(self) -> dict:
        n, p, b, c = C.c_int(), C.c_int(), C.c_longlong(), C.c_int()
        native.check(native.lib().amsc_ens_info(self._h, C.byref(n), C.byref(p),
                                                C.byref(b), C.byref(c)))
        return {"nscen": n.value, "precision": p.value, "bytes": b.value,
                "capacity": c.value}

    @staticmethod
    def bytes_estimate(prms, nps, precision: str = "fp64", capacity: int = 0) -> int:
        P = (native.AmscParams * len(prms))(*[native.AmscParams.from_dict(p) for p in prms])
        v = native.lib().amsc_ens_bytes_estimate(len(prms), P,
                                                 np.asarray(nps, np.int32),
                                                 PRECISION[precision], int(capacity))
        if v < 0:
            raise ValueError("bad arguments")
        return int(v)
