"""Host-side driver over the C ABI: the role of main/gpu_offload.cpp (load
DATA, run the time loop, read fields back), with all state resident on the
GPU between calls.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .data import load_data_json, params_from_data, particles_from_data

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
