"""ctypes binding of the C ABI in include/amsc_mpm.h (libamsc_mpm.so).

There is no fallback: if the gfx950 library is not built, `lib()` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libamsc_mpm.so")

AMSC_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP error", 3: "out of memory",
          4: "capacity exceeded", 5: "stale field"}


class AmscError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"amsc error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class AmscParams(C.Structure):
    """amsc_params (include/amsc_mpm.h)."""

    _fields_ = [
        ("nrows", C.c_int), ("ncols", C.c_int),
        ("hx", C.c_double), ("hy", C.c_double),
        ("g", C.c_double), ("rho", C.c_double), ("mu", C.c_double),
        ("tauy", C.c_double), ("xi", C.c_double), ("phi", C.c_double),
        ("bingham_on", C.c_double), ("friction_on", C.c_double),
        ("cfl", C.c_double), ("eq_level", C.c_double),
        ("dt_fixed", C.c_double), ("T", C.c_double),
        ("bc_flag", C.c_int), ("nsteps", C.c_int), ("max_steps", C.c_int),
        ("ms_on", C.c_int),
        ("ms_alpha", C.c_double), ("ms_beta", C.c_double),
        ("ms_split_hp_min", C.c_double), ("ms_shear_split", C.c_double),
        ("ms_hp_min", C.c_double), ("ms_max_dv", C.c_double),
        ("ms_max_ops", C.c_double),
        ("ms_min_level", C.c_int), ("ms_max_level", C.c_int),
        ("ms_min_ppc", C.c_int), ("ms_interval", C.c_int),
        ("ms_mode", C.c_int), ("ms_stretch", C.c_double),
    ]

    @classmethod
    def from_dict(cls, prm: dict):
        s = cls()
        for name, typ in cls._fields_:
            setattr(s, name, typ(prm[name]).value)
        return s


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Nullable:
    """An ndpointer argument that also accepts None (NULL)."""

    def __init__(self, base):
        self.base = base

    def from_param(self, obj):
        return None if obj is None else self.base.from_param(obj)


_dpn, _ipn = _Nullable(_dp), _Nullable(_ip)
_vp = C.c_void_p

# The 18 step kernels: argument kinds in gpu_kernels.h order.
#   i int, d double, b bool (passed as int), I/P/O device pointer,
#   H host int[5] (coloff), X ignored host diagnostic pointer.
DROPIN = {
    "reset_nodal_arrays": "i" + "O" * 25,
    "scatter_mass_momentum": "iiddIIIH" + "PP" + "PPP" + "OOO",
    "gradZ_to_particles": "iidd" + "PPI" + "P" + "OO",
    "gravity_force": "id" + "PPPPP" + "OO",
    "friction_particle_force": "i" + "PPP" + "OO",
    "scatter_gravity_force": "iiddIIIH" + "PP" + "PP" + "OO",
    "scatter_friction_force": "iiddIIIH" + "PP" + "PP" + "OO",
    "compute_F_ext": "i" + "PPPP" + "OO",
    "scatter_internal_forces": "iiddIIIH" + "PPP" + "PPPP" + "OO",
    "nodal_force_and_momentum": "id" + "PPPP" + "OOOO",
    "nodal_velocity_and_bc": "iiidb" + "PPP" + "OOOOOOOO",
    "g2p_velocity_and_advect": "iiiddd" + "I" + "PPPP" + "OOOOOOOOOO" + "P",
    "g2pd_gradients_and_height_update": "iiddd" + "PPI" + "PP" + "OOOOOOO" + "XX",
    "voellmy_friction": "iddddd" + "PPP" + "OO",
    "bingham_stress_update": "iddddddd" + "PPPPPPPP" + "OOOO" + "XXXXXXX",
    "g2p_height": "iidd" + "PPI" + "PP" + "OO",
    "scatter_plotvars": "iiddIIIH" + "PP" + "PPPPPP" + "OOOOOO",
    "normalize_plotvars": "i" + "OOOOO",
}
_KIND = {"i": C.c_int, "d": C.c_double, "b": C.c_int, "I": _vp, "P": _vp,
         "O": _vp, "H": _ip, "X": _vp}

_lib = None


def lib():
    """The loaded gfx950 library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -m amsc_amd.build or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.amsc_last_error.restype = C.c_char_p
    L.amsc_set_stream.argtypes = [_vp]
    for name, sig in DROPIN.items():
        fn = getattr(L, "amsc_" + name)
        fn.argtypes = [_KIND[k] for k in sig]
        fn.restype = C.c_int
    L.amsc_bin_particles.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _ip]
    L.amsc_bin_particles_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _ip]
    L.amsc_ms_decide.argtypes = [C.POINTER(AmscParams), C.c_int, C.c_double] + \
        [_vp] * 15 + [_ip]
    L.amsc_cfl_dt.argtypes = [C.c_int, _vp, _vp, _vp, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_int,
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.amsc_ms_execute.argtypes = [C.POINTER(AmscParams), C.c_int, _vp, _vp, C.c_int,
                                  C.POINTER(C.c_char_p), C.POINTER(_vp),
                                  C.POINTER(_vp), _vp, _vp, _vp, _vp, C.c_int,
                                  C.POINTER(C.c_int)]
    L.amsc_sim_create.argtypes = [C.POINTER(AmscParams), C.c_int] + [_dp] * 11 + \
        [C.c_int, C.POINTER(_vp)]
    L.amsc_sim_create_from_json.argtypes = [C.c_char_p, C.c_int, C.c_int,
                                            C.POINTER(_vp)]
    L.amsc_sim_destroy.argtypes = [_vp]
    L.amsc_sim_step.argtypes = [_vp, C.c_int, C.POINTER(C.c_int)]
    L.amsc_sim_set_aux.argtypes = [_vp, C.c_int]
    L.amsc_sim_set_timing.argtypes = [_vp, C.c_int]
    L.amsc_sim_kernel_times.argtypes = [_vp, C.POINTER(C.c_char_p),
                                        C.POINTER(C.c_double),
                                        C.POINTER(C.c_int), C.c_int,
                                        C.POINTER(C.c_int)]
    L.amsc_sim_reset_kernel_times.argtypes = [_vp]
    L.amsc_sim_num_particles.argtypes = [_vp, C.POINTER(C.c_int)]
    L.amsc_sim_state.argtypes = [_vp, C.POINTER(C.c_int),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.amsc_sim_get_field.argtypes = [_vp, C.c_char_p, _dp, C.c_int]
    L.amsc_sim_get_field_int.argtypes = [_vp, C.c_char_p, _ip, C.c_int]
    L.amsc_sim_set_field.argtypes = [_vp, C.c_char_p, _dp, C.c_int]
    L.amsc_sim_conservation.argtypes = [_vp, _dp]
    L.amsc_sim_params.argtypes = [_vp, C.POINTER(AmscParams)]
    L.amsc_sim_write_particles_csv.argtypes = [_vp, C.c_char_p]
    L.amsc_sim_write_grid_vts.argtypes = [_vp, C.c_char_p]
    L.amsc_sim_ms_counts.argtypes = [_vp, _ip]
    L.amsc_sim_particle_steps.argtypes = [_vp, C.POINTER(C.c_longlong)]
    L.amsc_device_alloc.argtypes = [C.c_size_t, C.POINTER(_vp)]
    L.amsc_device_free.argtypes = [_vp]
    L.amsc_copy_to_device.argtypes = [_vp, _vp, C.c_size_t]
    L.amsc_copy_to_host.argtypes = [_vp, _vp, C.c_size_t]
    L.amsc_copy_on_device.argtypes = [_vp, _vp, C.c_size_t]
    L.amsc_sim_set_graph.argtypes = [_vp, C.c_int]
    L.amsc_sim_set_tiling.argtypes = [_vp, C.c_int, C.c_int]
    L.amsc_sim_autotune.argtypes = [_vp, C.c_int, C.POINTER(C.c_int),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]
    L.amsc_sim_time_p2g.argtypes = [_vp, C.c_int, C.POINTER(C.c_double)]
    L.amsc_ens_repack.argtypes = [_vp, C.c_int, _ipn, C.c_int,
                                  C.POINTER(AmscParams), _ipn] + [_dpn] * 11 + \
        [C.c_longlong, C.POINTER(_vp)]
    L.amsc_ens_create.argtypes = [C.c_int, C.POINTER(AmscParams), _ip] + \
        [_dp] * 11 + [C.c_int, C.c_longlong, C.POINTER(_vp)]
    L.amsc_ens_status.argtypes = [_vp, _ip, _dp, _dp, _ip, _ip]
    L.amsc_ens_diverged.argtypes = [_vp, _ip]
    L.amsc_ens_get_field.argtypes = [_vp, C.c_int, C.c_char_p, _dp, C.c_int]
    L.amsc_ens_get_field_int.argtypes = [_vp, C.c_int, C.c_char_p, _ip, C.c_int]
    L.amsc_ens_set_field.argtypes = [_vp, C.c_int, C.c_char_p, _dp, C.c_int]
    L.amsc_ens_conservation.argtypes = [_vp, C.c_int, _dp]
    L.amsc_ens_ms_counts.argtypes = [_vp, C.c_int, _ip]
    L.amsc_ens_info.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
    L.amsc_ens_bytes_estimate.argtypes = [C.c_int, C.POINTER(AmscParams), _ip,
                                          C.c_int, C.c_longlong]
    L.amsc_ens_bytes_estimate.restype = C.c_longlong
    _lib = L
    return L


def check(rc: int):
    if rc != AMSC_OK:
        raise AmscError(rc, lib().amsc_last_error().decode(errors="replace"))


def dropin(name):
    """Callable for amsc_<name>; raises AmscError on a non-zero status."""
    fn = getattr(lib(), "amsc_" + name)

    def call(*args):
        check(fn(*args))
    return call


def bin_particles(np_, nrows, ncols, hx, hy, d_x, d_y, d_p2g, d_cstart,
                  d_cptcls, d_ccidx=None, noalloc=False):
    """amsc_bin_particles (noalloc: amsc_bin_particles_noalloc); returns
    coloff."""
    coloff = np.zeros(5, np.int32)
    fn = lib().amsc_bin_particles_noalloc if noalloc else lib().amsc_bin_particles
    check(fn(np_, nrows, ncols, hx, hy, d_x, d_y, d_p2g, d_cstart, d_cptcls,
             d_ccidx, coloff))
    return coloff


def ms_decide(prm: dict, np_, dt, ptrs, d_level, d_cstart, d_cptcls,
              d_action, d_partner):
    """ptrs: device pointers of x y Ap hp vpx_dx vpx_dy vpy_dx vpy_dy vpx vpy."""
    counts = np.zeros(2, np.int32)
    P = AmscParams.from_dict(prm)
    check(lib().amsc_ms_decide(C.byref(P), np_, dt, *ptrs, d_level, d_cstart,
                               d_cptcls, d_action, d_partner, counts))
    return counts


def cfl_dt(np_, d_vpx, d_vpy, d_hp, g, cfl, hx, hy, it, dt):
    """amsc_cfl_dt: (dt after the call, cel)."""
    d = C.c_double(dt)
    cel = C.c_double(0.0)
    check(lib().amsc_cfl_dt(np_, d_vpx, d_vpy, d_hp, g, cfl, hx, hy, it,
                            C.byref(d), C.byref(cel)))
    return d.value, cel.value


def ms_execute(prm: dict, np_, d_action, d_partner, names, d_in, d_out,
               d_label_in, d_level_in, d_label_out, d_level_out, capacity):
    """amsc_ms_execute; names/d_in/d_out: lists (field names, device
    pointers).  Returns the new particle count."""
    nf = len(names)
    cn = (C.c_char_p * nf)(*[n.encode() for n in names])
    pin = (_vp * nf)(*d_in)
    pout = (_vp * nf)(*d_out)
    nnew = C.c_int(0)
    P = AmscParams.from_dict(prm)
    check(lib().amsc_ms_execute(C.byref(P), np_, d_action, d_partner, nf, cn, pin,
                                pout, d_label_in, d_level_in, d_label_out,
                                d_level_out, capacity, C.byref(nnew)))
    return nnew.value


class DeviceArray:
    """A numpy array's copy in device memory (library runtime), for callers
    of the kernel-level entry points without another device allocator."""

    def __init__(self, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.dtype, self.shape = a.dtype, a.shape
        self.nbytes = a.nbytes
        p = C.c_void_p()
        check(lib().amsc_device_alloc(max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value
        if self.nbytes:
            check(lib().amsc_copy_to_device(self.ptr, a.ctypes.data, self.nbytes))

    @classmethod
    def zeros(cls, n, dtype):
        return cls(np.zeros(n, dtype))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().amsc_copy_to_host(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().amsc_device_free(self.ptr)
            self.ptr = None

    __del__ = free
