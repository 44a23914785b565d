"""Seeded small inputs for the 18 step kernels of gpu_kernels.h, shared by
the oracle-vs-reference tests (CPU) and the HIP drop-in parity tests (GPU).

Argument names follow main/gpu_kernels.h; ARGS[name] lists them in order,
KINDS (same letters as oracle.SIGNATURES) says which are scalars, device
int arrays, read-only / in-out double arrays, the host coloff array, or the
diagnostic pointers the reference never writes.
"""
from __future__ import annotations

import math

import numpy as np

ARGS = {
    "reset_nodal_arrays": ["nn", "Mv", "mom_vx", "mom_vy", "F_ext_vx", "F_ext_vy",
                           "F_int_vx", "F_int_vy", "Ftot_vx", "Ftot_vy", "avx",
                           "avy", "vvx", "vvy", "vvxL", "vvyL", "Fric_x",
                           "Fric_y", "FPxv", "FPyv", "HV", "pv_rho", "pv_vvx",
                           "pv_vvy", "pv_avx", "pv_avy"],
    "scatter_mass_momentum": ["np", "nrows", "hx", "hy", "cstart", "cptcls",
                              "ccidx", "coloff", "x", "y", "Mp", "mom_px",
                              "mom_py", "Mv", "mom_vx", "mom_vy"],
    "gradZ_to_particles": ["np", "nrows", "hx", "hy", "x", "y", "p2g", "Z",
                           "dZxp", "dZyp"],
    "gravity_force": ["np", "g", "hp", "H", "Mp", "dZxp", "dZyp", "Fpx", "Fpy"],
    "friction_particle_force": ["np", "Ap", "Fb_x", "Fb_y", "Fric_px", "Fric_py"],
    "scatter_gravity_force": ["np", "nrows", "hx", "hy", "cstart", "cptcls",
                              "ccidx", "coloff", "x", "y", "Fpx", "Fpy",
                              "FPxv", "FPyv"],
    "scatter_friction_force": ["np", "nrows", "hx", "hy", "cstart", "cptcls",
                               "ccidx", "coloff", "x", "y", "Fric_px",
                               "Fric_py", "Fric_x", "Fric_y"],
    "compute_F_ext": ["nn", "FPxv", "FPyv", "Fric_x", "Fric_y", "F_ext_vx",
                      "F_ext_vy"],
    "scatter_internal_forces": ["np", "nrows", "hx", "hy", "cstart", "cptcls",
                                "ccidx", "coloff", "x", "y", "Vp", "F11", "F12",
                                "F21", "F22", "F_int_vx", "F_int_vy"],
    "nodal_force_and_momentum": ["nn", "dt", "F_ext_vx", "F_ext_vy", "F_int_vx",
                                 "F_int_vy", "Ftot_vx", "Ftot_vy", "mom_vx",
                                 "mom_vy"],
    "nodal_velocity_and_bc": ["nn", "nrows", "ncols", "dt", "bc_flag", "Mv",
                              "Ftot_vx", "Ftot_vy", "mom_vx", "mom_vy", "avx",
                              "avy", "vvx", "vvy", "vvxL", "vvyL"],
    "g2p_velocity_and_advect": ["np", "nrows", "ncols", "hx", "hy", "dt", "p2g",
                                "vvxL", "vvyL", "avx", "avy", "x", "y", "vpx",
                                "vpy", "apx", "apy", "mom_px", "mom_py", "vpxL",
                                "vpyL", "Mp"],
    "g2pd_gradients_and_height_update": ["np", "nrows", "hx", "hy", "dt", "x",
                                         "y", "p2g", "vvxL", "vvyL", "vpx_dx",
                                         "vpx_dy", "vpy_dx", "vpy_dy", "hp",
                                         "Vp", "Ap", "min_sc_out", "min_hp_out"],
    "voellmy_friction": ["np", "rho", "g", "tan_fa", "xi", "fric_on", "vpx",
                         "vpy", "hp", "Fb_x", "Fb_y"],
    "bingham_stress_update": ["np", "rho", "g", "mu", "tau_Y", "cc", "A_coeff",
                              "C_coeff", "vpx", "vpy", "hp", "H", "vpx_dx",
                              "vpx_dy", "vpy_dx", "vpy_dy", "F11", "F12", "F21",
                              "F22", "max_coeff_out", "min_hp_out",
                              "max_nv_out", "max_F11_out", "max_dev_out",
                              "max_pres_out", "max_Dxx_out"],
    "g2p_height": ["np", "nrows", "hx", "hy", "x", "y", "p2g", "Z", "hp", "Zp",
                   "hpZ"],
    "scatter_plotvars": ["np", "nrows", "hx", "hy", "cstart", "cptcls", "ccidx",
                         "coloff", "x", "y", "Mp", "vpx", "vpy", "apx", "apy",
                         "hp", "pv_rho", "pv_vvx", "pv_vvy", "pv_avx", "pv_avy",
                         "HV"],
    "normalize_plotvars": ["nn", "pv_rho", "pv_vvx", "pv_vvy", "pv_avx", "pv_avy"],
}

INT_ARRAYS = {"cstart", "cptcls", "ccidx", "p2g"}
SCALARS_INT = {"np", "nn", "nrows", "ncols", "bc_flag"}

NODAL = {"Mv", "mom_vx", "mom_vy", "F_ext_vx", "F_ext_vy", "F_int_vx",
         "F_int_vy", "Ftot_vx", "Ftot_vy", "avx", "avy", "vvx", "vvy", "vvxL",
         "vvyL", "Fric_x", "Fric_y", "FPxv", "FPyv", "HV", "pv_rho", "pv_vvx",
         "pv_vvy", "pv_avx", "pv_avy", "Z"}
PARTICLE = {"x", "y", "Mp", "mom_px", "mom_py", "dZxp", "dZyp", "hp", "H",
            "Fpx", "Fpy", "Ap", "Fb_x", "Fb_y", "Fric_px", "Fric_py", "Vp",
            "F11", "F12", "F21", "F22", "vpx", "vpy", "apx", "apy", "vpxL",
            "vpyL", "vpx_dx", "vpx_dy", "vpy_dx", "vpy_dy", "Zp", "hpZ"}


def make_case(seed=0, nrows=7, ncols=9, npart=180, hx=0.7, hy=0.45):
    """Random state on an nrows x ncols grid with branch-covering values:
    particles on cell edges and domain edges, hp around the 1e-10/1e-3/1e-2
    thresholds, nodes around the Mv > 1e-2 threshold, zero velocities."""
    from oracle import oracle  # the checker bins the particles
    rng = np.random.default_rng(seed)
    Lx, Ly = ncols * hx, nrows * hy
    x = rng.uniform(0, Lx, npart)
    y = rng.uniform(0, Ly, npart)
    # exact cell edges, domain edges, near-edge values
    k = npart // 10
    x[:k] = rng.integers(0, ncols + 1, k) * hx
    y[k:2 * k] = rng.integers(0, nrows + 1, k) * hy
    x[2 * k] = 1e-10
    y[2 * k] = Ly - 1e-10
    x[2 * k + 1] = Lx - 1e-10
    x = np.clip(x, 1e-10, Lx - 1e-10)
    y = np.clip(y, 1e-10, Ly - 1e-10)
    p2g, cstart, cptcls, ccidx, coloff = oracle.bin_particles(x, y, nrows, ncols, hx, hy)
    nn = (nrows + 1) * (ncols + 1)
    c = {"np": npart, "nn": nn, "nrows": nrows, "ncols": ncols, "hx": hx, "hy": hy,
         "x": x, "y": y, "p2g": p2g, "cstart": cstart, "cptcls": cptcls,
         "ccidx": ccidx, "coloff": coloff}
    for name in PARTICLE - {"x", "y"}:
        c[name] = rng.normal(0, 2.0, npart)
    hp = rng.uniform(0.0, 3.0, npart)
    hp[3 * k:3 * k + 5] = [0.0, 1e-11, 5e-4, 1e-3, 1e-2]
    c["hp"] = hp
    c["H"] = np.where(rng.random(npart) < 0.5, 0.0, rng.uniform(0, 2, npart))
    c["Mp"] = rng.uniform(0.5, 5.0, npart)
    c["Vp"] = rng.uniform(0.1, 2.0, npart)
    c["Ap"] = rng.uniform(0.05, 1.0, npart)
    c["F21"] = c["F12"].copy()
    vz = rng.random(npart) < 0.1
    c["vpx"][vz] = 0.0
    c["vpy"][vz] = 0.0
    c["vpx"][:3] = [1e-12, 0.0, -3.0]
    for name in NODAL:
        c[name] = rng.normal(0, 1.0, nn)
    Mv = rng.uniform(0.0, 3.0, nn)
    Mv[:6] = [0.0, 1e-3, 1e-2, 1.0000001e-2, 5e-9, 2e-8]
    c["Mv"] = Mv
    c["pv_rho"] = Mv.copy()
    c["Z"] = 3.0 + 0.4 * np.repeat(np.arange(ncols + 1) * hx, nrows + 1) + \
        rng.normal(0, 0.1, nn)
    # scalars
    c.update(dict(g=9.81, dt=3.7e-3, bc_flag=1, rho=1800.0, mu=0.35, tau_Y=42.0,
                  cc=1.0, A_coeff=1.5, C_coeff=65.0 / 32.0,
                  tan_fa=math.tan(17.0 * math.pi / 180.0), xi=450.0, fric_on=1.0))
    for d in ("min_sc_out", "min_hp_out", "max_coeff_out", "max_nv_out",
              "max_F11_out", "max_dev_out", "max_pres_out", "max_Dxx_out"):
        c[d] = np.full(1, -7.0)
    return c


def kinds(name):
    from oracle.oracle import SIGNATURES
    return SIGNATURES[name]


def output_names(name):
    return [a for a, k in zip(ARGS[name], kinds(name)) if k == "O"
            and not a.endswith("_out")]


def host_args(name, case, override=None):
    """Argument list for the oracle / reference host builds (copies of the
    arrays, so outputs do not leak between calls).  Returns (args, outputs)."""
    override = override or {}
    args, outs = [], {}
    for a, k in zip(ARGS[name], kinds(name)):
        v = override.get(a, case[a])
        if k in "idb":
            args.append(v)
        elif k in "IH":
            args.append(np.ascontiguousarray(v, np.int32).copy())
        else:
            arr = np.ascontiguousarray(v, np.float64).copy()
            args.append(arr)
            if k == "O":
                outs[a] = arr
    return args, outs
