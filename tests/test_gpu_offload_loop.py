"""gpu_offload's time loop kept, its kernels swapped: amsc_offload_loop runs
the reference driver's step (main/gpu_offload.cpp:410-562: K1-K10 and the
CSR rebuild, fixed or adaptive dt, merge/split off) through the C++ drop-in header
include/gpu_kernels.h on the device.  The same loop restated here over the
oracle's kernels (each pinned bit for bit to the reference's compiled
main/gpu_kernels.cpp) and, where that build is present, over the
reference's own compiled kernels, must give the same bits for every
particle and nodal array and the same CSR after 25 steps of a slope with
friction and Bingham stress, and 40 steps of a dam break whose particles
change cells."""
import math
import os
import subprocess

import numpy as np
import pytest

from amsc_amd import offload_input, scenarios
from amsc_amd.data import params_from_data, particles_from_data
from kernel_cases import ARGS, host_args
from oracle import oracle
from test_oracle_step import small_slope_case

pytestmark = pytest.mark.gpu

LOOP = os.path.join(os.path.dirname(__file__), "..", "amsc_amd", "_lib",
                    "amsc_offload_loop")

# particle arrays in the driver's order (kernel argument names) and the
# oracle simulation's names for them
PARTICLE = offload_input.PARTICLE
ORACLE_NAME = {"F11": "F_11", "F12": "F_12", "F21": "F_21", "F22": "F_22"}
NODAL = ["Mv", "mom_vx", "mom_vy", "F_ext_vx", "F_ext_vy", "F_int_vx",
         "F_int_vy", "Ftot_vx", "Ftot_vy", "avx", "avy", "vvx", "vvy", "vvxL",
         "vvyL", "Fric_x", "Fric_y", "FPxv", "FPyv", "HV", "pv_rho", "pv_vvx",
         "pv_vvy", "pv_avx", "pv_avy"]
STEP = ["reset_nodal_arrays", "scatter_mass_momentum", "gradZ_to_particles",
        "gravity_force", "friction_particle_force", "scatter_gravity_force",
        "scatter_friction_force", "compute_F_ext", "scatter_internal_forces",
        "nodal_force_and_momentum", "nodal_velocity_and_bc",
        "g2p_velocity_and_advect", "g2pd_gradients_and_height_update",
        "voellmy_friction", "bingham_stress_update", "g2p_height",
        "scatter_plotvars", "normalize_plotvars"]
assert sorted(STEP) == sorted(ARGS)


def initial_state(d, dt):
    """The step-0 state of a case (the oracle's initialisation)."""
    prm = params_from_data(d)
    o = oracle.OracleSim(prm, particles_from_data(d), d["Z"])
    st = {k: o.get(ORACLE_NAME.get(k, k)) for k in PARTICLE}
    nrows, ncols = prm["nrows"], prm["ncols"]
    nn = (nrows + 1) * (ncols + 1)
    st.update({k: np.zeros(nn) for k in NODAL})
    st["Z"] = np.asarray(d["Z"], np.float64)
    st.update(np=len(st["x"]), nn=nn, nrows=nrows, ncols=ncols, bc_flag=1)
    st.update(offload_input.scalars(prm, dt))
    for k in ("min_sc_out", "min_hp_out", "max_coeff_out", "max_nv_out",
              "max_F11_out", "max_dev_out", "max_pres_out", "max_Dxx_out"):
        st[k] = np.zeros(1)
    return st


def rebin(st):
    p2g, cstart, cptcls, ccidx, coloff = oracle.bin_particles(
        st["x"], st["y"], st["nrows"], st["ncols"], st["hx"], st["hy"])
    st.update(p2g=p2g, cstart=cstart, cptcls=cptcls, ccidx=ccidx, coloff=coloff)


def cfl_dt(st, it):
    """gpu_offload.cpp:282-301: dt from the particle maxima (it > 0)."""
    hmax = st["hp"].max()
    sq = math.sqrt(st["g"] * hmax)
    cel = abs(max(sq + np.abs(st["vpx"]).max(), sq + np.abs(st["vpy"]).max()))
    if it > 0:
        st["dt"] = st["cfl"] * min(st["hx"], st["hy"]) / (1e-2 + cel)


def host_loop(st, nsteps, kernel, adaptive=False):
    rebin(st)
    if adaptive:
        st["dt"] = 1e-3  # gpu_offload.cpp:255
    for it in range(nsteps):
        if adaptive:
            cfl_dt(st, it)
        for name in STEP:
            args, outs = host_args(name, st)
            kernel(name)(*args)
            st.update(outs)
        rebin(st)
    return st


def device_loop(st, nsteps, tmp_path, enqueue_only=False):
    np_, nn, nc = st["np"], st["nn"], st["nrows"] * st["ncols"]
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    offload_input.write(src, st, nsteps)
    cmd = [os.path.abspath(LOOP), str(src), str(dst)]
    r = subprocess.run(cmd + (["--async"] if enqueue_only else []),
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = dst.read_bytes()
    nf = len(PARTICLE) * np_ + len(NODAL) * nn
    f = np.frombuffer(raw[:8 * nf], np.float64)
    i = np.frombuffer(raw[8 * nf:], np.int32)
    out = {}
    for j, k in enumerate(PARTICLE):
        out[k] = f[j * np_:(j + 1) * np_]
    base = len(PARTICLE) * np_
    for j, k in enumerate(NODAL):
        out[k] = f[base + j * nn:base + (j + 1) * nn]
    out["p2g"] = i[:np_]
    out["cstart"] = i[np_:np_ + nc + 1]
    out["cptcls"] = i[np_ + nc + 1:]
    return out


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


CASES = {
    # Voellmy friction and Bingham stress on, slow flow
    "slope": (small_slope_case, 25, 2e-3),
    # channel dam break (dambreak.m): fast enough that particles change cells
    "dambreak": (scenarios.dambreak_case, 40, 1e-3),
    # the same dam break with gpu_offload's adaptive dt (dt <= 0 selects it)
    "dambreak_adaptive": (scenarios.dambreak_case, 40, 0.0),
}


@pytest.mark.parametrize("checker", ["oracle", "reference"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_offload_loop_through_the_dropin_header(case, checker, tmp_path):
    if checker == "reference" and not oracle.ref_available():
        pytest.skip("the reference's gpu_kernels.cpp host build is not present")
    print(f"[offload loop] {case} {checker}", flush=True)
    make, nsteps, dt = CASES[case]
    d = make()
    st = initial_state(d, dt)
    p0 = oracle.bin_particles(st["x"], st["y"], st["nrows"], st["ncols"],
                              st["hx"], st["hy"])[0]
    got = device_loop(st, nsteps, tmp_path)
    kern = oracle.oracle_kernel if checker == "oracle" else oracle.ref_kernel
    want = host_loop(st, nsteps, kern, adaptive=dt <= 0.0)
    for k in PARTICLE + NODAL:
        assert np.array_equal(bits(got[k]), bits(want[k])), k
    for k in ("p2g", "cstart", "cptcls"):
        assert np.array_equal(got[k], np.asarray(want[k], np.int32)), k
    assert not np.isnan(got["x"]).any()
    if case.startswith("dambreak"):
        assert np.count_nonzero(got["p2g"] != p0) > 100  # particles changed cells
    else:
        assert np.any(got["Fric_x"] != 0.0)  # friction acted


def test_input_writer_state_matches_the_oracle(tmp_path):
    """amsc_amd.offload_input (which writes the loop's input from the device
    library's step-0 initialisation, for timing and profiling the loop)
    starts from the oracle's step-0 state in every array the loop reads
    before writing it, and the loop then gives the oracle's bits."""
    print("[offload input] slope", flush=True)
    d = small_slope_case()
    dev = offload_input.initial_state(d, 1e-3)
    ref = initial_state(d, 1e-3)
    read_first = ["x", "y", "Mp", "Vp", "Ap", "hp", "H", "vpx", "vpy", "mom_px",
                  "mom_py", "Fb_x", "Fb_y", "F11", "F12", "F21", "F22",
                  "Fric_px", "Fric_py"]
    for k in read_first:
        assert np.array_equal(bits(dev[k]), bits(ref[k])), k
    for k in offload_input.SCALARS + ["nrows", "ncols", "bc_flag"]:
        assert dev[k] == ref[k], k
    ref.update({k: dev[k] for k in PARTICLE if k not in read_first})
    dev.update(np=ref["np"], nn=ref["nn"])
    got = device_loop(dev, 5, tmp_path)
    want = host_loop(ref, 5, oracle.oracle_kernel)
    for k in ("x", "y", "hp", "vpx", "vpy", "F11", "F22", "Zp"):
        assert np.array_equal(bits(got[k]), bits(want[k])), k


def test_enqueue_only_calls_give_the_same_bits(tmp_path):
    """gpu_kernels::set_synchronous(false): the same loop with calls that only
    enqueue their kernels gives the oracle's bits."""
    print("[offload loop] dambreak enqueue-only", flush=True)
    make, nsteps, dt = CASES["dambreak"]
    st = initial_state(make(), dt)
    got = device_loop(st, nsteps, tmp_path, enqueue_only=True)
    want = host_loop(st, nsteps, oracle.oracle_kernel)
    for k in PARTICLE + NODAL:
        assert np.array_equal(bits(got[k]), bits(want[k])), k
    for k in ("p2g", "cstart", "cptcls"):
        assert np.array_equal(got[k], np.asarray(want[k], np.int32)), k
