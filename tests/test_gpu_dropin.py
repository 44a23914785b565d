"""HIP drop-ins of the 18 gpu_kernels entry points (amsc_* in
include/amsc_mpm.h) against the oracle on the same seeded inputs, and against
the committed vectors of the reference's own build: bit-identical outputs."""
import os

import numpy as np
import pytest

from kernel_cases import ARGS, host_args, kinds, make_case
from oracle import oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_kernels.npz")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def run_device(name, case):
    from amsc_amd import native
    args, outs, keep = [], {}, []
    for a, k in zip(ARGS[name], kinds(name)):
        v = case[a]
        if k in "id":
            args.append(v)
        elif k == "b":
            args.append(int(bool(v)))
        elif k == "H":
            args.append(np.ascontiguousarray(v, np.int32).copy())
        elif k == "I":
            t = native.DeviceArray(v, np.int32)
            keep.append(t)
            args.append(t.ptr)
        elif k in "PO":
            if a.endswith("_out"):
                args.append(None)
                continue
            t = native.DeviceArray(v, np.float64)
            keep.append(t)
            args.append(t.ptr)
            if k == "O":
                outs[a] = t
        else:
            args.append(None)
    print(f"[dropin] {name}", flush=True)
    native.dropin(name)(*args)
    native.check(native.lib().amsc_synchronize())
    return {k: t.numpy() for k, t in outs.items()}


@pytest.mark.parametrize("name", sorted(ARGS))
@pytest.mark.parametrize("seed", [0, 1])
def test_dropin_matches_oracle(name, seed):
    case = make_case(seed)
    a, o = host_args(name, case)
    oracle.oracle_kernel(name)(*a)
    got = run_device(name, case)
    assert set(got) == {k for k in o if not k.endswith("_out")}
    for k in got:
        assert np.array_equal(bits(got[k]), bits(o[k])), f"{name}: {k} differs"


def test_dropin_matches_reference_vectors():
    g = np.load(GOLDEN, allow_pickle=False)
    case = {k[len("in/"):]: g[k] for k in g.files if k.startswith("in/")}
    for k in ("np", "nn", "nrows", "ncols", "bc_flag"):
        case[k] = int(case[k])
    for k in ("hx", "hy", "g", "dt", "rho", "mu", "tau_Y", "cc", "A_coeff",
              "C_coeff", "tan_fa", "xi", "fric_on"):
        case[k] = float(case[k])
    for name in sorted(ARGS):
        got = run_device(name, case)
        for k, v in got.items():
            assert np.array_equal(bits(v), bits(g[f"out/{name}/{k}"])), f"{name}: {k}"


def test_scatter_accumulates_into_existing_values_in_colour_order():
    """Large grid with many particles per cell: the node gather adds to the
    current nodal values in the reference's colour order."""
    case = make_case(7, nrows=40, ncols=33, npart=30000, hx=0.25, hy=0.3)
    for name in ("scatter_mass_momentum", "scatter_internal_forces",
                 "scatter_plotvars"):
        a, o = host_args(name, case)
        oracle.oracle_kernel(name)(*a)
        got = run_device(name, case)
        for k in got:
            assert np.array_equal(bits(got[k]), bits(o[k])), f"{name}: {k}"


def test_bad_colouring_is_rejected():
    from amsc_amd import native
    case = make_case(0)
    name = "scatter_mass_momentum"
    bad = dict(case)
    bad["coloff"] = np.array([0, 5, 9, 13, case["nrows"] * case["ncols"] + 1], np.int32)
    with pytest.raises(native.AmscError):
        run_device(name, bad)


def test_cfl_dt_matches_the_oracle_adaptive_step():
    """amsc_cfl_dt on the state after step 1 gives the dt the oracle's
    adaptive step 2 takes (gpu_offload.cpp:282-301, optim_par.cpp:285-307),
    bit for bit; the first step keeps its dt; the maxima are of |v| and of
    hp, whatever the signs."""
    from amsc_amd import native
    from amsc_amd.data import params_from_data, particles_from_data
    from test_oracle_step import small_slope_case
    print("[dropin] cfl dt", flush=True)
    d = small_slope_case()
    d.update(DT_FIXED=0.0, CFL=0.4, T=1.0)
    prm = params_from_data(d)
    o = oracle.OracleSim(prm, particles_from_data(d), d["Z"])
    assert o.step()
    arrs = [native.DeviceArray(o.get(k)) for k in ("vpx", "vpy", "hp")]
    assert o.step()
    dt, cel = native.cfl_dt(len(o.get("x")), *[a.ptr for a in arrs], prm["g"],
                            prm["cfl"], prm["hx"], prm["hy"], 1, 7.0)
    assert np.float64(dt).view(np.int64) == np.float64(o.dt).view(np.int64)
    dt0, _ = native.cfl_dt(len(o.get("x")), *[a.ptr for a in arrs], prm["g"],
                           prm["cfl"], prm["hx"], prm["hy"], 0, 1e-3)
    assert dt0 == 1e-3
    rng = np.random.default_rng(4)
    for n in (1, 255, 70001):
        vx, vy, h = rng.normal(0, 3, n), rng.normal(0, 3, n), rng.uniform(0, 5, n)
        vx[n // 2] = -40.0
        a = [native.DeviceArray(v) for v in (vx, vy, h)]
        _, cel = native.cfl_dt(n, *[t.ptr for t in a], 9.81, 0.4, 1.0, 1.0, 3, 1.0)
        sq = np.sqrt(9.81 * h.max())
        assert cel == abs(max(sq + np.abs(vx).max(), sq + np.abs(vy).max()))
