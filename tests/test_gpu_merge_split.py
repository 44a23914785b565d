"""Device merge/split on the caller's arrays against the oracle:
decisions (amsc_ms_decide, decide_actions of merge_split_ops_cmes.h:116-253)
give identical action and partner arrays, with and without a binding
ms_max_ops cap; execution (amsc_ms_execute, execute_merge_split :258-427)
gives the oracle's new particle set bit for bit."""
import numpy as np
import pytest

from amsc_amd import scenarios
from amsc_amd.data import params_from_data
from oracle import oracle

pytestmark = pytest.mark.gpu


def snapshot(seed, n=6000, nrows=30, ncols=40, h=1.0):
    rng = np.random.default_rng(seed)
    # clusters so that cells hold many candidates
    cx = rng.uniform(0, ncols * h, 60)
    cy = rng.uniform(0, nrows * h, 60)
    k = rng.integers(0, 60, n)
    x = np.clip(cx[k] + rng.normal(0, 0.6, n), 1e-10, ncols * h - 1e-10)
    y = np.clip(cy[k] + rng.normal(0, 0.6, n), 1e-10, nrows * h - 1e-10)
    s = dict(x=x, y=y,
             Ap=rng.choice([0.01, 0.05, 0.2, 1.2], n) * rng.uniform(0.8, 1.2, n),
             hp=rng.uniform(0.0, 2.0, n),
             vpx_dx=rng.normal(0, 0.5, n), vpx_dy=rng.normal(0, 0.5, n),
             vpy_dx=rng.normal(0, 0.5, n), vpy_dy=rng.normal(0, 0.5, n),
             vpx=rng.normal(0, 0.02, n), vpy=rng.normal(0, 0.02, n),
             level=rng.integers(-3, 4, n).astype(np.int32))
    d = scenarios.dambreak_flat(1)
    prm = params_from_data(d)
    prm.update(nrows=nrows, ncols=ncols, hx=h, hy=h, ms_alpha=0.9, ms_beta=0.3,
               ms_shear_split=0.05, ms_split_hp_min=0.2, ms_hp_min=0.05,
               ms_max_dv=0.03, ms_min_ppc=2)
    return prm, s


def device_decide(prm, s, dt):
    from amsc_amd import native
    D = native.DeviceArray
    n = len(s["x"])
    keep = []

    def dev(a, dtype=np.float64):
        t = D(a, dtype)
        keep.append(t)
        return t.ptr
    _, cstart, cptcls, _, _ = oracle.bin_particles(s["x"], s["y"], prm["nrows"],
                                                   prm["ncols"], prm["hx"], prm["hy"])
    ptrs = [dev(s[k]) for k in ("x", "y", "Ap", "hp", "vpx_dx", "vpx_dy",
                                "vpy_dx", "vpy_dy", "vpx", "vpy")]
    act = D.zeros(n, np.int32)
    part = D.zeros(n, np.int32)
    counts = native.ms_decide(prm, n, dt, ptrs, dev(s["level"], np.int32),
                              dev(cstart, np.int32), dev(cptcls, np.int32),
                              act.ptr, part.ptr)
    return act.numpy(), part.numpy(), counts


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("max_ops", [1e9, 25, 0])
def test_decisions_identical(seed, max_ops):
    prm, s = snapshot(seed)
    prm["ms_max_ops"] = max_ops
    dt = 0.05
    print(f"[ms] seed={seed} max_ops={max_ops}", flush=True)
    a1, p1, c1 = device_decide(prm, s, dt)
    a2, p2, c2 = oracle.ms_decide(prm, dt, s["x"], s["y"], s["Ap"], s["hp"],
                                  s["vpx_dx"], s["vpx_dy"], s["vpy_dx"],
                                  s["vpy_dy"], s["vpx"], s["vpy"], s["level"])
    assert np.array_equal(a1, a2)
    assert np.array_equal(p1, p2)
    assert np.array_equal(c1, c2)
    if max_ops == 1e9:
        assert (a1 == 1).sum() > 50 and (a1 == 2).sum() > 50  # non-trivial


def exec_state(case, steps):
    """An oracle state just before a merge/split, its decisions, and the
    oracle's execute_merge_split of it."""
    from test_oracle_step import small_slope_case
    from amsc_amd.data import particles_from_data
    if case == "slope":
        d = small_slope_case()
        d.update(MERGE_SPLIT_ON=1, ms_call_interval=4, ms_alpha=0.3, ms_beta=0.52,
                 ms_max_dv=2.0, ms_hp_min=0.0)
    else:
        d = scenarios.dambreak_case()
        d.update(ms_alpha=0.45, ms_beta=0.55, ms_max_dv=0.5, ms_hp_min=0.01,
                 ms_shear_split=0.02)
    prm = params_from_data(d)
    o = oracle.OracleSim(prm, particles_from_data(d), d["Z"])
    o.run(steps)
    names = oracle.field_names()
    f = {k: o.get(k) for k in names}
    label, level = o.get_int("label"), o.get_int("level")
    act, partner, counts = oracle.ms_decide(
        prm, o.dt, f["x"], f["y"], f["Ap"], f["hp"], f["vpx_dx"], f["vpx_dy"],
        f["vpy_dx"], f["vpy_dy"], f["vpx"], f["vpy"], level)
    return prm, names, f, label, level, act, partner, counts


@pytest.mark.parametrize("case,steps", [("slope", 7), ("dambreak", 30)])
def test_execute_matches_oracle(case, steps):
    """amsc_ms_execute (execute_merge_split on the caller's arrays, all 35
    of the reference's particle fields) equals the oracle's execution of
    the same decisions bit for bit: fields, labels, levels, order."""
    from amsc_amd import native
    D = native.DeviceArray
    print(f"[ms] execute {case}", flush=True)
    prm, names, f, label, level, act, partner, counts = exec_state(case, steps)
    assert counts[0] > 0 and (case == "slope" or counts[1] > 0), counts
    want, wl, wv = oracle.ms_execute(prm, act, partner, f, label, level)
    n = len(act)
    cap = n + int(np.sum(act == 1)) + 7
    keep = [D(f[k]) for k in names]
    outs = [D.zeros(cap, np.float64) for _ in names]
    da, dp = D(act, np.int32), D(partner, np.int32)
    dl, dv = D(label, np.int32), D(level, np.int32)
    ol, ov = D.zeros(cap, np.int32), D.zeros(cap, np.int32)
    nnew = native.ms_execute(prm, n, da.ptr, dp.ptr, names, [a.ptr for a in keep],
                             [a.ptr for a in outs], dl.ptr, dv.ptr, ol.ptr, ov.ptr,
                             cap)
    assert nnew == len(want["x"])
    for k, a in zip(names, outs):
        got = a.numpy()[:nnew]
        assert np.array_equal(got.view(np.int64), want[k].view(np.int64)), k
    assert np.array_equal(ol.numpy()[:nnew], wl)
    assert np.array_equal(ov.numpy()[:nnew], wv)
    # a capacity one short is refused
    with pytest.raises(native.AmscError):
        native.ms_execute(prm, n, da.ptr, dp.ptr, names, [a.ptr for a in keep],
                          [a.ptr for a in outs], dl.ptr, dv.ptr, ol.ptr, ov.ptr,
                          nnew - 1)
