"""The legacy ELFS merge/split rule (ms_mode 1, merge_split_ops.h:86-755) on
the device against the oracle, bit for bit: the dambreak.m channel with
merge/split every 5 steps, without and with the operation cap (ms_max_ops 7
binds: the serial replay in cell order and the demotion of every later split
and candidate), alone and in a batch beside scenarios of the other rule, with
eager launches and graph replay."""
import numpy as np
import pytest

from amsc_amd import scenarios
from amsc_amd.data import params_from_data, particles_from_data
from oracle import oracle
from test_gpu_sim import advance, assert_same, pair
from test_gpu_ensemble import FIELDS, assert_scen_equal, bits, scen
from test_oracle_step import small_slope_case

pytestmark = pytest.mark.gpu


def elfs_case(max_ops=1e9):
    d = scenarios.dambreak_case()
    d.update(ms_mode=1, ms_alpha=2.0, ms_beta=3.0, ms_max_ops=max_ops,
             ms_call_interval=5, ms_hp_min=0.05, ms_max_dv=0.01)
    return d


@pytest.mark.parametrize("max_ops", [1e9, 7])
def test_elfs_single_bitwise(max_ops):
    print(f"[elfs] single max_ops={max_ops}", flush=True)
    d = elfs_case(max_ops)
    o, g, _ = pair(d, capacity=8 * len(d["x"]))
    splits = merges = 0
    for k in range(1, 21):
        advance(o, g, 1)
        if k % 5 == 0:
            assert np.array_equal(g.ms_counts(), o.ms_counts()), f"step {k}"
            s, m = (int(v) for v in o.ms_counts())
            splits += s
            merges += m
            if max_ops == 7:
                assert s <= 7 and m <= 7
            assert_same(o, g, f"step {k}")
    assert splits > 0 and merges > 0, (splits, merges)
    if max_ops == 7:  # the cap bound at least once
        assert splits + merges >= 14
    g.close()


def elfs_batch():
    b = small_slope_case(seed=9, nrows=14, ncols=11, hx=0.4, hy=0.5)
    b.update(MERGE_SPLIT_ON=1, ms_call_interval=4, ms_alpha=0.3, ms_beta=0.52,
             ms_max_dv=2.0, ms_hp_min=0.0)
    return [scen(elfs_case(1e9)), scen(b), scen(elfs_case(7)), scen(small_slope_case())]


def test_elfs_in_a_mixed_rule_batch():
    from amsc_amd.simulation import Ensemble
    print("[elfs] batch of both rules vs oracle", flush=True)
    cases = elfs_batch()
    e = Ensemble(cases, "fp64", capacity=8 * sum(len(p["x"]) for _, p, _ in cases))
    e.set_aux(True)
    orc = [oracle.OracleSim(prm, part, Z) for prm, part, Z in cases]
    for step in range(1, 21):
        e.step(1)
        for o in orc:
            o.step()
        if step % 5 == 0 or step % 4 == 0:
            for k, o in enumerate(orc):
                assert_scen_equal(e, k, o, f"step {step}")
                if step % (4 if k == 1 else 5) == 0 and k != 3:
                    assert np.array_equal(e.ms_counts(k), o.ms_counts()), (step, k)
    for k in (0, 1, 2):
        assert e.status()["np"][k] != len(cases[k][1]["x"]), k
    e.close()


def test_elfs_graph_replay_matches_eager():
    from amsc_amd.simulation import Ensemble
    print("[elfs] graph vs eager", flush=True)
    cases = elfs_batch()
    cap = 8 * sum(len(p["x"]) for _, p, _ in cases)
    a = Ensemble(cases, "fp64", capacity=cap)
    b = Ensemble(cases, "fp64", capacity=cap)
    b.set_graph(True)
    for s in (a, b):
        s.set_aux(True)
        s.step(20)
    for k in range(len(cases)):
        assert a.num(k) == b.num(k)
        for f in FIELDS:
            assert np.array_equal(bits(a.get(k, f)), bits(b.get(k, f))), (k, f)
        assert np.array_equal(a.ms_counts(k), b.ms_counts(k))
    a.close()
    b.close()


def test_elfs_bytes_estimate_covers_allocation():
    from amsc_amd.simulation import Ensemble
    cases = elfs_batch()
    est = Ensemble.bytes_estimate([p for p, _, _ in cases],
                                  [len(q["x"]) for _, q, _ in cases], "fp64")
    e = Ensemble(cases, "fp64")
    held = e.info()["bytes"]
    assert held <= est <= 1.2 * held + (1 << 20), (held, est)
    e.close()
