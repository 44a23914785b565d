"""The oracle on the reference's own test cases, at the thresholds of the
reference's scripts:

* data/well_balance/well_balance.m (lake at rest over a Gaussian bump):
  L-infinity of |hp - (eq_level - Zp)|, |hp vpx|, |hp vpy| below 1e-10
  after 100 steps (data/well_balance/scripts/wb_error.py:72);
* data/conservation_test/conservation_test.m (dam break on an inclined
  plane with an obstacle): err_mass, err_mom below 1e-10 at every step
  (data/conservation_test/scripts/conservation_check.py:36);
* data/fena/DATA.json (a real terrain, committed as tests/golden/
  DATA_fena.json) and data/dambreak/dambreak.m (a column released down a
  sloping channel): the same conservation bound.
"""
import os

import numpy as np

from amsc_amd import scenarios
from amsc_amd.data import load_data_json, params_from_data, particles_from_data
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _run(d, steps):
    o = oracle.OracleSim(params_from_data(d), particles_from_data(d), d["Z"])
    worst = 0.0
    done = 0
    for _ in range(steps):
        if not o.step():
            break
        done += 1
        worst = max(worst, *o.conservation()[:2])
    return o, done, worst


def test_well_balance_case_stays_at_rest():
    d = scenarios.well_balance_case()
    assert len(d["x"]) == 59 * 59
    o, done, worst = _run(d, 100)
    assert done == 100 and worst < 1e-10
    hp, zp = o.get("hp"), o.get("Zp")
    assert np.abs(hp - (d["eq_level"] - zp)).max() < 1e-10
    assert np.abs(hp * o.get("vpx")).max() < 1e-10
    assert np.abs(hp * o.get("vpy")).max() < 1e-10


def test_conservation_case_transfers_to_round_off():
    d = scenarios.conservation_case()
    # 44 x 44 lattice less the hp <= 1 columns and the boundary-cell rows
    assert len(d["x"]) == 1554 and d["hp"].min() > 1.0
    o, done, worst = _run(d, 150)
    assert done == 150 and worst < 1e-10


def test_fena_terrain_transfers_to_round_off():
    d = load_data_json(os.path.join(GOLDEN, "DATA_fena.json"))
    assert len(d["x"]) == 16430 and (d["Nex"], d["Ney"]) == (259, 350)
    o, done, worst = _run(d, 30)
    assert done == 30 and worst < 1e-10


def test_dambreak_case_transfers_to_round_off():
    d = scenarios.dambreak_case()
    assert len(d["x"]) == 99 * 61  # x in [1, 50], y in [0, 30] at 0.5 m
    o, done, worst = _run(d, 100)
    assert done == 100 and worst < 1e-10
