"""The restated scaling-study generator (amsc_amd.scenarios.bench_slope,
data/performance/scaling.m:60-123) reproduces the four input files the
reference ships, data/performance/DATA_bench_{1000,3000,10000,30000}.json
(committed under tests/golden/): every key and value exactly, except the
DATA gradient dZdx, which scaling.m takes from Octave's numerical gradient
of Z (round-off around -tan 2 deg) and the generator writes as the exact
-tan 2 deg (the physics never reads it: optim_par.cpp:246-248,387)."""
import math
import os

import numpy as np
import pytest

from amsc_amd import scenarios
from amsc_amd.data import load_data_json

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n,cells,particles", [(1000, 16, 784), (3000, 27, 2500),
                                               (10000, 50, 9216), (30000, 87, 28900)])
def test_bench_slope_reproduces_the_shipped_inputs(n, cells, particles):
    f = load_data_json(os.path.join(GOLDEN, f"DATA_bench_{n}.json"))
    g = scenarios.bench_slope(n)
    assert set(f) == set(g)
    assert int(f["Nex"]) == int(f["Ney"]) == cells and len(f["x"]) == particles
    for k in f:
        a, b = np.asarray(f[k], np.float64), np.asarray(g[k], np.float64)
        assert a.shape == b.shape, k
        if k == "dZdx":
            assert np.abs(a - b).max() <= 1e-13, k
            assert np.all(b == -math.tan(2 * math.pi / 180))
        else:
            assert np.array_equal(a, b), k
