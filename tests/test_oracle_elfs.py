"""The oracle's legacy ELFS merge/split (merge_split_ops.h:86-755, ms_mode 1).

1. The two cell passes against a direct restatement of their definitions
   on random occupancies: exterior = empty cells connected (4-neighbours,
   through empty cells) to an empty cell of the grid edge; distance = h
   times the fewest 4-neighbour steps through wet cells from a seed (a wet
   cell off the domain edge with an exterior cell left or right of it),
   DBL_MAX where no seed is reached.
2. Merge/split steps conserve mass and momentum exactly in count terms (a
   split adds 3 particles, a merge removes 1) and to round-off in sums, and
   keep the P2G transfer at round-off afterwards.
3. The operation caps bind in cell order: with a cap of k the first k
   splits (cell order, then id order) are kept.
"""
from collections import deque

import numpy as np
import pytest

from amsc_amd import scenarios
from amsc_amd.data import params_from_data, particles_from_data
from oracle import oracle

DBL_MAX = np.finfo(np.float64).max


def direct_cells(wet, h):
    nr, nc = wet.shape
    ext = np.zeros_like(wet, bool)
    q = deque()
    for r in range(nr):
        for c in range(nc):
            if (r in (0, nr - 1) or c in (0, nc - 1)) and not wet[r, c]:
                ext[r, c] = True
                q.append((r, c))
    while q:
        r, c = q.popleft()
        for rr, cc in ((r - 1, c), (r + 1, c), (r, c - 1), (r, c + 1)):
            if 0 <= rr < nr and 0 <= cc < nc and not wet[rr, cc] and not ext[rr, cc]:
                ext[rr, cc] = True
                q.append((rr, cc))
    hops = np.full(wet.shape, -1)
    q = deque()
    for r in range(1, nr - 1):
        for c in range(1, nc - 1):
            if wet[r, c] and (ext[r, c - 1] or ext[r, c + 1]):
                hops[r, c] = 0
                q.append((r, c))
    while q:
        r, c = q.popleft()
        for rr, cc in ((r - 1, c), (r + 1, c), (r, c - 1), (r, c + 1)):
            if 0 <= rr < nr and 0 <= cc < nc and wet[rr, cc] and hops[rr, cc] < 0:
                hops[rr, cc] = hops[r, c] + 1
                q.append((rr, cc))
    dist = np.where(hops >= 0, hops * h, DBL_MAX)
    return ext, dist


@pytest.mark.parametrize("seed,shape,p", [(1, (9, 13), 0.5), (2, (30, 17), 0.7),
                                          (3, (40, 40), 0.85), (4, (1, 7), 0.5)])
def test_cell_passes_match_their_definition(seed, shape, p):
    rng = np.random.default_rng(seed)
    counts = (rng.random(shape) < p) * rng.integers(1, 5, shape)
    # an enclosed dry hole and a wet blob touching the edge
    if shape[0] > 10:
        counts[4:9, 4:9] = 3
        counts[6, 6] = 0
    ext, dist = oracle.elfs_cells(counts, 0.7)
    e2, d2 = direct_cells(counts > 0, 0.7)
    assert np.array_equal(ext, e2)
    assert np.array_equal(dist, d2)


def elfs_case(max_ops=1e9):
    d = scenarios.dambreak_case()
    d.update(ms_mode=1, ms_alpha=2.0, ms_beta=3.0, ms_max_ops=max_ops,
             ms_call_interval=5, ms_hp_min=0.05, ms_max_dv=0.01)
    return d


def test_merge_split_steps_conserve_mass_and_momentum():
    d = elfs_case()
    prm = params_from_data(d)
    o = oracle.OracleSim(prm, particles_from_data(d), d["Z"])
    changed = 0
    for step in range(1, 21):
        n0 = o.np
        m0 = o.get("Mp").sum()
        assert o.step()
        s, m = o.ms_counts()
        if step % 5 == 0:
            assert o.np == n0 + 3 * s - m
            changed += s + m
        assert abs(o.get("Mp").sum() - m0) <= 1e-12 * m0
        em, emom, _ = o.conservation()
        assert em < 1e-10 and emom < 1e-10
    assert changed > 0


def test_split_cap_keeps_the_first_splits_in_cell_order():
    d = elfs_case(max_ops=7)
    prm = params_from_data(d)
    o = oracle.OracleSim(prm, particles_from_data(d), d["Z"])
    o.run(4)
    n4 = o.np
    o.run(1)
    s, m = o.ms_counts()
    assert s <= 7 and m <= 7
    assert o.np == n4 + 3 * s - m
