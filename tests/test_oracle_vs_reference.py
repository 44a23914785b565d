"""The oracle's restatement of the 18 step kernels against the reference's
own main/gpu_kernels.cpp, compiled for the host (oracle/_ref), on the same
seeded inputs: bit-identical outputs.  Also against the committed vectors
in tests/golden/ref_kernels.npz written from that build, so the check holds
where the reference sources are absent."""
import os

import numpy as np
import pytest

from kernel_cases import ARGS, host_args, make_case, output_names
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_kernels.npz")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


@pytest.mark.parametrize("name", sorted(ARGS))
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_matches_reference_build(name, seed):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    case = make_case(seed)
    a1, o1 = host_args(name, case)
    a2, o2 = host_args(name, case)
    oracle.oracle_kernel(name)(*a1)
    oracle.ref_kernel(name)(*a2)
    for k in o1:
        assert np.array_equal(bits(o1[k]), bits(o2[k])), f"{name}: {k} differs"


def test_oracle_matches_committed_reference_vectors():
    g = np.load(GOLDEN, allow_pickle=False)
    case = {k[len("in/"):]: g[k] for k in g.files if k.startswith("in/")}
    for k in ("np", "nn", "nrows", "ncols", "bc_flag"):
        case[k] = int(case[k])
    for k in ("hx", "hy", "g", "dt", "rho", "mu", "tau_Y", "cc", "A_coeff",
              "C_coeff", "tan_fa", "xi", "fric_on"):
        case[k] = float(case[k])
    checked = 0
    for name in sorted(ARGS):
        args, outs = host_args(name, case)
        oracle.oracle_kernel(name)(*args)
        for k, v in outs.items():
            ref = g[f"out/{name}/{k}"]
            assert np.array_equal(bits(v), bits(ref)), f"{name}: {k}"
            checked += 1
    assert checked >= 40


def test_scatter_colour_loop_touches_every_node_once_per_cell():
    """The colouring of or_bin_particles is proper (no two cells sharing a
    node have one colour): the colour-parallel scatter of the reference is
    race-free only under that property (gpu_kernels.cpp:104-126)."""
    nrows, ncols = 5, 8
    case = make_case(3, nrows=nrows, ncols=ncols, npart=50)
    ccidx, coloff = case["ccidx"], case["coloff"]
    colour = np.empty(nrows * ncols, int)
    for k in range(4):
        colour[ccidx[coloff[k]:coloff[k + 1]]] = k
    assert sorted(ccidx.tolist()) == list(range(nrows * ncols))
    for cell in range(nrows * ncols):
        r, c = cell % nrows, cell // nrows
        for dc in (-1, 0, 1):
            for dr in (-1, 0, 1):
                if (dc or dr) and 0 <= c + dc < ncols and 0 <= r + dr < nrows:
                    assert colour[cell] != colour[(r + dr) + (c + dc) * nrows]


@pytest.mark.parametrize("case", ["slope", "dambreak", "dambreak_adaptive"])
def test_offload_loop_composition_matches_reference_build(case):
    """gpu_offload's step (gpu_offload.cpp:410-562, fixed dt) composed from
    the oracle's kernels equals the same composition over the reference's
    own compiled kernels, bit for bit, at the end of the run: the loop the
    device drop-in is checked against in test_gpu_offload_loop.py."""
    if not oracle.ref_available():
        pytest.skip("reference build not present (make -C oracle ref)")
    from test_gpu_offload_loop import CASES, NODAL, PARTICLE, host_loop, initial_state
    make, nsteps, dt = CASES[case]
    d = make()
    a = host_loop(initial_state(d, dt), nsteps, oracle.oracle_kernel, dt <= 0.0)
    b = host_loop(initial_state(d, dt), nsteps, oracle.ref_kernel, dt <= 0.0)
    for k in PARTICLE + NODAL:
        assert np.array_equal(a[k].view(np.int64), b[k].view(np.int64)), k
    for k in ("p2g", "cstart", "cptcls"):
        assert np.array_equal(a[k], b[k]), k
