"""Device binning (amsc_bin_particles) against the oracle's counting sort:
identical integers (cell of each particle, cell starts, particle order, the
colouring), including empty grids, crowded cells and cell-edge positions."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def device_bin(x, y, nrows, ncols, hx, hy):
    from amsc_amd import native
    D = native.DeviceArray
    n = len(x)
    tx, ty = D(x, np.float64), D(y, np.float64)
    p2g = D.zeros(max(n, 1), np.int32)
    cptcls = D.zeros(max(n, 1), np.int32)
    cstart = D.zeros(nrows * ncols + 1, np.int32)
    ccidx = D.zeros(max(nrows * ncols, 1), np.int32)
    print(f"[bin] n={n} grid={nrows}x{ncols}", flush=True)
    coloff = native.bin_particles(n, nrows, ncols, hx, hy, tx.ptr, ty.ptr,
                                  p2g.ptr, cstart.ptr, cptcls.ptr, ccidx.ptr)
    return (p2g.numpy()[:n], cstart.numpy(), cptcls.numpy()[:n],
            ccidx.numpy()[:nrows * ncols], coloff)


def check(x, y, nrows, ncols, hx, hy):
    got = device_bin(x, y, nrows, ncols, hx, hy)
    exp = oracle.bin_particles(x, y, nrows, ncols, hx, hy)
    for name, a, b in zip(("p2g", "cstart", "cptcls", "ccidx", "coloff"), got, exp):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("n,nrows,ncols", [(1000, 17, 23), (200000, 300, 257),
                                           (1_000_000, 1000, 1000)])
def test_random(n, nrows, ncols):
    rng = np.random.default_rng(n)
    hx, hy = 0.5, 0.75
    x = rng.uniform(1e-10, ncols * hx - 1e-10, n)
    y = rng.uniform(1e-10, nrows * hy - 1e-10, n)
    x[: n // 20] = np.floor(x[: n // 20] / hx) * hx  # exactly on cell edges
    check(x, y, nrows, ncols, hx, hy)


def test_crowded_cells_and_sorted_input():
    rng = np.random.default_rng(1)
    nrows, ncols, hx, hy = 9, 7, 1.0, 1.0
    # 300 particles in one cell (global in-cell sort path), 40 in another
    x = np.concatenate([rng.uniform(2.0, 3.0, 300), rng.uniform(5.0, 6.0, 40),
                        rng.uniform(0, 7, 300)])
    y = np.concatenate([rng.uniform(4.0, 5.0, 300), rng.uniform(1.0, 2.0, 40),
                        rng.uniform(0, 9, 300)])
    perm = rng.permutation(len(x))
    check(x[perm], y[perm], nrows, ncols, hx, hy)
    xs, ys = np.sort(x), np.sort(y)
    check(xs, ys, nrows, ncols, hx, hy)


def test_pile_up_in_one_cell():
    """20000 particles in one cell (what clamping at the domain edge makes
    of a run whose velocities explode): the in-cell sort stays O(n log n)."""
    rng = np.random.default_rng(7)
    nrows, ncols, hx, hy = 12, 10, 1.0, 1.0
    x = np.concatenate([np.full(20000, 1e-10), rng.uniform(0, 10, 2000)])
    y = np.concatenate([rng.uniform(0.0, 1.0, 20000), rng.uniform(0, 12, 2000)])
    perm = rng.permutation(len(x))
    check(x[perm], y[perm], nrows, ncols, hx, hy)


def test_empty_and_clamped():
    check(np.zeros(0), np.zeros(0), 4, 5, 1.0, 1.0)
    # outside the domain: clamped to the border cells
    check(np.array([-1.0, 100.0, 2.5, 5.0]), np.array([2.0, -3.0, 50.0, 4.0]),
          4, 5, 1.0, 1.0)
