"""amsc_bin_particles_noalloc (the binning entry point that allocates
nothing: its workspace is a fixed array of the library's device memory)
against amsc_bin_particles through the C ABI: identical integers on two
calls in a row, the second with fewer particles on a smaller grid, so that
the shared workspace's leftovers from the first call must not leak into the
second; sizes beyond its fixed workspace are refused before any device
work."""
import numpy as np
import pytest

from amsc_amd import native

MAX = 1 << 22


def bin_both(x, y, nrows, ncols, hx, hy):
    D = native.DeviceArray
    n = len(x)
    tx, ty = D(x, np.float64), D(y, np.float64)
    outs = []
    for noalloc in (False, True):
        p2g = D.zeros(max(n, 1), np.int32)
        cptcls = D.zeros(max(n, 1), np.int32)
        cstart = D.zeros(nrows * ncols + 1, np.int32)
        ccidx = D.zeros(nrows * ncols, np.int32)
        print(f"[noalloc] n={n} grid={nrows}x{ncols} noalloc={noalloc}", flush=True)
        coloff = native.bin_particles(n, nrows, ncols, hx, hy, tx.ptr, ty.ptr, p2g.ptr,
                                      cstart.ptr, cptcls.ptr, ccidx.ptr, noalloc=noalloc)
        outs.append((p2g.numpy()[:n], cstart.numpy(), cptcls.numpy()[:n],
                     ccidx.numpy(), coloff))
    return outs


def case(seed, n, nrows, ncols, hx, hy):
    rng = np.random.default_rng(seed)
    x = rng.uniform(1e-10, ncols * hx - 1e-10, n)
    y = rng.uniform(1e-10, nrows * hy - 1e-10, n)
    x[: n // 10] = np.floor(x[: n // 10] / hx) * hx  # on cell edges
    # a crowded cell (heap-sort path) and particles stored out of cell order
    x[n // 10: n // 10 + 40] = 0.5 * hx
    y[n // 10: n // 10 + 40] = 0.5 * hy
    return x, y


@pytest.mark.gpu
def test_noalloc_matches_binning_twice_second_smaller():
    for seed, n, nrows, ncols, hx, hy in ((1, 3_000_000, 1500, 1400, 0.5, 0.75),
                                          (2, 20_000, 90, 110, 1.0, 1.0)):
        x, y = case(seed, n, nrows, ncols, hx, hy)
        ref, got = bin_both(x, y, nrows, ncols, hx, hy)
        for name, a, b in zip(("p2g", "cstart", "cptcls", "ccidx", "coloff"), ref, got):
            assert np.array_equal(a, b), (n, name)


@pytest.mark.gpu
def test_noalloc_at_its_maximum_size():
    nrows = ncols = 2048
    x, y = case(3, MAX, nrows, ncols, 0.5, 0.5)
    ref, got = bin_both(x, y, nrows, ncols, 0.5, 0.5)
    for name, a, b in zip(("p2g", "cstart", "cptcls", "ccidx", "coloff"), ref, got):
        assert np.array_equal(a, b), name


def test_noalloc_refuses_sizes_beyond_its_workspace():
    lib = native.lib()
    coloff = np.zeros(5, np.int32)
    rc = lib.amsc_bin_particles_noalloc(10, 2049, 2048, 1.0, 1.0, None, None, None,
                                        None, None, None, coloff)
    assert rc == 1 and b"4194304 cells" in lib.amsc_last_error()
    rc = lib.amsc_bin_particles_noalloc(MAX + 1, 64, 64, 1.0, 1.0, None, None, None,
                                        None, None, None, coloff)
    assert rc == 1 and b"4194304 particles" in lib.amsc_last_error()
