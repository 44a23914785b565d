"""The multiply-shift division used on the device to split a cell id into
(column, row) (FastDiv in amsc_amd/csrc/common.hpp): q = (n*m) >> k with
k = 31 + ceil(log2 d), m = ceil(2^k/d) equals n // d for every n < 2^31.
Checked here by brute force on edge values and random values."""
import numpy as np


def fastdiv(n, d):
    l = 0
    while (1 << l) < d:
        l += 1
    k = 31 + l
    m = ((1 << k) + d - 1) // d
    assert m < (1 << 64)
    q = (n * m) >> k
    assert n * m < (1 << 64)  # the 64-bit product never overflows
    return q


def test_fastdiv_exact():
    rng = np.random.default_rng(0)
    divisors = [1, 2, 3, 5, 7, 16, 17, 100, 999, 1000, 1024, 2048, 2047, 4096,
                65535, 65536, 1 << 20, 12345, (1 << 31) - 1]
    divisors += rng.integers(1, 1 << 16, 200).tolist()
    for d in divisors:
        ns = [0, 1, d - 1, d, d + 1, 2 * d - 1, (1 << 31) - 1, (1 << 31) - 2]
        ns += [((1 << 31) - 1) // d * d, ((1 << 31) - 1) // d * d - 1]
        ns += rng.integers(0, 1 << 31, 300).tolist()
        for n in ns:
            if 0 <= n < (1 << 31):
                assert fastdiv(int(n), int(d)) == int(n) // int(d), (n, d)
