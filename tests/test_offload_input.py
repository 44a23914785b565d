"""The input file of amsc_offload_loop (amsc_amd/offload_input.py): its
layout is the one amsc_amd/csrc/offload_loop.cpp reads (five int32, fourteen
float64 scalars, the particle arrays in the driver's order, Z), and the
scalars are gpu_offload's (main/gpu_offload.cpp:155-168)."""
import math

import numpy as np
import pytest

from amsc_amd import offload_input


def state(n=7, nrows=3, ncols=4):
    rng = np.random.default_rng(5)
    st = {k: rng.normal(size=n) for k in offload_input.PARTICLE}
    st["Z"] = rng.normal(size=(nrows + 1) * (ncols + 1))
    st.update(nrows=nrows, ncols=ncols, bc_flag=1)
    st.update({k: float(i + 1) for i, k in enumerate(offload_input.SCALARS)})
    return st


def test_layout(tmp_path):
    st = state()
    p = tmp_path / "in.bin"
    offload_input.write(p, st, 9)
    raw = p.read_bytes()
    n, nn = 7, 4 * 5
    assert len(raw) == 5 * 4 + 14 * 8 + 33 * n * 8 + nn * 8
    assert np.frombuffer(raw[:20], np.int32).tolist() == [n, 3, 4, 9, 1]
    assert np.frombuffer(raw[20:132], np.float64).tolist() == list(range(1, 15))
    body = np.frombuffer(raw[132:], np.float64)
    for i, k in enumerate(offload_input.PARTICLE):
        assert np.array_equal(body[i * n:(i + 1) * n], st[k]), k
    assert np.array_equal(body[33 * n:], st["Z"])


def test_size_mismatch_is_an_error(tmp_path):
    st = state()
    st["F21"] = st["F21"][:-1]
    with pytest.raises(ValueError, match="F21"):
        offload_input.write(tmp_path / "in.bin", st, 1)


def test_scalars():
    prm = {"hx": 0.5, "hy": 0.25, "g": 9.81, "rho": 2000.0, "mu": 0.5, "tauy": 200.0,
           "bingham_on": 1.0, "phi": 15.0, "xi": 500.0, "friction_on": 1.0, "cfl": 0.4}
    s = offload_input.scalars(prm, 2e-3)
    assert s["A_coeff"] == 1.5 and s["C_coeff"] == 65.0 / 32.0
    assert s["tan_fa"] == math.tan(15.0 * math.pi / 180.0)
    assert s["tau_Y"] == 200.0 and s["cc"] == 1.0 and s["dt"] == 2e-3
    assert sorted(s) == sorted(offload_input.SCALARS)
