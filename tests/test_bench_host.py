"""Host-side parts of bench.py (no GPU): the one-GPU guard, the CPU
baseline leg, and the error measures the mixed-precision comparison
reports (nodal depth field, label-matched per-particle errors, their
aggregation over scenarios)."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_cpu_baseline_on_a_small_case():
    from test_oracle_step import small_slope_case
    rate, seconds, np_end = bench.cpu_baseline(small_slope_case(), steps=5, threads=1)
    assert rate > 0 and seconds > 0 and np_end > 0


PRM = {"nrows": 6, "ncols": 8, "hx": 0.5, "hy": 0.25}


def particles(seed, n=300):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, PRM["ncols"] * PRM["hx"], n)
    y = rng.uniform(0, PRM["nrows"] * PRM["hy"], n)
    return {"x": x, "y": y, "Vp": rng.uniform(0.1, 1.0, n), "hp": rng.uniform(0.5, 2.0, n),
            "vpx": rng.normal(0, 1, n), "vpy": rng.normal(0, 1, n),
            "F_11": rng.normal(0, 50, n), "F_12": rng.normal(0, 20, n),
            "F_22": rng.normal(0, 50, n), "label": np.arange(n, dtype=np.int32)}


def test_nodal_depth_conserves_volume_and_is_continuous():
    p = particles(1)
    g = bench.nodal_depth(p["x"], p["y"], p["Vp"], PRM)
    area = PRM["hx"] * PRM["hy"]
    assert abs(g.sum() * area - p["Vp"].sum()) <= 1e-12 * p["Vp"].sum()
    # a particle moved across a cell edge by round-off changes the field by
    # round-off (a per-cell count would move its whole volume)
    x = p["x"].copy()
    x[0] = 2 * PRM["hx"] - 1e-13
    g1 = bench.nodal_depth(x, p["y"], p["Vp"], PRM)
    x[0] = 2 * PRM["hx"] + 1e-13
    g2 = bench.nodal_depth(x, p["y"], p["Vp"], PRM)
    assert np.abs(g1 - g2).max() < 1e-9


def test_field_errors_match_particles_by_label():
    a = particles(2)
    b = {k: v.copy() for k, v in a.items()}
    e = bench.field_errors(a, b, PRM)
    assert e["hp_max"] == 0.0 and e["matched_frac"] == 1.0 and e["node_depth_max"] == 0.0
    # b's particles stored in another order, one new particle (label -1)
    perm = np.random.default_rng(3).permutation(len(a["x"]))
    c = {k: v[perm] for k, v in a.items()}
    c = {k: np.append(v, v[:1]) for k, v in c.items()}
    c["label"][-1] = -1
    e = bench.field_errors(a, c, PRM)
    assert e["hp_max"] == 0.0
    assert e["matched_frac"] == pytest.approx(len(a["x"]) / (len(a["x"]) + 1))
    # a known depth error at one particle, relative to b's maximum depth
    d = {k: v.copy() for k, v in a.items()}
    d["hp"][5] += 1e-3
    e = bench.field_errors(d, a, PRM)
    assert e["hp_max"] == pytest.approx(1e-3 / a["hp"].max(), rel=1e-9)
    assert e["hp_frac_within_1e-4"] == pytest.approx(1 - 1 / len(a["x"]))
    # position error in cells, stress error against the largest component
    d = {k: v.copy() for k, v in a.items()}
    d["x"][7] += 0.1 * PRM["hy"]
    d["F_12"][9] += 2.0
    e = bench.field_errors(d, a, PRM)
    assert e["pos_max_over_h"] == pytest.approx(0.1 * PRM["hy"] / min(PRM["hx"], PRM["hy"]))
    fmax = max(np.abs(a[k]).max() for k in ("F_11", "F_12", "F_22"))
    assert e["F_max"] == pytest.approx(2.0 / fmax, rel=1e-9) and e["F_median"] == 0.0


def test_aggregate_errors():
    errs = [{"hp_max": 1e-5, "matched_frac": 1.0}, {"hp_max": 2e-3, "matched_frac": 0.5},
            {"matched_frac": 0.0}]
    out = bench.aggregate_errors(errs)
    assert out["hp_max"] == {"max": 2e-3, "median": pytest.approx(1.005e-3), "min": 1e-5,
                             "scenarios": 2}
    assert out["matched_frac"]["median"] == 0.5
    assert out["scenarios_hp_max_within_1e-4"] == 1
