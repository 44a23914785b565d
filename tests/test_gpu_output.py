"""The reference solver's output files written from the device state
(amsc_sim_write_particles_csv, amsc_sim_write_grid_vts, amsc_run --output),
read back the way the reference's scripts read them (columns by quoted
header name, data/conservation_test/scripts/compare_cpu_gpu.py:22-36 and
data/well_balance/scripts/wb_error.py:24-37; conservation_errors.csv with
header time,err_mass,err_mom, conservation_check.py:14-24) and compared
with the FP64 oracle: values written with 17 significant digits read back
bit for bit."""
import csv
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from amsc_amd import scenarios
from amsc_amd.data import load_data_json, params_from_data, particles_from_data
from oracle import oracle
from test_oracle_step import small_slope_case

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RUN = os.path.join(os.path.dirname(__file__), "..", "amsc_amd", "_lib", "amsc_run")


def read_csv(path):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    header = [c.strip().strip('"').strip() for c in rows[0]]
    data = np.array([[float(v) for v in r] for r in rows[1:] if r], np.float64)
    return {h: data[:, i] for i, h in enumerate(header)}


def read_vts(path, nrows, ncols):
    root = ET.parse(path).getroot()
    out = {}
    for da in root.iter("DataArray"):
        name = da.get("Name")
        if name is None:
            continue
        v = np.array(da.text.split(), np.float64).reshape(nrows + 1, ncols + 1)
        out[name] = v.T.ravel()  # VTK order (x fastest) -> node id r + c*(nrows+1)
    return out


def test_particles_csv_and_grid_vts(tmp_path):
    from amsc_amd.simulation import Simulation
    print("[out] csv + vts", flush=True)
    d = small_slope_case()
    prm = params_from_data(d)
    part = particles_from_data(d)
    o = oracle.OracleSim(prm, part, d["Z"])
    g = Simulation(prm, part, d["Z"])
    g.set_aux(True)
    o.run(7)
    g.step(7)
    g.write_particles_csv(tmp_path / "p.csv")
    g.write_grid_vts(tmp_path / "g.vts")
    c = read_csv(tmp_path / "p.csv")
    assert len(c["x"]) == o.np
    for k in ("x", "y", "hp", "vpx", "vpy", "mom_px", "mom_py", "Zp", "Ap", "F_11",
              "F_21", "Fb_x", "vpx_dx"):
        assert np.array_equal(c[k], o.get(k)), k
    assert np.array_equal(c["label"].astype(np.int32), o.get_int("label"))
    v = read_vts(tmp_path / "g.vts", prm["nrows"], prm["ncols"])
    assert np.array_equal(v["Z"], np.asarray(d["Z"], np.float64))
    for k in ("vvxL", "vvyL", "Mv", "mom_vx", "F_int_vx", "Ftot_vy", "avx"):
        assert np.array_equal(v[k], o.get(k)), k
    g.close()


def test_amsc_run_writes_the_reference_files(tmp_path):
    """amsc_run --output on data/performance/DATA_bench_1000.json for 20
    steps: GRID_forZ.vts, 20 conservation rows below the 1e-10 threshold,
    nc_particles_<it>.csv at the start of steps 10 and 20 (the state after
    steps 9 and 19) and nc_grid_<it>.vts after them."""
    print("[out] amsc_run --output", flush=True)
    d = load_data_json(os.path.join(GOLDEN, "DATA_bench_1000.json"))
    d["NSTEPS"] = 20
    import json
    src = os.path.join(GOLDEN, "DATA_bench_1000.json")
    with open(src) as f:
        raw = json.load(f)
    raw["NSTEPS"] = 20
    with open(tmp_path / "DATA.json", "w") as f:
        json.dump(raw, f)
    r = subprocess.run([os.path.abspath(RUN), "--output"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "PERF np=" in r.stdout
    for fn in ("GRID_forZ.vts", "nc_particles_10.csv", "nc_particles_20.csv",
               "nc_grid_10.vts", "nc_grid_20.vts", "conservation_errors.csv"):
        assert (tmp_path / fn).exists(), fn
    with open(tmp_path / "conservation_errors.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["time", "err_mass", "err_mom"]
    assert len(rows) == 21
    assert max(max(float(a), float(b)) for _, a, b in rows[1:]) < 1e-10
    prm = params_from_data(d)
    o = oracle.OracleSim(prm, particles_from_data(d), d["Z"])
    o.run(9)
    c = read_csv(tmp_path / "nc_particles_10.csv")
    for k in ("x", "hp", "vpx", "mom_py"):
        assert np.array_equal(c[k], o.get(k)), k
    assert abs(float(rows[10][0]) - o.t) <= 1e-6 * max(1.0, o.t)  # step 10 starts at t_9
    o.run(1)
    v = read_vts(tmp_path / "nc_grid_10.vts", prm["nrows"], prm["ncols"])
    for k in ("vvxL", "Mv"):
        assert np.array_equal(v[k], o.get(k)), k
