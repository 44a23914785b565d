"""DATA.json input: the schema of main/mpm_data.h:6-122, read with the
standard json module (floats round-trip exactly: the reference writers print
%.17g, e.g. data/performance/scaling.m:52).

`params_from_data` flattens a DATA dict into the solver parameter names used
by the C ABI (`amsc_params` in include/amsc_mpm.h).
"""
from __future__ import annotations

import json

import numpy as np

PARTICLE_KEYS = ("x", "y", "Mp", "Ap", "vpx", "vpy", "mom_px", "mom_py", "hp", "Vp")
NODAL_KEYS = ("Z", "dZdx", "dZdy")

# Optional keys and their defaults, main/mpm_data.h:99-112.
MS_DEFAULTS = {
    "ms_alpha": 0.9, "ms_beta": 0.30, "ms_split_hp_min": 0.0,
    "ms_shear_split": 1e30, "ms_hp_min": 0.05, "ms_max_dv": 0.01,
    "ms_min_level": -2, "ms_max_level": 2, "ms_min_particles_per_cell": 2,
    "ms_call_interval": 10, "ms_max_ops": 1e9,
    # this build's keys: merge/split rule (0: merge_split_ops_cmes.h, 1: the
    # legacy ELFS rule of merge_split_ops.h) and the ELFS divergence trigger
    # (ms_config::stretch_thresh, merge_split_ops.h:27)
    "ms_mode": 0, "ms_stretch_thresh": 1e7,
}


def load_data_json(path: str) -> dict:
    """Read DATA.json; arrays become float64 numpy arrays.  Missing required
    keys raise KeyError like nlohmann's j.at() (mpm_data.h:64-97)."""
    with open(path) as f:
        raw = json.load(f)
    return data_from_dict(raw)


def data_from_dict(raw: dict) -> dict:
    d = {}
    required = PARTICLE_KEYS + NODAL_KEYS + (
        "Nex", "Ney", "hx", "hy", "rho", "g", "T", "xi", "mu", "phi", "tauy",
        "BINGHAM_ON", "FRICTION_ON", "CFL", "BC_FLAG", "MERGE_SPLIT_ON",
        "eq_level")
    for k in required:
        if k not in raw:
            raise KeyError(f"DATA.json: missing required key '{k}'")
    for k, v in raw.items():
        if isinstance(v, list):
            d[k] = np.asarray(v, dtype=np.float64)
        else:
            d[k] = v
    d.setdefault("DT_FIXED", 0.0)
    d.setdefault("NSTEPS", 0)
    for k, v in MS_DEFAULTS.items():
        d.setdefault(k, v)
    n = len(d["x"])
    for k in PARTICLE_KEYS:
        if len(d[k]) != n:
            raise ValueError(f"DATA.json: '{k}' has {len(d[k])} entries, x has {n}")
    nn = (int(d["Nex"]) + 1) * (int(d["Ney"]) + 1)
    for k in NODAL_KEYS:
        if len(d[k]) != nn:
            raise ValueError(f"DATA.json: '{k}' has {len(d[k])} entries, expected {nn}")
    return d


def params_from_data(d: dict, max_steps: int = 0) -> dict:
    """Solver parameters (names of amsc_params) from a DATA dict.
    `max_steps` caps an adaptive (DT_FIXED == 0) run, which otherwise runs
    while t < T (optim_par.cpp:281)."""
    return {
        "nrows": int(d["Ney"]), "ncols": int(d["Nex"]),
        "hx": float(d["hx"]), "hy": float(d["hy"]),
        "g": float(d["g"]), "rho": float(d["rho"]), "mu": float(d["mu"]),
        "tauy": float(d["tauy"]), "xi": float(d["xi"]), "phi": float(d["phi"]),
        "bingham_on": float(d["BINGHAM_ON"]),
        "friction_on": float(d["FRICTION_ON"]), "cfl": float(d["CFL"]),
        "eq_level": float(d["eq_level"]), "dt_fixed": float(d["DT_FIXED"]),
        "T": float(d["T"]), "bc_flag": int(d["BC_FLAG"]),
        "nsteps": int(d["NSTEPS"]),
        "max_steps": int(max_steps or d.get("MAX_STEPS", 0)),
        "ms_on": int(d["MERGE_SPLIT_ON"]),
        "ms_alpha": float(d["ms_alpha"]), "ms_beta": float(d["ms_beta"]),
        "ms_split_hp_min": float(d["ms_split_hp_min"]),
        "ms_shear_split": float(d["ms_shear_split"]),
        "ms_hp_min": float(d["ms_hp_min"]), "ms_max_dv": float(d["ms_max_dv"]),
        "ms_max_ops": float(d["ms_max_ops"]),
        "ms_min_level": int(d["ms_min_level"]),
        "ms_max_level": int(d["ms_max_level"]),
        "ms_min_ppc": int(d["ms_min_particles_per_cell"]),
        "ms_interval": int(d["ms_call_interval"]),
        "ms_mode": int(d["ms_mode"]), "ms_stretch": float(d["ms_stretch_thresh"]),
    }


def particles_from_data(d: dict) -> dict:
    return {k: np.ascontiguousarray(d[k], np.float64) for k in PARTICLE_KEYS}


def write_data_json(path: str, d: dict) -> None:
    """Write a DATA dict with %.17g numbers (round-trips exactly)."""
    def fmt(v):
        if isinstance(v, np.ndarray):
            return "[" + ",".join("%.17g" % x for x in v.tolist()) + "]"
        if isinstance(v, (int, np.integer)):
            return str(int(v))
        return "%.17g" % float(v)
    with open(path, "w") as f:
        f.write("{\n")
        items = list(d.items())
        for i, (k, v) in enumerate(items):
            f.write('"%s": %s%s\n' % (k, fmt(v), "," if i + 1 < len(items) else ""))
        f.write("}\n")
