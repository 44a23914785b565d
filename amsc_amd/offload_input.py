"""Input file of amsc_offload_loop, the time loop of the reference's
gpu_offload driver (main/gpu_offload.cpp:276-575, merge/split off) run over
the kernel-level drop-in include/gpu_kernels.h.

Layout (little-endian):
    int32    np nrows ncols nsteps bc_flag
    float64  hx hy dt g rho mu tau_Y cc A_coeff C_coeff tan_fa xi fric_on cfl
             (dt > 0: fixed dt; dt <= 0: adaptive with CFL = cfl)
    float64  the particle arrays of PARTICLE, np each, then Z, (nrows+1)(ncols+1)

    python -m amsc_amd.offload_input [--config slope2m] [--steps 20] [--dt 2e-3] out.bin

writes the step-0 state of one of the bench workloads, initialised by the
device library (optim_par.cpp:235-264), so that the loop can be timed or
profiled on its own:  amsc_amd/_lib/amsc_offload_loop out.bin result.bin
"""
from __future__ import annotations

import argparse
import math

import numpy as np

# particle arrays in the driver's order (the kernels' argument names)
PARTICLE = ["x", "y", "Mp", "Vp", "Ap", "hp", "H", "vpx", "vpy", "mom_px",
            "mom_py", "Fb_x", "Fb_y", "F11", "F12", "F21", "F22", "Fric_px",
            "Fric_py", "dZxp", "dZyp", "Fpx", "Fpy", "apx", "apy", "vpxL",
            "vpyL", "vpx_dx", "vpx_dy", "vpy_dx", "vpy_dy", "Zp", "hpZ"]
SCALARS = ["hx", "hy", "dt", "g", "rho", "mu", "tau_Y", "cc", "A_coeff",
           "C_coeff", "tan_fa", "xi", "fric_on", "cfl"]


def scalars(prm: dict, dt: float) -> dict:
    """The driver's scalar arguments from run parameters (gpu_offload.cpp:
    155-168: A = 3/2, C = 65/32, tan of the friction angle in degrees)."""
    return dict(hx=prm["hx"], hy=prm["hy"], dt=dt, g=prm["g"], rho=prm["rho"],
                mu=prm["mu"], tau_Y=prm["tauy"], cc=prm["bingham_on"],
                A_coeff=3.0 / 2.0, C_coeff=65.0 / 32.0,
                tan_fa=math.tan(prm["phi"] * math.pi / 180.0), xi=prm["xi"],
                fric_on=prm["friction_on"], cfl=prm["cfl"])


def write(path, st: dict, nsteps: int) -> None:
    """Write st (PARTICLE arrays, Z, nrows, ncols, bc_flag and SCALARS) as
    the loop's input for nsteps steps."""
    n = len(st["x"])
    with open(path, "wb") as f:
        np.array([n, st["nrows"], st["ncols"], nsteps, st["bc_flag"]],
                 np.int32).tofile(f)
        np.array([st[k] for k in SCALARS], np.float64).tofile(f)
        for k in PARTICLE:
            a = np.ascontiguousarray(st[k], np.float64)
            if a.shape != (n,):
                raise ValueError(f"particle array {k}: shape {a.shape}, expected ({n},)")
            a.tofile(f)
        nn = (st["nrows"] + 1) * (st["ncols"] + 1)
        z = np.ascontiguousarray(st["Z"], np.float64)
        if z.shape != (nn,):
            raise ValueError(f"Z: shape {z.shape}, expected ({nn},)")
        z.tofile(f)


def initial_state(d: dict, dt: float) -> dict:
    """Step-0 state of a DATA case from the device library's initialisation
    (H, the hydrostatic stress, Ap, momentum); the arrays the loop writes
    before it reads them start at zero, F21 equals F12."""
    from .data import params_from_data, particles_from_data
    from .simulation import Simulation
    prm = params_from_data(d)
    prm["max_steps"] = 0
    sim = Simulation(prm, particles_from_data(d), d["Z"])
    try:
        n = sim.num_particles
        st = {k: np.zeros(n) for k in PARTICLE}
        for k in ("x", "y", "Mp", "Vp", "Ap", "hp", "H", "vpx", "vpy", "mom_px",
                  "mom_py", "Fb_x", "Fb_y", "Fric_px", "Fric_py"):
            st[k] = sim.get(k)
        for k, s in (("F11", "F_11"), ("F12", "F_12"), ("F21", "F_12"), ("F22", "F_22")):
            st[k] = sim.get(s)
    finally:
        sim.close()
    st["Z"] = np.asarray(d["Z"], np.float64)
    st.update(nrows=prm["nrows"], ncols=prm["ncols"], bc_flag=prm["bc_flag"])
    st.update(scalars(prm, dt))
    return st


def main(argv=None) -> None:
    from . import scenarios
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("out")
    ap.add_argument("--config", default="slope2m", choices=["slope2m", "large4m", "dambreak"])
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--dt", type=float, default=2e-3, help="fixed dt; <= 0: adaptive")
    a = ap.parse_args(argv)
    d = {"slope2m": lambda: scenarios.slope_landslide(max_steps=0),
         "large4m": scenarios.large_fractal,
         "dambreak": lambda: scenarios.dambreak_flat(100000)}[a.config]()
    st = initial_state(d, a.dt)
    write(a.out, st, a.steps)
    print(f"wrote {a.out}: {len(st['x'])} particles, grid {st['nrows']}x{st['ncols']}, "
          f"{a.steps} steps", flush=True)


if __name__ == "__main__":
    main()
