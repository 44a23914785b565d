"""Seeded DATA generators for the workloads of this repository.

Every generator returns a DATA dict (the DATA.json schema, main/mpm_data.h)
built with numpy's PCG64 generator from a stated seed, so the same call gives
the same bytes on every machine.  Nodal arrays are column-major with y
fastest, node id r + c*(Ney+1) (the Octave meshgrid + Z(:) order of e.g.
data/dambreak/dambreak.m and data/performance/scaling.m:76-80).

  dambreak_flat()      100x100 cells, hx=0.1, flat bed, 2x2 m column of
                       1600 particles, Bingham mu=0.1 tauy=50, dt=1e-3 x 500.
  slope_landslide()    1000x1000 cells, hx=1, Z = 0.3 x + 8 seeded Gaussian
                       bumps, elliptical release of ~2M particles (4 ppc),
                       hp = 5(1-r^2), Voellmy phi=15 deg xi=500 + Bingham
                       mu=0.5 tauy=200, rho=2000, adaptive CFL=0.4, merge/split
                       every 10 steps.
  large_fractal()      2048x2048 cells, hx=0.5, 25 deg mean slope + seeded
                       fractal noise, polygonal release of ~4M particles,
                       hp in [0.5, 8] smoothed, dt=2e-3 x 1000, merge/split on;
                       still_water=True gives the eq_level=10 lake at rest.
  bench_slope(N)       the data/performance/scaling.m layout at N particles.
  ensemble_scenario()  one member of the hazard ensemble (seeded sizes and
                       rheology).
"""
from __future__ import annotations

import math

import numpy as np

from .data import MS_DEFAULTS


def _grid_nodes(nex, ney, hx, hy):
    """Node coordinates in node-id order r + c*(ney+1)."""
    c = np.repeat(np.arange(nex + 1), ney + 1)
    r = np.tile(np.arange(ney + 1), nex + 1)
    return c * hx, r * hy


def _particles_in_cells(cells_c, cells_r, hx, hy, ppc_side=2):
    """ppc_side^2 particles per cell on a regular sub-grid at offsets
    (k + 0.5)/ppc_side of the cell, cells in the given order."""
    offs = (np.arange(ppc_side) + 0.5) / ppc_side
    oc, orr = np.meshgrid(offs, offs, indexing="ij")
    oc = oc.ravel()
    orr = orr.ravel()
    x = ((cells_c[:, None] + oc[None, :]) * hx).ravel()
    y = ((cells_r[:, None] + orr[None, :]) * hy).ravel()
    return x, y


def _base(nex, ney, hx, hy, x, y, hp, rho, Z, dZdx, dZdy, **kw):
    Ap = np.full(x.shape, hx * hy / 4.0)
    Vp = hp * Ap
    Mp = rho * Vp
    zeros = np.zeros_like(x)
    d = {
        "x": x, "y": y, "Mp": Mp, "Ap": Ap, "vpx": zeros.copy(),
        "vpy": zeros.copy(), "Nex": int(nex), "Ney": int(ney), "hx": float(hx),
        "hy": float(hy), "hp": hp, "mom_px": zeros.copy(),
        "mom_py": zeros.copy(), "g": 9.81, "T": 1e30, "rho": float(rho),
        "Vp": Vp, "xi": 500.0, "Z": Z, "dZdx": dZdx, "dZdy": dZdy, "mu": 0.5,
        "phi": 15.0, "tauy": 200.0, "BINGHAM_ON": 1.0, "FRICTION_ON": 1.0,
        "CFL": 0.4, "BC_FLAG": 1, "MERGE_SPLIT_ON": 1, "eq_level": 0.0,
        "DT_FIXED": 0.0, "NSTEPS": 0,
    }
    d.update(MS_DEFAULTS)
    d.update(kw)
    return d


def dambreak_flat(nsteps=500):
    nex = ney = 100
    hx = hy = 0.1
    cols = np.arange(40, 60)
    cc, rr = np.meshgrid(cols, cols, indexing="ij")
    x, y = _particles_in_cells(cc.ravel(), rr.ravel(), hx, hy)
    hp = np.ones_like(x)
    nn = (nex + 1) * (ney + 1)
    Z = np.zeros(nn)
    return _base(nex, ney, hx, hy, x, y, hp, 1800.0, Z, Z.copy(), Z.copy(),
                 mu=0.1, tauy=50.0, BINGHAM_ON=1.0, FRICTION_ON=0.0, phi=0.0,
                 xi=0.0, DT_FIXED=1e-3, NSTEPS=nsteps, T=nsteps * 1e-3,
                 MERGE_SPLIT_ON=0, BC_FLAG=1, eq_level=0.0)


def _bumps(rng, n, lo_xy, hi_xy, amp, width):
    cx = rng.uniform(lo_xy, hi_xy, n)
    cy = rng.uniform(lo_xy, hi_xy, n)
    A = rng.uniform(amp[0], amp[1], n)
    W = rng.uniform(width[0], width[1], n)
    return cx, cy, A, W


def _bump_field(X, Y, bumps):
    cx, cy, A, W = bumps
    Z = np.zeros_like(X)
    gx = np.zeros_like(X)
    gy = np.zeros_like(X)
    for k in range(len(cx)):
        e = A[k] * np.exp(-((X - cx[k]) ** 2 + (Y - cy[k]) ** 2) / W[k] ** 2)
        Z += e
        gx += e * (-2.0 * (X - cx[k]) / W[k] ** 2)
        gy += e * (-2.0 * (Y - cy[k]) / W[k] ** 2)
    return Z, gx, gy


def slope_landslide(seed=2, n_target=2_000_000, nex=1000, max_steps=2000):
    rng = np.random.default_rng(seed)
    ney = nex
    hx = hy = 1.0
    L = nex * hx
    X, Y = _grid_nodes(nex, ney, hx, hy)
    bumps = _bumps(rng, 8, 0.0, L, (2.0, 20.0), (20.0, 100.0))
    Zb, gxb, gyb = _bump_field(X, Y, bumps)
    Z = 0.3 * X + Zb
    dZdx = 0.3 + gxb
    dZdy = gyb
    # ellipse of area n_target/4 cells, semi-axes a (x) : b (y) = 0.77
    area = n_target / 4.0 * hx * hy
    a = math.sqrt(area / math.pi * 0.77)
    b = area / (math.pi * a)
    cx0, cy0 = 0.6 * L, 0.5 * L
    cc, rr = np.meshgrid(np.arange(nex), np.arange(ney), indexing="ij")
    xc = (cc + 0.5) * hx
    yc = (rr + 0.5) * hy
    inside = ((xc - cx0) / a) ** 2 + ((yc - cy0) / b) ** 2 < 1.0
    x, y = _particles_in_cells(cc[inside], rr[inside], hx, hy)
    r2 = ((x - cx0) / a) ** 2 + ((y - cy0) / b) ** 2
    hp = 5.0 * np.maximum(1.0 - r2, 0.02)
    return _base(nex, ney, hx, hy, x, y, hp, 2000.0, Z, dZdx, dZdy,
                 phi=15.0, xi=500.0, mu=0.5, tauy=200.0, CFL=0.4,
                 DT_FIXED=0.0, T=1e30, NSTEPS=0, MERGE_SPLIT_ON=1,
                 ms_call_interval=10, MAX_STEPS=max_steps)


def _fractal_noise(rng, X, Y, L, octaves=6, amp=6.0, modes=6):
    """Sum of random Fourier modes with amplitude ~ 2^-octave (fBm-like)."""
    Z = np.zeros_like(X)
    gx = np.zeros_like(X)
    gy = np.zeros_like(X)
    for o in range(octaves):
        k0 = 2.0 * math.pi * (2 ** o) * 2.0 / L
        for _ in range(modes):
            th = rng.uniform(0, 2 * math.pi)
            ph = rng.uniform(0, 2 * math.pi)
            kx, ky = k0 * math.cos(th), k0 * math.sin(th)
            a = amp * (0.5 ** o) / math.sqrt(modes)
            arg = kx * X + ky * Y + ph
            Z += a * np.cos(arg)
            s = -a * np.sin(arg)
            gx += s * kx
            gy += s * ky
    return Z, gx, gy


def large_fractal(seed=3, n_target=4_000_000, nex=2048, nsteps=1000,
                  still_water=False):
    rng = np.random.default_rng(seed)
    ney = nex
    hx = hy = 0.5
    L = nex * hx
    X, Y = _grid_nodes(nex, ney, hx, hy)
    slope = math.tan(math.radians(25.0))
    Zn, gxn, gyn = _fractal_noise(rng, X, Y, L)
    Z = slope * X + Zn
    dZdx = slope + gxn
    dZdy = gyn
    cc, rr = np.meshgrid(np.arange(nex), np.arange(ney), indexing="ij")
    xc = (cc + 0.5) * hx
    yc = (rr + 0.5) * hy
    if still_water:
        # lake at rest: eq_level = 10 m, particles where the bilinear depth
        # 10 - Z exceeds 0.05 m; hp set from the same bilinear Z.
        eq = 10.0
        sel = (xc < 40.0)
        x, y = _particles_in_cells(cc[sel], rr[sel], hx, hy)
        Zg = Z.reshape(nex + 1, ney + 1)
        c = np.minimum(np.floor(x / hx).astype(np.int64), nex - 1)
        r = np.minimum(np.floor(y / hy).astype(np.int64), ney - 1)
        xi_ = (x - c * hx) / hx
        eta = (y - r * hy) / hy
        zp = ((1 - xi_) * (1 - eta) * Zg[c, r] + (1 - xi_) * eta * Zg[c, r + 1]
              + xi_ * (1 - eta) * Zg[c + 1, r] + xi_ * eta * Zg[c + 1, r + 1])
        keep = (eq - zp) > 0.05
        x, y = x[keep], y[keep]
        hp = eq - zp[keep]
        # rheology and boundaries off as in data/well_balance/well_balance.m
        return _base(nex, ney, hx, hy, x, y, hp, 2000.0, Z, dZdx, dZdy,
                     phi=15.0, xi=500.0, mu=0.5, tauy=200.0, DT_FIXED=2e-3,
                     NSTEPS=nsteps, T=nsteps * 2e-3, MERGE_SPLIT_ON=0,
                     BC_FLAG=0, eq_level=eq, BINGHAM_ON=0.0, FRICTION_ON=0.0)
    # star-shaped polygon of area n_target/4 cells
    area = n_target / 4.0 * hx * hy
    nv = 11
    ang = np.sort(rng.uniform(0, 2 * math.pi, nv))
    rad = rng.uniform(0.75, 1.25, nv)
    cx0, cy0 = 0.62 * L, 0.5 * L
    th = np.arctan2(yc - cy0, xc - cx0) % (2 * math.pi)
    # piecewise-linear radius in angle
    angx = np.concatenate([ang - 2 * math.pi, ang, ang + 2 * math.pi])
    radx = np.concatenate([rad, rad, rad])
    rpoly = np.interp(th, angx, radx)
    dist = np.hypot(xc - cx0, yc - cy0)
    # scale so that the selected area matches the target
    ratio = np.sort((dist / rpoly).ravel())
    ncell = int(round(area / (hx * hy)))
    R = ratio[min(ncell, ratio.size - 1)]
    inside = dist / rpoly < R
    x, y = _particles_in_cells(cc[inside], rr[inside], hx, hy)
    # smoothed random depth in [0.5, 8]
    hn, _, _ = _fractal_noise(rng, x, y, L, octaves=3, amp=1.0, modes=4)
    hn = (hn - hn.min()) / max(hn.max() - hn.min(), 1e-12)
    hp = 0.5 + 7.5 * hn
    return _base(nex, ney, hx, hy, x, y, hp, 2000.0, Z, dZdx, dZdy,
                 phi=15.0, xi=500.0, mu=0.5, tauy=200.0, DT_FIXED=2e-3,
                 NSTEPS=nsteps, T=nsteps * 2e-3, MERGE_SPLIT_ON=1,
                 ms_call_interval=10)


def bench_slope(n=1000, dt_fixed=1e-3, nsteps=100):
    """data/performance/scaling.m:60-123 layout (4 ppc, 2 deg slope, h0=5)."""
    hx = hy = 1.0
    ncell = int(round(math.sqrt(n) / 2))
    nex = ney = ncell
    Lx, Ly = nex * hx, ney * hy
    X, Y = _grid_nodes(nex, ney, hx, hy)
    slope = math.tan(2 * math.pi / 180)
    Z = 100.0 - slope * X
    dZdx = np.full_like(Z, -slope)
    dZdy = np.zeros_like(Z)
    s = 0.5
    xp = np.arange(1.0 + s / 2, Lx - 1.0 + 1e-9, s)
    yp = np.arange(1.0 + s / 2, Ly - 1.0 + 1e-9, s)
    XP, YP = np.meshgrid(xp, yp, indexing="ij")
    x, y = XP.ravel(), YP.ravel()
    hp = np.full_like(x, 5.0)
    d = _base(nex, ney, hx, hy, x, y, hp, 1200.0, Z, dZdx, dZdy, g=9.81,
              T=20.0, xi=200.0, mu=2000.0, phi=20.0, tauy=30.0,
              BINGHAM_ON=1.0, FRICTION_ON=1.0, CFL=0.2, BC_FLAG=1,
              eq_level=0.0, DT_FIXED=dt_fixed, MERGE_SPLIT_ON=1,
              NSTEPS=nsteps)
    d["Ap"] = np.full_like(x, s * s)
    d["Vp"] = d["Ap"] * hp
    d["Mp"] = 1200.0 * d["Vp"]
    return d


def ensemble_scenario(seed, side_range=(64, 1024), np_range=(1e4, 1e6),
                      np_cap=None):
    """One hazard-ensemble member: grid side ~ logU(side_range), particles
    ~ logU(np_range) at 4 ppc clamped to 2*side^2 (half the cells), Gaussian
    bump slope, phi ~ atan(U(tan10, tan30)) deg, xi ~ U(200, 2000),
    mu ~ U(0.01, 1), tauy ~ U(1, 500) (tauy > 0: optim_par divides by it)."""
    rng = np.random.default_rng(seed)
    side = int(round(math.exp(rng.uniform(math.log(side_range[0]), math.log(side_range[1])))))
    npart = math.exp(rng.uniform(math.log(np_range[0]), math.log(np_range[1])))
    npart = min(npart, 2.0 * side * side)
    if np_cap:
        npart = min(npart, np_cap)
    hx = hy = 1.0
    L = side * hx
    X, Y = _grid_nodes(side, side, hx, hy)
    bumps = _bumps(rng, 4, 0.0, L, (1.0, 0.05 * L + 2.0), (0.05 * L + 5, 0.2 * L + 10))
    Zb, gxb, gyb = _bump_field(X, Y, bumps)
    sl = rng.uniform(0.1, 0.4)
    Z = sl * X + Zb
    ncell = max(int(npart / 4), 4)
    # square-ish release centred at 0.65 L
    a = int(math.ceil(math.sqrt(ncell)))
    c0 = int(0.65 * side) - a // 2
    r0 = side // 2 - a // 2
    c0 = max(1, min(c0, side - a - 1))
    r0 = max(1, min(r0, side - a - 1))
    cc, rr = np.meshgrid(np.arange(c0, c0 + a), np.arange(r0, r0 + a), indexing="ij")
    cc, rr = cc.ravel()[:ncell], rr.ravel()[:ncell]
    x, y = _particles_in_cells(cc, rr, hx, hy)
    hp = rng.uniform(1.0, 5.0) * np.ones_like(x)
    phi = math.degrees(math.atan(rng.uniform(math.tan(math.radians(10)), math.tan(math.radians(30)))))
    return _base(side, side, hx, hy, x, y, hp, 2000.0, Z, sl + gxb, gyb,
                 phi=phi, xi=float(rng.uniform(200, 2000)),
                 mu=float(rng.uniform(0.01, 1.0)), tauy=float(rng.uniform(1.0, 500.0)),
                 CFL=0.4, DT_FIXED=0.0, T=float(rng.uniform(20.0, 120.0)),
                 MERGE_SPLIT_ON=1, ms_call_interval=10)
