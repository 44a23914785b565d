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
  conservation_case()  data/conservation_test/conservation_test.m: dam break
                       on an inclined plane with a Gaussian obstacle, 75x30
                       cells over [0,150]x[0,20].
  well_balance_case()  data/well_balance/well_balance.m: lake at rest
                       (eq_level = 10) over a Gaussian bump, 50x50 cells.
  dambreak_case()      data/dambreak/dambreak.m: a 5 m column released down
                       a 200x60-cell channel sloping at 5 %.

The last three restate the reference's Octave generators, including Octave's
meshgrid/colon ordering, gradient() and interp2(..., 'linear') (the plane
form a + b s + c t + d s t); Octave is not available here, so equality
with the files Octave would write is not checked.
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


def _octave_range(a, inc, b):
    """Octave's a:inc:b (elements a + k*inc)."""
    n = int(math.floor((b - a) / inc * (1.0 + 4 * np.finfo(float).eps))) + 1
    return a + np.arange(n) * inc


def _octave_gradient(M, hx, hy):
    """gradient(M, hx, hy) of Octave for a matrix indexed (row = y, column =
    x): central differences inside, one-sided at the edges."""
    def along(A, h):  # along axis 1
        x = np.arange(A.shape[1]) * h
        out = np.empty_like(A)
        out[:, 0] = (A[:, 1] - A[:, 0]) / (x[1] - x[0])
        out[:, -1] = (A[:, -1] - A[:, -2]) / (x[-1] - x[-2])
        out[:, 1:-1] = (A[:, 2:] - A[:, :-2]) / (x[2:] - x[:-2])
        return out
    return along(M, hx), along(M.T, hy).T


def _octave_interp2(xv, yv, M, xi, yi):
    """interp2(X, Y, M, xi, yi, 'linear') of Octave on a meshgrid (xv along
    columns, yv along rows): the containing cell (points on the last grid
    line use the last cell), then a + b s + c t + d s t."""
    j = np.clip(np.searchsorted(xv, xi, side="right") - 1, 0, len(xv) - 2)
    i = np.clip(np.searchsorted(yv, yi, side="right") - 1, 0, len(yv) - 2)
    s = (xi - xv[j]) / (xv[j + 1] - xv[j])
    t = (yi - yv[i]) / (yv[i + 1] - yv[i])
    a = M[i, j]
    b = M[i, j + 1] - a
    c = M[i + 1, j] - a
    d = M[i + 1, j + 1] - a - b - c
    return a + b * s + c * t + d * s * t


def _colmajor(M):
    """M(:) of Octave: column-major flattening (node id r + c*(rows))."""
    return np.ascontiguousarray(M.T).ravel()


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


def conservation_case():
    """data/conservation_test/conservation_test.m:8-76 and its DATA.json keys
    (:107-146): domain [0,150]x[0,20], 75x30 cells, Z = 60 - x/3 +
    15 exp(-((x-50)^2 + (y-10)^2)/7), particles on a 0.45 m lattice in
    [0,20]^2 with hp = 60 - Z(xp) > 1 away from the boundary cells, equal
    masses, Bingham + Voellmy, adaptive CFL 0.2 to T = 20, merge/split on."""
    nex, ney, Lx, Ly = 75, 30, 150.0, 20.0
    hx, hy = Lx / nex, Ly / ney
    xv = np.linspace(0, Lx, nex + 1)
    yv = np.linspace(0, Ly, ney + 1)
    X, Y = np.meshgrid(xv, yv)
    Z = 60 - (1 / 3) * X + 15 * np.exp(-((X - 50) ** 2) / 7 - ((Y - 10) ** 2) / 7)
    gx, gy = _octave_gradient(Z, hx, hy)
    DX = DY = 0.45
    xp, yp = np.meshgrid(_octave_range(DX / 2, DX, 20 - DX / 2),
                         _octave_range(DY / 2, DY, 20 - DY / 2))
    xp, yp = _colmajor(xp), _colmajor(yp)
    hp = 60 - _octave_interp2(xv, yv, Z, xp, yp)
    keep = hp > 1.0
    xp, yp, hp = xp[keep], yp[keep], hp[keep]
    edge = (xp < hx) | (xp > Lx - hx) | (yp < hy) | (yp > Ly - hy)
    xp, yp, hp = xp[~edge], yp[~edge], hp[~edge]
    n = len(xp)
    rho = 1200.0
    Msys = np.sum(hp * DX * DY * rho)
    Mp = Msys / n * np.ones(n)
    Vp = Mp / rho
    Ap = Vp / hp
    d = _base(nex, ney, hx, hy, xp, yp, hp, rho, _colmajor(Z), _colmajor(gx),
              _colmajor(gy), xi=200.0, T=20.0, mu=10.0, phi=12.0, tauy=30.0,
              MERGE_SPLIT_ON=1, BINGHAM_ON=1.0, CFL=0.2, BC_FLAG=1,
              eq_level=0.0, DT_FIXED=0.0, NSTEPS=100, FRICTION_ON=1.0,
              ms_alpha=1.5, ms_split_hp_min=0.3, ms_beta=0.3)
    d.update(Mp=Mp, Vp=Vp, Ap=Ap)
    return d


def well_balance_case():
    """data/well_balance/well_balance.m:6-125 and its DATA.json keys
    (:186-226): 10x10 m, 50x50 cells, Z = 5 exp(-0.4((x-5)^2 + (y-5)^2)),
    eq_level = 10, particles on a 0.17 m lattice over [0,10]^2 with
    hp = interp2(10 - Z), Ap = 0.17^2, Vp = Ap hp, rho = 1200, no rheology,
    no boundary conditions, adaptive CFL 0.1 to T = 1, merge/split on."""
    Lx = Ly = 10.0
    nex = ney = 50
    hx, hy = Lx / nex, Ly / ney
    xv = _octave_range(0.0, hx, Lx)
    yv = _octave_range(0.0, hy, Ly)
    X, Y = np.meshgrid(xv, yv)
    Z = 5.0 * np.exp(-2 / 5 * ((X - 5) ** 2 + (Y - 5) ** 2))
    gx, gy = _octave_gradient(Z, hx, hy)
    h = 10 - Z
    h[h < 0] = 0
    DX = DY = 0.17
    xp, yp = np.meshgrid(_octave_range(0.0, DX, 10.0), _octave_range(0.0, DY, 10.0))
    hp = _colmajor(_octave_interp2(xv, yv, h, xp, yp))
    xp, yp = _colmajor(xp), _colmajor(yp)
    n = len(xp)
    rho = 1200.0
    Ap = DX * DY * np.ones(n)
    Vp = Ap * hp
    Mp = rho * Vp
    d = _base(nex, ney, hx, hy, xp, yp, hp, rho, _colmajor(Z), _colmajor(gx),
              _colmajor(gy), xi=200.0, T=1.0, mu=2000.0, phi=20.0, tauy=30.0,
              BC_FLAG=0, CFL=0.1, BINGHAM_ON=0.0, FRICTION_ON=0.0, DT_FIXED=0.0,
              NSTEPS=100, MERGE_SPLIT_ON=1, ms_alpha=1.3, ms_beta=0.3,
              eq_level=10.0)
    d.update(Mp=Mp, Vp=Vp, Ap=Ap)
    return d


def dambreak_case():
    """data/dambreak/dambreak.m:7-113 and its DATA.json keys (:171-204):
    [0,100]x[0,30] m, 200x60 cells, Z = max(5 - 0.05 x, 0), particles on a
    0.5 m lattice with hp = 5 where 1 <= x <= 50, equal masses, rho = 1000,
    no rheology, adaptive CFL 0.2 to T = 8, merge/split on."""
    nex, ney = 200, 60
    xv = np.linspace(0, 100, nex + 1)
    yv = np.linspace(0, 30, ney + 1)
    hx, hy = 100 / nex, 30 / ney
    X, Y = np.meshgrid(xv, yv)
    Z = 5 - 0.05 * X
    Z[Z < 0] = 0
    DX = DY = 0.5
    xp, yp = np.meshgrid(_octave_range(0.0, DX, 100.0), _octave_range(0.0, DY, 30.0))
    hp = 5.0 * np.ones_like(xp)
    hp[(xp < 1) | (xp > 50)] = 0
    hp, xp, yp = _colmajor(hp), _colmajor(xp), _colmajor(yp)
    keep = hp > 0.001
    hp, xp, yp = hp[keep], xp[keep], yp[keep]
    gx, gy = _octave_gradient(Z, hx, hy)
    n = len(xp)
    rho = 1000.0
    Msys = np.sum(hp * DX * DY * rho)
    Mp = Msys / n * np.ones(n)
    Vp = Mp / rho
    Ap = Vp / hp
    d = _base(nex, ney, hx, hy, xp, yp, hp, rho, _colmajor(Z), _colmajor(gx),
              _colmajor(gy), xi=200.0, T=8.0, mu=50.0, phi=37.0, tauy=2000.0,
              MERGE_SPLIT_ON=1, ms_alpha=1.3, ms_beta=0.3, BINGHAM_ON=0.0,
              FRICTION_ON=0.0, CFL=0.2, eq_level=0.0, BC_FLAG=1,
              DT_FIXED=0.0, NSTEPS=0)
    d.update(Mp=Mp, Vp=Vp, Ap=Ap)
    return d
