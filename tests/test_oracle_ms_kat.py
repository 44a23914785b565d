"""Known-answer tests of the oracle's merge/split against the reference's
rules, derived by hand from the reference source with exactly representable
inputs (no expected value is computed with the oracle's formulas).

The solvers' rule, merge_split_ops_cmes.h:
* decide_actions (:116-253): the per-particle marking at its thresholds
  (strict comparisons, the if/else order, the level guards, the split gate
  on thin material), the per-cell split cap in ascending cell order and
  ascending id within a cell, the greedy nearest-partner pairing (strict
  `<` on d2 and on |dv|^2, first-wins ties, only later unpaired candidates,
  the merge cap), unpaired candidates kept, and the min-particles-per-cell
  undo of the first primary in id order;
* execute_merge_split (:258-427) with principal_stretch_dir (:93-110):
  output order and offsets, the split children at +-0.25 sqrt(Ap) along the
  principal stretch direction (including the direction the reference's
  eigenvector formula returns when the larger strain is along y, see
  test_split_direction_quirk), clamping at the domain edge, halved
  extensive and copied intensive fields, Ap = Vp/hp, mom = Mp v, the merge
  centre of mass, summed extensive and mass-averaged intensive fields,
  hp = Vp/Ap, levels and labels, and the exact mass and momentum through a
  split.

The legacy ELFS rule, merge_split_ops.h:
* decide_actions (:349-510) with compute_elfs (:266-342): the elfs value of
  a particle from the cell distance and its refinement near exterior cells
  of its row, the split / merge thresholds at equality, the level guard
  without fall-through to merging, the divergence trigger, and the 1 -> 4
  split weight in the min-particles-per-cell rule;
* execute_merge_split (:513-716): four children at (x +- r/2, y +- r/2),
  quartered extensive fields, the primary's label kept on a merge.
"""
import math

import numpy as np
import pytest

from oracle import oracle

KEEP, SPLIT, PRIMARY, SECONDARY = 0, 1, 2, 3

# 8 x 8 cells of 1 m: min_h = 1, so r_split = alpha, r_merge = beta
BASE = {"nrows": 8, "ncols": 8, "hx": 1.0, "hy": 1.0, "g": 9.81, "rho": 2000.0,
        "mu": 0.1, "tauy": 50.0, "xi": 500.0, "phi": 15.0, "bingham_on": 1.0,
        "friction_on": 1.0, "cfl": 0.4, "eq_level": 0.0, "dt_fixed": 0.0, "T": 1.0,
        "bc_flag": 1, "nsteps": 1, "max_steps": 0, "ms_on": 1,
        "ms_alpha": 1.5, "ms_beta": 0.25, "ms_split_hp_min": 0.25,
        "ms_shear_split": 0.25, "ms_hp_min": 0.125, "ms_max_dv": 0.5,
        "ms_max_ops": 1e9, "ms_min_level": -2, "ms_max_level": 2, "ms_min_ppc": 2,
        "ms_interval": 10, "ms_mode": 0, "ms_stretch": 1e7}
DT = 0.5

NEUTRAL = {"Ap": 0.5625, "hp": 1.0}     # r = 0.75: neither stretched nor thick
CAND = {"Ap": 0.015625, "hp": 1.0}      # r = 0.125 < 0.25: a merge candidate


def P(cell, dx=0.5, dy=0.5, **kw):
    """A particle in cell (c, r) at offset (dx, dy) inside it."""
    c, r = cell
    p = {"x": c + dx, "y": r + dy, "Ap": NEUTRAL["Ap"], "hp": NEUTRAL["hp"],
         "vpx_dx": 0.0, "vpx_dy": 0.0, "vpy_dx": 0.0, "vpy_dy": 0.0,
         "vpx": 0.0, "vpy": 0.0, "level": 0}
    p.update(kw)
    return p


def decide(parts, **over):
    prm = dict(BASE, **over)
    a = {k: np.array([p[k] for p in parts], np.float64)
         for k in ("x", "y", "Ap", "hp", "vpx_dx", "vpx_dy", "vpy_dx", "vpy_dy", "vpx", "vpy")}
    lev = np.array([p["level"] for p in parts], np.int32)
    return oracle.ms_decide(prm, DT, a["x"], a["y"], a["Ap"], a["hp"], a["vpx_dx"],
                            a["vpx_dy"], a["vpy_dx"], a["vpy_dy"], a["vpx"], a["vpy"], lev)


# ---------------------------------------------------------------- marking
def test_split_marking_at_its_thresholds():
    parts = [
        P((0, 0), Ap=2.25),                          # 0: r = 1.5 = r_split: not >
        P((0, 1), Ap=2.25 + 1e-9),                   # 1: r > r_split: split
        P((0, 2), vpx_dy=0.25, vpy_dx=0.25),         # 2: gdot dt = 0.25 = shear_split
        P((0, 3), vpx_dy=0.5, vpy_dx=0.5),           # 3: gdot dt = 0.5 > 0.25
        P((0, 4), vpx_dy=1.0),                       # 4: exy = (1 + 0)/2: gdot dt 0.5
        P((0, 5), vpy_dx=0.5),                       # 5: exy = 0.25: gdot dt 0.25, not >
        P((0, 6), vpx_dx=0.25),                      # 6: gdot = sqrt(2/16): 0.177 dt-scaled
        P((0, 7), vpx_dx=0.5),                       # 7: gdot dt = sqrt(1/8) > 0.25
        P((1, 0), Ap=4.0, hp=0.25),                  # 8: hp = split_hp_min: not thin
        P((1, 1), Ap=4.0, hp=0.25 - 1e-12),          # 9: thin: no split, not thick: keep
        P((1, 2), Ap=4.0, level=-2),                 # 10: level = min_level: no split
        P((1, 3), Ap=4.0, level=-1),                 # 11: above min_level: split
        P((1, 4), Ap=CAND["Ap"], vpx_dy=1.0, vpy_dx=1.0),  # 12: sheared and thick: split first
        # the shear rate averages both off-diagonal gradients
        P((2, 0), vpy_dx=1.0),                       # 13: exy = 0.5: gdot dt 0.5
        P((2, 1), vpx_dy=0.5),                       # 14: exy = 0.25: not >
    ]
    act, partner, counts = decide(parts)
    want = [KEEP, SPLIT, KEEP, SPLIT, SPLIT, KEEP, KEEP, SPLIT, SPLIT, KEEP, KEEP, SPLIT, SPLIT,
            SPLIT, KEEP]
    assert act.tolist() == want
    assert counts.tolist() == [want.count(SPLIT), 0]
    assert (partner == -1).all()


def merge_probe_cell(cell, probe):
    """A neutral particle, the probe, then a sure candidate 0.25 m from it,
    co-moving: the probe merges (as primary, the lower id) iff it is a
    candidate; the cell keeps 2 particles either way."""
    c, r = cell
    return [P(cell), dict(P(cell, 0.25, 0.25), **probe),
            P(cell, 0.5, 0.25, **CAND)]


@pytest.mark.parametrize("probe,merges", [
    ({"Ap": 0.0625}, False),                       # r = 0.25 = r_merge: not <
    ({"Ap": 0.0625 - 1e-9}, True),                 # r < r_merge
    ({"Ap": CAND["Ap"], "hp": 0.125}, False),      # hp = hp_min: not >
    ({"Ap": CAND["Ap"], "hp": 0.125 + 1e-12}, True),
    ({"Ap": CAND["Ap"], "level": 2}, False),       # level = max_level: not <
    ({"Ap": CAND["Ap"], "level": 1}, True),
    # thin material is not split (hp < split_hp_min) and falls through to
    # the merge test: sheared + thick + thin -> merge candidate
    ({"Ap": CAND["Ap"], "hp": 0.2, "vpx_dy": 1.0, "vpy_dx": 1.0}, True),
    # a split blocked by the level guard also falls through to the merge test
    ({"Ap": CAND["Ap"], "level": -2, "vpx_dy": 1.0, "vpy_dx": 1.0}, True),
])
def test_merge_marking_at_its_thresholds(probe, merges):
    parts = merge_probe_cell((3, 3), probe)
    act, partner, counts = decide(parts)
    if merges:
        assert act.tolist() == [KEEP, PRIMARY, SECONDARY]
        assert partner.tolist() == [-1, 2, -1]
        assert counts.tolist() == [0, 1]
    else:
        assert act.tolist() == [KEEP, KEEP, KEEP]
        assert counts.tolist() == [0, 0]


# ---------------------------------------------------------------- pairing
def cell_with(cell, cands, neutrals=1):
    """neutrals kept particles (the lower ids), then merge candidates."""
    return [P(cell, 0.9, 0.9) for _ in range(neutrals)] + [
        P(cell, dx, dy, Ap=CAND["Ap"], hp=1.0, **kw) for dx, dy, kw in cands]


def test_pairing_first_wins_a_tie():
    # a at the centre, b and c 0.25 m right and left: equal d2 (0.0625); b
    # comes first, and a later equal d2 does not replace it (strict <)
    parts = cell_with((2, 2), [(0.5, 0.5, {}), (0.75, 0.5, {}), (0.25, 0.5, {})])
    act, partner, counts = decide(parts)
    assert act.tolist() == [KEEP, PRIMARY, SECONDARY, KEEP]
    assert partner.tolist() == [-1, 2, -1, -1]
    assert counts.tolist() == [0, 1]


@pytest.mark.parametrize("dvx,merges", [(0.5, False), (0.5 - 2.0 ** -20, True),
                                        (-0.5, False)])
def test_pairing_velocity_filter_is_strict(dvx, merges):
    # |dv|^2 = max_dv^2 = 0.25 exactly is rejected
    parts = cell_with((4, 1), [(0.25, 0.5, {"vpx": 0.0}), (0.75, 0.5, {"vpx": dvx})])
    act, _, counts = decide(parts)
    assert act.tolist() == ([KEEP, PRIMARY, SECONDARY] if merges else [KEEP] * 3)
    assert counts[1] == int(merges)


def test_pairing_takes_the_nearest_later_unpaired_candidate():
    # a (id 1) is nearest to c (id 3); then b (id 2) would prefer c, which
    # is paired, and takes d (id 4)
    parts = cell_with((5, 5), [(0.10, 0.10, {}),   # a: b at 0.2, c at 0.1
                               (0.30, 0.10, {}),   # b: c at 0.1, d at 0.6
                               (0.20, 0.10, {}),   # c
                               (0.90, 0.10, {})])  # d
    act, partner, counts = decide(parts)
    assert act.tolist() == [KEEP, PRIMARY, PRIMARY, SECONDARY, SECONDARY]
    assert partner.tolist() == [-1, 3, 4, -1, -1]
    assert counts.tolist() == [0, 2]


def test_pairing_skips_a_nearer_candidate_moving_apart():
    # b is nearer to a but moves at |dv| = 1; c is farther and co-moving
    parts = cell_with((6, 2), [(0.1, 0.5, {}), (0.2, 0.5, {"vpy": 1.0}), (0.6, 0.5, {})])
    act, partner, _ = decide(parts)
    assert act.tolist() == [KEEP, PRIMARY, KEEP, SECONDARY]
    assert partner.tolist() == [-1, 3, -1, -1]


def test_unpaired_candidate_is_kept_and_candidates_look_forward_only():
    # three candidates: a pairs with b; c has no later candidate left
    parts = cell_with((1, 6), [(0.2, 0.2, {}), (0.3, 0.2, {}), (0.8, 0.8, {})], neutrals=0)
    act, partner, counts = decide(parts)
    assert act.tolist() == [PRIMARY, SECONDARY, KEEP]
    assert counts.tolist() == [0, 1]


# ---------------------------------------------------------------- caps
def test_split_cap_counts_in_cell_order_then_id_order():
    stretched = {"Ap": 4.0}
    parts = [P((2, 4), **stretched),   # id 0: cell 4 + 2*8 = 20
             P((0, 3), **stretched),   # id 1: cell 3
             P((0, 3), 0.2, 0.2, **stretched),  # id 2: cell 3
             P((1, 2), **stretched)]   # id 3: cell 10
    act, _, counts = decide(parts, ms_max_ops=2)
    assert act.tolist() == [KEEP, SPLIT, SPLIT, KEEP]
    assert counts.tolist() == [2, 0]
    act, _, counts = decide(parts, ms_max_ops=3)
    assert act.tolist() == [KEEP, SPLIT, SPLIT, SPLIT]
    act, _, counts = decide(parts, ms_max_ops=0)
    assert act.tolist() == [KEEP] * 4 and counts.tolist() == [0, 0]


def test_merge_cap_counts_in_cell_order_and_is_separate_from_splits():
    pair_hi = cell_with((5, 6), [(0.2, 0.2, {}), (0.3, 0.2, {})])   # cell 46
    pair_lo = cell_with((1, 1), [(0.2, 0.2, {}), (0.3, 0.2, {})])   # cell 9
    split = [P((6, 6), Ap=4.0)]                                      # cell 54
    parts = pair_hi + pair_lo + split
    act, partner, counts = decide(parts, ms_max_ops=1)
    # the lower cell's pair merges; the higher cell's candidates are kept;
    # the split has its own count
    assert act.tolist() == [KEEP, KEEP, KEEP, KEEP, PRIMARY, SECONDARY, SPLIT]
    assert partner[4] == 5
    assert counts.tolist() == [1, 1]


# ---------------------------------------------------------------- min ppc
def test_min_ppc_undoes_a_lone_pair():
    parts = cell_with((3, 5), [(0.2, 0.2, {}), (0.3, 0.2, {})], neutrals=0)
    act, partner, counts = decide(parts)
    assert act.tolist() == [KEEP, KEEP] and (partner == -1).all()
    assert counts.tolist() == [0, 0]


def test_min_ppc_undoes_the_first_primary_in_id_order():
    # a-d nearest, b-c next: primaries a and b; with min_ppc 3 the cell
    # projects 2 particles, and a's merge (the first primary) is undone
    cands = [(0.10, 0.5, {}), (0.45, 0.5, {}), (0.55, 0.5, {}), (0.15, 0.5, {})]
    parts = cell_with((4, 4), cands, neutrals=0)
    act, partner, counts = decide(parts, ms_min_ppc=3)
    assert act.tolist() == [KEEP, PRIMARY, SECONDARY, KEEP]
    assert partner.tolist() == [-1, 2, -1, -1]
    assert counts.tolist() == [0, 1]
    # with min_ppc 2 both merges stand
    act, partner, counts = decide(parts)
    assert act.tolist() == [PRIMARY, PRIMARY, SECONDARY, SECONDARY]
    assert partner.tolist() == [3, 2, -1, -1]


def test_min_ppc_counts_a_split_as_two():
    parts = [P((2, 6), Ap=4.0)] + cell_with((2, 6), [(0.2, 0.2, {}), (0.3, 0.2, {})],
                                             neutrals=0)
    act, _, counts = decide(parts)
    assert act.tolist() == [SPLIT, PRIMARY, SECONDARY]   # projected 2 + 1
    act, _, counts = decide(parts, ms_min_ppc=4)
    assert act.tolist() == [SPLIT, KEEP, KEEP]            # 3 < 4: undone
    assert counts.tolist() == [1, 0]


# ---------------------------------------------------------------- execution
NAMES = oracle.field_names()
EXT = {"Mp", "Vp", "Ap", "mom_px", "mom_py"}


def snapshot(parts):
    """Every oracle field: given values, the rest distinct per particle."""
    n = len(parts)
    f = {k: np.arange(n, dtype=np.float64) * 0.5 + 1.0 + 0.01 * j
         for j, k in enumerate(NAMES)}
    for i, p in enumerate(parts):
        for k, v in p.items():
            f[k][i] = v
    return f


def execute(parts, act, partner, labels=None, levels=None, elfs=False, **over):
    prm = dict(BASE, **over)
    f = snapshot(parts)
    n = len(parts)
    lab = np.arange(100, 100 + n, dtype=np.int32) if labels is None else np.array(labels, np.int32)
    lev = np.zeros(n, np.int32) if levels is None else np.array(levels, np.int32)
    out, olab, olev = oracle.ms_execute(prm, np.array(act, np.int32),
                                        np.array(partner, np.int32), f, lab, lev, elfs=elfs)
    return f, out, olab, olev


def test_execution_order_offsets_labels_and_levels():
    parts = [{"x": 0.5, "y": 0.5}, {"x": 2.5, "y": 2.5, "vpx_dx": 2.0, "vpy_dy": 1.0,
                                    "vpx_dy": 0.0, "vpy_dx": 0.0, "Ap": 0.25},
             {"x": 4.25, "y": 4.5}, {"x": 6.5, "y": 6.5}, {"x": 4.75, "y": 4.5}]
    act = [KEEP, SPLIT, PRIMARY, KEEP, SECONDARY]
    partner = [-1, -1, 4, -1, -1]
    f, out, lab, lev = execute(parts, act, partner, levels=[0, 1, -1, 2, 0])
    assert len(out["x"]) == 5                       # weights 1 + 2 + 1 + 1 + 0
    for k in NAMES:                                 # kept particles copied as is
        assert out[k][0] == f[k][0] and out[k][4] == f[k][3], k
    assert lab.tolist() == [100, -1, -1, -1, 103]
    assert lev.tolist() == [0, 0, 0, 0, 2]


def test_split_children_along_a_diagonal_strain():
    # exx = 2 > eyy = 1, no shear: eigenvalue 2, eigenvector (1, 0);
    # sqrt(Ap) = 0.5, so the children sit 0.125 m either side along x
    p = {"x": 3.5, "y": 2.5, "Ap": 0.25, "Vp": 0.75, "hp": 1.5, "Mp": 1500.0,
         "vpx": 2.0, "vpy": -1.0, "vpx_dx": 2.0, "vpy_dy": 1.0, "vpx_dy": 0.0,
         "vpy_dx": 0.0, "F_11": -7.0, "H": 0.25}
    f, out, lab, lev = execute([p], [SPLIT], [-1], levels=[0])
    assert out["x"].tolist() == [3.625, 3.375] and out["y"].tolist() == [2.5, 2.5]
    assert out["xp"].tolist() == [3.625, 3.375] and out["yp"].tolist() == [2.5, 2.5]
    assert out["Mp"].tolist() == [750.0, 750.0] and out["Vp"].tolist() == [0.375, 0.375]
    assert out["Ap"].tolist() == [0.25, 0.25]            # (Vp/2) / hp
    assert out["mom_px"].tolist() == [1500.0, 1500.0]    # (Mp/2) vpx
    assert out["mom_py"].tolist() == [-750.0, -750.0]
    for k in ("hp", "vpx", "vpy", "F_11", "H", "vpx_dx", "Fb_x"):   # intensive
        assert out[k].tolist() == [f[k][0]] * 2, k
    assert lev.tolist() == [-1, -1] and lab.tolist() == [-1, -1]
    # mass and momentum exactly conserved by a split
    assert out["Mp"].sum() == p["Mp"]
    assert out["mom_px"].sum() == p["Mp"] * p["vpx"]


def test_split_direction_quirk():
    """exx = 1 < eyy = 3, no shear: the larger eigenvalue's direction is y,
    but the reference's eigenvector (l1 - eyy, exy) is (0, 0) there and its
    fallback for a vanishing vector is (1, 0) (merge_split_ops_cmes.h:
    102-109): the children are placed along x."""
    p = {"x": 3.5, "y": 2.5, "Ap": 0.25, "Vp": 0.25, "hp": 1.0, "Mp": 10.0,
         "vpx_dx": 1.0, "vpy_dy": 3.0, "vpx_dy": 0.0, "vpy_dx": 0.0}
    _, out, _, _ = execute([p], [SPLIT], [-1])
    assert out["x"].tolist() == [3.625, 3.375] and out["y"].tolist() == [2.5, 2.5]


@pytest.mark.parametrize("g", [0.5, -0.5])
def test_split_children_along_a_pure_shear(g):
    # exx = eyy = 0, exy = g: eigenvalue |g|, direction (1, sign(g))/sqrt 2
    p = {"x": 3.5, "y": 2.5, "Ap": 1.0, "Vp": 1.0, "hp": 1.0, "Mp": 10.0,
         "vpx_dx": 0.0, "vpy_dy": 0.0, "vpx_dy": g, "vpy_dx": g}
    _, out, _, _ = execute([p], [SPLIT], [-1])
    d = 0.25 / math.sqrt(2.0)
    s = 1.0 if g > 0 else -1.0
    np.testing.assert_allclose(out["x"], [3.5 + d, 3.5 - d], rtol=0, atol=1e-15)
    np.testing.assert_allclose(out["y"], [2.5 + s * d, 2.5 - s * d], rtol=0, atol=1e-15)


def test_split_children_clamped_to_the_domain():
    # Lx = Ly = 8; children 0.125 m either side along x
    lo = {"x": 0.05, "y": 4.0, "Ap": 0.25, "Vp": 0.25, "hp": 1.0, "Mp": 1.0,
          "vpx_dx": 2.0, "vpy_dy": 1.0, "vpx_dy": 0.0, "vpy_dx": 0.0}
    hi = dict(lo, x=7.95)
    _, out, _, _ = execute([lo, hi], [SPLIT, SPLIT], [-1, -1])
    assert out["x"][0] == 0.05 + 0.125 and out["x"][1] == 1e-10
    assert out["x"][2] == 8.0 - 1e-10 and out["x"][3] == 7.95 - 0.125
    # along y: a vertical shear direction pushes a child below y = 0
    py = {"x": 4.0, "y": 0.1, "Ap": 1.0, "Vp": 1.0, "hp": 1.0, "Mp": 1.0,
          "vpx_dx": 0.0, "vpy_dy": 0.0, "vpx_dy": 0.5, "vpy_dx": 0.5}
    _, out, _, _ = execute([py], [SPLIT], [-1])
    assert out["y"][1] == 1e-10


def test_merge_centre_of_mass_sums_and_averages():
    a = {"x": 2.25, "y": 3.25, "Mp": 1.0, "Vp": 0.25, "Ap": 0.125, "hp": 1.0,
         "vpx": 1.0, "vpy": -2.0, "F_11": 8.0, "H": 0.5}
    b = {"x": 2.75, "y": 3.75, "Mp": 3.0, "Vp": 0.5, "Ap": 0.25, "hp": 3.0,
         "vpx": 2.0, "vpy": 2.0, "F_11": 4.0, "H": 1.5}
    _, out, lab, lev = execute([a, b], [PRIMARY, SECONDARY], [1, -1], levels=[1, 0],
                               labels=[7, 8])
    assert len(out["x"]) == 1
    assert out["x"][0] == 2.625 and out["y"][0] == 3.625     # (M1 x1 + M2 x2) / Mt
    assert out["Mp"][0] == 4.0 and out["Vp"][0] == 0.75 and out["Ap"][0] == 0.375
    assert out["hp"][0] == 2.0          # Vp / Ap, not the mass average 2.5
    assert out["vpx"][0] == 1.75 and out["vpy"][0] == 1.0    # mass-averaged
    assert out["F_11"][0] == 5.0 and out["H"][0] == 1.25
    assert out["mom_px"][0] == 7.0 and out["mom_py"][0] == 4.0   # Mp v
    assert out["xp"][0] == 2.625 and out["yp"][0] == 3.625
    assert lev.tolist() == [2] and lab.tolist() == [-1]


# ---------------------------------------------------------------- ELFS rule
ELFS = dict(ms_mode=1, ms_alpha=1.0, ms_beta=1.0, ms_stretch=0.5)


def elfs_decide(parts, **over):
    prm = dict(BASE, **dict(ELFS, **over))
    a = {k: np.array([p[k] for p in parts], np.float64)
         for k in ("x", "y", "Ap", "hp", "vpx_dx", "vpy_dy", "vpx", "vpy")}
    lev = np.array([p["level"] for p in parts], np.int32)
    return oracle.elfs_decide(prm, a["x"], a["y"], a["Ap"], a["hp"], a["vpx_dx"],
                              a["vpy_dy"], a["vpx"], a["vpy"], lev)


def wet_block(extra, skip=()):
    """Wet cells at rows 0..3, columns 2..5 (one particle each with Ap = 4,
    r = 2, at the cell's upper right, id (c - 2) * 4 + r), plus extra
    particles.  Exterior: every dry cell.  Boundary distance: 0 at columns
    2 and 5 of rows 1..3 (an exterior cell left or right, off the domain
    edge), 1 m one cell inward; row 0 is the domain edge, its cells are no
    seeds: 1 m at columns 2 and 5 (from row 1), 2 m at columns 3 and 4."""
    parts = [P((c, r), 0.9, 0.9, Ap=4.0) for c in range(2, 6) for r in range(4)
             if (c, r) not in skip]
    return parts + extra


def test_elfs_value_and_thresholds():
    # alpha = beta = 1: split if elfs < r, merge if elfs > r (r = sqrt(Ap))
    extra = [
        P((3, 2), 0.5, 0.5, Ap=1.0),     # 16: elfs 1 = r: keep
        P((3, 2), 0.6, 0.5, Ap=0.25),    # 17: elfs 1 > r = 0.5: candidate
        P((3, 2), 0.7, 0.5, Ap=0.25),    # 18: candidate, pairs with 17
        P((2, 2), 0.5, 0.5, Ap=0.25),    # 19: elfs 0 (a seed cell) < 0.5: split
        # 20: cell (2, 0), distance 1 <= max(hx, hy): refined by the
        # exterior cell (1, 0) of its row to 0.25 m < r = 0.5: split
        P((2, 0), 0.25, 0.5, Ap=0.25),
        # 21, 22: cell (3, 0), distance 2 > max(hx, hy): not refined (the
        # exterior cell (1, 0) is 1.5 m away), elfs 2 > r = 1.75: candidates
        P((3, 0), 0.5, 0.5, Ap=3.0625),
        P((3, 0), 0.6, 0.5, Ap=3.0625),
    ]
    parts = wet_block(extra)
    act, partner, counts = elfs_decide(parts)
    assert act[16:].tolist() == [KEEP, PRIMARY, SECONDARY, SPLIT, SPLIT, PRIMARY,
                                 SECONDARY]
    assert partner[17] == 18 and partner[21] == 22
    # the block's particles (r = 2): elfs <= 1 < 2 splits them, except at
    # (3, 0) and (4, 0) (elfs 2, neither < 2 nor > 2); (2, 0) and (5, 0)
    # are refined to 0.9 and 0.1 m
    want = [KEEP if i in (4, 8) else SPLIT for i in range(16)]
    assert act[:16].tolist() == want
    assert counts.tolist() == [14 + 2, 2]


def test_elfs_divergence_trigger_and_level_guard():
    extra = [
        P((4, 2), 0.3, 0.3, Ap=0.25, vpx_dx=0.25, vpy_dy=0.25),   # div 0.5 = stretch
        P((4, 2), 0.4, 0.3, Ap=0.25, vpx_dx=0.5, vpy_dy=0.25),    # div 0.75: split
        # split-eligible (div) at level = min_level: kept, and unlike the
        # solvers' rule it does not fall through to merging (elfs 1 > 0.5
        # would make it a candidate), so its would-be partner stays unpaired
        P((4, 1), 0.3, 0.3, Ap=0.25, vpx_dx=1.0, level=-2),
        P((4, 1), 0.4, 0.3, Ap=0.25),
    ]
    act, _, _ = elfs_decide(wet_block(extra))
    # (4, 2): the first is a merge candidate (elfs 1 > 0.5) left without a
    # partner (the second splits): kept
    assert act[16:].tolist() == [KEEP, SPLIT, KEEP, KEEP]


def test_elfs_min_ppc_counts_a_split_as_four():
    # cell (3, 3) without its block particle: a lone pair is undone; a pair
    # beside a split stands (4 + 1 >= 2)
    extra = [P((3, 3), 0.3, 0.3, Ap=0.25), P((3, 3), 0.4, 0.3, Ap=0.25)]
    act, _, counts = elfs_decide(wet_block(extra, skip=[(3, 3)]))
    assert act[-2:].tolist() == [KEEP, KEEP]
    extra.append(P((3, 3), 0.8, 0.8, Ap=0.25, vpx_dx=1.0))   # div 1 > 0.5: split
    act, _, counts = elfs_decide(wet_block(extra, skip=[(3, 3)]))
    assert act[-3:].tolist() == [PRIMARY, SECONDARY, SPLIT]


def test_elfs_execution_four_children_and_kept_label():
    s = {"x": 3.5, "y": 2.5, "Ap": 0.25, "Vp": 0.5, "hp": 2.0, "Mp": 8.0,
         "vpx": 1.0, "vpy": 0.5}
    a = {"x": 5.25, "y": 5.5, "Mp": 1.0, "Vp": 0.25, "Ap": 0.125, "hp": 2.0, "vpx": 1.0}
    b = {"x": 5.75, "y": 5.5, "Mp": 3.0, "Vp": 0.25, "Ap": 0.125, "hp": 2.0, "vpx": 3.0}
    edge = {"x": 0.1, "y": 7.9, "Ap": 0.25, "Vp": 0.5, "hp": 2.0, "Mp": 8.0}
    f, out, lab, lev = execute([s, a, b, edge], [SPLIT, PRIMARY, SECONDARY, SPLIT],
                               [-1, 2, -1, -1], labels=[5, 6, 7, 9], levels=[0, 1, 1, 0],
                               elfs=True)
    assert len(out["x"]) == 9
    # children at (x +- r/2, y +- r/2), r = 0.5, in the order ++, +-, -+, --
    assert out["x"][:4].tolist() == [3.75, 3.75, 3.25, 3.25]
    assert out["y"][:4].tolist() == [2.75, 2.25, 2.75, 2.25]
    assert out["Mp"][:4].tolist() == [2.0] * 4 and out["Vp"][:4].tolist() == [0.125] * 4
    assert out["Ap"][:4].tolist() == [0.0625] * 4          # quartered, not Vp/hp
    assert out["mom_px"][:4].tolist() == [2.0] * 4 and out["hp"][:4].tolist() == [2.0] * 4
    assert lab[:4].tolist() == [-1] * 4 and lev[:4].tolist() == [-1] * 4
    # the merge keeps the primary's label (the solvers' rule sets -1)
    assert out["x"][4] == 5.625 and out["vpx"][4] == 2.5 and out["mom_px"][4] == 10.0
    assert out["Mp"][4] == 4.0 and out["Vp"][4] == 0.5
    assert lab[4] == 6 and lev[4] == 2
    # children of a particle at the corner: clamped to [1e-10, 8 - 1e-10]
    assert out["x"][5:].tolist() == [0.1 + 0.25, 0.1 + 0.25, 1e-10, 1e-10]
    assert out["y"][5:].tolist() == [8.0 - 1e-10, 7.9 - 0.25, 8.0 - 1e-10, 7.9 - 0.25]
