"""Pins the oracle's walk_volume (grid_traversal.pyx:26-140) without running
the reference, by the method of test_rays.py:17-67 done in exact rational
arithmetic:

  * the cells visited with exit_t > enter_t are exactly the cells a ray meets
    with positive length inside t in [max(0, t_entry), max_t] (slab test on
    Fractions of the same float64 inputs);
  * per-visit [enter_t, exit_t] equals the exact slab interval of that cell
    (to 4 ulps for general rays, bit-exact for dyadic rays whose
    intersections are representable);
  * sum(exit_t - enter_t) equals the clipped chord (test_rays.py:36);
  * t is non-decreasing and consecutive cells differ by one unit step.
"""

from fractions import Fraction as F

import numpy as np
import pytest

from oracle import oracle


def exact_cells(le, re, dims, pos, vdir, max_t):
    """Brute force: {(i,j,k): (t_in, t_out)} with positive length."""
    le = [F(float(x)) for x in le]
    re = [F(float(x)) for x in re]
    pos = [F(float(x)) for x in pos]
    vdir = [F(float(x)) for x in vdir]
    # dds exactly as the float64 the oracle uses: (re-le)/dims rounded.
    dds = [F(float((float(re[i]) - float(le[i])) / dims[i])) for i in range(3)]
    t0, t1 = F(0), F(float(min(max_t, 1.0)))
    # brick entry first (the DDA starts at the brick boundary or t=0)
    out = {}
    axes = []
    for a in range(3):
        lo_list = []
        for c in range(dims[a]):
            lo = le[a] + c * dds[a]
            hi = le[a] + (c + 1) * dds[a] if c + 1 < dims[a] else re[a]
            if vdir[a] == 0:
                if lo <= pos[a] < hi or (c == dims[a] - 1 and pos[a] == hi):
                    lo_list.append((c, None, None))
                continue
            ta = (lo - pos[a]) / vdir[a]
            tb = (hi - pos[a]) / vdir[a]
            lo_list.append((c, min(ta, tb), max(ta, tb)))
        axes.append(lo_list)
    for ci, ai, bi in axes[0]:
        for cj, aj, bj in axes[1]:
            for ck, ak, bk in axes[2]:
                tin = max([t0] + [x for x in (ai, aj, ak) if x is not None])
                tout = min([t1] + [x for x in (bi, bj, bk) if x is not None])
                if tin < tout:
                    out[(ci, cj, ck)] = (tin, tout)
    return out


def check_ray(le, re, dims, pos, vdir, max_t=1.0, exact_t=False):
    hit, idx, t = oracle.walk_log(le, re, dims, pos, vdir, max_t)
    assert hit == len(idx)
    ex = exact_cells(le, re, dims, pos, vdir, max_t)
    pos_len = {tuple(map(int, idx[n])): (t[n, 0], t[n, 1]) for n in range(hit)
               if t[n, 1] > t[n, 0]}
    assert set(pos_len) == set(ex), (sorted(set(pos_len) ^ set(ex)))[:5]
    for c, (a, b) in pos_len.items():
        ea, eb = ex[c]
        if exact_t:
            assert F(a) == ea and F(b) == eb, (c, a, b, float(ea), float(eb))
        else:
            assert abs(a - float(ea)) <= 4 * np.spacing(max(abs(a), 1e-300)) + 1e-15
            assert abs(b - float(eb)) <= 4 * np.spacing(max(abs(b), 1e-300)) + 1e-15
    if hit:
        assert np.all(np.diff(t[:, 0]) >= 0)
        assert np.all(t[:, 1] >= t[:, 0])
        steps = np.abs(np.diff(idx, axis=0)).sum(axis=1)
        assert np.all(steps == 1)
        chord = sum((b - a) for (a, b) in ex.values())
        assert abs(float(np.sum(t[:, 1] - t[:, 0])) - float(chord)) <= 1e-13 * max(1.0, float(chord))
    return hit, idx, t


@pytest.mark.parametrize("seed", range(40))
def test_random_rays(seed):
    rng = np.random.default_rng(seed)
    dims = rng.integers(1, 7, size=3).astype(np.int32)
    le = rng.uniform(-1, 1, size=3)
    re = le + rng.uniform(0.2, 2.0, size=3)
    # rays from outside, inside and grazing
    pos = rng.uniform(le - 1.0, re + 1.0)
    target = rng.uniform(le, re)
    vdir = (target - pos) * rng.uniform(1.0, 3.0)
    if seed % 5 == 0:
        vdir[rng.integers(0, 3)] = 0.0
    check_ray(le, re, dims, pos, vdir, max_t=1.0 if seed % 3 else rng.uniform(0.2, 1.5))


DYADIC = [
    # (le, re, dims, pos, dir): all intersections exactly representable
    ((0, 0, 0), (1, 1, 1), (4, 4, 4), (-0.5, 0.125, 0.375), (2.0, 0.5, 0.25)),
    # through cell edges and corners (ties between axes)
    ((0, 0, 0), (1, 1, 1), (4, 4, 4), (-0.5, -0.5, -0.5), (2.0, 2.0, 2.0)),
    ((0, 0, 0), (1, 1, 1), (8, 4, 2), (-0.5, -0.25, 0.5), (2.0, 1.0, 0.0)),
    # axis-parallel along a cell face (v_dir has two zeros)
    ((0, 0, 0), (1, 1, 1), (4, 4, 4), (0.25, 0.5, -1.0), (0.0, 0.0, 4.0)),
    # start inside, leave through a corner
    ((0, 0, 0), (2, 2, 2), (4, 4, 4), (0.75, 0.75, 0.75), (2.0, 2.0, 2.0)),
    # negative directions
    ((-1, -1, -1), (1, 1, 1), (8, 8, 8), (1.5, 1.25, 2.0), (-2.0, -2.0, -4.0)),
    # max_t cuts the ray inside the brick
    ((0, 0, 0), (1, 1, 1), (4, 4, 4), (-0.5, 0.5, 0.5), (2.0, 0.25, 0.125)),
]


@pytest.mark.parametrize("case", range(len(DYADIC)))
def test_dyadic_rays_bit_exact(case):
    le, re, dims, pos, vdir = DYADIC[case]
    max_t = 0.625 if case == len(DYADIC) - 1 else 1.0
    hit, idx, t = check_ray(le, re, np.array(dims, dtype=np.int32), pos, vdir, max_t,
                            exact_t=True)
    assert hit > 0


def test_corner_tie_produces_zero_length_visits():
    """A ray through cell corners meets three faces at one t: walk_volume steps
    one axis per iteration, so the tie shows up as zero-length visits, and the
    positive-length cells are still exactly the brute-force ones."""
    hit, idx, t = check_ray((0, 0, 0), (1, 1, 1), np.array([4, 4, 4], np.int32),
                            (-0.5, -0.5, -0.5), (2.0, 2.0, 2.0), exact_t=True)
    zero = t[:, 1] == t[:, 0]
    assert zero.sum() == 2 * 3  # 3 interior corners, 2 extra steps each
    assert (~zero).sum() == 4


def test_miss_and_behind():
    assert oracle.walk_log((0, 0, 0), (1, 1, 1), [4, 4, 4], (2, 2, 2), (1, 1, 1))[0] == 0
    assert oracle.walk_log((0, 0, 0), (1, 1, 1), [4, 4, 4], (-2, 0.5, 0.5), (1, 0, 0))[0] == 0
    # entering exactly at max_t: no visit (grid_traversal.pyx:81)
    assert oracle.walk_log((0, 0, 0), (1, 1, 1), [4, 4, 4], (-1, 0.5, 0.5), (2, 0, 0), 0.5)[0] == 0


def test_unit_box_chord_sums_to_one():
    """test_rays.py:36: a ray that spans the box from t=0 to t=1 has
    sum(dt) == 1 to 14 digits, over multi-brick splits too."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        p1 = rng.uniform(0, 1, 3)
        p2 = rng.uniform(0, 1, 3)
        total = 0.0
        for sx in range(2):
            for sy in range(2):
                le = np.array([sx * 0.5, sy * 0.5, 0.0])
                re = le + np.array([0.5, 0.5, 1.0])
                _, _, t = oracle.walk_log(le, re, [8, 8, 16], p1, p2 - p1)
                total += float(np.sum(t[:, 1] - t[:, 0]))
        assert abs(total - 1.0) < 1e-13
