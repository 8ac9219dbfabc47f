"""Pins the oracle's samplers, table lookups, interpolation and lens culling
by closed forms and mathematical invariants (no reference run):

  ProjectionSampler (image_samplers.pyx:330-343)
    * constant field c: pixel = c * chord length (exact slab chord);
    * sqrt(3) chord along (1,1,1) (test_offaxisprojection_pytestonly.py:518-552);
    * linearity in the data; weighted projection of a constant = constant.
  VolumeRenderSampler + FIT_eval_transfer (image_samplers.pyx:453-487,
  field_interpolation_tables.pxd:85-110)
    * constant field: rgba_c = 1 - prod_k (1 - dt_k e_c) (non-grey), with dt_k
      from the pinned visit log; grey: alpha from e_3 and colours from the
      same transmittance.
  FIT_get_value (:66-80): exact at nodes, linear between, 0 outside the open
    interval and for NaN / 0 / subnormal inputs.
  offset_interpolate (fixed_interpolator.cpp:34-52): exact on linear fields.
  calculate_extent_plane_parallel + clip (lenses.pyx:22-55,
    image_samplers.pyx:159-166): every pixel whose ray meets the brick with
    positive length lies inside the clipped extent.
"""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import render
from yt_amd.transfer_functions import ColorTransferFunction, TransferFunction


def _params(res, normal, width=1.0, center=(0.5, 0.5, 0.5), north=None):
    return render.camera_params(center, normal, width, res, north)


def _brick(n, data, le=(0, 0, 0), re=(1, 1, 1), mask=None):
    return dict(le=np.array(le, float), re=np.array(re, float), dims=np.array([n] * 3),
                data=data, mask=mask)


def _ray(p, vi, vj):
    w = p["width"]
    nv = p["nv"]
    px = w[0] * float(vi) / (float(nv[0]) - 1) - w[0] / 2.0
    py = w[1] * float(vj) / (float(nv[1]) - 1) - w[1] / 2.0
    vp = np.ravel(p["vp_pos"])
    pos = np.array([vp[0] * px + vp[3] * py + vp[9], vp[1] * px + vp[4] * py + vp[10],
                    vp[2] * px + vp[5] * py + vp[11]])
    return pos, np.ravel(p["vp_dir"]).copy()


def _chord(le, re, pos, d, max_t=1.0):
    t0, t1 = 0.0, max_t
    for a in range(3):
        if d[a] == 0:
            if not (le[a] <= pos[a] <= re[a]):
                return 0.0
            continue
        ta, tb = (le[a] - pos[a]) / d[a], (re[a] - pos[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return max(0.0, t1 - t0)


def test_projection_constant_field_chord():
    n = 8
    c = 2.5
    p = _params(17, (0.3, -0.5, 0.81), width=1.3)
    img = oracle.cast(0, p, [_brick(n, [np.full((n, n, n), c)])])
    for vi in range(17):
        for vj in range(17):
            pos, d = _ray(p, vi, vj)
            ch = _chord((0, 0, 0), (1, 1, 1), pos, d)
            assert abs(img[vi, vj, 0] - c * ch) <= 1e-12 * max(1.0, c * ch)


def test_projection_sqrt3():
    n = 16
    p = _params(65, (1.0, 1.0, 1.0), width=np.sqrt(3.0))
    img = oracle.cast(0, p, [_brick(n, [np.ones((n, n, n))])])
    out = render.finalize_projection(img, p["width"], False)
    assert abs(out.max() - np.sqrt(3.0)) < 1e-2
    assert out.max() <= np.sqrt(3.0) * (1 + 1e-12)


def test_projection_linear_in_data():
    rng = np.random.default_rng(0)
    n = 12
    f, g = rng.uniform(0, 1, (n,) * 3), rng.uniform(0, 1, (n,) * 3)
    p = _params(40, (0.2, 0.9, -0.3))
    a = oracle.cast(0, p, [_brick(n, [f])])[..., 0]
    b = oracle.cast(0, p, [_brick(n, [g])])[..., 0]
    ab = oracle.cast(0, p, [_brick(n, [3.0 * f + 0.5 * g])])[..., 0]
    assert np.allclose(ab, 3.0 * a + 0.5 * b, rtol=1e-12, atol=1e-14)


def test_weighted_projection_of_constant():
    rng = np.random.default_rng(1)
    n = 10
    w = rng.uniform(0.1, 2, (n,) * 3)
    p = _params(30, (0.5, 0.5, -0.7))
    img = oracle.cast(0, p, [_brick(n, [4.0 * w, w])])
    out = render.finalize_projection(img, p["width"], True)
    hit = img[..., 1] > 0
    assert hit.sum() > 100
    assert np.allclose(out[hit], 4.0, rtol=1e-13)
    assert np.all(out[~hit] == 0)


def _ctf_const(lo, hi, nbins, rgba):
    tf = ColorTransferFunction((lo, hi), nbins=nbins)
    for t, v in zip(tf.funcs, rgba):
        t.y[:] = v
    return tf


@pytest.mark.parametrize("grey", [False, True])
def test_vr_constant_field_closed_form(grey):
    """Constant field: every sub-sample sees the same table values, so each
    channel is 1 - prod(1 - dt e) over sub-samples (non-grey), which the
    visit log's dt values determine."""
    from tests.helpers import tf_dict

    n = 6
    v0 = 0.3
    rgba_tab = (0.7, 0.2, 0.4, 0.9)
    tf = _ctf_const(-1.0, 1.0, 32, rgba_tab)
    tf.grey_opacity = grey
    p = _params(9, (0.4, 0.3, 0.85), width=1.2)
    vdata = np.full((n + 1,) * 3, v0)
    img = oracle.cast(1, p, [_brick(n, [vdata], mask=np.ones((n,) * 3, np.uint8))],
                      tf=tf_dict(tf), n_samples=10)
    a = rgba_tab[3]
    e = [rgba_tab[0] * a, rgba_tab[1] * a, rgba_tab[2] * a, a]
    for vi in range(9):
        for vj in range(9):
            pos, d = _ray(p, vi, vj)
            _, _, t = oracle.walk_log((0, 0, 0), (1, 1, 1), [n] * 3, pos, d)
            dts = np.repeat((t[:, 1] - t[:, 0]) / 10.0, 10)
            if grey:
                trans = np.prod(np.maximum(1.0 - dts * e[3], 0.0))
                exp = [0.0] * 4
                # rgba = sum_k dt_k e * prod_{j>k} (1 - dt_j e_3)
                for c in range(4):
                    acc = 0.0
                    for dt in dts:
                        acc = dt * e[c] + max(1 - dt * e[3], 0.0) * acc
                    exp[c] = acc
                assert abs(exp[3] - (1 - trans)) < 1e-12
            else:
                exp = [1.0 - np.prod(1.0 - dts * e[c]) for c in range(3)] + [0.0]
            for c in range(4 if grey else 3):
                assert abs(img[vi, vj, c] - exp[c]) <= 1e-12, (vi, vj, c)
            if not grey:
                assert img[vi, vj, 3] == 0.0


def test_vr_mask_skips_cells():
    from tests.helpers import tf_dict

    n = 4
    tf = _ctf_const(-1.0, 1.0, 16, (1.0, 1.0, 1.0, 1.0))
    p = _params(12, (0.1, 0.2, 1.0))
    vdata = np.full((n + 1,) * 3, 0.5)
    zero_mask = np.zeros((n,) * 3, np.uint8)
    img = oracle.cast(1, p, [_brick(n, [vdata], mask=zero_mask)], tf=tf_dict(tf))
    assert np.all(img == 0)


def test_fit_values():
    nb = 11
    x = np.linspace(-2.0, 3.0, nb)
    y = np.random.default_rng(2).uniform(0, 1, nb)
    # nodes (the last node equals b1, which the open interval excludes)
    v = oracle.fit_values(y, -2.0, 3.0, x[1:-1])
    # x = 0 is a node here: isnormal(0) is false, so the table gives 0 there
    zero = x[1:-1] == 0.0
    assert zero.sum() == 1 and np.all(v[zero] == 0.0)
    assert np.allclose(v[~zero], y[1:-1][~zero], rtol=0, atol=1e-15)
    mid = 0.5 * (x[1:-1] + x[2:])
    v = oracle.fit_values(y, -2.0, 3.0, mid)
    assert np.allclose(v, 0.5 * (y[1:-1] + y[2:]), atol=1e-14)
    outside = np.array([-2.0, -5.0, 3.0, 7.0, np.nan, 0.0, 5e-324, -1e-310])
    assert np.all(oracle.fit_values(y, -2.0, 3.0, outside) == 0.0)


def test_trilinear_exact_on_linear_fields():
    rng = np.random.default_rng(4)
    dims = np.array([3, 4, 5], np.int32)
    a, b, c, d = rng.uniform(-1, 1, 4)
    i, j, k = np.meshgrid(*[np.arange(x + 1) for x in dims], indexing="ij")
    vd = a + b * i + c * j + d * k
    cells = np.stack([rng.integers(0, dims[q], 50) for q in range(3)], axis=1).astype(np.int32)
    dps = rng.uniform(0, 1, (50, 3))
    out = oracle.trilinear(dims, vd, cells, dps)
    pos = cells + dps
    exp = a + b * pos[:, 0] + c * pos[:, 1] + d * pos[:, 2]
    assert np.allclose(out, exp, atol=1e-13)
    corner = oracle.trilinear(dims, vd, cells, np.zeros((50, 3)))
    assert np.array_equal(corner, vd[cells[:, 0], cells[:, 1], cells[:, 2]])


@pytest.mark.parametrize("seed", range(6))
def test_extent_is_conservative(seed):
    rng = np.random.default_rng(seed)
    nv = int(rng.integers(8, 40))
    nrm = rng.normal(size=3)
    p = _params(nv, nrm / np.linalg.norm(nrm), width=float(rng.uniform(0.8, 1.5)))
    bricks = []
    for _ in range(5):
        le = rng.uniform(0, 0.8, 3)
        re = le + rng.uniform(0.05, 0.3, 3)
        bricks.append(dict(le=le, re=re, dims=np.array([4, 4, 4])))
    ext = oracle.extents(p, bricks)
    for b, e in zip(bricks, ext):
        for vi in range(nv):
            for vj in range(nv):
                pos, d = _ray(p, vi, vj)
                if _chord(b["le"], b["re"], pos, d) > 1e-9:
                    assert e[0] <= vi < e[1] and e[2] <= vj < e[3], (vi, vj, e)


def test_transfer_function_gaussian():
    tf = TransferFunction((-1.0, 1.0), nbins=5)
    tf.add_gaussian(0.0, 0.5, 2.0)
    assert np.allclose(tf.y, 2.0 * np.exp(-np.linspace(-1, 1, 5) ** 2 / 0.5))
    ctf = ColorTransferFunction((0.0, 1.0), nbins=8)
    assert ctf.n_field_tables == 4
    assert ctf.weight_table_ids[:4] == [3, 3, 3, -1]
    assert ctf.field_table_ids == [0, 1, 2, 3, 0, 0]


def test_interpolated_projection_integrates_linear_fields():
    """InterpolatedProjectionSampler (image_samplers.pyx:368-402): on a
    vertex-centred linear field the trilinear interpolant is the field, and
    the midpoint sub-samples integrate a linear function exactly, so each
    channel is the line integral of its field over the ray's chord in t."""
    n = 5
    coef = [(0.3, 1.0, -2.0, 0.5), (1.5, 0.0, 0.7, 0.2), (-0.4, 0.25, 0.5, -1.0)]
    g = np.linspace(0.0, 1.0, n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    fields = [a + bx * X + by * Y + bz * Z for a, bx, by, bz in coef]
    p = _params(11, (0.35, -0.5, 0.79), width=1.3)
    img = oracle.cast(2, p, [_brick(n, fields)], n_samples=7)
    for vi in range(11):
        for vj in range(11):
            pos, d = _ray(p, vi, vj)
            t0, t1 = 0.0, 1.0
            for a in range(3):
                ta, tb = (0.0 - pos[a]) / d[a], (1.0 - pos[a]) / d[a]
                t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
            for c, (a0, bx, by, bz) in enumerate(coef):
                want = 0.0
                if t1 > t0:
                    b = np.array([bx, by, bz])
                    mid = pos + 0.5 * (t0 + t1) * d
                    want = (a0 + b @ mid) * (t1 - t0)
                assert abs(img[vi, vj, c] - want) <= 1e-12, (vi, vj, c, img[vi, vj, c], want)
            assert img[vi, vj, 3] == 0.0


def test_interpolated_projection_ignores_mask_and_uses_three_channels():
    n = 3
    f = [np.full((n + 1,) * 3, float(k + 1)) for k in range(5)]
    p = _params(6, (0.2, 0.3, 0.93))
    a = oracle.cast(2, p, [_brick(n, f, mask=np.zeros((n,) * 3, np.uint8))], n_samples=4)
    b = oracle.cast(2, p, [_brick(n, f)], n_samples=4)
    assert np.array_equal(a, b)
    assert np.all(a[..., 3] == 0.0)
    hit = a[..., 0] > 0
    assert np.allclose(a[..., 1][hit], 2 * a[..., 0][hit]) and np.allclose(a[..., 2][hit], 3 * a[..., 0][hit])


@pytest.mark.parametrize("field", ["constant", "linear"])
def test_light_source_closed_form(field):
    """LightSourceRenderSampler (image_samplers.pyx:548-590) with a constant
    transfer function.  On a linear field the finite-difference gradient is
    minus the unit gradient everywhere, so dot = -l . g_hat.  On a constant
    field the differences are rounding noise that eval_gradient normalises
    to an arbitrary unit vector (as the reference does), so that case uses
    l_rgba = 0, which makes the light term 1.  Each channel then follows
    rgba' = (1 - ta) e (1 + dot l_c) + ta rgba with ta = max(1 - dt e, 0)
    over the visit log's sub-sample dt (field_interpolation_tables.pxd:115-143)."""
    from tests.helpers import tf_dict

    n = 5
    rgba_tab = (0.6, 0.3, 0.8, 0.5)
    tf = _ctf_const(-10.0, 10.0, 16, rgba_tab)
    g = np.linspace(0.0, 1.0, n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    b = np.array([0.8, -0.5, 0.3])
    vdata = np.full((n + 1,) * 3, 0.4) if field == "constant" else 0.4 + b[0] * X + b[1] * Y + b[2] * Z
    ldir = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    lrgba = (0.0, 0.0, 0.0, 0.0) if field == "constant" else (1.0, 0.5, 0.25, 1.0)
    td = dict(tf_dict(tf), light_dir=ldir, light_rgba=lrgba)
    p = _params(7, (0.3, 0.45, 0.84), width=1.1)
    img = oracle.cast(3, p, [_brick(n, [vdata], mask=np.zeros((n,) * 3, np.uint8))], tf=td,
                      n_samples=4)
    a = rgba_tab[3]
    e = [rgba_tab[c] * a for c in range(3)]
    dot = 0.0 if field == "constant" else float(-ldir @ (b / np.linalg.norm(b)))
    for vi in range(7):
        for vj in range(7):
            pos, d = _ray(p, vi, vj)
            _, _, t = oracle.walk_log((0, 0, 0), (1, 1, 1), [n] * 3, pos, d)
            dts = np.repeat((t[:, 1] - t[:, 0]) / 4.0, 4)
            for c in range(3):
                acc = 0.0
                for dt in dts:
                    ta = max(1.0 - dt * e[c], 0.0)
                    acc = (1.0 - ta) * e[c] * (1.0 + dot * lrgba[c]) + ta * acc
                assert abs(img[vi, vj, c] - acc) <= 1e-11, (vi, vj, c)


def test_oracle_sample_counts():
    """The oracle's per-pixel sample counts against its visit counts: the
    interpolated projection takes n_samples at every visit; the volume
    render n_samples at every visit of a mask-1 cell, so with all-ones masks
    it equals visits x n_samples and with child masks it is less where rays
    cross covered cells; the projection counts visits with exit_t > enter_t,
    all of them for these rays (no exact multi-axis ties)."""
    import numpy as np

    from oracle import oracle
    from tests.helpers import c0_params, small_amr, tf_dict, vr_tf_for
    from yt_amd.amr_kdtree import AMRKDTree

    ds = small_amr(root_n=16, levels=2, seed=41)
    kd = AMRKDTree(ds)
    kd.set_fields(["density"], [True], no_ghost=True)
    bricks = []
    for n in kd.leaves():
        pg = kd.get_brick_data(n)
        bricks.append(dict(le=pg.LeftEdge, re=pg.RightEdge, dims=pg.dims,
                           data=list(pg.my_data), mask=pg.source_mask))
    # kd leaves lie inside their finest grid, so their child masks are all
    # ones: mask out a slab of every brick to exercise the mask test
    unmasked = [dict(b, mask=None) for b in bricks]
    for b in bricks:
        m = np.ones(tuple(int(x) for x in b["dims"]), dtype=np.uint8)
        m[: max(1, m.shape[0] // 2)] = 0
        b["mask"] = m
    p = c0_params(24, normal=(0.3, 0.6, -0.7))
    order = kd.order(p["viewpoint"])
    tf = tf_dict(vr_tf_for(ds, nbins=32))

    def counts(kind, bs, ns):
        _, v, _, s = oracle.cast(kind, p, bs, order=order, tf=tf if kind == 1 else None,
                                 n_samples=ns, record=True, with_samples=True)
        return v, s

    v, s = counts(2, bricks, 3)
    assert v.sum() > 0 and np.array_equal(s, 3 * v)
    v, s = counts(1, unmasked, 7)
    assert np.array_equal(s, 7 * v)
    v, s = counts(1, bricks, 7)
    assert np.all(s % 7 == 0) and np.all(s <= 7 * v) and np.any(s < 7 * v)
    v, s = counts(0, unmasked, 1)
    assert np.array_equal(s, v)


def test_oracle_row_window():
    """cast(rows=(r0, r1)) computes exactly the full cast's pixels on those
    rows and leaves every other pixel at its input value (the timing sample
    of the CPU baseline)."""
    rng = np.random.default_rng(3)
    data = [np.exp(rng.normal(size=(12, 12, 12)))]
    p = _params(40, (0.3, 0.5, 0.81))
    full = oracle.cast(0, p, [_brick(12, data)])
    img0 = rng.normal(size=(40, 40, 4))
    part = oracle.cast(0, p, [_brick(12, data)], image=img0.copy(), rows=(13, 29))
    whole = oracle.cast(0, p, [_brick(12, data)], image=img0.copy())
    assert np.array_equal(part[13:29], whole[13:29])
    assert not np.array_equal(whole[13:29], img0[13:29])
    assert np.array_equal(part[:13], img0[:13]) and np.array_equal(part[29:], img0[29:])
    again = oracle.cast(0, p, [_brick(12, data)])
    assert np.array_equal(again, full)  # the window is reset after the call
