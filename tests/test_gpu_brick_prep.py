"""Device brick preparation (ytrc_dataset_build_vr) against the host path
(AMRKDTree.get_brick_data in numpy, uploaded with ytrc_dataset_upload).

Bar: the stored float32 vertex data are equal point for point.  The one
operation that may differ is log10 (the device's FP64 log10 against the
host libm's, both within an ulp of FP64), which can move a value across a
float32 rounding boundary: at most 4 vertices per dataset may then differ,
by one float32 ulp (each stored up to 8 times: a vertex on the faces of
vertex blocks is stored in every block that holds it, so at most 32
stored points).  NaN (log10 of v <= 0) must sit at the same points.
"""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, data, render

from .conftest import rel_l2
from .helpers import c0_params, oracle_vr_bricks, small_amr, tf_dict, uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu


def _compare(ctx, ds, field="density", log=True):
    dev = render.ResidentVolume(ctx, ds, vr_field=field, log_field=log, vr_prep="device")
    host = render.ResidentVolume(ctx, ds, vr_field=field, log_field=log, vr_prep="host")
    try:
        assert dev.n_vr_bricks == host.n_vr_bricks
        off_by_one, total = 0, 0
        for b in range(dev.n_vr_bricks):
            a = ctx.read_points(dev.vr_ds_id, b)
            h = ctx.read_points(host.vr_ds_id, b)
            assert a.shape == h.shape
            assert np.array_equal(np.isnan(a), np.isnan(h)), b
            ok = ~np.isnan(a)
            diff = a[ok] != h[ok]
            if diff.any():
                ulps = np.abs(a[ok][diff].view(np.int32).astype(np.int64)
                              - h[ok][diff].view(np.int32).astype(np.int64))
                assert ulps.max() <= 1, b
                off_by_one += int(diff.sum())
            total += a.size
        print(f"{dev.n_vr_bricks} bricks, {total} points, {off_by_one} one-ulp differences",
              flush=True)
        assert off_by_one <= 32
        assert ctx.dataset_bytes(dev.vr_ds_id) == ctx.dataset_bytes(host.vr_ds_id)
    finally:
        dev.free()
        host.free()


def test_uniform_grids_match_host(ctx):
    _compare(ctx, uniform_lognormal(40, seed=81, nprocs=4))


def test_float32_source_and_linear_field(ctx):
    rho = data.lognormal_density(36, seed=83, dtype=np.float32)
    ds = data.load_uniform_grid({"density": rho}, psize=(2, 1, 2))
    _compare(ctx, ds, log=True)
    _compare(ctx, ds, log=False)


def test_amr_leaves_match_host(ctx):
    _compare(ctx, small_amr(root_n=32, levels=2, seed=85))


def test_negative_values_give_nan(ctx):
    """log10 of v <= 0 is NaN (amr_kdtree.py:346-352) in both paths."""
    rng = np.random.default_rng(87)
    f = rng.normal(size=(24, 24, 24))
    ds = data.load_uniform_grid({"density": f}, nprocs=2)
    _compare(ctx, ds)


def test_device_prepared_vr_matches_oracle(ctx):
    ds = small_amr(root_n=32, levels=2, seed=89)
    tf = vr_tf_for(ds, nbins=128)
    vol = render.ResidentVolume(ctx, ds, vr_field="density", vr_prep="device")
    try:
        p = c0_params(96, normal=(0.2, 0.7, -0.68))
        spec = vol.vr_request(p, tf, record=True)
        b = _lib.Batch(ctx, [spec])
        try:
            b.run()
            img, (v, h, _) = b.read_image(0), b.read_record(0)
        finally:
            b.close()
        oimg, ov, oh = oracle.cast(1, p, oracle_vr_bricks(vol), order=spec.order, tf=tf_dict(tf),
                                   record=True)
        assert np.array_equal(v, ov) and np.array_equal(h, oh)
        for c in range(3):
            assert rel_l2(img[..., c], oimg[..., c]) <= 1e-5
    finally:
        vol.free()


def test_thin_grid_is_rejected(ctx):
    with pytest.raises(ValueError):
        ctx.build_vr([((0, 0, 0), (1, 1, 1), (1, 4, 4))], [[np.ones((1, 4, 4))]], [False],
                     [(0, (0, 0, 0), (1, 4, 4), (0, 0, 0), (1, 1, 1))])


def test_device_vertex_centring_equals_reference_bits(ctx):
    """k_vertex_centre against ghost_zone_interpolate's own output
    (interpolators.pyx:177-226, tests/golden/vertex_centre_vectors.npz):
    the device's FP64 vertex values, stored as float32, equal the
    reference's FP64 values uploaded as float32, point for point (no log,
    so the only rounding is the final float32 store in both paths)."""
    import os

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "vertex_centre_vectors.npz"))
    for i in range(int(z["n"])):
        a, want = z[f"in{i}"], z[f"out{i}"]
        dims = a.shape
        if min(dims) < 2:
            continue
        print(f"vertex centring vector {i}: dims {dims}", flush=True)
        le, re = (0.0, 0.0, 0.0), tuple(float(d) / max(dims) for d in dims)
        dev = ctx.build_vr([(le, re, dims)], [[a]], [False], [(0, (0, 0, 0), dims, le, re)])
        host = ctx.upload([(le, re, dims, True)], [[want]])
        try:
            got = ctx.read_points(dev, 0)
            ref = ctx.read_points(host, 0)
            assert got.shape == ref.shape
            assert np.array_equal(got.view(np.int32), ref.view(np.int32)), i
        finally:
            ctx.free(dev)
            ctx.free(host)
