"""GPU parity of InterpolatedProjectionSampler (YTRC_INTERPOLATED_PROJECTION,
image_samplers.pyx:344-402) against the FP64 oracle: visit sequences
bit-exact, images rel-L2 <= 1e-5.  The kernel sums each visit's sub-samples
in closed form (power sums of the per-visit cubic) where the reference adds
them one by one: the difference is rounding."""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, render
from yt_amd.image_samplers import InterpolatedProjectionSampler

from .conftest import rel_l2
from .helpers import c0_params, oracle_vr_bricks, small_amr, uniform_lognormal

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _one(ctx, spec, flags=_lib.BATCH_PERSISTENT):
    b = _lib.Batch(ctx, [spec], flags)
    try:
        b.run()
        return b.read_image(0), b.read_record(0) if spec.record else None
    finally:
        b.close()


def _check(img, rec, oimg, ov, oh, tag):
    """NaN (log10 of a non-positive extrapolated vertex value propagates
    through the sum, as in the reference) must sit at the same pixels; the
    finite pixels meet the tolerance."""
    v, h, _ = rec
    assert int(np.sum(v != ov)) == 0 and int(np.sum(h != oh)) == 0, tag
    a, b = img[..., 0], oimg[..., 0]
    assert np.array_equal(np.isnan(a), np.isnan(b)), tag
    ok = ~np.isnan(b)
    e = rel_l2(a[ok], b[ok])
    print(f"{tag}: visits {int(v.sum())}, NaN pixels {int((~ok).sum())}, rel-L2={e:.3e}", flush=True)
    assert e <= TOL
    assert np.all(img[..., 1:] == 0.0)


@pytest.mark.parametrize("n_samples", [10, 3])
def test_uniform_grids(ctx, n_samples):
    ds = uniform_lognormal(32, seed=111, nprocs=4)
    vol = render.ResidentVolume(ctx, ds, vr_field="density", log_field=False)
    try:
        p = c0_params(72, normal=(0.25, 0.6, -0.76))
        spec = vol.interpolated_projection_request(p, n_samples=n_samples, record=True)
        img, rec = _one(ctx, spec)
        oimg, ov, oh, os_ = oracle.cast(2, p, oracle_vr_bricks(vol), order=spec.order,
                                        n_samples=n_samples, record=True, with_samples=True)
        _check(img, rec, oimg, ov, oh, f"iproj uniform n={n_samples}")
        assert np.array_equal(rec[2], os_)
    finally:
        vol.free()


def test_amr_perspective(ctx):
    ds = small_amr(root_n=32, levels=2, seed=113)
    vol = render.ResidentVolume(ctx, ds, vr_field="density", log_field=True)
    try:
        p = render.perspective_params((1.7, -0.6, 1.2), (0.5, 0.5, 0.5), 0.8, (40, 36))
        spec = vol.interpolated_projection_request(p, record=True)
        img, rec = _one(ctx, spec, _lib.BATCH_PLAIN)
        oimg, ov, oh = oracle.cast(2, p, oracle_vr_bricks(vol), order=spec.order, record=True)
        _check(img, rec, oimg, ov, oh, "iproj AMR perspective")
    finally:
        vol.free()


def test_drop_in_sampler():
    from yt_amd.amr_kdtree import AMRKDTree

    ds = uniform_lognormal(24, seed=115, nprocs=2)
    kd = AMRKDTree(ds)
    kd.set_fields(["density"], [False])
    p = c0_params(30, normal=(0.5, 0.5, 0.7071))
    image = np.zeros((30, 30, 4))
    s = InterpolatedProjectionSampler(np.atleast_3d(p["vp_pos"]), np.atleast_3d(p["vp_dir"]),
                                      p["center"], p["bounds"], image, p["x_vec"], p["y_vec"],
                                      p["width"], "KDTree", 10, lens_type="plane-parallel")
    bricks = []
    for pg in kd.traverse(viewpoint=p["viewpoint"]):
        s(pg)
        bricks.append(dict(le=pg.LeftEdge, re=pg.RightEdge, dims=pg.dims,
                           data=[np.asarray(d, dtype=np.float64) for d in pg.my_data], mask=None))
    out = s.aimage
    oimg = oracle.cast(2, p, bricks, n_samples=10)
    assert out is image
    assert rel_l2(out[..., 0], oimg[..., 0]) <= TOL
