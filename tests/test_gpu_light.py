"""GPU parity of LightSourceRenderSampler (YTRC_LIGHT_SOURCE,
image_samplers.pyx:496-600) against the FP64 oracle: visit sequences
bit-exact, images rel-L2 <= 1e-5.  The kernel forms the gradient of the
trilinear interpolant analytically where the reference takes finite
differences of it (identical up to rounding: the interpolant is linear along
each axis); the fields here vary from cell to cell, so the normalised
gradient is well conditioned."""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, render

from .conftest import rel_l2
from .helpers import c0_params, oracle_vr_bricks, small_amr, tf_dict, uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu

TOL = 1e-5
LIGHT = dict(light_dir=(1.0, 0.5, -0.3), light_rgba=(0.8, 0.6, 0.4, 1.0))


def _td(tf, spec):
    return dict(tf_dict(tf), light_dir=spec.light_dir, light_rgba=spec.light_rgba)


def _close(a, b, tag):
    """A NaN corner (log10 of a non-positive extrapolated vertex) makes the
    gradient and the light term NaN, which the reference composites into
    the pixel: NaN must sit at the same pixels; finite pixels meet TOL."""
    assert np.array_equal(np.isnan(a), np.isnan(b)), tag
    ok = ~np.isnan(b)
    e = rel_l2(a[ok], b[ok])
    print(f"{tag}: NaN pixels {int((~ok).sum())}, rel-L2={e:.3e}", flush=True)
    assert e <= TOL


@pytest.mark.parametrize("grey", [False, True])
def test_light_source_uniform(ctx, grey):
    ds = uniform_lognormal(32, seed=121, nprocs=4)
    tf = vr_tf_for(ds, nbins=128)
    tf.grey_opacity = grey
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        p = c0_params(64, normal=(0.45, -0.3, 0.84))
        spec = vol.light_source_request(p, tf, record=True, n_samples=6, **LIGHT)
        b = _lib.Batch(ctx, [spec], _lib.BATCH_PERSISTENT)
        try:
            b.run()
            img, (v, h, sm) = b.read_image(0), b.read_record(0)
        finally:
            b.close()
        oimg, ov, oh, os_ = oracle.cast(3, p, oracle_vr_bricks(vol), order=spec.order,
                                        tf=_td(tf, spec), n_samples=6, record=True,
                                        with_samples=True)
        assert np.array_equal(v, ov) and np.array_equal(h, oh) and np.array_equal(sm, os_)
        for c in range(3):
            _close(img[..., c], oimg[..., c], f"light grey={grey} ch{c}")
    finally:
        vol.free()


def test_light_source_amr_perspective(ctx):
    ds = small_amr(root_n=32, levels=2, seed=123)
    tf = vr_tf_for(ds, nbins=64)
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        p = render.perspective_params((1.5, 1.6, -0.7), (0.5, 0.5, 0.5), 0.8, (40, 32))
        spec = vol.light_source_request(p, tf, **LIGHT)
        b = _lib.Batch(ctx, [spec], _lib.BATCH_PLAIN)
        try:
            b.run()
            img = b.read_image(0)
        finally:
            b.close()
        oimg = oracle.cast(3, p, oracle_vr_bricks(vol), order=spec.order, tf=_td(tf, spec))
        for c in range(3):
            _close(img[..., c], oimg[..., c], f"light AMR perspective ch{c}")
    finally:
        vol.free()
