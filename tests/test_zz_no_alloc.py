"""ytrc_render_noalloc (the entry point that allocates no device memory)
against the persistent batch path it shares its kernels with.

Bar: bit-exact images (same kernels, same per-ray arithmetic), on two calls
in a row, the second with a smaller image and a dataset of fewer bricks
(the fixed array keeps the first call's contents), for projection, volume
rendering and an input image; a request whose staging exceeds the fixed
array is refused with an error before anything runs.
"""

import numpy as np
import pytest

from yt_amd import _lib, render

from .helpers import c0_params, small_amr, uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu


def _batch(ctx, spec):
    b = _lib.Batch(ctx, [spec], _lib.BATCH_PERSISTENT)
    try:
        b.run()
        return b.read_image(0)
    finally:
        b.close()


def test_noalloc_equals_batch_on_two_calls(ctx):
    big = uniform_lognormal(48, seed=301, nprocs=8)
    small = small_amr(root_n=16, levels=1, seed=302)
    vb = render.ResidentVolume(ctx, big, proj_fields="density", vr_field="density")
    vs = render.ResidentVolume(ctx, small, proj_fields="density", vr_field="density")
    try:
        tfb, tfs = vr_tf_for(big, nbins=256), vr_tf_for(small, nbins=64)
        pb = c0_params(200, normal=(0.3, 0.5, 0.81))
        ps = c0_params(56, normal=(-0.6, 0.2, 0.77))
        rng = np.random.default_rng(303)
        calls = [
            ("VR 200^2, 8 bricks", vb.vr_request(pb, tfb)),
            ("VR 56^2, AMR", vs.vr_request(ps, tfs)),
            ("projection 200^2, 8 bricks", vb.projection_request(pb)),
            ("projection 56^2, AMR", vs.projection_request(ps)),
        ]
        spec = vs.vr_request(ps, tfs)
        spec.image_in = 0.1 + rng.uniform(0, 0.2, size=(56, 56, 4))
        calls.append(("VR 56^2 over an input image", spec))
        for name, spec in calls:
            print(f"noalloc: {name}", flush=True)
            got = ctx.render_noalloc(spec)
            want = _batch(ctx, spec)
            assert np.array_equal(got.view(np.int64), want.view(np.int64)), name
            assert np.any(got != 0), name
    finally:
        vb.free()
        vs.free()


def test_noalloc_refuses_what_does_not_fit(ctx):
    ds = uniform_lognormal(16, seed=304)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        tf = vr_tf_for(ds, nbins=32)
        # 3 float32 output planes of 4096^2 are 192 MiB: more than the array
        with pytest.raises(ValueError, match="fixed array"):
            ctx.render_noalloc(vol.vr_request(c0_params(4096), tf))
        with pytest.raises(ValueError, match="record"):
            ctx.render_noalloc(vol.projection_request(c0_params(16), record=True))
        # the context still works afterwards
        p = c0_params(24)
        spec = vol.projection_request(p)
        assert np.array_equal(ctx.render_noalloc(spec), _batch(ctx, spec))
    finally:
        vol.free()
