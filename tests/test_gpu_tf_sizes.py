"""ColorTransferFunction bin counts on the GPU.

The transmittance-form table holds 48 B per bin plus three zero entries in
LDS, so one launch takes 2 <= nbins <= max_lds / 48 - 2, with max_lds the
LDS one workgroup may allocate on the device (64 KiB: 1363 bins; 160 KiB:
3411 bins).  Sizes from 2 bins to that maximum are rendered in a plain
launch and in a persistent batch and compared with the FP64 oracle (rel-L2
<= 1e-5 per channel); one bin more is refused with a clear error instead
of being launched."""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, render
from yt_amd.transfer_functions import ColorTransferFunction

from .conftest import rel_l2
from .helpers import c0_params, oracle_vr_bricks, tf_dict, uniform_lognormal

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _max_bins(ctx):
    return ctx.max_lds // 48 - 2


def _tf(ds, nbins):
    rho = ds.grids[0]["density"]
    lo, hi = float(np.log10(rho.min())), float(np.log10(rho.max()))
    tf = ColorTransferFunction((lo, hi), nbins=nbins)
    for j, x in enumerate(np.linspace(0.1, 0.9, 5)):
        tf.add_gaussian(lo + x * (hi - lo), 0.004 * (hi - lo) ** 2,
                        [0.9, 0.2 + 0.15 * j, 0.5, 2.0 + j])
    return tf


@pytest.fixture(scope="module")
def vol(ctx):
    ds = uniform_lognormal(24, seed=41)
    v = render.ResidentVolume(ctx, ds, vr_field="density", log_field=True)
    yield v
    v.free()


@pytest.mark.parametrize("nbins", [2, 3, 1024, 1363, 2048, "max"])
@pytest.mark.parametrize("mode", ["plain", "persistent"])
def test_ctf_bin_counts(ctx, vol, nbins, mode):
    if nbins == "max":
        nbins = _max_bins(ctx)
    if nbins > _max_bins(ctx):
        pytest.skip(f"{nbins} bins exceed the {ctx.max_lds} B of LDS of a workgroup here")
    tf = _tf(vol.ds, nbins)
    p = c0_params(64, normal=(0.3, -0.45, 0.84))
    spec = vol.vr_request(p, tf, n_samples=10)
    b = _lib.Batch(ctx, [spec], _lib.BATCH_PLAIN if mode == "plain" else _lib.BATCH_PERSISTENT)
    try:
        b.run()
        img = b.read_image(0)
    finally:
        b.close()
    oimg = oracle.cast(1, p, oracle_vr_bricks(vol), order=spec.order, tf=tf_dict(tf),
                       n_samples=10)
    assert oimg[..., :3].max() > 0.0
    for c in range(3):
        e = rel_l2(img[..., c], oimg[..., c])
        print(f"nbins={nbins} {mode} ch{c}: rel-L2={e:.3e}", flush=True)
        assert e <= TOL


def test_ctf_too_many_bins_refused(ctx, vol):
    print(f"max_lds = {ctx.max_lds} B: at most {_max_bins(ctx)} bins", flush=True)
    tf = _tf(vol.ds, _max_bins(ctx) + 1)
    spec = vol.vr_request(c0_params(16), tf, n_samples=10)
    for mode in (_lib.BATCH_PLAIN, _lib.BATCH_PERSISTENT):
        with pytest.raises(ValueError, match="B of LDS"):
            _lib.Batch(ctx, [spec], mode).close()
