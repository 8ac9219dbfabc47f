"""GPU parity at the measured workloads' full sizes, against the FP64 oracle
on the same inputs (bit-exact visit sequences and per-pixel sample counts;
image rel-L2 <= 1e-5):

  * 256^3 lognormal density + temperature, density-weighted projection of
    temperature, 1024^2;
  * AMR hierarchy (128^3 root + 3 levels, ~640 grids), VolumeRenderSampler
    with an 8-Gaussian ColorTransferFunction over log10 density, 1024^2;
  * 512^3 float32 field in 2x2x2 bricks: one projection and one VR image.
"""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, data, render

from .conftest import rel_l2
from .helpers import c0_params, oracle_proj_bricks, oracle_vr_bricks, tf_dict, vr_tf_for

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _one(ctx, spec):
    b = _lib.Batch(ctx, [spec], _lib.BATCH_PERSISTENT)
    try:
        b.run()
        return b.read_image(0), b.read_record(0)
    finally:
        b.close()


def _timed_kernels(ctx, spec):
    """The same request without record mode, as the timed kernels run it
    (persistent launch in a hipGraph)."""
    spec.record = False
    b = _lib.Batch(ctx, [spec], _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH)
    try:
        b.run()
        return b.read_image(0)
    finally:
        b.close()
        spec.record = True


def _visits_equal(rec, ov, oh, os_, tag):
    v, h, s = rec
    nd, nh, ns = int(np.sum(v != ov)), int(np.sum(h != oh)), int(np.sum(s != os_))
    print(f"{tag}: visits {int(v.sum())}, samples {int(s.sum())} (oracle {int(os_.sum())}), "
          f"count-mismatch={nd} hash-mismatch={nh} sample-mismatch={ns}", flush=True)
    assert nd == 0 and nh == 0 and ns == 0


def test_weighted_projection_256_at_1024(ctx):
    rho = data.lognormal_density(256, seed=1001)
    T = data.temperature_from_density(rho, seed=1002)
    ds = data.load_uniform_grid({"density": rho, "temperature": T})
    vol = render.ResidentVolume(ctx, ds, proj_fields="temperature", weight="density")
    try:
        p = c0_params(1024)
        img, rec = _one(ctx, vol.projection_request(p, record=True))
        oimg, ov, oh, os_ = oracle.cast(
            0, p, oracle_proj_bricks(ds, ["temperature"], weight="density"), record=True,
            with_samples=True)
        _visits_equal(rec, ov, oh, os_, "256^3 weighted 1024^2")
        ref = render.finalize_projection(oimg, p["width"], True)
        for tag, im in (("record kernels", img),
                        ("timed kernels", _timed_kernels(ctx, vol.projection_request(p, record=True)))):
            out = render.finalize_projection(im, p["width"], True)
            e = rel_l2(out, ref)
            m = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
            print(f"256^3 weighted 1024^2 ({tag}): rel-L2={e:.3e} max-abs/max={m:.3e}", flush=True)
            assert e <= TOL and m <= TOL
    finally:
        vol.free()


def test_amr_vr_at_1024(ctx):
    gd = data.make_amr_hierarchy(128, 3)
    ds = data.load_amr_grids(gd, [128] * 3)
    tf = vr_tf_for(ds)
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        p = c0_params(1024, normal=(0.3, 0.5, 0.81))
        spec = vol.vr_request(p, tf, record=True)
        img, rec = _one(ctx, spec)
        oimg, ov, oh, os_ = oracle.cast(1, p, oracle_vr_bricks(vol), order=spec.order,
                                        tf=tf_dict(tf), record=True, with_samples=True)
        _visits_equal(rec, ov, oh, os_, f"AMR ({len(gd)} grids, {vol.n_vr_bricks} bricks) VR 1024^2")
        timed = _timed_kernels(ctx, spec)
        for tag, im in (("record kernels", img), ("timed kernels", timed)):
            for c in range(3):
                e = rel_l2(im[..., c], oimg[..., c])
                m = float(np.max(np.abs(im[..., c] - oimg[..., c])))
                print(f"AMR VR ch{c} ({tag}): rel-L2={e:.3e} max-abs={m:.3e} "
                      f"(max {oimg[..., c].max():.3f})", flush=True)
                assert e <= TOL and m <= TOL * oimg[..., c].max()
    finally:
        vol.free()


def test_512_bricks_projection_and_vr(ctx):
    rho = data.lognormal_density(512, seed=1001, dtype=np.float32)
    ds = data.load_uniform_grid({"density": rho}, psize=(2, 2, 2))
    del rho
    tf = vr_tf_for(ds)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        pp = c0_params(256, normal=(-0.62, 0.33, 0.71))
        img, rec = _one(ctx, vol.projection_request(pp, record=True))
        oimg, ov, oh, os_ = oracle.cast(0, pp, oracle_proj_bricks(ds, ["density"]), record=True,
                                        with_samples=True)
        _visits_equal(rec, ov, oh, os_, "512^3 projection 256^2")
        e = rel_l2(img[..., 0], oimg[..., 0])
        m = float(np.max(np.abs(img[..., 0] - oimg[..., 0])) / np.max(np.abs(oimg[..., 0])))
        print(f"512^3 projection: rel-L2={e:.3e} max-abs/max={m:.3e}", flush=True)
        assert e <= TOL
        pv = c0_params(192, normal=(0.45, -0.8, 0.39))
        spec = vol.vr_request(pv, tf, record=True)
        img, rec = _one(ctx, spec)
        oimg, ov, oh, os_ = oracle.cast(1, pv, oracle_vr_bricks(vol), order=spec.order,
                                        tf=tf_dict(tf), record=True, with_samples=True)
        _visits_equal(rec, ov, oh, os_, "512^3 VR 192^2")
        for c in range(3):
            e = rel_l2(img[..., c], oimg[..., c])
            m = float(np.max(np.abs(img[..., c] - oimg[..., c])))
            print(f"512^3 VR ch{c}: rel-L2={e:.3e} max-abs={m:.3e} (max {oimg[..., c].max():.3f})",
                  flush=True)
            assert e <= TOL
    finally:
        vol.free()
