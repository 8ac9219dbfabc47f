"""Multi-field kd volumes on the GPU path against the FP64 oracle.

* VolumeRenderSampler over two-field bricks with a MultiVariateTransferFunction
  whose tables read different fields, one weighted by a raw field value
  (weight_field_id) and two by other tables (weight_table_id, a chain), grey
  and not (field_interpolation_tables.pxd:56-110).
* A VolumeSource with a weight field and a ProjectionTransferFunction: two
  linear fields, InterpolatedProjectionSampler (render_source.py:104-134).
* ProjectionSampler handed vertex-centred kd bricks (set_sampler(camera,
  interpolated=False), render_source.py:464-488): it reads each brick's
  (d+1)^3 vertex array with the cell strides of d (image_samplers.pyx:341);
  the oracle is given the whole vertex arrays and indexes them the same way.

Bars: cell-visit sequences bit-exact; images rel-L2 <= 1e-5 against FP64.
"""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, render, scene
from yt_amd.camera import Camera
from yt_amd.transfer_functions import (
    MultiVariateTransferFunction,
    ProjectionTransferFunction,
    TransferFunction,
)

from .conftest import rel_l2
from .helpers import c0_params, oracle_vr_bricks, small_amr, tf_dict, uniform_lognormal

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run(ctx, spec):
    b = _lib.Batch(ctx, [spec], _lib.BATCH_PERSISTENT)
    try:
        b.run()
        img = b.read_image(0)
        rec = b.read_record(0) if spec.record else None
        return img, rec
    finally:
        b.close()


def _check_visits(rec, ovis, ohash, tag):
    v, h, _ = rec
    nd = int(np.sum(v != ovis))
    nh = int(np.sum(h != ohash))
    print(f"{tag}: visits gpu={int(v.sum())} oracle={int(ovis.sum())} "
          f"count-mismatch={nd} hash-mismatch={nh}", flush=True)
    assert nd == 0 and nh == 0


def _range(ds, field, log):
    v = np.concatenate([np.ravel(g[field]) for g in ds.grids])
    v = np.log10(v) if log else v
    return float(v.min()), float(v.max())


def _multivariate_tf(ds, grey):
    """Tables on log10 density (field 0) and linear temperature (field 1):
    the weight field stays finite where no-ghost vertex centring
    extrapolates a density vertex to <= 0 (log10 -> NaN, table value 0)."""
    lr, hr = _range(ds, "density", True)
    lt, ht = _range(ds, "temperature", False)
    tf = MultiVariateTransferFunction(grey_opacity=grey)
    t0 = TransferFunction((lr, hr), nbins=64)
    t0.add_gaussian(0.6 * lr + 0.4 * hr, 0.05 * (hr - lr) ** 2, 0.9)
    t1 = TransferFunction((lt, ht), nbins=48)
    t1.add_gaussian(0.5 * (lt + ht), 0.08 * (ht - lt) ** 2, 0.7)
    t2 = TransferFunction((lr, hr), nbins=32)
    t2.add_line((lr, 0.1), (hr, 0.8))
    t3 = TransferFunction((lt, ht), nbins=64)
    t3.add_gaussian(0.3 * lt + 0.7 * ht, 0.1 * (ht - lt) ** 2, 2.0)
    t3.add_step(lt, 0.5 * (lt + ht), 0.3)
    tf.add_field_table(t0, 0, weight_field_id=1)   # x raw T
    tf.add_field_table(t1, 1, weight_table_id=3)   # x alpha table
    tf.add_field_table(t2, 0, weight_table_id=0)   # x table 0 (already weighted)
    tf.add_field_table(t3, 1)
    for c in range(4):
        tf.link_channels(c, c)
    return tf


@pytest.mark.parametrize("grey", [False, True])
def test_vr_two_fields_multivariate(ctx, grey):
    ds = uniform_lognormal(32, seed=301, nprocs=8, with_temperature=True)
    tf = _multivariate_tf(ds, grey)
    vol = render.ResidentVolume(ctx, ds, vr_field=["density", "temperature"],
                                log_field=[True, False])
    try:
        p = c0_params(96, normal=(0.45, -0.3, 0.7))
        spec = vol.vr_request(p, tf, n_samples=6, record=True)
        img, rec = _run(ctx, spec)
        oimg, ovis, ohash = oracle.cast(1, p, oracle_vr_bricks(vol), order=spec.order,
                                        tf=tf_dict(tf), n_samples=6, record=True)
        assert np.all(np.isfinite(oimg))
        _check_visits(rec, ovis, ohash, f"vr two fields grey={grey}")
        for c in range(4 if grey else 3):
            e = rel_l2(img[..., c], oimg[..., c])
            print(f"vr two fields grey={grey} ch{c}: rel-L2={e:.3e} "
                  f"max={np.abs(oimg[..., c]).max():.3e}", flush=True)
            assert e <= TOL
    finally:
        vol.free()


def _camera(ds, res=(40, 36)):
    cam = Camera(None, ds)
    cam.resolution = res
    cam.width = 0.95
    cam.position = (1.3, -0.4, 1.5)
    cam.focus = (0.5, 0.5, 0.5)
    return cam


def test_weighted_projection_source(ctx):
    """ProjectionTransferFunction + weight field: the kd bricks carry
    [temperature, density] without log, and the interpolated projection
    integrates both."""
    ds = small_amr(root_n=32, levels=2, seed=311)
    for g in ds.grids:
        g.fields["temperature"] = g["density"] ** 0.6
    src = scene.VolumeSource(ds, "temperature", weight_field="density",
                             transfer_function=ProjectionTransferFunction(n_fields=2), ctx=ctx)
    assert src.sampler_type == "projection" and src.log_field is False
    try:
        cam = _camera(ds)
        im = src.render(cam)
        p = cam._get_sampler_params(src)
        p["viewpoint"] = np.array(cam.lens.viewpoint)
        order = src.volume.vr_order(p["viewpoint"])
        oimg = oracle.cast(2, p, oracle_vr_bricks(src.volume), order=order)
        for c in range(2):
            e = rel_l2(im[..., c], oimg[..., c])
            print(f"weighted projection source ch{c}: rel-L2={e:.3e}", flush=True)
            assert e <= TOL
        assert np.all(im[..., 3] == 1.0)
    finally:
        src.free()


def test_cell_indexed_projection_of_vertex_bricks(ctx):
    ds = small_amr(root_n=32, levels=2, seed=313)
    vol = render.ResidentVolume(ctx, ds, vr_field="density", log_field=False)
    try:
        p = c0_params(80, normal=(-0.3, 0.55, 0.62))
        spec = vol.cell_indexed_projection_request(p, record=True)
        img, rec = _run(ctx, spec)
        oimg, ovis, ohash = oracle.cast(0, p, oracle_vr_bricks(vol), order=spec.order,
                                        record=True)
        _check_visits(rec, ovis, ohash, "cell-indexed projection of vertex bricks")
        e = rel_l2(img[..., 0], oimg[..., 0])
        print(f"cell-indexed projection: rel-L2={e:.3e}", flush=True)
        assert e <= TOL
    finally:
        vol.free()


def test_set_sampler_not_interpolated(ctx):
    """The drop-in sampler from set_sampler(camera, interpolated=False),
    called with the source's kd bricks in viewpoint order, gives the
    oracle's ProjectionSampler image over those vertex arrays."""
    ds = small_amr(root_n=32, levels=2, seed=317)
    src = scene.VolumeSource(ds, "density", transfer_function=ProjectionTransferFunction(),
                             ctx=ctx)
    try:
        cam = _camera(ds, res=(32, 32))
        sampler = src.set_sampler(cam, interpolated=False)
        vp = np.array(cam.lens.viewpoint)
        bricks = list(src.volume.kd.traverse(vp))
        for b in bricks:
            sampler(b, num_threads=0)
        img = sampler.aimage
        p = cam._get_sampler_params(src)
        ob = [dict(le=b.LeftEdge, re=b.RightEdge, dims=b.dims,
                   data=[np.asarray(d, dtype=np.float64) for d in b.my_data], mask=None)
              for b in bricks]
        oimg = oracle.cast(0, p, ob)
        e = rel_l2(img[..., 0], oimg[..., 0])
        print(f"set_sampler(interpolated=False): rel-L2={e:.3e}", flush=True)
        assert e <= TOL
    finally:
        src.free()
