"""Partial-map requests (YTRC_FLAG_PARTIAL) on the GPU: the VR of each kd
subtree as an affine map rgba -> C + T rgba, composed in the viewpoint's
order, against the FP64 oracle's render of the whole tree (rel-L2 <= 1e-5
per channel): one GPU's share when kd subtrees are split over GPUs, and the
exchange that replaces yt's receive_and_reduce."""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, render
from yt_amd.transfer_functions import ColorTransferFunction

from .conftest import rel_l2
from .helpers import c0_params, oracle_vr_bricks, small_amr, tf_dict, vr_tf_for

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _grey_tf(ds):
    lo = min(np.log10(g["density"].min()) for g in ds.grids)
    hi = max(np.log10(g["density"].max()) for g in ds.grids)
    tf = ColorTransferFunction((lo, hi), nbins=64, grey_opacity=True)
    tf.add_gaussian(0.5 * (lo + hi), 0.1 * (hi - lo) ** 2, [0.8, 0.4, 0.2, 6.0])
    tf.add_gaussian(0.3 * lo + 0.7 * hi, 0.05 * (hi - lo) ** 2, [0.1, 0.6, 0.9, 9.0])
    return tf


def _maps(ctx, specs):
    b = _lib.Batch(ctx, specs, _lib.BATCH_PERSISTENT)
    try:
        b.run()
        return [b.read_partial(i) for i in range(len(specs))]
    finally:
        b.close()


@pytest.mark.parametrize("grey", [False, True])
def test_partial_maps_compose_to_whole_render(ctx, grey):
    ds = small_amr(root_n=32, levels=2, seed=511)
    tf = _grey_tf(ds) if grey else vr_tf_for(ds, nbins=128)
    vol = render.ResidentVolume(ctx, ds, vr_field="density", log_field=True)
    try:
        p = c0_params(112, normal=(-0.4, 0.35, 0.75))
        bricks = oracle_vr_bricks(vol)
        rng = np.random.default_rng(7)
        image_in = rng.uniform(0.0, 0.4, size=(112, 112, 4))
        whole = oracle.cast(1, p, bricks, order=vol.vr_order(p["viewpoint"]), tf=tf_dict(tf),
                            image=image_in.copy())
        nch = 4 if grey else 3
        for depth in (1, 2, 3):
            specs = vol.vr_partial_requests(p, tf, depth)
            assert len(specs) > 1
            assert np.array_equal(np.concatenate([s.order for s in specs]),
                                  vol.vr_order(p["viewpoint"]))
            maps = _maps(ctx, specs)
            img = render.compose_partials(maps, image_in=image_in)
            for c in range(nch):
                e = rel_l2(img[..., c], whole[..., c])
                print(f"grey={grey} depth={depth} parts={len(specs)} ch{c}: rel-L2={e:.3e}",
                      flush=True)
                assert e <= TOL
            if not grey:
                # alpha is left alone by a non-grey render: C = 0, T = 1
                assert all(np.all(m[..., 3] == 0.0) and np.all(m[..., 7] == 1.0) for m in maps)
                assert np.array_equal(img[..., 3], image_in[..., 3])
    finally:
        vol.free()


def test_partial_grey_reduce_tree_images(ctx):
    """The subtrees' images from zeros are the maps' C; merged pairwise with
    yt's over-operator (grey opacity) they give the whole render."""
    ds = small_amr(root_n=32, levels=2, seed=513)
    tf = _grey_tf(ds)
    vol = render.ResidentVolume(ctx, ds, vr_field="density", log_field=True)
    try:
        p = c0_params(96, normal=(0.5, 0.2, -0.6))
        whole = oracle.cast(1, p, oracle_vr_bricks(vol), order=vol.vr_order(p["viewpoint"]),
                            tf=tf_dict(tf))
        depth = 2
        vp = p["viewpoint"]
        # one image per subtree, in kd.subtrees(depth) index order
        parts = vol.kd.subtrees(depth)
        specs, which = [], []
        tfc = _lib.TF.from_tf_obj(tf)
        for i, part in enumerate(parts):
            order = np.array([vol._leaf_index[id(n)] for n in part.kd_traverse(viewpoint=vp)],
                             dtype=np.int32)
            if len(order):
                specs.append(_lib.RequestSpec(_lib.VOLUME_RENDER, p, vol.vr_ds_id, order,
                                              tf=tfc, partial=True))
                which.append(i)
        maps = _maps(ctx, specs)
        imgs = [np.zeros((96, 96, 4)) for _ in parts]
        for i, m in zip(which, maps):
            imgs[i] = m[..., :4]
        red = render.reduce_tree_images(vol.kd, imgs, vp, depth, use_opacity=True)
        for c in range(4):
            e = rel_l2(red[..., c], whole[..., c])
            print(f"grey reduce ch{c}: rel-L2={e:.3e}", flush=True)
            assert e <= TOL
    finally:
        vol.free()


def test_partial_request_errors(ctx):
    ds = small_amr(root_n=16, levels=1, seed=515)
    tf = vr_tf_for(ds, nbins=32)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        p = c0_params(16)
        proj = vol.projection_request(p)
        proj.partial = True
        with pytest.raises(NotImplementedError):
            _lib.Batch(ctx, [proj]).close()
        vr = vol.vr_request(p, tf)
        vr.partial = True
        vr.image_in = np.zeros((16, 16, 4))
        with pytest.raises(ValueError):
            _lib.Batch(ctx, [vr]).close()
        vr.image_in = None
        b = _lib.Batch(ctx, [vr])
        try:
            b.run()
            with pytest.raises(ValueError):
                b.read_image(0)
            assert b.read_partial(0).shape == (16, 16, 8)
        finally:
            b.close()
    finally:
        vol.free()
