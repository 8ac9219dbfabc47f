"""Composition of partial renders on the CPU, with the FP64 oracle as the
renderer (no GPU).

* compose_partials: the affine maps rgba -> C + T rgba of consecutive parts
  of a brick order (C = the part rendered from zeros, T = from ones minus C)
  applied in call order give the whole order's image, for non-grey and
  grey transfer functions and a non-zero starting image.
* reduce_tree_images (amr_kdtree.py:283-323 / amr_kdtools.py:6-39): for a
  grey-opacity transfer function, the pairwise over-operator merge of the
  subtrees' images rendered from zeros is the whole render; without
  opacity the merge is the sum of the images.
"""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import render
from yt_amd.amr_kdtree import AMRKDTree
from yt_amd.transfer_functions import ColorTransferFunction

from .helpers import c0_params, small_amr, tf_dict, vr_tf_for


def _setup(grey):
    ds = small_amr(root_n=16, levels=2, seed=501)
    kd = AMRKDTree(ds)
    kd.set_fields(["density"], [True], no_ghost=True)
    leaves = kd.leaves()
    index = {id(n): i for i, n in enumerate(leaves)}
    bricks = []
    for n in leaves:
        pg = kd.get_brick_data(n)
        bricks.append(dict(le=pg.LeftEdge, re=pg.RightEdge, dims=pg.dims,
                           data=list(pg.my_data), mask=pg.source_mask))
    if grey:
        lo = min(np.log10(g["density"].min()) for g in ds.grids)
        hi = max(np.log10(g["density"].max()) for g in ds.grids)
        tf = ColorTransferFunction((lo, hi), nbins=64, grey_opacity=True)
        tf.add_gaussian(0.5 * (lo + hi), 0.1 * (hi - lo) ** 2, [0.8, 0.4, 0.2, 6.0])
        tf.add_gaussian(0.3 * lo + 0.7 * hi, 0.05 * (hi - lo) ** 2, [0.1, 0.6, 0.9, 9.0])
    else:
        tf = vr_tf_for(ds, nbins=64)
    p = c0_params(40, normal=(0.3, -0.5, 0.8))
    return kd, index, bricks, tf_dict(tf), p


def _order(kd, index, node, vp):
    return np.array([index[id(n)] for n in node.kd_traverse(viewpoint=vp)], dtype=np.int32)


@pytest.mark.parametrize("grey", [False, True])
@pytest.mark.parametrize("depth", [1, 2, 3])
def test_compose_partials_is_the_whole_render(grey, depth):
    kd, index, bricks, tf, p = _setup(grey)
    vp = p["viewpoint"]
    rng = np.random.default_rng(depth)
    image_in = rng.uniform(0.0, 0.5, size=(40, 40, 4))
    full = oracle.cast(1, p, bricks, order=_order(kd, index, kd.trunk, vp), tf=tf,
                       image=image_in.copy())
    maps = []
    parts = kd.subtrees(depth, viewpoint=vp)
    assert len(parts) > 1
    for part in parts:
        order = _order(kd, index, part, vp)
        if len(order) == 0:
            continue
        c = oracle.cast(1, p, bricks, order=order, tf=tf, image=np.zeros((40, 40, 4)))
        t = oracle.cast(1, p, bricks, order=order, tf=tf, image=np.ones((40, 40, 4))) - c
        maps.append(np.concatenate([c, t], axis=2))
    got = render.compose_partials(maps, image_in=image_in)
    err = np.abs(got - full).max()
    print(f"grey={grey} depth={depth} parts={len(maps)}: max |compose - whole| = {err:.2e}")
    assert err <= 1e-12


def test_subtrees_partition_the_leaves():
    kd, index, _, _, p = _setup(False)
    vp = p["viewpoint"]
    whole = [id(n) for n in kd.trunk.kd_traverse(viewpoint=vp)]
    for depth in range(5):
        parts = kd.subtrees(depth, viewpoint=vp)
        cat = [id(n) for part in parts for n in part.kd_traverse(viewpoint=vp)]
        assert cat == whole
        assert len(parts) <= 2 ** depth


def test_reduce_tree_images():
    for grey in (True, False):
        kd, index, bricks, tf, p = _setup(grey)
        vp = p["viewpoint"]
        full = oracle.cast(1, p, bricks, order=_order(kd, index, kd.trunk, vp), tf=tf)
        for depth in (1, 2):
            imgs = []
            for part in kd.subtrees(depth):
                order = _order(kd, index, part, vp)
                imgs.append(oracle.cast(1, p, bricks, order=order, tf=tf) if len(order)
                            else np.zeros((40, 40, 4)))
            red = render.reduce_tree_images(kd, imgs, vp, depth, use_opacity=grey)
            if grey:
                err = np.abs(red - full).max()
                print(f"grey depth={depth}: max |reduce - whole| = {err:.2e}")
                assert err <= 1e-12
            else:
                assert np.allclose(red, np.sum(imgs, axis=0), rtol=0, atol=1e-15)
