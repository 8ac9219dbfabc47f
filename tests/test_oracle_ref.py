"""The oracle's trilinear interpolation against yt's own offset_interpolate
(fixed_interpolator.cpp:34-52), compiled from the reference tree into
oracle/_ref/ by `make -C oracle ref`: bit-equal FP64 results at random and
boundary positions.  Skipped where the reference tree was not available to
build it."""

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None,
                                reason="oracle/_ref not built (no reference tree here)")


@pytest.mark.parametrize("dims,seed", [((5, 7, 3), 1), ((16, 16, 16), 2), ((1, 2, 9), 3)])
def test_trilinear_bit_equal_to_reference(dims, seed):
    rng = np.random.default_rng(seed)
    v = np.exp(rng.normal(size=tuple(d + 1 for d in dims)))
    n = 400
    cells = np.stack([rng.integers(0, d, n) for d in dims], axis=1).astype(np.int32)
    dps = rng.random((n, 3))
    dps[:40] = rng.integers(0, 2, (40, 3))          # corners and faces
    dps[40:60, 1] = 0.5
    got = oracle.trilinear(dims, v, cells, dps)
    want = oracle.ref_offset_interpolate(dims, v, cells, dps)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("dims,seed", [((5, 7, 3), 11), ((8, 8, 8), 12)])
def test_gradient_bit_equal_to_reference(dims, seed):
    """eval_gradient (fixed_interpolator.cpp:116-148), including the
    one-sided differences within 0.05 of a face and the thresholds
    themselves."""
    rng = np.random.default_rng(seed)
    v = np.exp(rng.normal(size=tuple(d + 1 for d in dims)))
    n = 300
    cells = np.stack([rng.integers(0, d, n) for d in dims], axis=1).astype(np.int32)
    dps = rng.random((n, 3))
    dps[:30] = rng.choice([0.0, 0.05, 0.95, 1.0, 0.04999999, 0.95000001], size=(30, 3))
    v[0, 0, 0] = v[0, 0, 1]  # a few flat corners
    got = oracle.gradients(dims, v, cells, dps)
    want = oracle.ref_eval_gradient(dims, v, cells, dps)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    flat = np.ones(tuple(d + 1 for d in dims))
    assert np.all(oracle.gradients(dims, flat, cells[:10], dps[:10]) == 0.0)
