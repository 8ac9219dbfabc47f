"""The fast walks' address arithmetic against the storage layouts.

The projection fast walk addresses a cell of a 4x4x4-blocked brick as
16 c0 + 4 c1 + c2 + (c0>>2) K0 + (c1>>2) K1 + (c2>>2) 60 with
K0 = 64 (nbj nbk - 1), K1 = 64 nbk - 16; the volume-render fast walk
addresses corner (0, 0, 0) of a cell of a vertex-blocked brick (blocks of
4^3 cells holding 5^3 vertices padded to 128 points) in bytes as
100 c0 + 20 c1 + 4 c2 + (c0>>2) K0 + (c1>>2) K1 + (c2>>2) 496 with
K0 = 4 (128 nbj nbk - 100), K1 = 4 (128 nbk - 20).  Both must equal the
layouts' definitions (blocks in C order, points in C order inside a block)
for every cell of bricks of several shapes, including ragged edges.
"""

import itertools

import numpy as np
import pytest

SHAPES = [(1, 1, 1), (3, 5, 7), (4, 4, 4), (9, 2, 13), (16, 17, 5), (33, 8, 31), (64, 64, 64)]


def _cells(shape):
    return np.meshgrid(*(np.arange(n, dtype=np.int64) for n in shape), indexing="ij")


def _blocked(i, j, k, nbj, nbk):
    """Cell-centred layout: 4^3-cell blocks in C order, 64 cells per block."""
    blk = ((i // 4) * nbj + j // 4) * nbk + k // 4
    return blk * 64 + (i % 4) * 16 + (j % 4) * 4 + (k % 4)


def _vertex(i, j, k, nbj, nbk):
    """Vertex layout: 4^3-cell blocks in C order, each the 5^3 vertices of
    its cells in C order padded to 128 points."""
    blk = ((i // 4) * nbj + j // 4) * nbk + k // 4
    return blk * 128 + ((i % 4) * 5 + (j % 4)) * 5 + (k % 4)


@pytest.mark.parametrize("shape", SHAPES)
def test_projection_fast_offset(shape):
    i, j, k = _cells(shape)
    nbj, nbk = -(-shape[1] // 4), -(-shape[2] // 4)
    K0, K1 = 64 * (nbj * nbk - 1), 64 * nbk - 16
    fast = 16 * i + 4 * j + k + (i >> 2) * K0 + (j >> 2) * K1 + (k >> 2) * 60
    assert np.array_equal(fast, _blocked(i, j, k, nbj, nbk))


@pytest.mark.parametrize("shape", SHAPES)
def test_vertex_fast_offset(shape):
    i, j, k = _cells(shape)
    nbj, nbk = -(-shape[1] // 4), -(-shape[2] // 4)
    K0, K1 = 4 * (128 * nbj * nbk - 100), 4 * (128 * nbk - 20)
    fast = 100 * i + 20 * j + 4 * k + (i >> 2) * K0 + (j >> 2) * K1 + (k >> 2) * 496
    base = _vertex(i, j, k, nbj, nbk)
    assert np.array_equal(fast, 4 * base)
    # the 8 corners at fixed byte offsets from corner (0, 0, 0), all inside
    # the cell's block; the fast walk reads them as pairs at 0, 20, 100, 120
    for a, b, c in itertools.product((0, 1), repeat=3):
        off = 4 * ((a * 5 + b) * 5 + c)
        assert off in (0, 4, 20, 24, 100, 104, 120, 124)
        corner = base + (a * 5 + b) * 5 + c
        assert np.all(corner // 128 == base // 128)


def test_vertex_offset_mutations_fail():
    """A wrong block stride or in-block weight is caught by the shapes above."""
    i, j, k = _cells((9, 6, 10))
    nbj, nbk = 2, 3
    want = 4 * _vertex(i, j, k, nbj, nbk)
    for K0, K1, K2, W in [(4 * (128 * nbj * nbk - 96), 4 * (128 * nbk - 20), 496, (100, 20, 4)),
                          (4 * (128 * nbj * nbk - 100), 4 * (128 * nbk - 16), 496, (100, 20, 4)),
                          (4 * (128 * nbj * nbk - 100), 4 * (128 * nbk - 20), 512, (100, 20, 4)),
                          (4 * (128 * nbj * nbk - 100), 4 * (128 * nbk - 20), 496, (100, 16, 4))]:
        fast = W[0] * i + W[1] * j + W[2] * k + (i >> 2) * K0 + (j >> 2) * K1 + (k >> 2) * K2
        assert not np.array_equal(fast, want)
