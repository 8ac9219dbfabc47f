"""The kd-tree build, its viewpoint order and the no-ghost vertex centring
against golden vectors written by the reference's own Cython
(tests/golden/make_kdtree_vectors.py: amr_kdtools.pyx and interpolators.pyx
compiled from the reference tree in the build container).

Bar: exact.  Every node's id, grid, split dimension, split position and
edges; every brick order (node ids, farthest from the viewpoint first) for
32 viewpoints per hierarchy, 8 of them lying on split planes; and every bit
of ghost_zone_interpolate's output.  Checked for the restatement in
oracle/kdtree.py, for the native tree (ytrc_kdtree_*, host code of
libytrc.so) and for the host vertex centring (yt_amd.data).
"""

import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KD = os.path.join(HERE, "golden", "kdtree_vectors.npz")
VC = os.path.join(HERE, "golden", "vertex_centre_vectors.npz")


def _kd():
    return np.load(KD, allow_pickle=False)


CASES = ["amr16", "amr32", "amr64x3", "uniform8", "uniform64", "psize343", "c2"]


def _grids(z, case):
    return (z[f"{case}/gle"], z[f"{case}/gre"], z[f"{case}/gid"], z[f"{case}/level"])


def test_fixture_lists_every_case():
    assert sorted(_kd()["cases"].tolist()) == sorted(CASES)


def _check_nodes(z, case, nodes):
    """nodes: depth-first list of (node_id, grid, split_dim, split_pos,
    left_edge, right_edge, is_leaf)."""
    assert len(nodes) == len(z[f"{case}/node_id"])
    nid = np.array([n[0] for n in nodes], dtype=np.int64)
    grid = np.array([n[1] for n in nodes], dtype=np.int64)
    sdim = np.array([n[2] for n in nodes], dtype=np.int32)
    spos = np.array([n[3] for n in nodes], dtype=np.float64)
    le = np.array([n[4] for n in nodes], dtype=np.float64)
    re = np.array([n[5] for n in nodes], dtype=np.float64)
    leaf = np.array([n[6] for n in nodes], dtype=np.int8)
    assert np.array_equal(nid, z[f"{case}/node_id"])
    assert np.array_equal(grid, z[f"{case}/grid"])
    assert np.array_equal(sdim, z[f"{case}/split_dim"])
    internal = sdim >= 0
    # split positions bit for bit where a node is split
    assert np.array_equal(spos[internal].view(np.int64),
                          z[f"{case}/split_pos"][internal].view(np.int64))
    assert np.array_equal(le.view(np.int64), z[f"{case}/left_edge"].view(np.int64))
    assert np.array_equal(re.view(np.int64), z[f"{case}/right_edge"].view(np.int64))
    assert np.array_equal(leaf, z[f"{case}/leaf"])


@pytest.mark.parametrize("case", CASES)
def test_oracle_kdtree_equals_reference(case):
    from oracle import kdtree as okd

    z = _kd()
    gle, gre, gid, lvl = _grids(z, case)
    trunk = okd.build([(gle[i], gre[i], int(gid[i]), int(lvl[i])) for i in range(len(gid))])
    nodes = [(n.node_id, n.grid, n.split_dim, n.split_pos, n.left_edge, n.right_edge,
              int(n.is_leaf())) for n in trunk.depth_traverse()]
    _check_nodes(z, case, nodes)
    assert np.array_equal([n.node_id for n in trunk.kd_traverse()], z[f"{case}/leaves"])
    for vp, want in zip(z[f"{case}/viewpoints"], z[f"{case}/orders"]):
        got = [n.node_id for n in trunk.kd_traverse(viewpoint=vp)]
        assert np.array_equal(got, want), vp


@pytest.mark.parametrize("case", CASES)
def test_native_kdtree_equals_reference(case):
    from yt_amd import _lib

    z = _kd()
    gle, gre, gid, lvl = _grids(z, case)
    kd = _lib.KdTree(gle, gre, gid, lvl)
    try:
        raw = kd.nodes()
        pre, stack = [], [0]  # the native array is in creation order: walk it depth first
        while stack:
            i = stack.pop()
            pre.append(raw[i])
            if raw[i].right >= 0:
                stack.append(raw[i].right)
            if raw[i].left >= 0:
                stack.append(raw[i].left)
        nodes = [(n.node_id, n.grid, n.split_dim, n.split_pos if n.split_dim >= 0 else np.nan,
                  tuple(n.left_edge), tuple(n.right_edge), int(n.left < 0 or n.right < 0))
                 for n in pre]
        _check_nodes(z, case, nodes)
        leaves = z[f"{case}/leaves"]
        assert kd.n_leaves == len(leaves)
        for vp, want in zip(z[f"{case}/viewpoints"], z[f"{case}/orders"]):
            assert np.array_equal(leaves[kd.order(vp)], want), vp
    finally:
        kd.close()


def test_vertex_centring_equals_reference_bits():
    from yt_amd import data

    z = np.load(VC, allow_pickle=False)
    for i in range(int(z["n"])):
        got = data.ghost_zone_interpolate_no_ghost(z[f"in{i}"])
        want = z[f"out{i}"]
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), i


def test_a_wrong_median_fails_the_vectors():
    """The vectors catch a plausible mistake in the split rule: the median
    index rounded up instead of down changes the tree."""
    from oracle import kdtree as okd

    z = _kd()
    gle, gre, gid, lvl = _grids(z, "amr16")
    orig = okd.kdtree_get_choices

    def rounded_up(gles, gres, l_corner, r_corner):
        best_dim, my_max, uniq_best = -1, 0, None
        for dim in range(3):
            v = np.concatenate([gles[:, dim], gres[:, dim]])
            u = np.unique(v[(l_corner[dim] < v) & (v < r_corner[dim])])
            if len(u) > my_max:
                best_dim, my_max, uniq_best = dim, len(u), u
        if best_dim == -1:
            return -1, 0.0, None, None
        split = uniq_best[len(uniq_best) // 2]
        return best_dim, split, gles[:, best_dim] < split, gres[:, best_dim] > split

    okd.kdtree_get_choices = rounded_up
    try:
        trunk = okd.build([(gle[i], gre[i], int(gid[i]), int(lvl[i])) for i in range(len(gid))])
        spos = [n.split_pos for n in trunk.depth_traverse()]
    finally:
        okd.kdtree_get_choices = orig
    want = z["amr16/split_pos"]
    assert len(spos) != len(want) or not np.array_equal(np.nan_to_num(spos), np.nan_to_num(want))
