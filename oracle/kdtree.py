"""TEST INFRASTRUCTURE ONLY -- the kd-tree build and viewpoint order of
AMRKDTree, restated in Python as the checker of the native build
(ytrc_kdtree_*, yt_amd/csrc/ytrc_kdtree.hip).

Restates, for one process (rank 0 of 1: should_i_build is always true and
should_i_split never fires):
  Node, add_grids, insert_grids, split_grids, divide  amr_kdtools.pyx:193-532
  kdtree_get_choices                                  amr_kdtools.pyx:760-823
  viewpoint_traverse / step_viewpoint (farthest first) amr_kdtools.pyx:612-622,708-755
  Tree.build (levels coarse -> fine, blocks order)    amr_kdtree.py:63-118

Pinned in tests/test_host.py by what the tree must satisfy: the leaves tile
the domain (kd_sum_volume, amr_kdtools.pyx:534-543, equals the domain
volume); each leaf's grid covers it and is the finest grid that does; the
viewpoint order is painter's order (along any line of sight toward the
viewpoint, the leaves it crosses come in increasing t).
"""

import numpy as np


class ONode:
    def __init__(self, parent, left_edge, right_edge, grid, node_id):
        self.parent = parent
        self.left = None
        self.right = None
        self.left_edge = np.array(left_edge, dtype="float64")
        self.right_edge = np.array(right_edge, dtype="float64")
        self.grid = int(grid)
        self.node_id = int(node_id)
        self.split_dim = -1
        self.split_pos = np.nan

    def is_leaf(self):
        return self.left is None or self.right is None

    # amr_kdtools.pyx:193-262
    def add_grids(self, gles, gres, gids):
        if self.is_leaf():
            self.insert_grids(gles, gres, gids)
            return
        d, p = self.split_dim, self.split_pos
        less = gles[:, d] < p
        greater = gres[:, d] > p
        if less.any():
            self.left.add_grids(gles[less], gres[less], gids[less])
        if greater.any():
            self.right.add_grids(gles[greater], gres[greater], gids[greater])

    # amr_kdtools.pyx:275-313
    def insert_grids(self, gles, gres, gids):
        n = len(gids)
        if n == 0:
            return
        if n == 1:
            if np.all(gles[0] <= self.left_edge) and np.all(gres[0] >= self.right_edge):
                self.grid = int(gids[0])
                return
        if self.split_grids(gles, gres, gids) == -1:
            self.grid = -1

    # amr_kdtools.pyx:369-458
    def split_grids(self, gles, gres, gids):
        if len(gids) == 0:
            return 0
        best_dim, split, less_ids, greater_ids = kdtree_get_choices(
            gles, gres, self.left_edge, self.right_edge)
        if best_dim == -1:
            return -1
        self.divide(best_dim, split)
        if less_ids.any():
            self.left.insert_grids(gles[less_ids], gres[less_ids], gids[less_ids])
        if greater_ids.any():
            self.right.insert_grids(gles[greater_ids], gres[greater_ids], gids[greater_ids])
        return 0

    # amr_kdtools.pyx:509-532
    def divide(self, dim, pos):
        self.split_dim = int(dim)
        self.split_pos = float(pos)
        le = self.left_edge.copy()
        re = self.right_edge.copy()
        re[dim] = pos
        self.left = ONode(self, le, re, self.grid, self.node_id << 1)
        re[dim] = self.right_edge[dim]
        le[dim] = pos
        self.right = ONode(self, le, re, self.grid, (self.node_id << 1) + 1)

    def depth_traverse(self):
        yield self
        if not self.is_leaf():
            yield from self.left.depth_traverse()
            yield from self.right.depth_traverse()

    def viewpoint_traverse(self, viewpoint):
        yield self
        if not self.is_leaf():
            if viewpoint[self.split_dim] <= self.split_pos:
                first, second = self.right, self.left
            else:
                first, second = self.left, self.right
            yield from first.viewpoint_traverse(viewpoint)
            yield from second.viewpoint_traverse(viewpoint)

    def kd_traverse(self, viewpoint=None):
        it = self.depth_traverse() if viewpoint is None else self.viewpoint_traverse(viewpoint)
        for node in it:
            if node.is_leaf() and node.grid != -1:
                yield node


def kdtree_get_choices(gles, gres, l_corner, r_corner):
    """amr_kdtools.pyx:760-823: per dimension, the unique grid faces strictly
    inside the node; the dimension with the most wins (first on ties); the
    split is the ((n-1)//2)-th smallest."""
    best_dim, my_max, my_split = -1, 0, 0
    uniq_best = None
    for dim in range(3):
        v = np.concatenate([gles[:, dim], gres[:, dim]])
        inside = (l_corner[dim] < v) & (v < r_corner[dim])
        u = np.unique(v[inside])
        if len(u) > my_max:
            best_dim, my_max, my_split = dim, len(u), (len(u) - 1) // 2
            uniq_best = u
    if best_dim == -1:
        return -1, 0.0, None, None
    split = np.sort(uniq_best)[my_split]
    less_ids = gles[:, best_dim] < split
    greater_ids = gres[:, best_dim] > split
    return best_dim, split, less_ids, greater_ids


def build(grids):
    """grids: (left_edge, right_edge, id, level) in blocks order.  Returns
    the root ONode (amr_kdtree.py:63-118)."""
    trunk = ONode(None, [-np.inf] * 3, [np.inf] * 3, -1, 1)
    for lvl in sorted({g[3] for g in grids}):
        sel = [g for g in grids if g[3] == lvl]
        gles = np.array([g[0] for g in sel], dtype="float64")
        gres = np.array([g[1] for g in sel], dtype="float64")
        gids = np.array([g[2] for g in sel], dtype="int64")
        trunk.add_grids(gles, gres, gids)
    return trunk


def order(trunk, viewpoint):
    """Indices into the depth-first list of leaves with a grid, farthest
    from the viewpoint first."""
    index = {id(n): i for i, n in enumerate(trunk.kd_traverse())}
    return np.array([index[id(n)] for n in trunk.kd_traverse(viewpoint=viewpoint)], dtype=np.int32)
