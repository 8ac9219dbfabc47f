"""AMRKDTree: the brick set and brick order of VolumeRenderSampler.

Restates:
  Node, add_grids, insert_grids, split_grids, divide  amr_kdtools.pyx:21-532
  kdtree_get_choices                                  amr_kdtools.pyx:760-823
  viewpoint_traverse / step_viewpoint (farthest first) amr_kdtools.pyx:612-622,708-755
  Tree.build (levels coarse -> fine, blocks order)    amr_kdtree.py:63-118
  AMRKDTree.set_fields / traverse / get_brick_data    amr_kdtree.py:161-374
for one process (rank 0 of 1: should_i_build is always true and
should_i_split never fires).
"""

import numpy as np

from .partitioned_grid import PartitionedGrid


class Node:
    __slots__ = ("parent", "left", "right", "left_edge", "right_edge", "grid", "node_id",
                 "split_dim", "split_pos", "data", "dirty")

    def __init__(self, parent, left, right, left_edge, right_edge, grid, node_id):
        self.parent = parent
        self.left = left
        self.right = right
        self.left_edge = np.array(left_edge, dtype="float64")
        self.right_edge = np.array(right_edge, dtype="float64")
        self.grid = int(grid)
        self.node_id = int(node_id)
        self.split_dim = -1
        self.split_pos = np.nan
        self.data = None
        self.dirty = False

    def is_leaf(self):
        return self.left is None or self.right is None

    def get_left_edge(self):
        return self.left_edge.copy()

    def get_right_edge(self):
        return self.right_edge.copy()

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
        self.left = Node(self, None, None, le, re, self.grid, self.node_id << 1)
        re[dim] = self.right_edge[dim]
        le[dim] = pos
        self.right = Node(self, None, None, le, re, self.grid, (self.node_id << 1) + 1)

    def depth_traverse(self):
        yield self
        if not self.is_leaf():
            yield from self.left.depth_traverse()
            yield from self.right.depth_traverse()

    def viewpoint_traverse(self, viewpoint):
        """Farthest-from-viewpoint first (step_viewpoint, amr_kdtools.pyx:708-755):
        at a split, the right child first when viewpoint[dim] <= pos."""
        yield self
        if not self.is_leaf():
            if viewpoint[self.split_dim] <= self.split_pos:
                first, second = self.right, self.left
            else:
                first, second = self.left, self.right
            yield from first.viewpoint_traverse(viewpoint)
            yield from second.viewpoint_traverse(viewpoint)

    def kd_traverse(self, viewpoint=None):
        """Leaves that carry a grid (amr_kdtools.pyx:102-111)."""
        it = self.depth_traverse() if viewpoint is None else self.viewpoint_traverse(viewpoint)
        for node in it:
            if node.is_leaf() and node.grid != -1:
                yield node

    def kd_sum_volume(self):
        if self.is_leaf():
            if self.grid == -1:
                return 0.0
            return float(np.prod(self.right_edge - self.left_edge))
        return self.left.kd_sum_volume() + self.right.kd_sum_volume()


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


class AMRKDTree:
    """amr_kdtree.py:161-374 (single process)."""

    def __init__(self, ds, min_level=None, max_level=None, data_source=None):
        self.ds = ds
        self.data_source = ds if data_source is None else data_source
        self.fields = None
        self.log_fields = None
        self.no_ghost = True
        self._vcd_cache = {}
        self.trunk = Node(None, None, None, [-np.inf] * 3, [np.inf] * 3, -1, 1)
        min_level = 0 if min_level is None else min_level
        max_level = ds.max_level if max_level is None else max_level
        blocks = list(self.data_source.blocks())
        for lvl in range(min_level, max_level + 1):
            grids = [b for b, _ in blocks if b.Level == lvl]
            if not grids:
                continue
            gles = np.array([g.LeftEdge for g in grids])
            gres = np.array([g.RightEdge for g in grids])
            gids = np.array([g.id for g in grids], dtype="int64")
            self.trunk.add_grids(gles, gres, gids)
        self._grid_by_id = {g.id: g for g in ds.grids}

    def set_fields(self, fields, log_fields, no_ghost=True, force=False):
        if not np.iterable(log_fields) or isinstance(log_fields, (str, bytes)):
            log_fields = [log_fields]
        self.fields = list(fields)
        self.log_fields = list(log_fields)
        self.no_ghost = no_ghost
        self._vcd_cache = {}
        for node in self.trunk.depth_traverse():
            node.data = None

    def leaves(self, viewpoint=None):
        return list(self.trunk.kd_traverse(viewpoint=viewpoint))

    def data_root(self):
        """The smallest subtree holding every leaf that carries a grid: the
        trunk spans all space, so its first splits, at the domain faces,
        leave one side without grids."""
        node = self.trunk
        while not node.is_leaf():
            has_l = next(node.left.kd_traverse(), None) is not None
            has_r = next(node.right.kd_traverse(), None) is not None
            if has_l and has_r:
                break
            node = node.left if has_l else node.right
        return node

    def subtrees(self, depth, viewpoint=None):
        """The kd subtrees `depth` levels below data_root() (a branch that
        ends higher up is a part of its own): the shares of 2^depth GPUs
        when kd subtrees are split over GPUs, as yt's parallel volume
        render gives one subtree to each rank (nodes nprocs .. 2 nprocs - 1,
        amr_kdtree.py:283-323; yt builds that top of the tree by geometric
        halving, amr_kdtools.pyx:460-500, this splits the serial tree).
        With a viewpoint, in the order its traversal reaches them, so that
        their leaves, each in viewpoint order, concatenate to the whole
        tree's viewpoint order."""
        parts = []

        def rec(n, d):
            if d == depth or n.is_leaf():
                parts.append(n)
                return
            rec(n.left, d + 1)
            rec(n.right, d + 1)

        root = self.data_root()
        rec(root, 0)
        if viewpoint is not None:
            rank = {id(n): i for i, n in enumerate(root.viewpoint_traverse(viewpoint))}
            parts.sort(key=lambda n: rank[id(n)])
        return parts

    def traverse(self, viewpoint=None):
        for node in self.trunk.kd_traverse(viewpoint=viewpoint):
            yield self.get_brick_data(node)

    def _vertex_data(self, grid):
        key = grid.id
        if key not in self._vcd_cache:
            vcd = grid.get_vertex_centered_data(self.fields, no_ghost=self.no_ghost)
            dds = []
            for i, f in enumerate(self.fields):
                if self.log_fields[i]:
                    v = vcd[f].astype("float64")
                    v[v <= 0] = np.nan
                    with np.errstate(invalid="ignore"):
                        dds.append(np.log10(v))
                else:
                    dds.append(vcd[f].astype("float64"))
            self._vcd_cache[key] = dds
        return self._vcd_cache[key]

    def leaf_slice(self, node):
        """The grid, cell slice [li, ri), edges and child-mask slice of a
        leaf, as get_brick_data (amr_kdtree.py:325-374) derives them, without
        the vertex data: the inputs of device brick preparation."""
        grid = self._grid_by_id[node.grid]
        gle = grid.LeftEdge
        nle = node.get_left_edge()
        nre = node.get_right_edge()
        li = np.rint((nle - gle) / grid.dds).astype("int32")
        ri = np.rint((nre - gle) / grid.dds).astype("int32")
        assert np.all(grid.LeftEdge <= nle)
        assert np.all(grid.RightEdge >= nre)
        mask = grid.child_mask[li[0]:ri[0], li[1]:ri[1], li[2]:ri[2]].astype("uint8")
        return grid, li, ri, nle.copy(), nre.copy(), mask

    def get_brick_data(self, node):
        """amr_kdtree.py:325-374: the leaf clipped to its grid, vertex data
        sliced [li:ri+1], mask = child mask slice."""
        if node.data is not None and not node.dirty:
            return node.data
        grid = self._grid_by_id[node.grid]
        dds = grid.dds
        gle = grid.LeftEdge
        nle = node.get_left_edge()
        nre = node.get_right_edge()
        li = np.rint((nle - gle) / dds).astype("int32")
        ri = np.rint((nre - gle) / dds).astype("int32")
        dims = ri - li
        assert np.all(grid.LeftEdge <= nle)
        assert np.all(grid.RightEdge >= nre)
        vd = self._vertex_data(grid)
        mask = grid.child_mask[li[0]:ri[0], li[1]:ri[1], li[2]:ri[2]].astype("uint8")
        data = [d[li[0]:ri[0] + 1, li[1]:ri[1] + 1, li[2]:ri[2] + 1].copy() for d in vd]
        brick = PartitionedGrid(grid.id, data, mask, nle.copy(), nre.copy(), dims.astype("int64"))
        node.data = brick
        node.dirty = False
        return brick
