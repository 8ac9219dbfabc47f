"""AMRKDTree: the brick set and brick order of VolumeRenderSampler.

The tree is built, and ordered for each viewpoint, by the library's native
kd-tree (ytrc_kdtree_*, restating Tree.build, amr_kdtree.py:63-118, over
Node.add_grids / insert_grids / split_grids / divide and kdtree_get_choices,
amr_kdtools.pyx:193-532, 760-823, and step_viewpoint, amr_kdtools.pyx:
708-755).  This module mirrors its nodes as Node objects for the Python
side: traversal, subtrees, and the bricks of AMRKDTree.set_fields /
traverse / get_brick_data (amr_kdtree.py:161-374), single process.
"""

import numpy as np

from . import _lib
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


class AMRKDTree:
    """amr_kdtree.py:161-374 (single process).  The grids of
    data_source.blocks() between min_level and max_level go to the native
    build level by level, in blocks order; the tree's nodes are mirrored as
    Node objects (self.trunk), and order(viewpoint) asks the native tree."""

    def __init__(self, ds, min_level=None, max_level=None, data_source=None):
        self.ds = ds
        self.data_source = ds if data_source is None else data_source
        self.fields = None
        self.log_fields = None
        self.no_ghost = True
        self._vcd_cache = {}
        min_level = 0 if min_level is None else min_level
        max_level = ds.max_level if max_level is None else max_level
        grids = [g for g, _ in self.data_source.blocks() if min_level <= g.Level <= max_level]
        self._native = _lib.KdTree(
            np.array([g.LeftEdge for g in grids], dtype="float64").reshape(-1, 3),
            np.array([g.RightEdge for g in grids], dtype="float64").reshape(-1, 3),
            np.array([g.id for g in grids], dtype="int64"),
            np.array([g.Level for g in grids], dtype="int32"))
        nodes = self._native.nodes()
        mirror = [Node(None, None, None, n.left_edge, n.right_edge, n.grid, n.node_id)
                  for n in nodes]
        for m, n in zip(mirror, nodes):
            m.split_dim = int(n.split_dim)
            m.split_pos = float(n.split_pos)
            m.parent = mirror[n.parent] if n.parent >= 0 else None
            m.left = mirror[n.left] if n.left >= 0 else None
            m.right = mirror[n.right] if n.right >= 0 else None
        self.trunk = mirror[0]
        self._grid_by_id = {g.id: g for g in ds.grids}

    def order(self, viewpoint):
        """The bricks farthest from the viewpoint first, as indices into
        leaves() (kd_traverse(viewpoint), amr_kdtools.pyx:612-622)."""
        return self._native.order(viewpoint)

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
