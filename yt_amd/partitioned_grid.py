"""PartitionedGrid: one brick handed to an ImageSampler.

Restates yt/utilities/lib/partitioned_grid.pyx:23-70 (and the
VolumeContainer of volume_container.pxd:13-33): edges, dims, dds, idds,
field arrays and a uint8 mask.  The arrays are held, not copied: the sampler
uploads them to the GPU when it runs.
"""

import numpy as np


class PartitionedGrid:
    def __init__(self, parent_grid_id, data, mask, left_edge, right_edge, dims, n_fields=-1):
        self.parent_grid_id = int(parent_grid_id)
        self.LeftEdge = np.asarray(left_edge, dtype="float64")
        self.RightEdge = np.asarray(right_edge, dtype="float64")
        dims = np.asarray(dims, dtype="int64")
        if dims.shape != (3,):
            raise ValueError("dims must have 3 entries")
        self.dims = dims
        if n_fields == -1:
            n_fields = len(data)
        self.n_fields = int(n_fields)
        self.my_data = [np.asarray(d) for d in data]
        self.source_mask = None if mask is None else np.asarray(mask, dtype=np.uint8)
        # partitioned_grid.pyx:50-55
        self.dds = (self.RightEdge - self.LeftEdge) / dims
        self.idds = 1.0 / self.dds
        shapes = {tuple(d.shape) for d in self.my_data}
        if len(shapes) > 1:
            raise ValueError("all fields of a brick must share one shape")
        self.vertex_centred = bool(self.my_data) and tuple(self.my_data[0].shape) == tuple(dims + 1)
        if self.my_data and not self.vertex_centred and tuple(self.my_data[0].shape) != tuple(dims):
            raise ValueError(f"field shape {self.my_data[0].shape} matches neither dims {tuple(dims)} "
                             "nor dims+1")

    @property
    def container(self):
        """The VolumeContainer fields, as a dict (for tests and the oracle)."""
        return {
            "le": self.LeftEdge, "re": self.RightEdge, "dims": self.dims, "dds": self.dds,
            "idds": self.idds, "data": self.my_data, "mask": self.source_mask,
        }
