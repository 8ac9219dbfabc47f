"""yt_amd: an MI355X-native implementation of yt's brick ray-caster.

The hot path -- walk_volume's DDA feeding ProjectionSampler and
VolumeRenderSampler -- runs as hand-written HIP kernels for gfx950 in
libytrc.so (C ABI: include/ytrc.h, bound with ctypes in yt_amd._lib).  The
Python modules mirror the reference's host-side interface for that path:
off_axis_projection, the ImageSampler classes, PartitionedGrid, AMRKDTree,
the camera/lens parameters and the transfer functions.
"""

from . import _lib
from .amr_kdtree import AMRKDTree
from .camera import Camera, Orientation, PerspectiveLens, PlaneParallelLens, off_axis_camera
from .data import load_amr_grids, load_uniform_grid
from .image_samplers import (ImageSampler, InterpolatedProjectionSampler,
                             LightSourceRenderSampler, ProjectionSampler, VolumeRenderSampler)
from .off_axis_projection import (
    compute_stddev_image,
    off_axis_projection,
    off_axis_projection_buffer,
)
from .partitioned_grid import PartitionedGrid
from .scene import OpaqueSource, Scene, VolumeSource, ZBuffer, create_scene
from .transfer_functions import (
    ColorTransferFunction,
    MultiVariateTransferFunction,
    ProjectionTransferFunction,
    TransferFunction,
)

__all__ = [
    "AMRKDTree",
    "Camera",
    "ColorTransferFunction",
    "ImageSampler",
    "InterpolatedProjectionSampler",
    "LightSourceRenderSampler",
    "MultiVariateTransferFunction",
    "Orientation",
    "PartitionedGrid",
    "PerspectiveLens",
    "PlaneParallelLens",
    "ProjectionSampler",
    "ProjectionTransferFunction",
    "OpaqueSource",
    "Scene",
    "VolumeSource",
    "ZBuffer",
    "create_scene",
    "TransferFunction",
    "VolumeRenderSampler",
    "load_amr_grids",
    "load_uniform_grid",
    "off_axis_camera",
    "off_axis_projection",
    "off_axis_projection_buffer",
    "compute_stddev_image",
]
