"""off_axis_projection for grid datasets (the ProjectionSampler caller).

Restates yt/visualization/volume_rendering/off_axis_projection.py:352-547
for grid indices: camera set-up (:385-411), a ProjectionSampler over the
child-masked blocks in index order (:505-524), alpha := 1 (:526), and the
final normalisation (:537-547).  Returns image[:, :, 0] as float64
(res0, res1).  Lengths are in code units.
"""

import numpy as np

from .camera import off_axis_camera
from .image_samplers import ProjectionSampler
from .partitioned_grid import PartitionedGrid
from .render import finalize_projection


def off_axis_projection(data_source, center, normal_vector, width, resolution, item, weight=None,
                        volume=None, no_ghost=False, interpolated=False, north_vector=None,
                        depth=None, num_threads=1, method="integrate"):
    if method not in ("integrate", "sum"):
        raise NotImplementedError("Only 'integrate' or 'sum' methods are valid for off-axis-projections")
    if interpolated:
        raise NotImplementedError("interpolated projections are not part of this ray-caster")
    ds = getattr(data_source, "ds", data_source)
    cam = off_axis_camera(center, normal_vector, width, resolution, north_vector,
                          ds.domain_left_edge, ds.domain_right_edge)
    params = cam._get_sampler_params()
    image = params["image"]
    zbuffer = np.ones(image.shape[:2], "float64")
    sampler = ProjectionSampler(
        np.atleast_3d(params["vp_pos"]), np.atleast_3d(params["vp_dir"]), params["center"],
        params["bounds"], image, params["x_vec"], params["y_vec"], params["width"], "KDTree", 10,
        lens_type=params["lens_type"], zbuffer=zbuffer)
    for grid, mask in data_source.blocks():
        # off_axis_projection.py:510-514: the fields times the child mask.
        # The bricks are stored as float32, so the products are formed in the
        # fields' own precision (a float32 product rounds once, as the float64
        # product rounded to float32 does) and an all-ones mask is skipped.
        full = bool(mask.all())
        a = np.asarray(grid[item])
        if weight is None:
            data = [a if full else a * mask]
        else:
            w = np.asarray(grid[weight])
            dt = np.result_type(a.dtype, w.dtype, np.float32)
            aw = a.astype(dt, copy=False) * w.astype(dt, copy=False)
            w = w.astype(dt, copy=False)
            data = [aw, w] if full else [aw * mask, w * mask]
        m8 = mask.view(np.uint8) if mask.dtype == bool else mask.astype("uint8")
        pg = PartitionedGrid(grid.id, data, m8, grid.LeftEdge, grid.RightEdge,
                             grid.ActiveDimensions.astype("int64"))
        sampler(pg, num_threads=num_threads)
    return finalize_projection(sampler.aimage, cam.width, weight is not None, method)


def compute_stddev_image(buff2, buff):
    """math_utils.py:1546-1583: sigma = sqrt(<v^2> - <v>^2) from the
    weighted projections of v^2 and v; negative variances are set to 0 when
    every one of them is round-off (np.isclose of the two terms), otherwise
    ValueError."""
    buff1 = buff * buff
    y = buff2 - buff1
    neg = y < 0.0
    if np.isclose(np.asarray(buff2), np.asarray(buff1))[neg].all():
        y[neg] = 0.0
    else:
        raise ValueError("Something went wrong, there are significant negative "
                         "variances in the standard deviation image!")
    return np.sqrt(y)


def off_axis_projection_buffer(data_source, center, normal_vector, width, buff_size, item,
                               weight=None, volume=None, no_ghost=False, interpolated=False,
                               north_vector=None, depth=None, method="integrate", moment=1):
    """OffAxisProjectionFixedResolutionBuffer._generate_image_and_mask
    (fixed_resolution.py:622-692): the off-axis projection at the buffer's
    size; for moment 2 the weighted standard deviation, from a second
    projection of item**2 with the same weight (compute_stddev_image).
    Returned as the buffer stores it, transposed (swapaxes(0, 1)): shape
    (buff_size[1], buff_size[0]), rows along the north vector."""
    if moment not in (1, 2):
        raise ValueError("moment must be 1 or 2")
    kw = dict(weight=weight, volume=volume, no_ghost=no_ghost, interpolated=interpolated,
              north_vector=north_vector, depth=depth, method=method)
    buff = off_axis_projection(data_source, center, normal_vector, width, buff_size, item, **kw)
    if moment == 2:
        ds = getattr(data_source, "ds", data_source)
        item_sq = f"tmp_{item}_squared"
        grids = list(ds.grids)
        for g in grids:
            g.fields[item_sq] = np.asarray(g[item], dtype="float64") ** 2
        try:
            buff2 = off_axis_projection(data_source, center, normal_vector, width, buff_size,
                                        item_sq, **kw)
        finally:
            for g in grids:
                g.fields.pop(item_sq, None)
        buff = compute_stddev_image(buff2, buff)
    return buff.swapaxes(0, 1)
