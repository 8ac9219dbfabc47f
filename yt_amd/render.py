"""Device-resident volumes and batched requests.

A ResidentVolume uploads a dataset's bricks to HBM once: the child-masked
cell-centred blocks the projection path walks (off_axis_projection.py:506-522)
and/or the vertex-centred log10 kd-tree bricks of the volume-render path
(amr_kdtree.py:325-374).  Requests then only name a camera; the brick order
is the reference's call order: blocks order for projections, the kd-tree's
farthest-from-viewpoint order for volume renders (render_source.py:594-597).
"""

import numpy as np

from . import _lib
from .amr_kdtree import AMRKDTree
from .camera import off_axis_camera


class ResidentVolume:
    """vr_field: a field name or a list of them (a VolumeSource's
    [field, weight_field], render_source.py:121-134); log_field: one flag for
    all of them or one per field.

    vr_prep: "device" computes the kd bricks' vertex data on the GPU from
    each grid's cells (ytrc_dataset_build_vr); "host" computes it with numpy
    and uploads the bricks (ytrc_dataset_upload).  Both store the same
    bricks."""

    def __init__(self, ctx, ds, proj_fields=None, weight=None, vr_field=None, log_field=True,
                 vr_prep="device"):
        self.ctx = ctx
        self.ds = ds
        self.proj_ds_id = None
        self.vr_ds_id = None
        self._cell_indexed_ds_id = None
        self.weighted = weight is not None
        self.n_proj_bricks = 0
        self.n_vr_bricks = 0
        self.proj_cells = 0
        if proj_fields is not None:
            if isinstance(proj_fields, str):
                proj_fields = [proj_fields]
            bricks, fields = [], []
            for grid, mask in ds.blocks():
                # The child-masked fields (off_axis_projection.py:510-514),
                # multiplied in the fields' own precision: float32 products
                # round once, as the float64 product rounded to float32 would
                # (the device stores float32); an all-ones mask is skipped.
                arrs = {f: np.asarray(grid[f]) for f in list(proj_fields) + ([weight] if weight else [])}
                dt = np.result_type(*[a.dtype for a in arrs.values()], np.float32)
                full = bool(np.all(mask))
                m = None if full else mask.astype(dt)

                def masked(a):
                    a = a.astype(dt, copy=False)
                    return a if m is None else a * m

                if weight is None:
                    data = [masked(arrs[f]) for f in proj_fields]
                else:
                    f0 = proj_fields[0]
                    w = arrs[weight].astype(dt, copy=False)
                    data = [masked(arrs[f0].astype(dt, copy=False) * w), masked(w)]
                bricks.append((grid.LeftEdge, grid.RightEdge, grid.ActiveDimensions, False))
                fields.append(data)
                self.proj_cells += int(np.prod(grid.ActiveDimensions))
            self.proj_ds_id = ctx.upload(bricks, fields)
            self.n_proj_bricks = len(bricks)
        if vr_field is not None:
            vr_fields = [vr_field] if isinstance(vr_field, str) else list(vr_field)
            if np.ndim(log_field) == 0:
                log_fields = [bool(log_field)] * len(vr_fields)
            else:
                log_fields = [bool(x) for x in log_field]
            if len(log_fields) != len(vr_fields):
                raise ValueError("one log_field flag per field")
            self.vr_fields = vr_fields
            self.kd = AMRKDTree(ds)
            self.kd.set_fields(vr_fields, log_fields, no_ghost=True)
            self.leaves = self.kd.leaves()
            self._leaf_index = {id(n): i for i, n in enumerate(self.leaves)}
            if vr_prep == "device":
                gindex, grids, cells, leaves, masks = {}, [], [], [], []
                for node in self.leaves:
                    grid, li, ri, nle, nre, mask = self.kd.leaf_slice(node)
                    if grid.id not in gindex:
                        gindex[grid.id] = len(grids)
                        grids.append((grid.LeftEdge, grid.RightEdge, grid.ActiveDimensions))
                        cells.append([np.asarray(grid[f]) for f in vr_fields])
                    leaves.append((gindex[grid.id], li, ri, nle, nre))
                    masks.append(None if mask.all() else mask)
                self.vr_ds_id = ctx.build_vr(grids, cells, log_fields, leaves, masks)
                self.n_vr_bricks = len(leaves)
            elif vr_prep == "host":
                bricks, fields, masks = [], [], []
                for node in self.leaves:
                    pg = self.kd.get_brick_data(node)
                    bricks.append((pg.LeftEdge, pg.RightEdge, pg.dims, True))
                    fields.append(pg.my_data)
                    masks.append(pg.source_mask)
                    node.data = None  # keep only the device copy
                self.vr_ds_id = ctx.upload(bricks, fields, masks)
                self.n_vr_bricks = len(bricks)
                self.kd._vcd_cache = {}
            else:
                raise ValueError("vr_prep must be 'device' or 'host'")

    def device_bytes(self):
        b = 0
        for i in (self.proj_ds_id, self.vr_ds_id, self._cell_indexed_ds_id):
            if i is not None:
                b += self.ctx.dataset_bytes(i)
        return b

    def vr_order(self, viewpoint):
        """The kd bricks farthest from the viewpoint first (the native
        tree's order; indices into self.leaves)."""
        return self.kd.order(viewpoint)

    def projection_request(self, params, record=False, zbuffer=None):
        if self.proj_ds_id is None:
            raise ValueError("volume was not uploaded for projections")
        return _lib.RequestSpec(_lib.PROJECTION, params, self.proj_ds_id,
                                np.arange(self.n_proj_bricks, dtype=np.int32), record=record,
                                zbuffer=zbuffer)

    def vr_request(self, params, tf, n_samples=10, record=False, ert=False, zbuffer=None):
        if self.vr_ds_id is None:
            raise ValueError("volume was not uploaded for volume rendering")
        tfc = tf if isinstance(tf, _lib.TF) else _lib.TF.from_tf_obj(tf)
        return _lib.RequestSpec(_lib.VOLUME_RENDER, params, self.vr_ds_id,
                                self.vr_order(params["viewpoint"]), n_samples=n_samples, tf=tfc,
                                record=record, ert=ert, zbuffer=zbuffer)

    def vr_partial_requests(self, params, tf, depth, n_samples=10, zbuffer=None, ert=False):
        """One partial-map request (YTRC_FLAG_PARTIAL) per kd subtree
        `depth` levels below the root, in the viewpoint's order: the share
        of each of 2^depth GPUs when kd subtrees are split over GPUs.
        compose_partials of their maps, in this order, is the whole
        volume render.  Parts without leaves that carry a grid are dropped."""
        if self.vr_ds_id is None:
            raise ValueError("volume was not uploaded for volume rendering")
        tfc = tf if isinstance(tf, _lib.TF) else _lib.TF.from_tf_obj(tf)
        vp = params["viewpoint"]
        specs = []
        for part in self.kd.subtrees(depth, viewpoint=vp):
            order = np.array([self._leaf_index[id(n)] for n in part.kd_traverse(viewpoint=vp)],
                             dtype=np.int32)
            if len(order) == 0:
                continue
            specs.append(_lib.RequestSpec(_lib.VOLUME_RENDER, params, self.vr_ds_id, order,
                                          n_samples=n_samples, tf=tfc, zbuffer=zbuffer, ert=ert,
                                          partial=True))
        return specs

    def interpolated_projection_request(self, params, n_samples=10, record=False, zbuffer=None):
        """InterpolatedProjectionSampler over the kd bricks (the sampler of a
        "projection" VolumeSource, render_source.py:464-488), in the same
        brick order as a volume render."""
        if self.vr_ds_id is None:
            raise ValueError("volume was not uploaded with vertex-centred bricks")
        return _lib.RequestSpec(_lib.INTERPOLATED_PROJECTION, params, self.vr_ds_id,
                                self.vr_order(params["viewpoint"]), n_samples=n_samples,
                                record=record, zbuffer=zbuffer)

    def cell_indexed_projection_request(self, params, record=False, zbuffer=None):
        """ProjectionSampler over the kd bricks, the sampler that
        set_sampler(camera, interpolated=False) gives a "projection"
        VolumeSource (render_source.py:464-488, utils.py:115-139).  The
        bricks hold (d+1)^3 vertex values, but ProjectionSampler indexes them
        with the cell strides of d (image_samplers.pyx:341): it reads the
        first d0*d1*d2 vertex values in C order as if they were the d-shaped
        cell array.  That reading is reproduced exactly: on first use the
        host slices those values out of each brick's vertex data and uploads
        them as cell-centred bricks (no mask: ProjectionSampler has none)."""
        if self.vr_ds_id is None:
            raise ValueError("volume was not uploaded with vertex-centred bricks")
        if self._cell_indexed_ds_id is None:
            bricks, fields = [], []
            for node in self.leaves:
                pg = self.kd.get_brick_data(node)
                d = tuple(int(x) for x in pg.dims)
                n = d[0] * d[1] * d[2]
                bricks.append((pg.LeftEdge, pg.RightEdge, d, False))
                fields.append([np.ascontiguousarray(a).ravel()[:n].reshape(d) for a in pg.my_data])
                node.data = None
            self.kd._vcd_cache = {}
            self._cell_indexed_ds_id = self.ctx.upload(bricks, fields)
        return _lib.RequestSpec(_lib.PROJECTION, params, self._cell_indexed_ds_id,
                                self.vr_order(params["viewpoint"]), record=record,
                                zbuffer=zbuffer)

    def light_source_request(self, params, tf, light_dir=(1.0, 1.0, 1.0),
                             light_rgba=(1.0, 1.0, 1.0, 1.0), n_samples=10, record=False,
                             zbuffer=None):
        """LightSourceRenderSampler over the kd bricks (old_camera.py:688)."""
        if self.vr_ds_id is None:
            raise ValueError("volume was not uploaded with vertex-centred bricks")
        tfc = tf if isinstance(tf, _lib.TF) else _lib.TF.from_tf_obj(tf)
        return _lib.RequestSpec(_lib.LIGHT_SOURCE, params, self.vr_ds_id,
                                self.vr_order(params["viewpoint"]), n_samples=n_samples, tf=tfc,
                                record=record, zbuffer=zbuffer, light_dir=light_dir,
                                light_rgba=light_rgba)

    def free(self):
        for i in (self.proj_ds_id, self.vr_ds_id, self._cell_indexed_ds_id):
            if i is not None:
                self.ctx.free(i)
        self.proj_ds_id = self.vr_ds_id = self._cell_indexed_ds_id = None


def camera_params(center, normal, width, resolution, north=None):
    """Sampler parameters (+ kd viewpoint) of an off-axis plane-parallel camera."""
    cam = off_axis_camera(center, normal, width, resolution, north)
    p = cam._get_sampler_params()
    p["viewpoint"] = cam.lens.viewpoint.copy()
    return p


def lens_params(lens_type, position, focus, width, resolution, north=None,
                domain_left_edge=(0.0, 0.0, 0.0), domain_right_edge=(1.0, 1.0, 1.0)):
    """Sampler parameters (+ the lens's kd viewpoint) of a camera at
    `position` looking at `focus` through any of yt's lenses (lens.py,
    camera.py:275-281)."""
    from .camera import Camera

    cam = Camera(lens_type=lens_type, domain_left_edge=domain_left_edge,
                 domain_right_edge=domain_right_edge)
    if np.ndim(resolution) == 0:
        resolution = (resolution, resolution)
    cam.resolution = resolution
    cam.width = width
    if np.array_equal(np.asarray(position, dtype="float64"), cam.focus):
        cam.focus = focus
        cam.position = position
    else:
        cam.position = position
        cam.focus = focus
    if north is not None:
        cam.switch_orientation(north_vector=north)
    p = cam._get_sampler_params()
    p["viewpoint"] = np.array(cam.lens.viewpoint, dtype="float64").copy()
    return p


def perspective_params(position, focus, width, resolution, north=None,
                       domain_left_edge=(0.0, 0.0, 0.0), domain_right_edge=(1.0, 1.0, 1.0)):
    """Sampler parameters (+ kd viewpoint, the front centre) of a perspective
    camera at `position` looking at `focus` (lens.py:145-240, camera.py:275-281)."""
    return lens_params("perspective", position, focus, width, resolution, north,
                       domain_left_edge, domain_right_edge)


def finalize_projection(image, width, weighted, method="integrate"):
    """off_axis_projection.py:526-547 + VolumeSource.finalize_image
    (render_source.py:522-537): alpha := 1, then scale or divide."""
    image = image.copy()
    image[:, :, 3] = 1
    if method == "integrate":
        if not weighted:
            image *= float(np.asarray(width, dtype="float64").ravel()[-1] if np.ndim(width) else width)
        else:
            mask = image[:, :, 1] == 0
            nmask = np.logical_not(mask)
            image[:, :, 0][nmask] /= image[:, :, 1][nmask]
            image[mask] = 0
    return image[:, :, 0]


def compose_partials(maps, image_in=None):
    """The image of a brick order from the partial maps (C, T) of its
    consecutive parts, given in call order (farthest part first): each part
    composites over the image so far, rgba <- C + T rgba, from image_in (or
    zeros).  maps: (nv0, nv1, 8) arrays, C[0..3] then T[0..3]."""
    maps = list(maps)
    if not maps:
        raise ValueError("compose_partials: no maps")
    shape = maps[0].shape[:2] + (4,)
    img = np.zeros(shape) if image_in is None else np.array(image_in, dtype=np.float64)
    for m in maps:
        img = m[..., :4] + m[..., 4:] * img
    return img


def receive_and_reduce(image, arr2, add_to_front, use_opacity=True):
    """amr_kdtools.py:6-39 without the transport: merges the image received
    from the other subtree, `arr2`, into `image`.  Without opacity (every
    transfer function but grey-opacity ones, render_source.py:604-611) the
    two are added; otherwise the front one is composited over the back one
    with the front alpha, front + max(1 - front_alpha, 0) * back."""
    if not use_opacity:
        return image + arr2
    front, back = (arr2, image) if add_to_front else (image, arr2)
    if image.shape[2] == 3:
        return image + front
    ta = np.maximum(1.0 - front[:, :, 3], 0.0)
    return back * ta[:, :, None] + front


def reduce_tree_images(kd, part_images, viewpoint, depth, use_opacity=True):
    """AMRKDTree.reduce_tree_images (amr_kdtree.py:283-323) on one host: the
    images rendered, each from zeros, over the subtrees of
    `kd.subtrees(depth)` (in that function's index order, i.e. without a
    viewpoint) are merged pairwise up to kd.data_root(), a right subtree's
    image going to the front when viewpoint[split_dim] >= split_pos.  The
    empty subtrees above data_root() would merge zeros, which changes
    nothing."""
    parts = kd.subtrees(depth)
    if len(part_images) != len(parts):
        raise ValueError("one image per subtree")
    img = {id(n): np.asarray(im, dtype=np.float64) for n, im in zip(parts, part_images)}

    def rec(n):
        if id(n) in img:
            return img[id(n)]
        left, right = rec(n.left), rec(n.right)
        add_to_front = viewpoint[n.split_dim] >= n.split_pos
        return receive_and_reduce(left, right, add_to_front, use_opacity)

    return rec(kd.data_root())


def finalize_vr(image, grey_opacity=False):
    """render_source.py:522-537 (alpha := 1 unless grey) and the 180-degree
    rotation of Scene.composite (scene.py:625)."""
    image = image.copy()
    if not grey_opacity:
        image[:, :, 3] = 1
    return np.rot90(image, k=2)
