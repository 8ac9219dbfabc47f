"""Drop-in ImageSampler classes backed by the HIP ray-caster.

Same constructors, call protocol and result attribute as
yt/utilities/lib/image_samplers.pyx:
  ImageSampler(vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width,
               volume_method, *args, lens_type=..., zbuffer=..., camera_data=...)
                                                      :59-130
  ProjectionSampler(... volume_method, num_samples)   :327-343
  VolumeRenderSampler(... volume_method, tf_obj, n_samples=10)  :405-494
  sampler(pg, num_threads=0) once per brick, in compositing order  :132-197
  sampler.aimage -> the (nv0, nv1, 4) float64 image, accumulated in place
Calls are recorded; reading `aimage` casts every recorded brick in one GPU
batch (in call order, which fixes the compositing order) and writes the
result into the caller's image array.  Every caller of the reference reads
aimage only after its brick loop (render_source.py:594-599,
off_axis_projection.py:506-526).
"""

import numpy as np

from . import _lib


class ImageSampler:
    kind = None

    def __init__(self, vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width, volume_method,
                 *args, **kwargs):
        self.volume_method = volume_method
        self.camera_data = kwargs.pop("camera_data", None)
        zbuffer = kwargs.pop("zbuffer", None)
        image = np.asarray(image)
        if image.ndim != 3 or image.shape[2] != 4 or image.dtype != np.float64:
            raise ValueError("image must be a float64 (nv0, nv1, 4) array")
        if zbuffer is None:
            zbuffer = np.ones((image.shape[0], image.shape[1]), "float64")
        self.lens_type = kwargs.pop("lens_type", None)
        vp_pos = np.asarray(vp_pos, dtype="float64")
        vp_dir = np.asarray(vp_dir, dtype="float64")
        if self.lens_type != "plane-parallel":
            # image_samplers.pyx:89-96: per-pixel lenses must match the image.
            if not (vp_pos.shape[0] == vp_dir.shape[0] == image.shape[0]) or not (
                    vp_pos.shape[1] == vp_dir.shape[1] == image.shape[1]):
                raise RuntimeError(
                    f"Bad lens shape / direction for {self.lens_type}\n"
                    f"Shapes: ({vp_pos.shape[0]} - {vp_dir.shape[0]} - {image.shape[0]}) and "
                    f"({vp_pos.shape[1]} - {vp_dir.shape[1]} - {image.shape[1]})")
        elif vp_pos.size != 12 or vp_dir.size != 3:
            raise ValueError("plane-parallel lens needs 12 vp_pos and 3 vp_dir values")
        self.vp_pos = vp_pos
        self.vp_dir = vp_dir
        self.acenter = np.asarray(center, dtype="float64")
        self.ax_vec = np.asarray(x_vec, dtype="float64")
        self.ay_vec = np.asarray(y_vec, dtype="float64")
        self._image = image
        self.azbuffer = np.asarray(zbuffer, dtype="float64")
        self.nv = (image.shape[0], image.shape[1])
        self.bounds = tuple(float(b) for b in bounds)
        self.pdx = (self.bounds[1] - self.bounds[0]) / self.nv[0]
        self.pdy = (self.bounds[3] - self.bounds[2]) / self.nv[1]
        self.width = np.asarray(width, dtype="float64")[:3].copy()
        self._pending = []
        self.last_stats = None

    def __call__(self, pg, **kwa):
        if self.volume_method == "KDTree":
            return self.cast_through_kdtree(pg, **kwa)
        elif self.volume_method == "Octree":
            raise NotImplementedError("the octree caster is not part of this ray-caster")
        else:
            raise NotImplementedError(
                f'Volume rendering has not been implemented for method: "{self.volume_method}"')

    def cast_through_kdtree(self, pg, num_threads=0):
        self._pending.append(pg)
        return 0

    @property
    def image(self):
        return self.aimage

    @property
    def aimage(self):
        self.flush()
        return self._image

    def _params(self):
        return {
            "vp_pos": self.vp_pos, "vp_dir": self.vp_dir, "center": self.acenter,
            "bounds": self.bounds, "x_vec": self.ax_vec, "y_vec": self.ay_vec,
            "width": self.width, "nv": self.nv,
            "lens_type": self.lens_type if self.lens_type is not None else "null",
            "camera_data": self.camera_data,
        }

    def _tf(self):
        return None

    def _light(self):
        return {}

    def flush(self, record=False):
        """Casts every recorded brick on the GPU, in call order."""
        if not self._pending:
            return None
        pgs, self._pending = self._pending, []
        ctx = _lib.default_context()
        vc = self.kind != _lib.PROJECTION
        bricks = [(pg.LeftEdge, pg.RightEdge, pg.dims, vc) for pg in pgs]
        fields = []
        for pg in pgs:
            if pg.vertex_centred == vc:
                fields.append(pg.my_data[:pg.n_fields])
            elif not vc:
                # ProjectionSampler handed a vertex-centred kd brick (a
                # "projection" VolumeSource after set_sampler(camera,
                # interpolated=False), render_source.py:464-488): it indexes
                # the (d+1)^3 array with the cell strides of d
                # (image_samplers.pyx:341), reading its first d0*d1*d2 values
                # in C order as the cell array.  Reproduced as is.
                d = tuple(int(x) for x in pg.dims)
                n = d[0] * d[1] * d[2]
                fields.append([np.ascontiguousarray(a).ravel()[:n].reshape(d)
                               for a in pg.my_data[:pg.n_fields]])
            else:
                raise ValueError("VolumeRenderSampler and InterpolatedProjectionSampler need "
                                 "vertex-centred bricks")
        masks = [pg.source_mask for pg in pgs] if self.kind == _lib.VOLUME_RENDER else None
        ds_id = ctx.upload(bricks, fields, masks)
        try:
            spec = _lib.RequestSpec(
                self.kind, self._params(), ds_id, np.arange(len(pgs), dtype=np.int32),
                n_samples=getattr(self, "n_samples", 10), tf=self._tf(), zbuffer=self.azbuffer,
                image_in=self._image if np.any(self._image) else None, record=record,
                **self._light())
            # persistent: a stop request (Ctrl-C, below) reaches its waves
            batch = _lib.Batch(ctx, [spec], _lib.BATCH_PERSISTENT)
            try:
                batch.launch()
                batch.wait_interruptibly()
                self._image[...] = batch.read_image(0)
                rec = batch.read_record(0) if record else None
            finally:
                batch.close()
        finally:
            ctx.free(ds_id)
        return rec


class ProjectionSampler(ImageSampler):
    """image_samplers.pyx:327-343: line integral of min(4, n_fields) fields."""

    kind = _lib.PROJECTION


class InterpolatedProjectionSampler(ImageSampler):
    """image_samplers.pyx:344-402: vertex-centred bricks, n_samples trilinear
    sub-samples per cell, each field times dt into rgba[j], j < min(3,
    n_fields); no mask test."""

    kind = _lib.INTERPOLATED_PROJECTION

    def __init__(self, vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width, volume_method,
                 n_samples=10, **kwargs):
        super().__init__(vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width,
                         volume_method, **kwargs)
        self.n_samples = int(n_samples)


class VolumeRenderSampler(ImageSampler):
    """image_samplers.pyx:405-494: transfer function + compositing."""

    kind = _lib.VOLUME_RENDER

    def __init__(self, vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width, volume_method,
                 tf_obj, n_samples=10, **kwargs):
        super().__init__(vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width,
                         volume_method, **kwargs)
        n_fits = tf_obj.n_field_tables
        assert n_fits <= 6
        self.tf_obj = tf_obj
        self.n_samples = int(n_samples)
        self.my_field_tables = [(tf_obj.tables[i], tf_obj.tables[i].y) for i in range(n_fits)]
        self._tfc = _lib.TF.from_tf_obj(tf_obj)

    def _tf(self):
        return self._tfc


class LightSourceRenderSampler(VolumeRenderSampler):
    """image_samplers.pyx:496-600: the volume render's tables with a
    directional light; no mask test."""

    kind = _lib.LIGHT_SOURCE

    def __init__(self, vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width, volume_method,
                 tf_obj, n_samples=10, light_dir=(1.0, 1.0, 1.0), light_rgba=(1.0, 1.0, 1.0, 1.0),
                 **kwargs):
        super().__init__(vp_pos, vp_dir, center, bounds, image, x_vec, y_vec, width,
                         volume_method, tf_obj, n_samples, **kwargs)
        self.light_dir = np.asarray(light_dir, dtype="float64")
        self.light_rgba = np.asarray(light_rgba, dtype="float64")

    def _light(self):
        return {"light_dir": self.light_dir, "light_rgba": self.light_rgba}
