"""Scene, VolumeSource, ZBuffer and image output on top of the GPU
ray-caster.

Restates the reference's host-side composition of a volume rendering:
  ZBuffer                          volume_rendering/zbuffer_array.py:1-80
  VolumeSource.render/set_sampler  volume_rendering/render_source.py:464-616
  RenderSource.finalize_image      volume_rendering/render_source.py:522-537
  Scene.add_source/add_camera/composite/render/save
                                   volume_rendering/scene.py:113-625
  create_scene                     volume_rendering/volume_rendering.py:10-79
  ImageArray.rescale/add_background_color/_clipping_value/write_png
                                   data_objects/image_array.py:138-432
  write_bitmap                     visualization/image_writer.py:101-165
The rays are cast by the persistent GPU kernels through a ResidentVolume
(one per source, uploaded once); the brick order is the kd-tree's order for
the camera's viewpoint, as VolumeSource.render walks it.
"""

import struct
import zlib

import numpy as np

from . import _lib, render
from .camera import Camera
from .transfer_functions import (
    ColorTransferFunction,
    ProjectionTransferFunction,
    TransferFunction,
)


class ZBuffer:
    """An image with a depth per pixel (zbuffer_array.py:5-80).  Adding two
    keeps, pixel by pixel, the colour of the nearer one: with 2-d depth maps
    a pixel at equal depth in both becomes 0; with one-column depth maps the
    two are blended with weights 1 and 0, the second winning ties
    (zbuffer_array.py:44-57)."""

    def __init__(self, rgba, z):
        if rgba.shape[: z.ndim] != z.shape:
            raise AssertionError("ZBuffer: the image and its depths differ in shape")
        self.rgba = rgba
        self.z = z
        self.shape = z.shape

    def __add__(self, other):
        if self.shape != other.shape:
            raise AssertionError("ZBuffer: adding buffers of different shapes")
        mine = self.z < other.z
        if self.z.shape[1] == 1:
            w = mine[:, None, :].astype(np.float64)
            rgba = self.rgba * w + other.rgba * (1.0 - w)
        else:
            extra = (1,) * (self.rgba.ndim - mine.ndim)  # per-pixel masks over the channels
            rgba = np.zeros(self.rgba.shape)
            np.copyto(rgba, self.rgba, where=mine.reshape(mine.shape + extra))
            np.copyto(rgba, other.rgba, where=(other.z < self.z).reshape(mine.shape + extra))
        return ZBuffer(rgba, np.minimum(self.z, other.z))

    def __iadd__(self, other):
        merged = self + other
        self.rgba, self.z = merged.rgba, merged.z
        return self

    def __eq__(self, other):
        return bool((self.rgba == other.rgba).all() and (self.z == other.z).all())

    def paint(self, ind, value, z):
        """Draws value at pixel ind if z is nearer than what is there."""
        if self.z[ind] > z:
            self.z[ind] = z
            self.rgba[ind] = value


# ---------------------------------------------------------------------------
# Image output (image_array.py, image_writer.py)

def rescale(image, cmax=None, amax=None):
    """ImageArray.rescale (image_array.py:191-250), on a copy."""
    out = np.array(image, dtype="float64", copy=True)
    if cmax is None:
        cmax = image[:, :, :3].sum(axis=2).max()
    if cmax > 0.0:
        np.multiply(image[:, :, :3], 1.0 / cmax, out[:, :, :3])
    if image.shape[2] == 4:
        if amax is None:
            amax = image[:, :, 3].max()
        if amax > 0.0:
            np.multiply(image[:, :, 3], 1.0 / amax, out[:, :, 3])
    np.clip(out, 0.0, 1.0, out)
    return out


_BACKGROUNDS = {"white": (1.0, 1.0, 1.0, 1.0), "black": (0.0, 0.0, 0.0, 1.0)}


def add_background_color(image, background="black"):
    """ImageArray.add_background_color (image_array.py:138-189), on a copy:
    the image, its colours weighted by its alpha, over an opaque or
    transparent background colour ("white", "black", None for transparent
    black, or an RGBA tuple)."""
    bg = (0.0, 0.0, 0.0, 0.0) if background is None else _BACKGROUNDS.get(background, background)
    alpha = image[:, :, 3]
    behind = 1.0 - alpha
    out = np.array(image, copy=True)
    for c in range(3):
        out[:, :, c] = image[:, :, c] * alpha + (bg[c] * bg[3]) * behind
    out[:, :, 3] = alpha + bg[3] * behind
    return out


def clipping_value(image, sigma_clip):
    """ImageArray._clipping_value (image_array.py:426-432): mean + sigma_clip
    standard deviations of the non-zero colour values."""
    nz = image[:, :, :3][image[:, :, :3].nonzero()]
    return nz.mean() + sigma_clip * nz.std()


def _png_bytes(rgba8):
    """An 8-bit RGBA PNG (signature, IHDR, one zlib IDAT with filter-0 rows,
    IEND), the file format the reference's png_writer produces."""
    h, w, _ = rgba8.shape

    def chunk(tag, payload):
        return (struct.pack(">I", len(payload)) + tag + payload
                + struct.pack(">I", zlib.crc32(tag + payload) & 0xFFFFFFFF))

    raw = b"".join(b"\x00" + rgba8[y].tobytes() for y in range(h))
    return (b"\x89PNG\r\n\x1a\n"
            + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 6))
            + chunk(b"IEND", b""))


def write_bitmap(bitmap_array, filename, max_val=None, transpose=False):
    """image_writer.py:101-165: colour channels scaled by max_val (default the
    colour maximum) and clipped to [0, 1], alpha times 255, written as PNG.
    Returns the uint8 array (or the PNG bytes when filename is None)."""
    if len(bitmap_array.shape) != 3 or bitmap_array.shape[-1] not in (3, 4):
        raise RuntimeError("Expecting image array of shape (N,M,3) or "
                           f"(N,M,4), received {str(bitmap_array.shape)}")
    if bitmap_array.dtype != np.uint8:
        s1, s2 = bitmap_array.shape[:2]
        if bitmap_array.shape[-1] == 3:
            alpha_channel = np.full((s1, s2, 1), 255, dtype="uint8")
        else:
            alpha_channel = (255 * bitmap_array[:, :, 3]).astype("uint8").reshape(s1, s2, 1)
        if max_val is None:
            max_val = bitmap_array[:, :, :3].max()
            if max_val == 0.0:
                max_val = 1.0
        bitmap_array = np.clip(bitmap_array[:, :, :3] / max_val, 0.0, 1.0) * 255
        bitmap_array = np.concatenate([bitmap_array.astype("uint8"), alpha_channel], axis=-1)
    if transpose:
        bitmap_array = bitmap_array.swapaxes(0, 1).copy(order="C")
    data = _png_bytes(np.ascontiguousarray(bitmap_array))
    if filename is None:
        return data
    with open(filename, "wb") as f:
        f.write(data)
    return bitmap_array


def write_png(image, filename, sigma_clip=None, background="black", rescale_image=True):
    """ImageArray.write_png (image_array.py:252-319)."""
    scaled = rescale(image) if rescale_image else image
    out = add_background_color(scaled, background) if image.shape[-1] == 4 else scaled
    if filename is not None and filename[-4:] != ".png":
        filename += ".png"
    if sigma_clip is not None:
        return write_bitmap(out.swapaxes(0, 1), filename, clipping_value(out, sigma_clip))
    return write_bitmap(out.swapaxes(0, 1), filename)


# ---------------------------------------------------------------------------
# Sources and scenes

class OpaqueSource:
    """render_source.py:163-175: the base of sources that write into the
    scene's z-buffer (lines, boxes, ...).  Scene.composite renders them
    first, each given the shared ZBuffer to draw into."""

    opaque = True

    def __init__(self):
        self.zbuffer = None

    def set_zbuffer(self, zbuffer):
        self.zbuffer = zbuffer

    def render(self, camera, zbuffer=None):
        raise NotImplementedError


class VolumeSource:
    """KDTreeVolumeSource (render_source.py:563-616) on the GPU path.

    sampler_type "volume-render" casts with VolumeRenderSampler,
    "projection" with InterpolatedProjectionSampler (render() calls
    set_sampler with its `interpolated=True` default, render_source.py:
    464-488, 581).  A ProjectionTransferFunction switches the source to
    "projection" with linear fields (render_source.py:104-116, 295-302).  A
    weight field becomes the kd bricks' second field, with the same log
    flag (render_source.py:121-134)."""

    volume_method = "KDTree"

    def __init__(self, ds, field="density", log_field=True, transfer_function=None,
                 sampler_type="volume-render", n_samples=10, ctx=None, weight_field=None):
        self.ds = ds
        self.field = field
        self.log_field = log_field
        self.weight_field = weight_field
        self.sampler_type = sampler_type
        self._transfer_function = None
        self.transfer_function = transfer_function
        self.num_samples = n_samples
        self._ctx = ctx
        self.zbuffer = None
        self.current_image = None
        self.sampler = None
        self._volume = None
        self._volume_key = None

    @property
    def ctx(self):
        if self._ctx is None:
            self._ctx = _lib.default_context()
        return self._ctx

    @property
    def transfer_function(self):
        return self._transfer_function

    @transfer_function.setter
    def transfer_function(self, value):
        """render_source.py:280-302."""
        valid_types = (TransferFunction, ColorTransferFunction, ProjectionTransferFunction,
                       type(None))
        if not isinstance(value, valid_types):
            raise RuntimeError("transfer_function not a valid type, "
                               f"received object of type {type(value)}")
        if isinstance(value, ProjectionTransferFunction):
            self.sampler_type = "projection"
            self.log_field = False
        self._transfer_function = value

    def set_transfer_function(self, transfer_function):
        self.transfer_function = transfer_function
        return self

    def set_log(self, log_field):
        self.log_field = log_field
        return self

    def set_weight_field(self, weight_field):
        self.weight_field = weight_field
        return self

    def _fields(self):
        """validate_volume (render_source.py:121-134): [field] or
        [field, weight_field], one log flag for both."""
        fields = [self.field]
        if self.weight_field is not None:
            fields.append(self.weight_field)
        return fields, [bool(self.log_field)] * len(fields)

    @property
    def volume(self):
        fields, logs = self._fields()
        key = (tuple(fields), tuple(logs))
        if self._volume is not None and self._volume_key != key:
            self._volume.free()
            self._volume = None
        if self._volume is None:
            self._volume = render.ResidentVolume(self.ctx, self.ds, vr_field=fields,
                                                 log_field=logs)
            self._volume_key = key
        return self._volume

    def _params(self, camera):
        p = camera._get_sampler_params(self)
        p["viewpoint"] = np.array(camera.lens.viewpoint, dtype="float64")
        return p

    def set_sampler(self, camera, interpolated=True):
        """render_source.py:464-488 with utils.py:55-139: the drop-in sampler
        for this source and camera, to be called with the bricks of
        `self.volume.kd.traverse(camera.lens.viewpoint)`.  "projection" with
        interpolated=False is ProjectionSampler on vertex-centred bricks,
        which reads each brick's vertex array with cell strides (see
        image_samplers.ImageSampler.flush)."""
        from . import image_samplers as ims

        p = camera._get_sampler_params(self)
        image = p["image"]
        args = (np.atleast_3d(p["vp_pos"]), np.atleast_3d(p["vp_dir"]), p["center"],
                p["bounds"], image, p["x_vec"], p["y_vec"], p["width"], self.volume_method)
        kw = {"lens_type": p["lens_type"]}
        kw["zbuffer"] = (self.zbuffer.z if self.zbuffer is not None
                         else np.ones(image.shape[:2], "float64"))
        if self.sampler_type == "volume-render":
            if "camera_data" in p:
                kw["camera_data"] = p["camera_data"]
            if self.zbuffer is not None:
                image[:] = np.reshape(self.zbuffer.rgba, (*camera.resolution, 4))
            sampler = ims.VolumeRenderSampler(*args, self.transfer_function, self.num_samples,
                                              **kw)
        elif self.sampler_type == "projection" and interpolated:
            sampler = ims.InterpolatedProjectionSampler(*args, self.num_samples, **kw)
        elif self.sampler_type == "projection":
            sampler = ims.ProjectionSampler(*args, self.num_samples, **kw)
        else:
            raise NotImplementedError(f"{self.sampler_type} not implemented yet")
        self.sampler = sampler
        return sampler

    def render(self, camera, zbuffer=None):
        """render_source.py:563-603: cast every kd brick in viewpoint order,
        starting from zbuffer.rgba and stopping at zbuffer.z (utils.py:78-86),
        then finalize_image."""
        self.zbuffer = zbuffer
        p = self._params(camera)
        image_in = zb = None
        if zbuffer is not None:
            image_in = np.reshape(zbuffer.rgba, (*camera.resolution, 4))
            zb = zbuffer.z
        if self.sampler_type == "volume-render":
            if self.transfer_function is None:
                raise RuntimeError("a volume-render source needs a transfer function")
            spec = self.volume.vr_request(p, self.transfer_function, n_samples=self.num_samples,
                                          zbuffer=zb)
        elif self.sampler_type == "projection":
            spec = self.volume.interpolated_projection_request(p, n_samples=self.num_samples,
                                                               zbuffer=zb)
        else:
            raise NotImplementedError(f"{self.sampler_type} not implemented yet")
        spec.image_in = None if image_in is None or not np.any(image_in) else np.ascontiguousarray(
            image_in, dtype=np.float64)
        b = _lib.Batch(self.ctx, [spec], _lib.BATCH_PERSISTENT)
        try:
            b.run()
            image = b.read_image(0)
        finally:
            b.close()
        self.current_image = self.finalize_image(camera, image)
        if zbuffer is None:
            self.zbuffer = ZBuffer(self.current_image,
                                   np.full(self.current_image.shape[:2], np.inf))
        return self.current_image

    def finalize_image(self, camera, image):
        """render_source.py:522-537: alpha := 1 unless grey opacity."""
        image = image.reshape(*camera.resolution, 4)
        grey = getattr(self.transfer_function, "grey_opacity", False)
        if not grey:
            image[:, :, 3] = 1
        return image

    def free(self):
        if self._volume is not None:
            self._volume.free()
            self._volume = None


class Scene:
    """scene.py:42-625, the parts that render: sources, one camera,
    composite (opaque sources into the z-buffer, then transparent ones from
    it, then a 180-degree rotation) and save."""

    def __init__(self):
        self.sources = {}
        self.camera = None
        self._last_render = None

    def add_source(self, render_source, keyname=None):
        if keyname is None:
            keyname = f"source_{len(self.sources):02d}"
        self.sources[keyname] = render_source
        return self

    def __getitem__(self, item):
        return self.sources[item]

    @property
    def opaque_sources(self):
        return [(k, s) for k, s in self.sources.items() if getattr(s, "opaque", False)]

    @property
    def transparent_sources(self):
        return [(k, s) for k, s in self.sources.items() if not getattr(s, "opaque", False)]

    def add_camera(self, data_source=None, lens_type="plane-parallel", auto=False):
        """scene.py:627-683: a camera looking at the data's centre."""
        self.camera = Camera(self, data_source, lens_type=lens_type)
        return self.camera

    def composite(self, camera=None):
        """scene.py:581-625."""
        if camera is None:
            camera = self.camera
        empty = camera.lens.new_image(camera)
        opaque = ZBuffer(empty, np.full(empty.shape[:2], np.inf))
        im = empty
        for _, source in self.opaque_sources:
            source.render(camera, zbuffer=opaque)
            im = source.zbuffer.rgba
        for _, source in self.transparent_sources:
            im = source.render(camera, zbuffer=opaque)
            opaque.rgba = im
        return np.rot90(im, k=2)

    def render(self, camera=None):
        """scene.py:215-262."""
        self._last_render = self.composite(camera=camera)
        return self._last_render

    def save(self, fname=None, sigma_clip=None, render_scene=True):
        """scene.py:271-355 (PNG output)."""
        if render_scene or self._last_render is None:
            self.render()
        fname = fname or "rendering.png"
        write_png(self._last_render, fname, sigma_clip=sigma_clip)
        return fname


def create_scene(ds, field="density", lens_type="plane-parallel", transfer_function=None,
                 log_field=True, ctx=None):
    """volume_rendering.py:10-79: a scene with one volume source of `field`
    and a camera; the transfer function defaults to 8 Gaussians over the
    field's (log) range, a ColorTransferFunction as the reference's
    TransferFunctionHelper builds by default."""
    from .transfer_functions import gaussian_ctf

    sc = Scene()
    if transfer_function is None:
        # the field's extrema (of its positive values for a log field), in
        # log10 for a log field
        lo, hi = np.inf, -np.inf
        for g in ds.grids:
            v = np.asarray(g[field])
            if log_field:
                v = v[v > 0]
            if v.size:
                lo, hi = min(lo, float(v.min())), max(hi, float(v.max()))
        if log_field:
            lo, hi = float(np.log10(lo)), float(np.log10(hi))
        transfer_function = gaussian_ctf((lo, hi), n_gauss=8, nbins=256)
    src = VolumeSource(ds, field=field, log_field=log_field, transfer_function=transfer_function,
                       ctx=ctx)
    sc.add_source(src)
    sc.add_camera(ds, lens_type=lens_type)
    return sc
