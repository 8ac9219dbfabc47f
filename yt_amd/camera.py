"""Camera basis and plane-parallel lens parameters.

Restates, in numpy and without units (all lengths in code units):
  Orientation._setup_normalized_vectors   yt/utilities/orientation.py:59-85
  Camera (position/focus/width setters)   volume_rendering/camera.py:112-281,409-428
  Lens.setup_box_properties               volume_rendering/lens.py:45-61
  PlaneParallelLens._get_sampler_params   volume_rendering/lens.py:83-112
  PlaneParallelLens.set_viewpoint         volume_rendering/lens.py:114-121
  PerspectiveLens                         volume_rendering/lens.py:145-294
The sampler parameters (vp_pos, vp_dir, center, bounds, x_vec, y_vec, width)
are exactly what ImageSampler's constructor receives (utils.py:115-139).
"""

import numpy as np


def _aligned(a, b):
    comp = np.abs(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b))
    return np.isclose(comp, 1.0, 1.0e-13)


def _validate_unit_vectors(normal_vector, north_vector):
    """orientation.py:12-24."""
    if north_vector is not None:
        north_vector = np.array(north_vector, dtype="float64")
    if normal_vector is not None:
        normal_vector = np.array(normal_vector, dtype="float64")
    if not np.dot(normal_vector, normal_vector) > 0:
        raise ValueError("normal_vector cannot be the zero vector.")
    if north_vector is not None and _aligned(north_vector, normal_vector):
        raise ValueError("normal_vector and north_vector cannot be aligned.")
    return normal_vector, north_vector


class Orientation:
    """orientation.py:27-85."""

    def __init__(self, normal_vector, north_vector=None, steady_north=False):
        normal_vector, north_vector = _validate_unit_vectors(normal_vector, north_vector)
        self.steady_north = steady_north
        if north_vector is not None:
            self.steady_north = True
        self.north_vector = north_vector
        self._setup_normalized_vectors(normal_vector, north_vector)
        if self.north_vector is None:
            self.north_vector = self.unit_vectors[1]

    def _setup_normalized_vectors(self, normal_vector, north_vector):
        normal_vector, north_vector = _validate_unit_vectors(normal_vector, north_vector)
        normal_vector /= np.sqrt(np.dot(normal_vector, normal_vector))
        if north_vector is None:
            vecs = np.identity(3)
            t = np.cross(normal_vector, vecs).sum(axis=1)
            ax = t.argmax()
            east_vector = np.cross(vecs[ax, :], normal_vector).ravel()
            north_vector = np.cross(normal_vector, east_vector).ravel()
        else:
            if self.steady_north or (np.dot(north_vector, normal_vector) != 0.0):
                north_vector = north_vector - np.dot(north_vector, normal_vector) * normal_vector
            east_vector = np.cross(north_vector, normal_vector).ravel()
        north_vector /= np.sqrt(np.dot(north_vector, north_vector))
        east_vector /= np.sqrt(np.dot(east_vector, east_vector))
        self.normal_vector = normal_vector
        self.north_vector = north_vector
        self.unit_vectors = np.array([east_vector, north_vector, normal_vector])
        self.inv_mat = np.linalg.pinv(self.unit_vectors)


def default_north(normal):
    """The north vector off_axis_projection derives when none is given
    (off_axis_projection.py:402-407)."""
    vecs = np.identity(3)
    t = np.cross(vecs, normal).sum(axis=1)
    ax = t.argmax()
    east_vector = np.cross(vecs[ax, :], normal).ravel()
    return np.cross(normal, east_vector).ravel()


def _as3(value):
    """Scalar or 3-sequence to a float64 3-vector (camera.py:15-44, code units)."""
    a = np.asarray(value, dtype="float64")
    if a.ndim == 0:
        return np.array([float(a)] * 3)
    if a.size == 1:
        return np.array([float(a.ravel()[0])] * 3)
    if a.size != 3:
        raise RuntimeError(f"Cannot set camera property to invalid value '{value}'")
    return a.astype("float64").copy()


class PlaneParallelLens:
    """lens.py:72-142 (orthographic rays)."""

    lens_type = "plane-parallel"

    def setup_box_properties(self, camera):
        """lens.py:45-61."""
        uv = camera.unit_vectors
        width = camera.width
        center = camera.focus
        self.box_vectors = np.array([uv[0] * width[0], uv[1] * width[1], uv[2] * width[2]])
        self.origin = center - 0.5 * width.dot(uv)
        self.back_center = center - 0.5 * width[2] * uv[2]
        self.front_center = center + 0.5 * width[2] * uv[2]
        self.set_viewpoint(camera)

    def set_viewpoint(self, camera):
        """lens.py:114-121: a viewpoint 1e6 widths beyond the front plane."""
        self.viewpoint = self.front_center + camera.unit_vectors[2] * 1.0e6 * camera.width[2]

    def new_image(self, camera):
        return np.zeros((camera.resolution[0], camera.resolution[1], 4), dtype="float64",
                        order="C")

    def _get_sampler_params(self, camera, render_source=None):
        """lens.py:83-112."""
        zb = getattr(render_source, "zbuffer", None) if render_source is not None else None
        image = zb.rgba if zb is not None else self.new_image(camera)
        vp_pos = np.concatenate([camera.inv_mat.ravel("F"), self.back_center.ravel()])
        return {
            "vp_pos": vp_pos,
            "vp_dir": self.box_vectors[2],
            "center": self.back_center,
            "bounds": (-camera.width[0] / 2.0, camera.width[0] / 2.0,
                       -camera.width[1] / 2.0, camera.width[1] / 2.0),
            "x_vec": camera.unit_vectors[0],
            "y_vec": camera.unit_vectors[1],
            "width": np.array(camera.width, dtype="float64"),
            "image": image,
            "lens_type": "plane-parallel",
        }


class PerspectiveLens(PlaneParallelLens):
    """lens.py:145-294: rays from the camera position through a plane of
    width x height at distance width[2], one origin and direction per pixel
    (generate_vector_info_null, lenses.pyx:196-207)."""

    lens_type = "perspective"

    def set_viewpoint(self, camera):
        """lens.py:236-240: the front centre."""
        self.viewpoint = self.front_center

    def _get_sampler_params(self, camera, render_source=None):
        """lens.py:166-234, in the same numpy operations and order."""
        # Enforce width[1] / width[0] = resolution[1] / resolution[0]
        camera.width[1] = camera.width[0] * (camera.resolution[1] / camera.resolution[0])
        res0, res1 = camera.resolution
        east_vec = camera.unit_vectors[0]
        north_vec = camera.unit_vectors[1]
        normal_vec = camera.unit_vectors[2]
        px = np.linspace(-0.5, 0.5, res0)[np.newaxis, :]
        py = np.linspace(-0.5, 0.5, res1)[np.newaxis, :]
        sample_x = camera.width[0] * np.array(east_vec.reshape(3, 1) * px)
        sample_x = sample_x.transpose()
        sample_y = camera.width[1] * np.array(north_vec.reshape(3, 1) * py)
        sample_y = sample_y.transpose()
        sample_x = np.repeat(sample_x.reshape(res0, 1, 3), res1, axis=1)
        sample_y = np.repeat(sample_y.reshape(1, res1, 3), res0, axis=0)
        normal_vecs = np.tile(normal_vec, res0 * res1).reshape(res0, res1, 3)
        # The maximum possible length of ray
        max_length = np.linalg.norm(camera.position - camera._domain_center) + 0.5 * np.linalg.norm(
            camera._domain_width)
        # Rescale the ray to be long enough to cover the entire domain
        vectors = (sample_x + sample_y + normal_vecs * camera.width[2]) * (
            max_length / camera.width[2])
        positions = np.tile(camera.position, res0 * res1).reshape(res0, res1, 3)
        uv = np.ones(3, dtype="float64")
        return {
            "vp_pos": positions,
            "vp_dir": vectors,
            "center": self.back_center,
            "bounds": (0.0, 1.0, 0.0, 1.0),
            "x_vec": uv,
            "y_vec": uv,
            "width": np.zeros(3, dtype="float64"),
            "image": self.new_image(camera),
            "lens_type": "perspective",
        }


LENSES = {"plane-parallel": PlaneParallelLens, "perspective": PerspectiveLens}


class Camera(Orientation):
    """camera.py:47-428, restricted to what the ray-caster reads.

    Defaults follow camera.py:126-141: focus at the domain centre, position
    at the domain's right edge, width 1.5x the largest domain width."""

    def __init__(self, scene=None, data_source=None, lens_type="plane-parallel",
                 domain_left_edge=(0.0, 0.0, 0.0), domain_right_edge=(1.0, 1.0, 1.0)):
        if lens_type not in LENSES:
            raise NotImplementedError(f"lens type {lens_type!r} is not implemented")
        self.scene = scene
        self.lens = None
        self.north_vector = None
        self.normal_vector = None
        self._resolution = (512, 512)
        ds = getattr(data_source, "ds", data_source)
        if ds is not None and hasattr(ds, "domain_left_edge"):
            dle = np.asarray(ds.domain_left_edge, dtype="float64")
            dre = np.asarray(ds.domain_right_edge, dtype="float64")
        else:
            dle = np.asarray(domain_left_edge, dtype="float64")
            dre = np.asarray(domain_right_edge, dtype="float64")
        self._domain_center = 0.5 * (dle + dre)
        self._domain_width = dre - dle
        self._focus = 0.5 * (dle + dre)
        self._position = dre.copy()
        self._width = np.array([1.5 * (dre - dle).max()] * 3)
        super().__init__(self._focus - self._position, self.north_vector, steady_north=False)
        self.lens = LENSES[lens_type]()
        self.lens.setup_box_properties(self)

    def set_lens(self, lens_type):
        """camera.py:283-303."""
        if lens_type not in LENSES:
            raise RuntimeError(f"Lens type {lens_type!r} not in available list {sorted(LENSES)}")
        self.lens = LENSES[lens_type]()
        self.lens.setup_box_properties(self)

    @property
    def position(self):
        return self._position

    @position.setter
    def position(self, value):
        position = _as3(value)
        if np.array_equal(position, self._focus):
            raise RuntimeError("Cannot set the camera focus and position to the same value")
        self._position = position
        self.switch_orientation(normal_vector=self._focus - self._position,
                                north_vector=self.north_vector)

    @property
    def focus(self):
        return self._focus

    @focus.setter
    def focus(self, value):
        focus = _as3(value)
        if np.array_equal(focus, self._position):
            raise RuntimeError("Cannot set the camera focus and position to the same value")
        self._focus = focus
        self.switch_orientation(normal_vector=self._focus - self._position, north_vector=None)

    @property
    def width(self):
        return self._width

    @width.setter
    def width(self, value):
        self._width = _as3(value)
        self.switch_orientation()

    def set_width(self, width):
        self.width = width
        self.switch_orientation()

    @property
    def resolution(self):
        return self._resolution

    @resolution.setter
    def resolution(self, value):
        if np.ndim(value) > 0:
            if len(value) != 2:
                raise RuntimeError
            value = (int(value[0]), int(value[1]))
        else:
            value = (int(value), int(value))
        self._resolution = value

    def switch_orientation(self, normal_vector=None, north_vector=None):
        """camera.py:409-428."""
        if north_vector is None:
            north_vector = self.north_vector
        if normal_vector is None:
            normal_vector = self.normal_vector
        self._setup_normalized_vectors(normal_vector, north_vector)
        self.lens.setup_box_properties(self)

    def _get_sampler_params(self, render_source=None):
        """camera.py:275-281."""
        p = self.lens._get_sampler_params(self, render_source)
        p.update(width=np.array(self.width, dtype="float64"))
        p.update(camera_data=np.vstack((self.position, self.width, self.unit_vectors)))
        p["nv"] = tuple(self.resolution)
        return p


def off_axis_camera(center, normal_vector, width, resolution, north_vector=None,
                    domain_left_edge=(0.0, 0.0, 0.0), domain_right_edge=(1.0, 1.0, 1.0)):
    """The camera off_axis_projection builds (off_axis_projection.py:385-411)."""
    cam = Camera(domain_left_edge=domain_left_edge, domain_right_edge=domain_right_edge)
    cam.set_width(width)
    if np.ndim(resolution) == 0:
        resolution = [resolution] * 2
    cam.resolution = resolution
    width = _as3(width)
    center = np.asarray(center, dtype="float64")
    normal = np.array(normal_vector, dtype="float64")
    normal = normal / np.linalg.norm(normal)
    cam.position = center - width[2] * normal
    cam.focus = center
    if north_vector is None:
        north = default_north(normal)
    else:
        north = np.array(north_vector, dtype="float64")
        north = north / np.linalg.norm(north)
    cam.switch_orientation(normal, north)
    return cam


def sampler_params(center, normal_vector, width, resolution, north_vector=None):
    """Sampler parameters of the off-axis-projection camera, as a dict."""
    cam = off_axis_camera(center, normal_vector, width, resolution, north_vector)
    p = cam._get_sampler_params()
    p["viewpoint"] = cam.lens.viewpoint.copy()
    return p
