"""Camera basis and plane-parallel lens parameters.

Restates, in numpy and without units (all lengths in code units):
  Orientation._setup_normalized_vectors   yt/utilities/orientation.py:59-85
  Camera (position/focus/width setters)   volume_rendering/camera.py:112-281,409-428
  Lens.setup_box_properties               volume_rendering/lens.py:45-61
  PlaneParallelLens._get_sampler_params   volume_rendering/lens.py:83-112
  PlaneParallelLens.set_viewpoint         volume_rendering/lens.py:114-121
  PerspectiveLens                         volume_rendering/lens.py:145-294
  StereoPerspective/Fisheye/Spherical/StereoSpherical lenses  lens.py:295-836
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
        # the second width follows from the first and the image's aspect ratio
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
        # no point of the domain lies farther from the camera than this
        max_length = np.linalg.norm(camera.position - camera._domain_center) + 0.5 * np.linalg.norm(
            camera._domain_width)
        # scaled so that t = 1 reaches beyond the domain's farthest point
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


def get_rotation_matrix(theta, rot_vector):
    """math_utils.py:1140-1206: rotation by theta radians about rot_vector
    (Rodrigues' formula, in the reference's term order)."""
    ux, uy, uz = rot_vector[0], rot_vector[1], rot_vector[2]
    cost = np.cos(theta)
    sint = np.sin(theta)
    return np.array([
        [cost + ux**2 * (1 - cost), ux * uy * (1 - cost) - uz * sint, ux * uz * (1 - cost) + uy * sint],
        [uy * ux * (1 - cost) + uz * sint, cost + uy**2 * (1 - cost), uy * uz * (1 - cost) - ux * sint],
        [uz * ux * (1 - cost) - uy * sint, uz * uy * (1 - cost) + ux * sint, cost + uz**2 * (1 - cost)],
    ])


def arr_fisheye_vectors(resolution, fov, nimx=1, nimy=1, nimi=0, nimj=0, off_theta=0.0,
                        off_phi=0.0):
    """grid_traversal.pyx:322-350 (Bourke's dome fisheye): a unit vector per
    pixel at angle r * fov / 2 from +z, zero outside radius 1.01."""
    fov_rad = fov * np.pi / 180.0
    nx = resolution // nimx
    ny = resolution // nimy
    px = (2.0 * (nimi * nx + np.arange(nx))) / resolution - 1.0
    py = (2.0 * (nimj * ny + np.arange(ny))) / resolution - 1.0
    PX, PY = np.meshgrid(px, py, indexing="ij")
    r = np.sqrt(PX * PX + PY * PY)
    phi = np.arctan2(PY, PX) + off_phi
    theta = r * fov_rad / 2.0 + off_theta
    vp = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], axis=-1)
    vp[r > 1.01] = 0.0
    return vp


def _max_length(camera):
    return np.linalg.norm(camera.position - camera._domain_center) + 0.5 * np.linalg.norm(
        camera._domain_width)


class StereoPerspectiveLens(PerspectiveLens):
    """lens.py:295-426: left and right perspective images side by side
    along the first image axis; disparity 3 pixels by default."""

    lens_type = "stereo-perspective"

    def __init__(self):
        self.disparity = None

    def _get_sampler_params(self, camera, render_source=None):
        camera.width[1] = camera.width[0] * (2.0 * camera.resolution[1] / camera.resolution[0])
        if self.disparity is None:
            self.disparity = 3.0 * camera.width[0] / camera.resolution[0]
        vl, pl = self._get_positions_vectors(camera, -self.disparity)
        vr, pr = self._get_positions_vectors(camera, self.disparity)
        uv = np.ones(3, dtype="float64")
        return {
            "vp_pos": np.vstack([pl, pr]).reshape(*camera.resolution, 3),
            "vp_dir": np.vstack([vl, vr]).reshape(*camera.resolution, 3),
            "center": self.back_center,
            "bounds": (0.0, 1.0, 0.0, 1.0),
            "x_vec": uv,
            "y_vec": uv,
            "width": np.zeros(3, dtype="float64"),
            "image": self.new_image(camera),
            "lens_type": "stereo-perspective",
        }

    def _get_positions_vectors(self, camera, disparity):
        """lens.py:368-426."""
        nx = int(np.floor(camera.resolution[0]) / 2)
        ny = camera.resolution[1]
        east_vec, north_vec, normal_vec = camera.unit_vectors
        angle_disparity = -np.arctan2(disparity, camera.width[2])
        R = get_rotation_matrix(angle_disparity, north_vec)
        east_vec_rot = np.dot(R, east_vec)
        normal_vec_rot = np.dot(R, normal_vec)
        px = np.linspace(-0.5, 0.5, nx)[np.newaxis, :]
        py = np.linspace(-0.5, 0.5, ny)[np.newaxis, :]
        sample_x = (camera.width[0] * np.array(east_vec_rot.reshape(3, 1) * px)).transpose()
        sample_y = (camera.width[1] * np.array(north_vec.reshape(3, 1) * py)).transpose()
        sample_x = np.repeat(sample_x.reshape(nx, 1, 3), ny, axis=1)
        sample_y = np.repeat(sample_y.reshape(1, ny, 3), nx, axis=0)
        normal_vecs = np.tile(normal_vec_rot, nx * ny).reshape(nx, ny, 3)
        east_vecs = np.tile(east_vec_rot, nx * ny).reshape(nx, ny, 3)
        max_length = _max_length(camera) + np.abs(self.disparity)
        vectors = (sample_x + sample_y + normal_vecs * camera.width[2]) * (
            max_length / camera.width[2])
        positions = np.tile(camera.position, nx * ny).reshape(nx, ny, 3)
        positions = positions + east_vecs * disparity
        return vectors, positions


class FisheyeLens(PlaneParallelLens):
    """lens.py:515-574: dome fisheye, field of view `fov` degrees, square
    image of resolution[0] pixels; rays from the camera position."""

    lens_type = "fisheye"

    def __init__(self):
        self.fov = 180.0
        self.radius = 1.0
        self.center = None
        self.rotation_matrix = np.eye(3)

    def setup_box_properties(self, camera):
        self.radius = camera.width.max()
        super().setup_box_properties(camera)
        self.set_viewpoint(camera)

    def set_viewpoint(self, camera):
        self.viewpoint = camera.position

    def new_image(self, camera):
        return np.zeros((camera.resolution[0], camera.resolution[0], 4), dtype="float64")

    def _get_sampler_params(self, camera, render_source=None):
        shape = (camera.resolution[0], camera.resolution[0], 3)
        vp = -arr_fisheye_vectors(camera.resolution[0], self.fov).reshape(shape)
        vp = vp.dot(np.linalg.inv(self.rotation_matrix))
        vp *= self.radius
        uv = np.ones(3, dtype="float64")
        return {
            "vp_pos": np.full(shape, camera.position, dtype="float64"),
            "vp_dir": vp,
            "center": self.center if self.center is not None else np.zeros(3),
            "bounds": (0.0, 1.0, 0.0, 1.0),
            "x_vec": uv,
            "y_vec": uv,
            "width": np.zeros(3, dtype="float64"),
            "image": self.new_image(camera),
            "lens_type": "fisheye",
        }


class SphericalLens(PlaneParallelLens):
    """lens.py:626-688: equirectangular (longitude along image axis 0,
    latitude along axis 1) rays from the camera position."""

    lens_type = "spherical"

    def __init__(self):
        self.radius = 1.0
        self.center = None
        self.rotation_matrix = np.eye(3)

    def setup_box_properties(self, camera):
        self.radius = camera.width.max()
        super().setup_box_properties(camera)
        self.set_viewpoint(camera)

    def set_viewpoint(self, camera):
        self.viewpoint = camera.position

    def _uv(self, camera):
        R1 = get_rotation_matrix(0.5 * np.pi, [1, 0, 0])
        R2 = get_rotation_matrix(0.5 * np.pi, [0, 0, 1])
        return np.dot(R2, np.dot(R1, camera.unit_vectors))

    def _get_sampler_params(self, camera, render_source=None):
        px = np.linspace(-np.pi, np.pi, camera.resolution[0], endpoint=True)[:, None]
        py = np.linspace(-np.pi / 2.0, np.pi / 2.0, camera.resolution[1], endpoint=True)[None, :]
        positions = np.tile(camera.position, np.prod(camera.resolution)).reshape(
            *camera.resolution, 3)
        uv = self._uv(camera)
        vectors = np.empty((*camera.resolution, 3), dtype="float64", order="C")
        vectors[:, :, 0] = np.cos(px) * np.cos(py)
        vectors[:, :, 1] = np.sin(px) * np.cos(py)
        vectors[:, :, 2] = np.sin(py)
        vectors = np.dot((vectors * _max_length(camera)), uv).reshape(*camera.resolution, 3)
        dummy = np.ones(3, dtype="float64")
        return {
            "vp_pos": positions,
            "vp_dir": vectors,
            "center": self.back_center,
            "bounds": (0.0, 1.0, 0.0, 1.0),
            "x_vec": dummy,
            "y_vec": dummy,
            "width": np.zeros(3, dtype="float64"),
            "image": self.new_image(camera).reshape(*camera.resolution, 4),
            "lens_type": "spherical",
        }


class StereoSphericalLens(SphericalLens):
    """lens.py:731-836: two spherical images (left, right eye) stacked along
    image axis 1, eye offset `disparity` along the local east direction."""

    lens_type = "stereo-spherical"

    def __init__(self):
        super().__init__()
        self.disparity = None

    def _get_sampler_params(self, camera, render_source=None):
        if self.disparity is None:
            self.disparity = camera.width[0] / 1000.0
        ny = int(np.floor(camera.resolution[1]) / 2)
        nx = camera.resolution[0]
        px = np.linspace(-np.pi, np.pi, nx, endpoint=True)[:, None]
        py = np.linspace(-np.pi / 2.0, np.pi / 2.0, ny, endpoint=True)[None, :]
        vectors = np.zeros((nx, ny, 3), dtype="float64", order="C")
        vectors[:, :, 0] = np.cos(px) * np.cos(py)
        vectors[:, :, 1] = np.sin(px) * np.cos(py)
        vectors[:, :, 2] = np.sin(py)
        max_length = _max_length(camera) + np.abs(self.disparity)
        vectors = vectors * max_length
        uv = self._uv(camera)
        vectors = np.dot(vectors, uv)
        vectors2 = np.zeros((nx, ny, 3), dtype="float64", order="C")
        vectors2[:, :, 0] = -np.sin(px) * np.ones((1, ny))
        vectors2[:, :, 1] = np.cos(px) * np.ones((1, ny))
        vectors2[:, :, 2] = 0
        vectors2 = np.dot(vectors2, uv)
        positions = np.tile(camera.position, nx * ny).reshape(nx, ny, 3)
        # The left and right are switched here since VR is in LHS (lens.py:808).
        positions_left = positions + vectors2 * self.disparity
        positions_right = positions + vectors2 * (-self.disparity)
        dummy = np.ones(3, dtype="float64")
        return {
            "vp_pos": np.hstack([positions_left, positions_right]).reshape(*camera.resolution, 3),
            "vp_dir": np.hstack([vectors, vectors]).reshape(*camera.resolution, 3),
            "center": self.back_center,
            "bounds": (0.0, 1.0, 0.0, 1.0),
            "x_vec": dummy,
            "y_vec": dummy,
            "width": np.zeros(3, dtype="float64"),
            "image": self.new_image(camera).reshape(*camera.resolution, 4),
            "lens_type": "stereo-spherical",
        }


LENSES = {
    "plane-parallel": PlaneParallelLens,
    "perspective": PerspectiveLens,
    "stereo-perspective": StereoPerspectiveLens,
    "fisheye": FisheyeLens,
    "spherical": SphericalLens,
    "stereo-spherical": StereoSphericalLens,
}


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

    def set_position(self, position, north_vector=None):
        """camera.py:367-385."""
        if north_vector is not None:
            self.north_vector = north_vector
        self.position = position

    def get_position(self):
        return self.position

    def set_focus(self, new_focus):
        """camera.py:391-403."""
        self.focus = new_focus

    def get_focus(self):
        return self.focus

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
