"""The oracle's per-pixel lens path (generate_vector_info_null,
lenses.pyx:196-207) and its perspective extent (calculate_extent_perspective,
lenses.pyx:60-161), pinned on the CPU:

  * the clipped perspective extent of a brick in front of the camera holds
    every pixel whose ray meets the brick (brute-force slab test on the
    lens's own rays), for non-square images and off-centre bricks, so a
    swapped east/north, a sign error or a wrong pixel scale fails;
  * a per-pixel cast whose rays are the plane-parallel lens's rays equals the
    plane-parallel cast bit for bit (the walk and samplers are shared; only
    where the rays come from changes);
  * the PerspectiveLens sampler parameters follow lens.py:166-234 (ray
    origins at the camera, directions through the image plane scaled to the
    maximum ray length).
"""

import numpy as np
import pytest

from oracle import oracle
from yt_amd.camera import Camera

from .helpers import c0_params, oracle_proj_bricks, uniform_lognormal


def _persp(res, position, focus=(0.5, 0.5, 0.5), width=0.6, north=None):
    cam = Camera(lens_type="perspective")
    cam.resolution = res
    cam.position = position
    cam.focus = focus
    cam.width = width
    if north is not None:
        cam.switch_orientation(north_vector=north)
    p = cam._get_sampler_params()
    return cam, p


def _chord(le, re, pos, d):
    """Length in t in [0, 1] of the ray inside [le, re] (slab test)."""
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if d[a] == 0.0:
            if not (le[a] <= pos[a] <= re[a]):
                return 0.0
            continue
        ta, tb = (le[a] - pos[a]) / d[a], (re[a] - pos[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return max(0.0, t1 - t0)


@pytest.mark.parametrize("seed", range(6))
def test_perspective_extent_holds_every_hit_pixel(seed):
    rng = np.random.default_rng(seed)
    res = (int(rng.integers(12, 40)), int(rng.integers(12, 40)))
    d = rng.normal(size=3)
    pos = 0.5 + 1.6 * d / np.linalg.norm(d)
    cam, p = _persp(res, pos, width=float(rng.uniform(0.5, 1.2)))
    normal = cam.unit_vectors[2]
    bricks = []
    while len(bricks) < 5:
        le = rng.uniform(0.1, 0.8, 3)
        re = le + rng.uniform(0.05, 0.25, 3)
        corners = np.array([[x, y, z] for x in (le[0], re[0]) for y in (le[1], re[1])
                            for z in (le[2], re[2])])
        if np.all((corners - pos) @ normal > 0):  # in front of the camera
            bricks.append(dict(le=le, re=re, dims=np.array([4, 4, 4])))
    ext = oracle.extents_perspective(p["camera_data"], res, bricks)
    hits = 0
    for b, e in zip(bricks, ext):
        hv, hj = [], []
        for vi in range(res[0]):
            for vj in range(res[1]):
                if _chord(b["le"], b["re"], p["vp_pos"][vi, vj], p["vp_dir"][vi, vj]) > 1e-9:
                    hits += 1
                    hv.append(vi)
                    hj.append(vj)
                    assert e[0] <= vi < e[1] and e[2] <= vj < e[3], (vi, vj, e)
        if hv:
            # and tight: within the +-1 clip and one pixel of index mapping
            assert e[0] >= min(hv) - 3 and e[1] <= max(hv) + 4, (e, min(hv), max(hv))
            assert e[2] >= min(hj) - 3 and e[3] <= max(hj) + 4, (e, min(hj), max(hj))
    assert hits > 0


def test_per_pixel_rays_equal_plane_parallel_cast():
    ds = uniform_lognormal(24, seed=3, nprocs=2)
    bricks = oracle_proj_bricks(ds, ["density"])
    p = c0_params(20, normal=(0.4, -0.3, 0.87))
    ref, rv, rh = oracle.cast(0, p, bricks, record=True)
    nv0, nv1 = p["nv"]
    w = p["width"]
    pos = np.zeros((nv0, nv1, 3))
    dirs = np.zeros((nv0, nv1, 3))
    vp = np.asarray(p["vp_pos"]).ravel()
    for vi in range(nv0):
        for vj in range(nv1):
            px = w[0] * vi / (nv0 - 1.0) - w[0] / 2.0
            py = w[1] * vj / (nv1 - 1.0) - w[1] / 2.0
            for c in range(3):
                pos[vi, vj, c] = vp[c] * px + vp[3 + c] * py + vp[9 + c]
            dirs[vi, vj] = np.ravel(p["vp_dir"])
    q = dict(p, lens_type="null", vp_pos=pos, vp_dir=dirs)
    img, v, h = oracle.cast(0, q, bricks, record=True)
    assert np.array_equal(img, ref) and np.array_equal(v, rv) and np.array_equal(h, rh)


def test_perspective_sampler_params():
    cam, p = _persp((16, 8), (2.0, 0.4, 0.6), width=0.8)
    assert np.allclose(cam.width[1], cam.width[0] * 8 / 16)
    assert np.all(p["vp_pos"] == cam.position)
    max_length = np.linalg.norm(cam.position - 0.5) + 0.5 * np.sqrt(3.0)
    centre = 0.5 * (p["vp_dir"][7, 3] + p["vp_dir"][8, 4])
    assert np.allclose(centre / np.linalg.norm(centre), cam.unit_vectors[2])
    corner = p["vp_dir"][0, 0] / (max_length / cam.width[2])
    want = (-0.5 * cam.width[0] * cam.unit_vectors[0] - 0.5 * cam.width[1] * cam.unit_vectors[1]
            + cam.width[2] * cam.unit_vectors[2])
    assert np.allclose(corner, want)
    assert np.allclose(cam.lens.viewpoint, cam.lens.front_center)


def _cam(lens_type, res, position, focus):
    cam = Camera(lens_type=lens_type)
    cam.resolution = res
    cam.focus = focus
    cam.position = position
    return cam, cam._get_sampler_params()


def test_fisheye_directions():
    """lens.py:515-574 with arr_fisheye_vectors (grid_traversal.pyx:322-350):
    directions of length radius at angle r * fov / 2 from the view axis, zero
    outside r = 1.01."""
    cam, p = _cam("fisheye", (24, 24), (0.55, 0.45, 0.5), (0.2, 0.6, 0.4))
    d = p["vp_dir"]
    n = np.linalg.norm(d, axis=2)
    i = np.arange(24)
    px = 2.0 * i / 24 - 1.0
    r = np.sqrt(px[:, None] ** 2 + px[None, :] ** 2)
    assert np.all(n[r > 1.01] == 0.0)
    assert np.allclose(n[r <= 1.01], cam.width.max())
    cos_t = -d[..., 2] / np.where(n > 0, n, 1.0)  # the lens negates the vectors
    assert np.allclose(cos_t[r <= 1.01], np.cos(r[r <= 1.01] * np.pi / 2.0))
    assert np.all(p["vp_pos"] == cam.position)


def test_spherical_and_stereo_lenses():
    cam, p = _cam("spherical", (20, 10), (0.55, 0.45, 0.5), (0.2, 0.6, 0.4))
    ml = np.linalg.norm(cam.position - 0.5) + 0.5 * np.sqrt(3.0)
    assert np.allclose(np.linalg.norm(p["vp_dir"], axis=2), ml)
    cam, p = _cam("stereo-spherical", (20, 10), (0.55, 0.45, 0.5), (0.2, 0.6, 0.4))
    left, right = p["vp_pos"][:, :5], p["vp_pos"][:, 5:]
    assert np.allclose(np.linalg.norm(left - right, axis=2), 2 * cam.width[0] / 1000.0)
    assert np.array_equal(p["vp_dir"][:, :5], p["vp_dir"][:, 5:])
    cam, p = _cam("stereo-perspective", (16, 12), (1.8, 1.2, 0.9), (0.5, 0.5, 0.5))
    assert p["vp_dir"].shape == (16, 12, 3)
    # each eye sits `disparity` along its own east vector, rotated by
    # -+arctan(disparity / width[2]) about north (lens.py:368-426)
    eye = p["vp_pos"][:8, 0] - p["vp_pos"][8:, 0]
    disp = 3.0 * cam.width[0] / 16
    want = 2 * disp * np.cos(np.arctan2(disp, cam.width[2]))
    assert np.allclose(np.linalg.norm(eye, axis=1), want)


def test_camera_inside_cast_equals_brute_force_visits():
    """Per-pixel rays starting inside a brick (spherical lens at the domain
    centre): the oracle's visit count per ray equals the number of cells the
    exact segment crosses with positive length (slab test per cell)."""
    from yt_amd import data

    ds = data.load_uniform_grid({"density": np.ones((6, 6, 6))}, nprocs=1)
    cam, p = _cam("spherical", (8, 6), (0.41, 0.53, 0.47), (0.2, 0.6, 0.4))
    bricks = oracle_proj_bricks(ds, ["density"])
    img, v, h = oracle.cast(0, p, bricks, record=True)
    for vi in range(8):
        for vj in range(6):
            pos, d = p["vp_pos"][vi, vj], p["vp_dir"][vi, vj]
            _, idx, t = oracle.walk_log((0, 0, 0), (1, 1, 1), [6] * 3, pos, d)
            assert v[vi, vj] == len(idx)
            # a unit field integrates to the chord in t
            assert abs(img[vi, vj, 0] - _chord((0, 0, 0), (1, 1, 1), pos, d)) < 1e-12
