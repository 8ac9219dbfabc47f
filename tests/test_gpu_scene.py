"""Scene / VolumeSource / create_scene on the GPU path against the oracle:
the rendered image is the oracle's VolumeRenderSampler (or interpolated
projection) image with alpha set to 1 (render_source.py:522-537) and
rotated by 180 degrees (scene.py:625); a second transparent source starts
from the first one's image (scene.py:612-619)."""

import os

import numpy as np
import pytest

from oracle import oracle
from yt_amd import scene
from yt_amd.transfer_functions import gaussian_ctf

from .conftest import rel_l2
from .helpers import oracle_vr_bricks, small_amr, tf_dict, uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _setup(ds, sc):
    cam = sc.camera
    cam.resolution = (48, 40)
    cam.width = 0.9
    cam.position = (1.4, 1.1, 1.6)
    cam.focus = (0.5, 0.5, 0.5)
    p = cam._get_sampler_params(sc.sources["source_00"])
    p["viewpoint"] = np.array(cam.lens.viewpoint)
    return p


@pytest.mark.parametrize("lens", ["plane-parallel", "perspective"])
def test_create_scene_render(ctx, lens, tmp_path):
    ds = small_amr(root_n=32, levels=2, seed=141)
    tf = vr_tf_for(ds, nbins=128)
    sc = scene.create_scene(ds, "density", lens_type=lens, transfer_function=tf, ctx=ctx)
    src = sc.sources["source_00"]
    try:
        p = _setup(ds, sc)
        im = sc.render()
        order = src.volume.vr_order(p["viewpoint"])
        oimg = oracle.cast(1, p, oracle_vr_bricks(src.volume), order=order, tf=tf_dict(tf))
        oimg[:, :, 3] = 1.0
        want = np.rot90(oimg, k=2)
        assert im.shape == want.shape
        for c in range(4):
            assert rel_l2(im[..., c], want[..., c]) <= TOL, c
        fn = sc.save(str(tmp_path / "render.png"), sigma_clip=4.0, render_scene=False)
        with open(fn, "rb") as f:
            assert f.read(8) == b"\x89PNG\r\n\x1a\n"
        assert os.path.getsize(fn) > 100
    finally:
        src.free()


def test_two_transparent_sources_and_projection_source(ctx):
    ds = small_amr(root_n=32, levels=2, seed=143)
    tf1 = vr_tf_for(ds, nbins=64)
    rho = np.concatenate([np.ravel(g["density"]) for g in ds.grids])
    tf2 = gaussian_ctf((float(np.log10(rho.min())), float(np.log10(rho.max()))), n_gauss=3, nbins=64)
    sc = scene.Scene()
    s1 = scene.VolumeSource(ds, "density", transfer_function=tf1, ctx=ctx)
    s2 = scene.VolumeSource(ds, "density", transfer_function=tf2, ctx=ctx)
    sc.add_source(s1).add_source(s2)
    sc.add_camera(ds)
    try:
        p = _setup(ds, sc)
        im = sc.render()
        bricks = oracle_vr_bricks(s1.volume)
        order = s1.volume.vr_order(p["viewpoint"])
        first = oracle.cast(1, p, bricks, order=order, tf=tf_dict(tf1))
        first[:, :, 3] = 1.0
        second = oracle.cast(1, p, bricks, order=order, tf=tf_dict(tf2), image=first.copy())
        second[:, :, 3] = 1.0
        want = np.rot90(second, k=2)
        for c in range(3):
            assert rel_l2(im[..., c], want[..., c]) <= TOL, c
        # a "projection" source: InterpolatedProjectionSampler
        s3 = scene.VolumeSource(ds, "density", log_field=False, sampler_type="projection", ctx=ctx)
        pim = s3.render(sc.camera)
        oproj = oracle.cast(2, p, oracle_vr_bricks(s3.volume), order=s3.volume.vr_order(p["viewpoint"]))
        assert rel_l2(pim[..., 0], oproj[..., 0]) <= TOL
        s3.free()
    finally:
        s1.free()
        s2.free()


@pytest.mark.parametrize("lens", ["perspective", "stereo-perspective", "fisheye", "spherical",
                                  "stereo-spherical", "plane-parallel"])
def test_scene_lenses_grey_opacity(ctx, lens, tmp_path):
    """The reference's lens tests (volume_rendering/tests/test_lenses.py):
    a Scene with one grey-opacity volume source and a camera of each lens,
    at the positions those tests use, rendered and saved; here the image is
    also checked against the oracle (rel-L2 <= 1e-5, alpha included)."""
    ds = uniform_lognormal(32, seed=171, nprocs=8)
    tf = vr_tf_for(ds, nbins=128)
    tf.grey_opacity = True
    sc = scene.Scene()
    cam = sc.add_camera(ds, lens_type=lens)
    if lens in ("stereo-perspective", "spherical"):
        cam.resolution = (64, 32)
    else:
        cam.resolution = (48, 48)
    if lens == "fisheye":
        cam.lens.fov = 360.0
        cam.set_width(1.0)
        cam.set_position(np.array([0.45, 0.52, 0.48]))
    elif lens == "plane-parallel":
        cam.set_width(0.8)
    elif lens in ("spherical", "stereo-spherical"):
        cam.position = (0.6, 0.5, 0.5)
    elif lens == "stereo-perspective":
        cam.position = (0.7, 0.7, 0.7)
    else:
        cam.position = (1.0, 1.0, 1.0)
    src = scene.VolumeSource(ds, "density", transfer_function=tf, ctx=ctx)
    sc.add_source(src)
    try:
        im = sc.render()
        p = cam._get_sampler_params(src)
        p["viewpoint"] = np.array(cam.lens.viewpoint)
        order = src.volume.vr_order(p["viewpoint"])
        oimg = oracle.cast(1, p, oracle_vr_bricks(src.volume), order=order, tf=tf_dict(tf))
        want = np.rot90(oimg, k=2)
        assert im.shape == want.shape
        for c in range(4):
            e = rel_l2(im[..., c], want[..., c])
            print(f"scene {lens} ch{c}: rel-L2={e:.3e}", flush=True)
            assert e <= TOL
        fn = sc.save(str(tmp_path / f"test_{lens}.png"), sigma_clip=6.0, render_scene=False)
        assert os.path.getsize(fn) > 100
    finally:
        src.free()


def test_opaque_source_zbuffer(ctx):
    """An opaque source draws a blue plane at depth 0.45 into the scene's
    z-buffer (Scene.composite renders opaque sources first, scene.py:
    612-619); the transparent volume source then starts every ray from that
    plane's colour and stops it at max_t = clip(z, 0, 1) (utils.py:78-86,
    image_samplers.pyx:173-176): the oracle with that image and z-buffer."""

    class Plane(scene.OpaqueSource):
        def render(self, camera, zbuffer=None):
            zbuffer.rgba[:, :, 2] = 1.0
            zbuffer.rgba[:, :, 3] = 1.0
            zbuffer.z[:] = 0.45
            self.zbuffer = zbuffer
            return zbuffer.rgba

    ds = small_amr(root_n=32, levels=2, seed=177)
    tf = vr_tf_for(ds, nbins=96)
    sc = scene.Scene()
    cam = sc.add_camera(ds)
    cam.resolution = (40, 40)
    cam.set_width(1.2)
    src = scene.VolumeSource(ds, "density", transfer_function=tf, ctx=ctx)
    sc.add_source(src)
    sc.add_source(Plane())
    try:
        im = sc.render()
        p = cam._get_sampler_params(None)
        p["viewpoint"] = np.array(cam.lens.viewpoint)
        start = np.zeros((40, 40, 4))
        start[:, :, 2] = 1.0
        start[:, :, 3] = 1.0
        oimg = oracle.cast(1, p, oracle_vr_bricks(src.volume), order=src.volume.vr_order(p["viewpoint"]),
                           tf=tf_dict(tf), image=start, zbuffer=np.full((40, 40), 0.45))
        oimg[:, :, 3] = 1.0
        want = np.rot90(oimg, k=2)
        for c in range(3):
            e = rel_l2(im[..., c], want[..., c])
            print(f"opaque plane + VR ch{c}: rel-L2={e:.3e}", flush=True)
            assert e <= TOL
        # the plane's blue stays (non-grey compositing of an opaque channel)
        # and the volume in front of it adds red
        assert np.allclose(im[..., 2], 1.0) and im[..., 0].max() > 0.0
    finally:
        src.free()
