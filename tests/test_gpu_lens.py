"""GPU parity of the per-pixel lens path (YTRC_LENS_PER_PIXEL: the
perspective lens and any lens that hands the sampler per-pixel origins and
directions, generate_vector_info_null, lenses.pyx:196-207) against the FP64
oracle: visit sequences bit-exact, images rel-L2 <= 1e-5."""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, render
from yt_amd.image_samplers import ProjectionSampler

from .conftest import rel_l2
from .helpers import c0_params, oracle_proj_bricks, oracle_vr_bricks, small_amr, tf_dict, \
    uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run(ctx, specs, flags):
    b = _lib.Batch(ctx, specs, flags)
    try:
        b.run()
        return [(b.read_image(i), b.read_record(i) if s.record else None)
                for i, s in enumerate(specs)]
    finally:
        b.close()


def _same_visits(rec, ov, oh, tag):
    v, h, _ = rec
    nd, nh = int(np.sum(v != ov)), int(np.sum(h != oh))
    print(f"{tag}: visits {int(v.sum())}, count-mismatch={nd} hash-mismatch={nh}", flush=True)
    assert nd == 0 and nh == 0


@pytest.mark.parametrize("flags", [_lib.BATCH_PLAIN, _lib.BATCH_PERSISTENT])
def test_perspective_projection(ctx, flags):
    ds = uniform_lognormal(32, seed=91, nprocs=4)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density")
    try:
        p = render.perspective_params((1.6, 1.2, 0.9), (0.5, 0.5, 0.5), 0.7, (40, 28))
        (img, rec), = _run(ctx, [vol.projection_request(p, record=True)], flags)
        oimg, ov, oh = oracle.cast(0, p, oracle_proj_bricks(ds, ["density"]), record=True)
        _same_visits(rec, ov, oh, "perspective projection")
        assert int(ov.sum()) > 0
        e = rel_l2(img[..., 0], oimg[..., 0])
        print(f"perspective projection rel-L2={e:.3e}", flush=True)
        assert e <= TOL
    finally:
        vol.free()


def test_perspective_vr_amr(ctx):
    ds = small_amr(root_n=32, levels=2, seed=93)
    tf = vr_tf_for(ds, nbins=128)
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        p = render.perspective_params((-0.9, 1.4, 1.7), (0.5, 0.5, 0.5), 0.9, (36, 44))
        spec = vol.vr_request(p, tf, record=True)
        (img, rec), = _run(ctx, [spec], _lib.BATCH_PERSISTENT)
        oimg, ov, oh = oracle.cast(1, p, oracle_vr_bricks(vol), order=spec.order, tf=tf_dict(tf),
                                   record=True)
        _same_visits(rec, ov, oh, "perspective VR")
        for c in range(3):
            e = rel_l2(img[..., c], oimg[..., c])
            print(f"perspective VR ch{c} rel-L2={e:.3e}", flush=True)
            assert e <= TOL
    finally:
        vol.free()


def test_per_pixel_rays_without_camera_data(ctx):
    """Rays fanning out from one point in every direction (as a fisheye or
    spherical lens produces), some axis-parallel: whole-image extents."""
    ds = uniform_lognormal(24, seed=95, nprocs=2)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density")
    try:
        nv = (30, 34)
        rng = np.random.default_rng(97)
        dirs = rng.normal(size=nv + (3,))
        dirs[0, :, 1:] = 0.0  # axis-parallel rays
        dirs[1, :, 0] = 0.0
        dirs *= 2.0 / np.linalg.norm(dirs, axis=2, keepdims=True)
        pos = np.broadcast_to(np.array([0.45, 0.52, 0.61]), nv + (3,)).copy()
        p = dict(c0_params(8), nv=nv, lens_type="fisheye", vp_pos=pos, vp_dir=dirs,
                 camera_data=None)
        (img, rec), = _run(ctx, [vol.projection_request(p, record=True)], _lib.BATCH_PERSISTENT)
        oimg, ov, oh = oracle.cast(0, p, oracle_proj_bricks(ds, ["density"]), record=True)
        _same_visits(rec, ov, oh, "per-pixel rays")
        assert rel_l2(img[..., 0], oimg[..., 0]) <= TOL
    finally:
        vol.free()


def test_mixed_lenses_in_one_batch(ctx):
    ds = uniform_lognormal(32, seed=99)
    tf = vr_tf_for(ds, nbins=64)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        pp = c0_params(40, normal=(0.3, -0.2, 0.93))
        pv = render.perspective_params((1.8, 0.3, 1.1), (0.5, 0.5, 0.5), 0.8, (33, 40))
        specs = [vol.projection_request(pp), vol.projection_request(pv), vol.vr_request(pv, tf),
                 vol.vr_request(pp, tf)]
        alone = [_run(ctx, [s], _lib.BATCH_PLAIN)[0][0] for s in specs]
        mixed = _run(ctx, specs, _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH)
        for i in range(len(specs)):
            assert np.array_equal(mixed[i][0], alone[i]), i
    finally:
        vol.free()


def test_drop_in_sampler_perspective():
    """ProjectionSampler(..., lens_type='perspective', camera_data=...) as
    utils.py:115-139 builds it, through the recorded-call protocol."""
    ds = uniform_lognormal(24, seed=101, nprocs=2)
    p = render.perspective_params((1.5, 1.5, 1.2), (0.5, 0.5, 0.5), 0.8, (26, 26))
    image = np.zeros((26, 26, 4))
    s = ProjectionSampler(np.atleast_3d(p["vp_pos"]), np.atleast_3d(p["vp_dir"]), p["center"],
                          p["bounds"], image, p["x_vec"], p["y_vec"], p["width"], "KDTree", 1,
                          lens_type="perspective", camera_data=p["camera_data"])
    from yt_amd.partitioned_grid import PartitionedGrid

    for i, (g, m) in enumerate(ds.blocks()):
        s(PartitionedGrid(g.id, [np.asarray(g["density"], dtype="float64") * m], m.astype("uint8"),
                          g.LeftEdge, g.RightEdge, g.ActiveDimensions))
    out = s.aimage
    oimg = oracle.cast(0, p, oracle_proj_bricks(ds, ["density"]))
    assert out is image
    assert rel_l2(out[..., 0], oimg[..., 0]) <= TOL


@pytest.mark.parametrize("lens", ["fisheye", "spherical", "stereo-spherical", "stereo-perspective"])
def test_other_lenses_vr(ctx, lens):
    """Cameras inside the AMR volume (rays start inside bricks; the fisheye
    has zero-length directions outside its disc) and the stereo
    perspective: volume renders and their visit sequences equal the oracle's."""
    ds = small_amr(root_n=32, levels=2, seed=131)
    tf = vr_tf_for(ds, nbins=64)
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        inside = lens != "stereo-perspective"
        pos = (0.47, 0.52, 0.44) if inside else (1.7, 1.1, 1.3)
        res = (32, 32) if lens == "fisheye" else (40, 24)
        p = render.lens_params(lens, pos, (0.2, 0.7, 0.3) if inside else (0.5, 0.5, 0.5), 0.9, res)
        spec = vol.vr_request(p, tf, record=True)
        (img, rec), = _run(ctx, [spec], _lib.BATCH_PERSISTENT)
        oimg, ov, oh = oracle.cast(1, p, oracle_vr_bricks(vol), order=spec.order, tf=tf_dict(tf),
                                   record=True)
        _same_visits(rec, ov, oh, f"{lens} VR")
        for c in range(3):
            e = rel_l2(img[..., c], oimg[..., c])
            print(f"{lens} VR ch{c} rel-L2={e:.3e}", flush=True)
            assert e <= TOL
    finally:
        vol.free()
