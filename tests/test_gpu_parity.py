"""GPU parity: the HIP ray-caster (through the C ABI) against the FP64 oracle.

Bars (written into each test):
  * cell-visit sequences: per-ray visit count and order-sensitive hash of
    (brick, i, j, k, bits(enter_t), bits(exit_t)) bit-exact;
  * projection images: relative L2 <= 1e-5 and max-abs/max <= 1e-5 vs FP64;
  * volume-render images: relative L2 <= 1e-5 vs FP64.
"""

import json
import sys

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, render
from yt_amd.transfer_functions import ColorTransferFunction

from .conftest import rel_l2
from .helpers import (
    c0_params,
    oracle_proj_bricks,
    oracle_vr_bricks,
    small_amr,
    tf_dict,
    uniform_lognormal,
    vr_tf_for,
)

pytestmark = pytest.mark.gpu

PROJ_TOL = 1e-5
VR_TOL = 1e-5


def _run(ctx, specs, flags=_lib.BATCH_PLAIN):
    b = _lib.Batch(ctx, specs, flags)
    try:
        b.run()
        out = []
        for i, s in enumerate(specs):
            img = b.read_image(i)
            rec = b.read_record(i) if s.record else None
            out.append((img, rec))
        return out
    finally:
        b.close()


def _check_visits(rec, ovis, ohash, tag, osamp=None):
    v, h, s = rec
    nd = int(np.sum(v != ovis))
    nh = int(np.sum(h != ohash))
    ns = 0 if osamp is None else int(np.sum(s != osamp))
    print(f"{tag}: total visits gpu={int(v.sum())} oracle={int(ovis.sum())} "
          f"count-mismatch={nd} hash-mismatch={nh} sample-mismatch={ns}", flush=True)
    assert nd == 0 and nh == 0 and ns == 0


@pytest.mark.parametrize("res", [64, 256])
def test_c0_projection(ctx, res):
    """64^3 lognormal, normal (0.3,0.5,0.81), unweighted projection."""
    ds = uniform_lognormal(64)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density")
    try:
        p = c0_params(res)
        (img, rec), = _run(ctx, [vol.projection_request(p, record=True)])
        oimg, ovis, ohash, osamp = oracle.cast(0, p, oracle_proj_bricks(ds, ["density"]),
                                               record=True, with_samples=True)
        _check_visits(rec, ovis, ohash, f"c0 res={res}", osamp)
        e = rel_l2(img[..., 0], oimg[..., 0])
        m = np.abs(img[..., 0] - oimg[..., 0]).max() / np.abs(oimg[..., 0]).max()
        print(f"c0 res={res}: rel-L2={e:.3e} max-abs/max={m:.3e}", flush=True)
        assert e <= PROJ_TOL and m <= PROJ_TOL
    finally:
        vol.free()


def test_weighted_projection_multibrick(ctx):
    """Weighted projection over an 8-grid uniform split (load_uniform_grid nprocs=8)."""
    ds = uniform_lognormal(64, nprocs=8, with_temperature=True)
    assert len(ds.grids) == 8
    vol = render.ResidentVolume(ctx, ds, proj_fields="temperature", weight="density")
    try:
        p = c0_params(200, normal=(-0.41, 0.27, 0.6))
        (img, rec), = _run(ctx, [vol.projection_request(p, record=True)])
        oimg, ovis, ohash = oracle.cast(
            0, p, oracle_proj_bricks(ds, ["temperature"], weight="density"), record=True)
        _check_visits(rec, ovis, ohash, "weighted 8 grids")
        for c in (0, 1):
            e = rel_l2(img[..., c], oimg[..., c])
            print(f"weighted ch{c}: rel-L2={e:.3e}", flush=True)
            assert e <= PROJ_TOL
        g = render.finalize_projection(img, p["width"], True)
        og = render.finalize_projection(oimg, p["width"], True)
        assert rel_l2(g, og) <= PROJ_TOL
    finally:
        vol.free()


def test_axis_aligned_and_zbuffer(ctx):
    """Axis-aligned normal (zero v_dir components) and a per-pixel max_t."""
    ds = uniform_lognormal(32, seed=11)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density")
    try:
        p = c0_params(96, normal=(0.0, 0.0, 1.0))
        rng = np.random.default_rng(5)
        zb = rng.uniform(-0.2, 1.3, size=(96, 96))
        spec = vol.projection_request(p, record=True, zbuffer=zb)
        (img, rec), = _run(ctx, [spec])
        oimg, ovis, ohash = oracle.cast(0, p, oracle_proj_bricks(ds, ["density"]), zbuffer=zb,
                                        record=True)
        _check_visits(rec, ovis, ohash, "axis-aligned+zbuffer")
        assert rel_l2(img[..., 0], oimg[..., 0]) <= PROJ_TOL
    finally:
        vol.free()


def test_vr_amr(ctx):
    """VolumeRenderSampler over kd bricks of a 2-level AMR set, 8-Gaussian CTF."""
    ds = small_amr()
    tf = vr_tf_for(ds)
    vol = render.ResidentVolume(ctx, ds, vr_field="density", log_field=True)
    try:
        p = c0_params(128, normal=(0.2, -0.7, 0.5))
        spec = vol.vr_request(p, tf, n_samples=10, record=True)
        (img, rec), = _run(ctx, [spec])
        bricks = oracle_vr_bricks(vol)
        oimg, ovis, ohash, osamp = oracle.cast(1, p, bricks, order=spec.order, tf=tf_dict(tf),
                                               n_samples=10, record=True, with_samples=True)
        _check_visits(rec, ovis, ohash, f"vr amr ({len(bricks)} bricks)", osamp)
        for c in range(3):
            e = rel_l2(img[..., c], oimg[..., c])
            print(f"vr ch{c}: rel-L2={e:.3e} max={oimg[..., c].max():.3e}", flush=True)
            assert e <= VR_TOL
        assert np.array_equal(img[..., 3], oimg[..., 3])
    finally:
        vol.free()


def test_vr_grey_opacity(ctx):
    ds = uniform_lognormal(24, seed=3)
    lo, hi = np.log10(ds.grids[0]["density"].min()), np.log10(ds.grids[0]["density"].max())
    tf = ColorTransferFunction((lo, hi), nbins=64, grey_opacity=True)
    tf.add_gaussian(0.5 * (lo + hi), 0.05 * (hi - lo) ** 2, [0.8, 0.4, 0.2, 3.0])
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        p = c0_params(64, normal=(0.6, 0.3, -0.2))
        spec = vol.vr_request(p, tf, n_samples=7, record=True)
        (img, rec), = _run(ctx, [spec])
        oimg, ovis, ohash, osamp = oracle.cast(1, p, oracle_vr_bricks(vol), order=spec.order,
                                               tf=tf_dict(tf), n_samples=7, record=True,
                                               with_samples=True)
        _check_visits(rec, ovis, ohash, "vr grey", osamp)
        for c in range(4):
            assert rel_l2(img[..., c], oimg[..., c]) <= VR_TOL
    finally:
        vol.free()


def test_vr_zero_and_nan_vertex_values(ctx):
    """FIT_get_value returns 0 for values that are not normal
    (field_interpolation_tables.pxd:74): a block of density exactly 1
    (log10 -> 0, inside the table range) and cells <= 0 (log10 -> NaN)
    contribute nothing, in the ColorTransferFunction path (10 sub-samples)."""
    ds = uniform_lognormal(32, seed=29)
    rho = ds.grids[0]["density"]
    rho[4:14, 6:20, 3:17] = 1.0
    rho[20:23, 20:24, 8:11] = -1.0
    tf = ColorTransferFunction((-1.0, 1.0), nbins=128)
    for j, x in enumerate(np.linspace(-0.9, 0.9, 6)):
        tf.add_gaussian(x, 0.02, [0.9, 0.3 + 0.1 * j, 0.2, 4.0])
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        p = c0_params(96, normal=(0.25, 0.6, -0.4))
        spec = vol.vr_request(p, tf, n_samples=10, record=True)
        (img, rec), = _run(ctx, [spec])
        bricks = oracle_vr_bricks(vol)
        assert np.isnan(bricks[0]["data"][0]).any()
        assert (bricks[0]["data"][0] == 0.0).sum() > 1000
        oimg, ovis, ohash = oracle.cast(1, p, bricks, order=spec.order, tf=tf_dict(tf),
                                        n_samples=10, record=True)
        _check_visits(rec, ovis, ohash, "vr zero/NaN values")
        for c in range(3):
            assert np.all(np.isfinite(img[..., c]))
            e = rel_l2(img[..., c], oimg[..., c])
            print(f"vr zero/NaN ch{c}: rel-L2={e:.3e}", flush=True)
            assert e <= VR_TOL
    finally:
        vol.free()


def test_batch_many_images(ctx):
    """Several requests of different sizes and kinds in one batch give the
    same images as each run alone."""
    ds = uniform_lognormal(32, seed=21)
    tf = vr_tf_for(ds, nbins=128)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        rng = np.random.default_rng(1004)
        specs = []
        for i in range(6):
            n = rng.normal(size=3)
            res = int(rng.choice([16, 40, 64, 100]))
            p = c0_params(res, normal=tuple(n / np.linalg.norm(n)))
            specs.append(vol.projection_request(p) if i % 2 == 0 else vol.vr_request(p, tf))
        together = _run(ctx, specs)
        for i, s in enumerate(specs):
            (alone, _), = _run(ctx, [s])
            assert np.array_equal(alone, together[i][0]), i
        # The persistent tile queue and its hipGraph replay give the same
        # images as one launch per request, on every replay.
        pers = _run(ctx, specs, _lib.BATCH_PERSISTENT)
        for i in range(len(specs)):
            assert np.array_equal(pers[i][0], together[i][0]), i
        b = _lib.Batch(ctx, specs, _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH)
        try:
            for rep in range(3):
                b.run()
                for i in range(len(specs)):
                    assert np.array_equal(b.read_image(i), together[i][0]), (rep, i)
        finally:
            b.close()
    finally:
        vol.free()


def test_persistent_record_matches_oracle(ctx):
    """Record mode through the persistent kernel: visit sequences of a mixed
    batch equal the oracle's."""
    ds = uniform_lognormal(32, seed=31, nprocs=4)
    tf = vr_tf_for(ds, nbins=64)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        pp = c0_params(72, normal=(0.1, 0.9, -0.4))
        pv = c0_params(56, normal=(-0.5, 0.2, 0.8))
        specs = [vol.projection_request(pp, record=True), vol.vr_request(pv, tf, record=True)]
        out = _run(ctx, specs, _lib.BATCH_PERSISTENT)
        _, ov, oh = oracle.cast(0, pp, oracle_proj_bricks(ds, ["density"]), record=True)
        _check_visits(out[0][1], ov, oh, "persistent proj")
        _, ov, oh = oracle.cast(1, pv, oracle_vr_bricks(vol), order=specs[1].order,
                                tf=tf_dict(tf), record=True)
        _check_visits(out[1][1], ov, oh, "persistent vr")
    finally:
        vol.free()


def test_drop_in_off_axis_projection():
    """yt_amd.off_axis_projection (ImageSampler protocol) against the oracle."""
    import yt_amd

    ds = uniform_lognormal(32, seed=5)
    img = yt_amd.off_axis_projection(ds, (0.5, 0.5, 0.5), (0.3, 0.5, 0.81), 1.0, 64, "density")
    p = c0_params(64)
    oimg = oracle.cast(0, p, oracle_proj_bricks(ds, ["density"]))
    ref = render.finalize_projection(oimg, p["width"], False)
    assert img.shape == (64, 64)
    assert rel_l2(img, ref) <= PROJ_TOL


def test_sqrt3_chord(ctx):
    """Closed form (test_offaxisprojection_pytestonly.py:518-552): projecting
    ones along (1,1,1) through the unit cube peaks at sqrt(3)."""
    ds = uniform_lognormal(16, seed=1)
    ds.grids[0].fields["ones"] = np.ones((16, 16, 16))
    vol = render.ResidentVolume(ctx, ds, proj_fields="ones")
    try:
        p = c0_params(65, normal=(1.0, 1.0, 1.0), width=np.sqrt(3.0))
        (img, _), = _run(ctx, [vol.projection_request(p)])
        out = render.finalize_projection(img, p["width"], False)
        assert abs(out.max() - np.sqrt(3.0)) < 1e-2
    finally:
        vol.free()


def test_early_ray_termination(ctx):
    """Front-to-back compositing with early termination (YTRC_FLAG_ERT):
    with an opaque transfer function over 4x4x4 bricks, waves stop early
    (fewer recorded visits than the full walk) and every pixel stays within
    2e-4 of the FP64 oracle: a lane stops only once its transmittance is
    below 1e-4, which bounds what is left out."""
    ds = uniform_lognormal(64, seed=51, psize=(4, 4, 4))
    rho = np.concatenate([g["density"].ravel() for g in ds.grids])
    lo, hi = np.log10(rho.min()), np.log10(rho.max())
    tf = ColorTransferFunction((lo, hi), nbins=128)
    tf.add_gaussian(0.5 * (lo + hi), 0.5 * (hi - lo) ** 2, [0.9, 0.6, 0.3, 1.0])
    for t in tf.funcs:
        t.y[:] = np.maximum(t.y, 0.0) * 40.0
    vol = render.ResidentVolume(ctx, ds, vr_field="density")
    try:
        p = c0_params(96, normal=(0.3, 0.5, 0.81))
        full = vol.vr_request(p, tf, record=True)
        ert = vol.vr_request(p, tf, record=True, ert=True)
        (img_f, rec_f), (img_e, rec_e) = _run(ctx, [full, ert])
        oimg, ov, oh = oracle.cast(1, p, oracle_vr_bricks(vol), order=full.order, tf=tf_dict(tf),
                                   record=True)
        _check_visits(rec_f, ov, oh, "ERT reference walk")
        assert rec_e[0].sum() < 0.8 * rec_f[0].sum()
        err = np.abs(img_e[..., :3] - oimg[..., :3]).max()
        print(f"ERT: visits {int(rec_e[0].sum())} of {int(rec_f[0].sum())}, max-abs={err:.2e}",
              flush=True)
        assert err <= 2e-4
        assert rel_l2(img_f[..., :3], oimg[..., :3]) <= VR_TOL
    finally:
        vol.free()


@pytest.mark.parametrize("tile", [(4, 16), (16, 4)])
def test_tile_shapes(ctx, tile):
    """Every tile shape casts the same rays with the same per-ray code: images
    and recorded visit sequences equal the 8x8 tiling's exactly, for an image
    size that leaves ragged tiles at both edges."""
    ds = uniform_lognormal(32, seed=61, nprocs=4)
    tf = vr_tf_for(ds, nbins=64)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        specs = [vol.projection_request(c0_params(70, normal=(0.2, -0.7, 0.6)), record=True),
                 vol.vr_request(c0_params(37, normal=(0.6, 0.3, -0.74)), tf, record=True)]
        base = _run(ctx, specs)
        for flags in (_lib.BATCH_PLAIN, _lib.BATCH_PERSISTENT):
            b = _lib.Batch(ctx, specs, flags, tile=tile)
            try:
                b.run()
                for i in range(len(specs)):
                    assert np.array_equal(b.read_image(i), base[i][0]), (flags, i)
                    for got, want in zip(b.read_record(i), base[i][1]):
                        assert np.array_equal(got, want), (flags, i)
            finally:
                b.close()
    finally:
        vol.free()


def test_autotune_picks_a_working_shape(ctx, tmp_path):
    """The autotuner measures candidates (a full grid over a subset of the
    knobs, and the default coordinate descent), caches the fastest, returns
    the cached entry unmeasured the second time, and every launch shape it
    can choose -- tile shape, rays per work item (1-8 tiles), waves per
    workgroup, workgroups per CU, table bytes per launch, vertex blocks in
    LDS -- renders the same images as the default."""
    from yt_amd import autotune

    ds = uniform_lognormal(32, seed=71)
    tf = vr_tf_for(ds, nbins=64)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        specs = [vol.projection_request(c0_params(96, normal=(0.3, 0.5, 0.81))),
                 vol.vr_request(c0_params(80, normal=(-0.4, 0.1, 0.9)), tf)]
        path = str(tmp_path / "tune.json")
        best, table = autotune.tune(ctx, specs, item_tiles=(1, 4), waves_per_group=(1, 4),
                                    groups_per_cu=(0, 2), tf_lds=("all", "one"), reps=1, path=path,
                                    search="grid")
        assert len(table) == 3 * 2 * 2 * 2 * 2 * 2 and all(ms > 0 for _, ms in table)
        # one table per variant: both table-size values give the same
        # launches and share one measurement
        by = {json.dumps(c, sort_keys=True): ms for c, ms in table}
        for c, ms in table:
            assert by[json.dumps(dict(c, tf_lds="all"), sort_keys=True)] == ms
        again, table2 = autotune.tune(ctx, specs, path=path)
        assert again == best and table2 == []
        # the cache key comes from the library (ytrc_ctx_device_name): no
        # second HIP runtime (torch's) is loaded into the process
        assert "torch" not in sys.modules
        assert "gfx" in ctx.device_name
        d_best, d_table = autotune.tune(ctx, specs, reps=1, use_cache=False)
        assert 0 < len(d_table) < 3 * 4 * 4 * 2 * 3 * 5
        base = _run(ctx, specs)
        for cfg in [best, d_best, dict(autotune.DEFAULT, item_tiles=2),
                    dict(autotune.DEFAULT, item_tiles=8, tile=[4, 16], waves_per_group=1),
                    dict(autotune.DEFAULT, vertex_lds=True, waves_per_group=2)]:
            b = autotune.make_batch(ctx, specs, cfg)
            try:
                b.run()
                for i in range(len(specs)):
                    assert np.array_equal(b.read_image(i), base[i][0]), (cfg, i)
            finally:
                b.close()
    finally:
        vol.free()


def test_vr_child_masks_below_max_level(ctx):
    """AMRKDTree(ds, max_level=...) leaves finer grids out, so its bricks'
    child masks are not all ones (amr_kdtree.py:346-367): VolumeRenderSampler
    skips mask-0 cells (image_samplers.pyx:467-470).  Visits, per-pixel
    sample counts and the image against the oracle."""
    from yt_amd.amr_kdtree import AMRKDTree

    ds = small_amr(root_n=32, levels=2, seed=61)
    assert ds.max_level == 2
    kd = AMRKDTree(ds, max_level=1)
    kd.set_fields(["density"], [True], no_ghost=True)
    leaves = kd.leaves()
    pgs = [kd.get_brick_data(n) for n in leaves]
    assert any(not pg.source_mask.all() for pg in pgs)
    tf = vr_tf_for(ds, nbins=128)
    ds_id = ctx.upload([(pg.LeftEdge, pg.RightEdge, pg.dims, True) for pg in pgs],
                       [pg.my_data for pg in pgs], [pg.source_mask for pg in pgs])
    try:
        p = c0_params(96, normal=(-0.5, 0.4, 0.77))
        order = kd.order(p["viewpoint"])
        spec = _lib.RequestSpec(_lib.VOLUME_RENDER, p, ds_id, order, n_samples=10,
                                tf=_lib.TF.from_tf_obj(tf), record=True)
        (img, rec), = _run(ctx, [spec], _lib.BATCH_PERSISTENT)
        bricks = [dict(le=pg.LeftEdge, re=pg.RightEdge, dims=pg.dims,
                       data=[np.asarray(d, dtype=np.float64) for d in pg.my_data],
                       mask=pg.source_mask) for pg in pgs]
        oimg, ovis, ohash, osamp = oracle.cast(1, p, bricks, order=order, tf=tf_dict(tf),
                                               n_samples=10, record=True, with_samples=True)
        _check_visits(rec, ovis, ohash, "vr masked bricks", osamp)
        assert np.any(osamp < 10 * ovis)
        for c in range(3):
            e = rel_l2(img[..., c], oimg[..., c])
            print(f"vr masked ch{c}: rel-L2={e:.3e}", flush=True)
            assert e <= VR_TOL
    finally:
        ctx.free(ds_id)
