"""Transfer-function table bytes per persistent launch (YTRC_BATCH_TF_LDS_*).

Every workgroup of a persistent launch stages the distinct transfer-function
tables of the launch's requests in LDS.  With a cap, a kernel variant's
requests are split over launches by table.  Six volume renders over three
distinct tables (about 12 KB each) and one projection: the launch plans
follow the cap, and every cap renders the same images, bit for bit, as one
launch per variant and as one plain launch per request.
"""

import numpy as np
import pytest

from yt_amd import _lib, render

from .helpers import c0_params, uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu

NORMALS = [(0.3, 0.5, 0.81), (-0.6, 0.2, 0.77), (0.1, -0.9, 0.42), (0.7, 0.7, -0.14),
           (-0.2, -0.3, 0.93), (0.5, -0.5, 0.71)]


def _images(ctx, specs, flags, **kw):
    b = _lib.Batch(ctx, specs, flags, **kw)
    try:
        b.run()
        return [b.read_image(i) for i in range(len(specs))], b.launch_plan()
    finally:
        b.close()


def test_tf_lds_cap_splits_launches_and_keeps_images(ctx):
    ds = uniform_lognormal(48, seed=91, nprocs=4)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    try:
        tfs = [vr_tf_for(ds, nbins=n) for n in (256, 250, 240)]
        specs = [vol.vr_request(c0_params(64 + 8 * k, normal=NORMALS[k]), tfs[k % 3]) for k in range(6)]
        specs.append(vol.projection_request(c0_params(72)))
        base, plain_plan = _images(ctx, specs, _lib.BATCH_PLAIN)
        assert plain_plan == (0, 0)
        plans = {}
        for tl in ("all", "32k", "16k", "one"):
            imgs, plan = _images(ctx, specs, _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH, tf_lds=tl)
            plans[tl] = plan
            for i in range(len(specs)):
                assert np.array_equal(imgs[i], base[i]), (tl, i)
        print(f"launch plans (launches, largest table bytes): {plans}", flush=True)
        one_table = plans["one"][1]
        assert 11000 < one_table < 16384
        assert plans["all"] == (2, plans["all"][1]) and plans["all"][1] > 2 * one_table
        assert plans["32k"][0] == 3 and 2 * 11000 < plans["32k"][1] <= 32768
        assert plans["16k"] == (4, one_table)
        assert plans["one"][0] == 4
    finally:
        vol.free()


def test_tf_lds_flag_rejects_unknown():
    with pytest.raises(ValueError):
        _lib.Batch(None, [], _lib.BATCH_PERSISTENT, tf_lds="8k")


def test_vertex_blocks_in_lds_keep_images(ctx):
    """YTRC_BATCH_VERTEX_LDS: volume renders with a ColorTransferFunction,
    one field and 10 sub-samples read their corners from a per-wave LDS copy
    of the vertex block where all lanes of a wave are in one block.  The
    values are the same, so the images equal those of the L2 path bit for
    bit: a uniform field in 8 bricks at image sides from 48 (rays several
    cells apart) to 320 (rays closer than cells), and an AMR set whose child
    masks skip cells; every launch shape the flag combines with."""
    from .helpers import small_amr

    for ds, sides in ((uniform_lognormal(64, seed=93, nprocs=8), (48, 160, 320)),
                      (small_amr(root_n=32, levels=2, seed=94), (96, 256))):
        vol = render.ResidentVolume(ctx, ds, vr_field="density")
        try:
            tf = vr_tf_for(ds, nbins=256)
            specs = [vol.vr_request(c0_params(n, normal=NORMALS[k]), tf) for k, n in enumerate(sides)]
            assert all(s.n_samples == 10 for s in specs)
            flags = _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH
            base, _ = _images(ctx, specs, flags)
            for kw in ({}, {"tile": (4, 16)}, {"item_tiles": 4}):
                for wpg in (1, 4):
                    b = _lib.Batch(ctx, specs, flags, vertex_lds=True, **kw)
                    try:
                        b.set_launch(wpg, 0)
                        b.run()
                        b.run()
                        for i in range(len(specs)):
                            assert np.array_equal(b.read_image(i), base[i]), (kw, wpg, i)
                    finally:
                        b.close()
        finally:
            vol.free()
