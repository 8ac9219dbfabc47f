"""Stopping a launch and bounding waits: the analog of the reference's
PyErr_CheckSignals every 1000 rays (image_samplers.pyx:188-190), which lets
a user interrupt a long cast.

* Batch.cancel() sets the batch's host-pinned stop flag; every wave of the
  running launch reads it once per tile and exits at its next tile, so the
  kernels end normally and early; sync() then raises YtrcCancelled and the
  images stay unreadable until a later launch completes, which renders
  them exactly as before.
* sync(timeout) / Server.wait(timeout) raise YtrcTimeout when the deadline
  passes first; the work keeps running and a later wait returns it.
* The drop-in samplers wait in slices and, on KeyboardInterrupt, stop the
  launch before re-raising; the caller's image is left as it was.
The stop is requested from the host; no kernel faults, traps or spins.
"""

import numpy as np
import pytest

from yt_amd import _lib, render
from yt_amd.image_samplers import VolumeRenderSampler

from .helpers import c0_params, uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu

NORMALS = [(0.3, 0.5, 0.81), (-0.6, 0.2, 0.77), (0.1, -0.9, 0.42), (0.7, 0.7, -0.14)]


@pytest.fixture(scope="module")
def big(ctx):
    ds = uniform_lognormal(128, seed=401, nprocs=8)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    tf = vr_tf_for(ds, nbins=256)
    specs = [vol.vr_request(c0_params(1024, normal=n), tf) for n in NORMALS]
    yield vol, specs, ds, tf
    vol.free()


def test_cancel_stops_a_persistent_launch(ctx, big):
    _, specs, _, _ = big
    b = _lib.Batch(ctx, specs, _lib.BATCH_PERSISTENT)
    try:
        full_ms = b.run()
        want = [b.read_image(i) for i in range(len(specs))]
        b.launch()
        b.cancel()
        with pytest.raises(_lib.YtrcCancelled):
            b.sync()
        with pytest.raises(_lib.YtrcCancelled):
            b.read_image(0)
        stopped_ms = b.kernel_ms()  # cancelled launches are not counted
        assert stopped_ms[1] == 1
        # a later launch renders the same images
        again_ms = b.run()
        print(f"full launch {full_ms:.2f} ms, again {again_ms:.2f} ms", flush=True)
        for i in range(len(specs)):
            assert np.array_equal(b.read_image(i), want[i]), i
    finally:
        b.close()


def test_cancel_stops_concurrent_launches(ctx, big):
    """A batch of volume renders and projections runs its two launches
    concurrently on two streams; a stop request ends both, sync() reports
    it, and a later launch renders the same images."""
    vol, specs, _, _ = big
    mixed = []
    for i, s in enumerate(specs):
        mixed.append(s)
        mixed.append(vol.projection_request(c0_params(1024, normal=NORMALS[i])))
    b = _lib.Batch(ctx, mixed, _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH)
    try:
        assert b.launch_plan()[0] == 2
        b.run()
        want = [b.read_image(i) for i in range(len(mixed))]
        b.launch()
        b.cancel()
        with pytest.raises(_lib.YtrcCancelled):
            b.sync()
        b.run()
        for i in range(len(mixed)):
            assert np.array_equal(b.read_image(i), want[i]), i
    finally:
        b.close()


def test_cancelled_launch_ends_early(ctx, big):
    """The stop request reaches the waves: a launch stopped at once takes a
    fraction of the GPU time of a full one (32 1024^2 volume renders; every
    wave reads the flag at one of every 16 items it takes, so the stopped
    launch's time does not grow with the batch)."""
    _, specs, _, _ = big
    b = _lib.Batch(ctx, specs * 8, _lib.BATCH_PERSISTENT)
    try:
        full_ms = b.run()
        b.launch()
        b.cancel()
        ms = _lib.C.c_float()
        rc = _lib.lib().ytrc_batch_sync(b._h, _lib.C.byref(ms))
        assert rc == _lib.YTRC_ECANCELLED
        print(f"full {full_ms:.2f} ms, stopped {ms.value:.2f} ms", flush=True)
        assert ms.value < 0.5 * full_ms
    finally:
        b.close()


def test_sync_deadline(ctx, big):
    _, specs, _, _ = big
    b = _lib.Batch(ctx, specs * 2, _lib.BATCH_PERSISTENT)
    try:
        b.launch()
        with pytest.raises(_lib.YtrcTimeout):
            b.sync(timeout=0.0)
        assert b.sync(timeout=60.0) > 0.0
        b.read_image(0)
    finally:
        b.close()


def test_drop_in_sampler_interrupt(ctx, big, monkeypatch):
    """A KeyboardInterrupt during the sampler's wait stops the launch and
    reaches the caller; the caller's image is untouched.  The interrupt is
    raised from the first timed wait slice (where Python would deliver a
    SIGINT), so the test does not depend on signal timing."""
    vol, _, ds, tf = big
    kd = vol.kd
    p = c0_params(512, normal=NORMALS[0])
    image = np.zeros((512, 512, 4))
    s = VolumeRenderSampler(np.atleast_3d(p["vp_pos"]), np.atleast_3d(p["vp_dir"]), p["center"],
                            p["bounds"], image, p["x_vec"], p["y_vec"], p["width"], "KDTree",
                            tf, 10, lens_type="plane-parallel")
    for pg in kd.traverse(viewpoint=p["viewpoint"]):
        s(pg)
    orig = _lib.Batch.sync
    seen = {"cancel": 0, "timed": 0}

    def sync(self, timeout=None):
        if timeout is not None and seen["timed"] == 0:
            seen["timed"] += 1
            raise KeyboardInterrupt
        return orig(self, timeout)

    orig_cancel = _lib.Batch.cancel

    def cancel(self):
        seen["cancel"] += 1
        orig_cancel(self)

    monkeypatch.setattr(_lib.Batch, "sync", sync)
    monkeypatch.setattr(_lib.Batch, "cancel", cancel)
    with pytest.raises(KeyboardInterrupt):
        s.aimage
    assert seen["cancel"] == 1
    assert not np.any(image)


def test_server_wait_deadline(ctx, big):
    _, specs, _, _ = big
    srv = _lib.Server(ctx, max_requests=8, budget_bytes=512 << 20, chunk_tiles=4096)
    try:
        assert srv.wait(timeout=0.01) == []  # nothing queued: returns at once
        rid = srv.submit(specs[0])
        assert rid is not None
        with pytest.raises(_lib.YtrcTimeout):
            srv.wait(timeout=0.0)
        got = []
        while not got:
            got = srv.wait(timeout=60.0)
        assert got == [rid]
        srv.read(rid)
    finally:
        srv.close()
