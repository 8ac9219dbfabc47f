"""GPU tests of the continuous-batching server (ytrc_server_*).

Each request's image is bit-identical whether it runs alone (one plain
launch) or mixed with other requests of other sizes, kinds and datasets in
the server's chunks: every pixel is computed by the same per-ray code, so
the tolerance written here is exact equality.
"""

import numpy as np
import pytest

from yt_amd import _lib, render

from .helpers import c0_params, small_amr, uniform_lognormal, vr_tf_for

pytestmark = pytest.mark.gpu


def _alone(ctx, spec):
    b = _lib.Batch(ctx, [spec], _lib.BATCH_PLAIN)
    try:
        b.run()
        return b.read_image(0)
    finally:
        b.close()


@pytest.fixture(scope="module")
def volumes(ctx):
    a = uniform_lognormal(40, seed=41)
    b = small_amr(root_n=16, levels=2, seed=43)
    va = render.ResidentVolume(ctx, a, proj_fields="density", vr_field="density")
    vb = render.ResidentVolume(ctx, b, proj_fields="density", vr_field="density")
    tfs = {id(va): vr_tf_for(a, nbins=128), id(vb): vr_tf_for(b, nbins=256)}
    yield [va, vb], tfs
    va.free()
    vb.free()


def _specs(volumes, tfs, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        vol = volumes[i % len(volumes)]
        nrm = rng.normal(size=3)
        p = c0_params(int(rng.integers(8, 140)), normal=tuple(nrm / np.linalg.norm(nrm)))
        out.append(vol.vr_request(p, tfs[id(vol)]) if rng.random() < 0.5
                   else vol.projection_request(p))
    return out


def test_server_mix_equals_alone(ctx, volumes):
    vols, tfs = volumes
    specs = _specs(vols, tfs, 10, seed=7)
    alone = [_alone(ctx, s) for s in specs]
    srv = _lib.Server(ctx, max_requests=16, budget_bytes=256 << 20, chunk_tiles=256)
    try:
        ids = {srv.submit(s): i for i, s in enumerate(specs)}
        assert None not in ids
        got = {}
        while len(got) < len(specs):
            for rid in srv.wait():
                got[ids[rid]] = srv.read(rid)
        for i in range(len(specs)):
            nch = 1 if specs[i].kind == _lib.PROJECTION else 3
            assert np.array_equal(got[i][..., :nch], alone[i][..., :nch]), i
        st = srv.stats()
        assert st["held"] == 0 and st["used"] == 0 and st["chunks"] > 1
    finally:
        srv.close()


def test_server_admission_waits_for_budget(ctx, volumes):
    vols, tfs = volumes
    specs = _specs(vols, tfs, 6, seed=9)
    # a budget that holds one 140^2 request region at a time
    srv = _lib.Server(ctx, max_requests=4, budget_bytes=1 << 20, chunk_tiles=128)
    try:
        pending = list(range(len(specs)))
        out = {}
        live = {}
        rejected = 0
        while pending or live:
            while pending:
                rid = srv.submit(specs[pending[0]])
                if rid is None:
                    rejected += 1
                    break
                live[rid] = pending.pop(0)
            for rid in srv.wait():
                out[live.pop(rid)] = srv.read(rid)
        assert rejected > 0
        for i, s in enumerate(specs):
            nch = 1 if s.kind == _lib.PROJECTION else 3
            assert np.array_equal(out[i][..., :nch], _alone(ctx, s)[..., :nch]), i
    finally:
        srv.close()


def test_server_cancel(ctx, volumes):
    vols, tfs = volumes
    specs = _specs(vols, tfs, 5, seed=11)
    srv = _lib.Server(ctx, max_requests=8, budget_bytes=64 << 20, chunk_tiles=64)
    try:
        ids = [srv.submit(s) for s in specs]
        srv.cancel(ids[2])
        done = set()
        while len(done) < 4:
            for rid in srv.wait():
                done.add(rid)
                srv.read_planar(rid)
        assert ids[2] not in done
        for _ in range(10):
            srv.progress()
            if srv.stats()["held"] == 0:
                break
        assert srv.stats()["used"] == 0
    finally:
        srv.close()


@pytest.mark.parametrize("policy,age_ms", [("fifo", 0.0), ("shortest_first", 1e7),
                                           ("shortest_first", 0.0)])
def test_server_policy_fill_order(ctx, volumes, policy, age_ms):
    """One 256^2 volume render (1024 tiles) then six 16^2 requests, chunks of
    64 tiles: in arrival order the large request finishes first; shortest
    first lets the small ones overtake it, unless they are past the age
    limit (0 ms: every request is, so arrival order again).  Every image
    equals the request run alone."""
    vols, tfs = volumes
    rng = np.random.default_rng(13)
    big = vols[0].vr_request(c0_params(256, normal=(0.3, -0.2, 0.93)), tfs[id(vols[0])])
    small = []
    for i in range(6):
        nrm = rng.normal(size=3)
        p = c0_params(16, normal=tuple(nrm / np.linalg.norm(nrm)))
        v = vols[i % 2]
        small.append(v.vr_request(p, tfs[id(v)]) if i % 3 else v.projection_request(p))
    specs = [big] + small
    alone = [_alone(ctx, s) for s in specs]
    srv = _lib.Server(ctx, max_requests=16, budget_bytes=64 << 20, chunk_tiles=64)
    try:
        srv.set_policy(policy, age_limit_ms=age_ms)
        ids = [srv.submit(s) for s in specs]
        assert None not in ids
        got = {}
        pos = {}  # completion rank: finished ids come back in chunk order
        while len(got) < len(specs):
            for rid in srv.wait():
                pos[rid] = len(pos)
                got[rid] = srv.read(rid)
        last_small = max(pos[r] for r in ids[1:])
        first_small = min(pos[r] for r in ids[1:])
        print(f"{policy} age={age_ms}: big finished {pos[ids[0]]}th, small "
              f"{first_small}..{last_small}", flush=True)
        if policy == "shortest_first" and age_ms > 0:
            assert last_small < pos[ids[0]]
        else:
            assert pos[ids[0]] < first_small
        for i, rid in enumerate(ids):
            nch = 1 if specs[i].kind == _lib.PROJECTION else 3
            assert np.array_equal(got[rid][..., :nch], alone[i][..., :nch]), i
        assert srv.stats()["used"] == 0
    finally:
        srv.close()


def test_server_policy_errors_and_cancel(ctx, volumes):
    vols, tfs = volumes
    srv = _lib.Server(ctx, max_requests=8, budget_bytes=64 << 20, chunk_tiles=64)
    try:
        with pytest.raises(ValueError):
            srv.set_policy("lifo")
        with pytest.raises(ValueError):
            _lib._check(_lib.lib().ytrc_server_set_policy(srv._h, 7, 1.0))
        with pytest.raises(ValueError):
            srv.set_policy("shortest_first", age_limit_ms=-1.0)
        srv.set_policy("shortest_first", age_limit_ms=1e7)
        specs = _specs(vols, tfs, 5, seed=17)
        ids = [srv.submit(s) for s in specs]
        srv.cancel(ids[3])
        done = set()
        while len(done) < 4:
            for rid in srv.wait():
                done.add(rid)
                srv.read_planar(rid)
        assert ids[3] not in done
        for _ in range(10):
            srv.progress()
            if srv.stats()["held"] == 0:
                break
        assert srv.stats()["used"] == 0
    finally:
        srv.close()


def test_server_all_sampler_kinds_and_lenses(ctx, volumes):
    """Requests of every sampler kind and lens family, mixed in the server's
    chunks, equal each request run alone."""
    vols, tfs = volumes
    va = vols[0]
    tf = tfs[id(va)]
    pp = c0_params(48, normal=(0.3, -0.4, 0.87))
    pv = render.perspective_params((1.6, 0.4, 1.2), (0.5, 0.5, 0.5), 0.8, (40, 36))
    pf = render.lens_params("fisheye", (0.45, 0.5, 0.55), (0.1, 0.7, 0.3), 0.9, (24, 24))
    specs = [va.projection_request(pv), va.vr_request(pf, tf), va.interpolated_projection_request(pp),
             va.light_source_request(pv, tf, light_dir=(0.3, 1.0, -0.2)),
             va.interpolated_projection_request(pf), va.vr_request(pp, tf)]
    alone = [_alone(ctx, s) for s in specs]
    srv = _lib.Server(ctx, max_requests=8, budget_bytes=128 << 20, chunk_tiles=64)
    try:
        ids = {srv.submit(s): i for i, s in enumerate(specs)}
        assert None not in ids
        got = {}
        while len(got) < len(specs):
            for rid in srv.wait():
                got[ids[rid]] = srv.read(rid)
        for i in range(len(specs)):
            a, b = got[i][..., :3], alone[i][..., :3]
            assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)],
                                                                               b[~np.isnan(b)]), i
    finally:
        srv.close()


def test_server_reuses_released_transfer_functions(ctx, volumes):
    """Transfer-function tables stay in the server's table after their
    requests are read, are found again by content, and are forgotten when a
    new table is written over them; the launch always copies every table in
    use into LDS.  Sequence: T1 (256 bins) and T2 (128 bins) placed and
    released; T3 (64 bins) is written over T1; T2, found again at its old
    place beyond T3, must still render exactly as alone."""
    vols, _ = volumes
    va = vols[0]
    t1 = vr_tf_for(va.ds, nbins=256)
    t2 = vr_tf_for(va.ds, nbins=128)
    t3 = vr_tf_for(va.ds, nbins=64)
    p = c0_params(72, normal=(0.3, -0.4, 0.87))
    srv = _lib.Server(ctx, max_requests=8, budget_bytes=64 << 20, chunk_tiles=64)

    def run(specs):
        ids = [srv.submit(s) for s in specs]
        assert None not in ids
        got = {}
        while len(got) < len(ids):
            for rid in srv.wait():
                got[rid] = srv.read(rid)
        return [got[i] for i in ids]

    try:
        run([va.vr_request(p, t1), va.vr_request(p, t2)])
        run([va.vr_request(p, t3)])
        s2 = va.vr_request(p, t2)
        img, = run([s2])
        assert np.array_equal(img[..., :3], _alone(ctx, s2)[..., :3])
        s1 = va.vr_request(p, t1)
        img1, = run([s1])
        assert np.array_equal(img1[..., :3], _alone(ctx, s1)[..., :3])
    finally:
        srv.close()


def test_server_refuses_request_larger_than_budget(ctx, volumes):
    """A request whose image and descriptors exceed the server's whole
    budget can never be admitted: submit raises ValueError instead of
    returning None (wait and retry)."""
    vols, tfs = volumes
    srv = _lib.Server(ctx, max_requests=4, budget_bytes=1 << 20, chunk_tiles=64)
    try:
        small = vols[0].projection_request(c0_params(32))
        assert srv.submit(small) is not None
        with pytest.raises(ValueError, match="whole budget"):
            srv.submit(vols[0].projection_request(c0_params(1024)))
        while srv.stats()["held"]:
            for rid in srv.wait():
                srv.read(rid)
        assert srv.stats()["used"] == 0
    finally:
        srv.close()


def test_server_large_transfer_function(ctx, volumes):
    """The server's table holds as much as a workgroup's LDS: a 2048-bin
    ColorTransferFunction (98 KiB of LDS) is served, mixed with a small one,
    and renders exactly as alone; a table larger than the LDS is refused
    (ValueError), not left waiting for room that never comes."""
    vols, _ = volumes
    va = vols[0]
    max_bins = ctx.max_lds // 48 - 2
    if max_bins < 2048:
        pytest.skip(f"a workgroup has only {ctx.max_lds} B of LDS here")
    big = vr_tf_for(va.ds, nbins=2048)
    small = vr_tf_for(va.ds, nbins=64)
    p = c0_params(64, normal=(0.5, 0.1, -0.86))
    specs = [va.vr_request(p, big), va.vr_request(p, small)]
    srv = _lib.Server(ctx, max_requests=8, budget_bytes=64 << 20, chunk_tiles=64)
    try:
        ids = [srv.submit(s) for s in specs]
        assert None not in ids
        got = {}
        while len(got) < len(ids):
            for rid in srv.wait():
                got[rid] = srv.read(rid)
        for rid, spec in zip(ids, specs):
            assert np.array_equal(got[rid][..., :3], _alone(ctx, spec)[..., :3])
        with pytest.raises(ValueError, match="B of LDS"):
            srv.submit(va.vr_request(p, vr_tf_for(va.ds, nbins=max_bins + 1)))
        assert srv.stats()["used"] == 0
    finally:
        srv.close()


def test_dataset_free_refused_while_in_use(ctx):
    """ytrc_dataset_free refuses a dataset that a live batch or an unread
    server request still reads (YTRC_EBUSY), so no kernel can read freed
    memory; destroying the batch, reading the request or destroying the
    server gives it back."""
    ds = uniform_lognormal(16, seed=5)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density")
    p = c0_params(24)
    b = _lib.Batch(ctx, [vol.projection_request(p)])
    with pytest.raises(_lib.YtrcError):
        ctx.free(vol.proj_ds_id)
    b.run()
    b.close()
    srv = _lib.Server(ctx, max_requests=4, budget_bytes=16 << 20, chunk_tiles=64)
    try:
        rid = srv.submit(vol.projection_request(p))
        assert rid is not None
        with pytest.raises(_lib.YtrcError):
            ctx.free(vol.proj_ds_id)
        done = []
        while rid not in done:
            done += srv.wait()
        srv.read(rid)
        assert srv.submit(vol.projection_request(p)) is not None  # left unread
    finally:
        srv.close()
    vol.free()
    assert vol.proj_ds_id is None
