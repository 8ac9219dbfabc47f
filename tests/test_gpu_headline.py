"""The timed kernels at the headline and serving workloads, against the FP64
oracle.

Headline batch (bench.py's step, built by bench.py's own functions): a 512^3
float32 lognormal field in 2x2x2 bricks of 256^3, 64 cameras with sides in
{128, 256, 512, 1024}, alternating projection and VR (8-Gaussian
ColorTransferFunction, 10 sub-samples).  The batch runs exactly as timed:
non-record requests, persistent tile-queue kernels captured in a hipGraph,
the default launch shape; the graph is replayed and every replay must equal
the first bit for bit, as must a second launch shape.  Every one of the 64
images is then compared with the oracle's FP64 image of the same request
(rel-L2 and max-abs / max of each written channel <= 1e-5), and a record
pass of the same specs must reproduce the oracle's per-pixel visit counts,
visit-sequence hashes and sample counts exactly.

Serving workload: the three resident datasets (256^3 and 512^3 uniform, the
643-grid AMR hierarchy), 16 requests with sides in [256, 2048], projection
and VR, served under both chunk fill orders; every served image equals the
same request run alone bit for bit, and 4 of them (both kinds, all three
datasets) are checked against the oracle.
"""

import numpy as np
import pytest

import bench
from oracle import oracle
from yt_amd import _lib, autotune, data, render, serve

from .helpers import oracle_proj_bricks, oracle_vr_bricks, tf_dict

pytestmark = pytest.mark.gpu

TOL = 1e-5
N_CAMS = 64


def _errors(img, oimg, nch):
    out = []
    for c in range(nch):
        a, b = img[..., c].astype(np.float64), oimg[..., c]
        den = float(np.sqrt(np.sum(b * b)))
        e = float(np.sqrt(np.sum((a - b) ** 2))) / den if den > 0 else float(np.sqrt(np.sum(a * a)))
        mx = float(np.max(np.abs(b)))
        m = float(np.max(np.abs(a - b))) / mx if mx > 0 else float(np.max(np.abs(a)))
        out.append((e, m))
    return out


@pytest.fixture(scope="module")
def headline(ctx):
    print("headline: building the 512^3 workload", flush=True)
    rho, ds, tf, cams = bench.build_workload(512, 2, N_CAMS)
    del rho
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    specs = bench.make_specs(vol, tf, cams)
    flags = _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH
    images = []
    b = autotune.make_batch(ctx, specs, autotune.DEFAULT, flags)
    try:
        b.run()
        images = [b.read_image(i).astype(np.float32) for i in range(N_CAMS)]
        for rep in range(2):
            b.run()
            for i in range(N_CAMS):
                assert np.array_equal(b.read_image(i).astype(np.float32), images[i]), (rep, i)
    finally:
        b.close()
    alt = dict(autotune.DEFAULT, waves_per_group=2)
    b = autotune.make_batch(ctx, specs, alt, flags)
    try:
        b.run()
        for i in range(N_CAMS):
            assert np.array_equal(b.read_image(i).astype(np.float32), images[i]), ("w2", i)
    finally:
        b.close()
    print("headline: 3 graph replays and a second launch shape give identical images", flush=True)
    rec = _lib.Batch(ctx, bench.make_specs(vol, tf, cams, record=True), _lib.BATCH_PERSISTENT)
    try:
        rec.run()
        records = [rec.read_record(i) for i in range(N_CAMS)]
    finally:
        rec.close()
    st = dict(cams=cams, specs=specs, images=images, records=records, tf=tf_dict(tf),
              pbricks=oracle_proj_bricks(ds, ["density"]), vbricks=oracle_vr_bricks(vol))
    vol.free()
    yield st


@pytest.mark.parametrize("i", range(N_CAMS))
def test_headline_image_against_oracle(headline, i):
    kind, p = headline["cams"][i]
    spec = headline["specs"][i]
    if kind == 0:
        oimg, ov, oh, os_ = oracle.cast(0, p, headline["pbricks"], record=True, with_samples=True)
        nch = 1
    else:
        oimg, ov, oh, os_ = oracle.cast(1, p, headline["vbricks"], order=spec.order,
                                        tf=headline["tf"], n_samples=10, record=True,
                                        with_samples=True)
        nch = 3
    v, h, s = headline["records"][i]
    errs = _errors(headline["images"][i], oimg, nch)
    print(f"headline image {i} ({'projection' if kind == 0 else 'VR'}, {p['nv'][0]}^2): "
          + ", ".join(f"ch{c} rel-L2 {e:.2e} max-abs/max {m:.2e}" for c, (e, m) in enumerate(errs))
          + f"; visits {int(v.sum())}", flush=True)
    assert np.array_equal(v, ov) and np.array_equal(h, oh) and np.array_equal(s, os_)
    for e, m in errs:
        assert e <= TOL and m <= TOL


@pytest.fixture(scope="module")
def serving_sets(ctx):
    print("serving: building the three resident datasets", flush=True)
    from yt_amd.transfer_functions import gaussian_ctf

    def tf_for(ds):
        lo = min(float(np.log10(np.min(g["density"]))) for g in ds.grids)
        hi = max(float(np.log10(np.max(g["density"]))) for g in ds.grids)
        return gaussian_ctf((lo, hi), n_gauss=8, nbins=256)

    d256 = data.load_uniform_grid({"density": data.lognormal_density(256, seed=1001,
                                                                      dtype=np.float32)})
    d512 = data.load_uniform_grid({"density": data.lognormal_density(512, seed=1001,
                                                                      dtype=np.float32)},
                                  psize=(2, 2, 2))
    amr = data.load_amr_grids(data.make_amr_hierarchy(128, 3), [128] * 3)
    vols = [render.ResidentVolume(ctx, d, proj_fields="density", vr_field="density")
            for d in (d256, d512, amr)]
    tfs = {id(v): tf_for(v.ds) for v in vols}
    yield vols, tfs
    for v in vols:
        v.free()


def _alone(ctx, spec):
    b = _lib.Batch(ctx, [spec], _lib.BATCH_PLAIN)
    try:
        b.run()
        return b.read_image(0)
    finally:
        b.close()


def _serve(ctx, specs, policy, stagger):
    """Serves specs under one fill policy.  With stagger, a new request is
    submitted only after the server has advanced, so requests arrive while
    others are on the GPU."""
    srv = _lib.Server(ctx, max_requests=256, budget_bytes=8 << 30, chunk_tiles=16384)
    srv.set_policy(policy)
    out, ids, nxt = {}, {}, 0
    try:
        while len(out) < len(specs):
            while nxt < len(specs) and (not stagger or nxt <= len(out) + 2):
                rid = srv.submit(specs[nxt])
                if rid is None:
                    break
                ids[rid] = nxt
                nxt += 1
            for rid in (srv.progress() if stagger else srv.wait()):
                out[ids.pop(rid)] = srv.read(rid)
        st = srv.stats()
        assert st["held"] == 0 and st["used"] == 0
        return out
    finally:
        srv.close()


def test_serving_mix_equals_alone_and_oracle(ctx, serving_sets):
    vols, tfs = serving_sets
    specs = serve.make_requests(vols, 16, seed=2006, tfs=tfs, rng_cam_seed=2004,
                                side_range=(256, 2048))
    sides = [s.nv[0] for s in specs]
    kinds = [s.kind for s in specs]
    print(f"serving: sides {sides}, kinds {kinds}", flush=True)
    assert min(sides) >= 256 and max(sides) <= 2048 and max(sides) > 1024
    assert _lib.PROJECTION in kinds and _lib.VOLUME_RENDER in kinds
    alone = [_alone(ctx, s) for s in specs]
    for policy, stagger in (("shortest_first", False), ("fifo", True)):
        got = _serve(ctx, specs, policy, stagger)
        for i, s in enumerate(specs):
            nch = 1 if s.kind == _lib.PROJECTION else 3
            assert np.array_equal(got[i][..., :nch], alone[i][..., :nch]), (policy, i)
        print(f"serving ({policy}{', staggered' if stagger else ''}): 16 images equal alone",
              flush=True)
    # the oracle on the two smallest requests of each kind, over different datasets
    picked = []
    for kind in (_lib.PROJECTION, _lib.VOLUME_RENDER):
        idx = sorted((i for i in range(16) if kinds[i] == kind), key=lambda i: sides[i])
        seen = set()
        for i in idx:
            v = vols[i % len(vols)]
            if id(v) not in seen:
                picked.append(i)
                seen.add(id(v))
            if len(seen) == 2:
                break
    assert len(picked) == 4
    for i in picked:
        vol, s = vols[i % len(vols)], specs[i]
        p = s.params
        if s.kind == _lib.PROJECTION:
            oimg = oracle.cast(0, p, oracle_proj_bricks(vol.ds, ["density"]))
            nch = 1
        else:
            oimg = oracle.cast(1, p, oracle_vr_bricks(vol), order=s.order,
                               tf=tf_dict(tfs[id(vol)]), n_samples=10)
            nch = 3
        errs = _errors(alone[i], oimg, nch)
        print(f"serving request {i} ({'projection' if nch == 1 else 'VR'}, {sides[i]}^2, "
              f"dataset {i % len(vols)}): "
              + ", ".join(f"ch{c} rel-L2 {e:.2e} max-abs/max {m:.2e}"
                          for c, (e, m) in enumerate(errs)), flush=True)
        for e, m in errs:
            assert e <= TOL and m <= TOL
