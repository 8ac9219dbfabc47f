"""The timed kernels (walk_projection_fast, walk_vr_ctf_fast and their
general-walk fallback) on edge geometry, against the FP64 oracle.

Bricks whose sides are not multiples of the 4-cell blocks (12/13 x 18/19 x
7/8 cells), bricks one cell thick (1 x 9 x 9; projection only: vertex
centring needs two cells per grid axis, interpolators.pyx:177-226), the
kd bricks of an AMR set, whose child masks skip cells, and one brick of
512 x 6 x 5 and of 600 x 6 x 5 cells (the persistent projection walk
tabulates its offset terms in LDS for sides up to 512 cells and forms them
arithmetically beyond); rays along an axis (two
direction components 0: tMax 1e60, no steps on those axes), along a face
diagonal and the body diagonal (ties between two and three axes at every
corner the rays pass), a camera wider than the volume (lanes that miss it)
and one whose rays start and end inside the volume (walk_volume's
"inside" start, grid_traversal.pyx:49-53).  Non-record requests in a
persistent hipGraph batch, as timed; bars: rel-L2 and max-abs/max <= 1e-5
per written channel.
"""

import numpy as np
import pytest

from oracle import oracle
from yt_amd import _lib, data, render

from .helpers import (c0_params, oracle_proj_bricks, oracle_vr_bricks, small_amr, tf_dict, uniform_lognormal,
                      vr_tf_for)

pytestmark = pytest.mark.gpu

TOL = 1e-5
CAMERAS = [("axis", (1.0, 0.0, 0.0), 1.3), ("face_diagonal", (1.0, 1.0, 0.0), 1.5),
           ("body_diagonal", (1.0, 1.0, 1.0), 1.8), ("inside", (0.3, 0.5, 0.81), 0.6),
           ("oblique", (-0.6, 0.2, 0.77), 1.2)]


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


def _dataset(name):
    if name == "uniform37":
        return uniform_lognormal(37, seed=311, psize=(3, 2, 5)), True
    if name == "slabs9":
        return uniform_lognormal(9, seed=312, psize=(9, 1, 1)), False
    if name.startswith("long"):
        n0 = int(name[4:])
        rho = np.random.default_rng(314).lognormal(0.0, 1.0, size=(n0, 6, 5))
        return data.load_uniform_grid({"density": rho}), True
    return small_amr(root_n=24, levels=2, seed=313), True


@pytest.mark.parametrize("dname", ["uniform37", "slabs9", "amr", "long512", "long600"])
def test_timed_kernels_on_edge_geometry(ctx, dname):
    ds, with_vr = _dataset(dname)
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density" if with_vr else None)
    try:
        tf = vr_tf_for(ds, nbins=128)
        params = [c0_params(40, normal=nv, width=w) for _, nv, w in CAMERAS]
        specs = []
        for p in params:
            specs.append(vol.projection_request(p))
            if with_vr:
                specs.append(vol.vr_request(p, tf))
        step = 2 if with_vr else 1
        b = _lib.Batch(ctx, specs, _lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH)
        try:
            b.run()
            imgs = [b.read_image(i) for i in range(len(specs))]
        finally:
            b.close()
        pbricks = oracle_proj_bricks(ds, ["density"])
        vbricks = oracle_vr_bricks(vol) if with_vr else None
        for k, (name, _, _) in enumerate(CAMERAS):
            p = params[k]
            checks = [("projection", imgs[step * k], oracle.cast(0, p, pbricks), 1)]
            if with_vr:
                ov = oracle.cast(1, p, vbricks, order=specs[2 * k + 1].order, tf=tf_dict(tf), n_samples=10)
                checks.append(("VR", imgs[2 * k + 1], ov, 3))
            for kind, img, oimg, nch in checks:
                errs = _errors(img, oimg, nch)
                print(f"{dname}, {name}, {kind}: "
                      + ", ".join(f"ch{c} rel-L2 {e:.2e} max-abs/max {m:.2e}" for c, (e, m) in enumerate(errs)),
                      flush=True)
                assert float(np.abs(oimg[..., :nch]).max()) > 0, (name, kind)
                for e, m in errs:
                    assert e <= TOL and m <= TOL, (name, kind, e, m)
    finally:
        vol.free()
