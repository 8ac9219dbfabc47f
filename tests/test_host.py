"""Host-side logic that feeds the kernels (no GPU): camera basis and sampler
parameters, vertex centring, child masks, the kd-tree's brick set and
viewpoint order, the uniform-grid split, and the C ABI's exported symbols."""

import ctypes
import os
import re

import numpy as np
import pytest

from yt_amd import _lib, amr_kdtree, camera, data, render

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_default_north_matches_on_axis():
    """off_axis_projection.py:402-407: for normal +z the image x axis is +x
    and y is +y (the on-axis projection convention)."""
    cam = camera.off_axis_camera((0.5, 0.5, 0.5), (0, 0, 1), 1.0, 32)
    assert np.allclose(cam.unit_vectors[0], [1, 0, 0])
    assert np.allclose(cam.unit_vectors[1], [0, 1, 0])
    assert np.allclose(cam.unit_vectors[2], [0, 0, 1])


@pytest.mark.parametrize("seed", range(5))
def test_basis_and_sampler_params(seed):
    rng = np.random.default_rng(seed)
    n = rng.normal(size=3)
    c = rng.uniform(0, 1, 3)
    w = float(rng.uniform(0.5, 2))
    p = render.camera_params(c, n, w, (31, 17), rng.normal(size=3))
    uv = np.array([p["x_vec"], p["y_vec"], np.ravel(p["vp_dir"]) / w])
    assert np.allclose(uv @ uv.T, np.eye(3), atol=1e-12)
    vp = np.ravel(p["vp_pos"])
    inv = vp[:9].reshape(3, 3, order="F")
    assert np.allclose(inv, np.linalg.pinv(uv), atol=1e-12)
    assert np.allclose(vp[9:], c - 0.5 * w * uv[2])  # back_center (lens.py:57)
    assert np.allclose(p["center"], vp[9:])
    assert p["bounds"] == (-w / 2, w / 2, -w / 2, w / 2)
    # the viewpoint is 1e6 widths beyond the front plane (lens.py:114-121)
    assert np.allclose(p["viewpoint"], c + 0.5 * w * uv[2] + uv[2] * 1e6 * w)


def test_vertex_centring_exact_on_linear_fields():
    """test_interpolators.py:107-157: no-ghost vertex centring reproduces a
    linear field exactly at every vertex."""
    n = (6, 5, 7)
    i, j, k = np.meshgrid(*[np.arange(x) + 0.5 for x in n], indexing="ij")
    f = 1.5 + 2.0 * i - 0.75 * j + 0.25 * k
    v = data.ghost_zone_interpolate_no_ghost(f)
    I, J, K = np.meshgrid(*[np.arange(x + 1.0) for x in n], indexing="ij")
    assert v.shape == tuple(x + 1 for x in n)
    assert np.array_equal(v, 1.5 + 2.0 * I - 0.75 * J + 0.25 * K)


def test_uniform_split_matches_decompose():
    rho = np.arange(64 ** 3, dtype=float).reshape((64,) * 3)
    ds = data.load_uniform_grid({"density": rho}, nprocs=8)
    assert len(ds.grids) == 8
    assert all(tuple(g.ActiveDimensions) == (32, 32, 32) for g in ds.grids)
    # split_array order: i-major (decompose.py:120-150)
    assert np.allclose(ds.grids[1].LeftEdge, [0, 0, 0.5])
    assert np.allclose(ds.grids[4].LeftEdge, [0.5, 0, 0])
    assert np.array_equal(ds.grids[5]["density"], rho[32:, :32, 32:])


def _small_amr():
    from tests.helpers import small_amr

    return small_amr(root_n=16, levels=2, seed=11)


def test_child_mask_and_blocks():
    ds = _small_amr()
    total = 0.0
    for g, m in ds.blocks():
        total += m.sum() * np.prod(g.dds)
    assert abs(total - 1.0) < 1e-12  # unmasked cells tile the domain once


def test_kdtree_invariants():
    ds = _small_amr()
    kd = amr_kdtree.AMRKDTree(ds)
    kd.set_fields(["density"], [True])
    leaves = kd.leaves()
    assert abs(kd.trunk.kd_sum_volume() - 1.0) < 1e-12
    by_id = {g.id: g for g in ds.grids}
    for node in leaves:
        g = by_id[node.grid]
        assert np.all(g.LeftEdge <= node.left_edge) and np.all(g.RightEdge >= node.right_edge)
        # finest covering grid: no finer grid overlaps the leaf interior
        for h in ds.grids:
            if h.Level > g.Level:
                lo = np.maximum(h.LeftEdge, node.left_edge)
                hi = np.minimum(h.RightEdge, node.right_edge)
                assert np.any(hi <= lo)
        pg = kd.get_brick_data(node)
        assert pg.my_data[0].shape == tuple(pg.dims + 1)
        assert pg.source_mask.shape == tuple(pg.dims)


def test_kdtree_viewpoint_order_is_front_to_back_along_rays():
    """Farthest-from-viewpoint first: along a ray travelling toward the
    viewpoint the leaves it crosses come in increasing t."""
    ds = _small_amr()
    kd = amr_kdtree.AMRKDTree(ds)
    rng = np.random.default_rng(5)
    for _ in range(10):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        vp = 0.5 + n * 1e6
        order = list(kd.trunk.kd_traverse(viewpoint=vp))
        pos = rng.uniform(0.2, 0.8, 3) - n
        tin = []
        for node in order:
            t0, t1 = -np.inf, np.inf
            for a in range(3):
                ta = (node.left_edge[a] - pos[a]) / n[a]
                tb = (node.right_edge[a] - pos[a]) / n[a]
                t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
            if t1 > t0:
                tin.append(t0)
        assert len(tin) > 1
        assert all(a <= b + 1e-12 for a, b in zip(tin, tin[1:]))


def _same_tree(a, b):
    na, nb = list(a.depth_traverse()), list(b.depth_traverse())
    assert len(na) == len(nb)
    for x, y in zip(na, nb):
        assert x.split_dim == y.split_dim and x.grid == y.grid and x.node_id == y.node_id
        assert (x.split_pos == y.split_pos) or (np.isnan(x.split_pos) and np.isnan(y.split_pos))
        assert np.array_equal(x.left_edge, y.left_edge) and np.array_equal(x.right_edge, y.right_edge)
        assert x.is_leaf() == y.is_leaf()


@pytest.mark.parametrize("case", ["amr16", "amr32", "amr64x3", "uniform8", "psize343"])
def test_native_kdtree_equals_oracle(case):
    """The native build (ytrc_kdtree_*) against the Python restatement in
    oracle/kdtree.py: the same nodes, splits, grids and node ids, and the
    same brick order from 24 viewpoints, inside and outside the domain."""
    from oracle import kdtree as okd
    from tests.helpers import small_amr, uniform_lognormal

    ds = {"amr16": lambda: small_amr(root_n=16, levels=2, seed=11),
          "amr32": lambda: small_amr(root_n=32, levels=2, seed=7),
          "amr64x3": lambda: small_amr(root_n=64, levels=3, seed=3),
          "uniform8": lambda: uniform_lognormal(16, nprocs=8),
          "psize343": lambda: uniform_lognormal(21, psize=(3, 4, 3))}[case]()
    kd = amr_kdtree.AMRKDTree(ds)
    ot = okd.build([(g.LeftEdge, g.RightEdge, g.id, g.Level) for g, _ in ds.blocks()])
    _same_tree(kd.trunk, ot)
    rng = np.random.default_rng(17)
    vps = list(rng.uniform(-1.0, 2.0, size=(20, 3))) + [
        np.array([0.5, 0.5, 0.5]), np.array([0.25, 0.5, 0.75]),
        np.array([1e6, -1e6, 3.0]), np.array([0.0, 0.0, 0.0])]
    for vp in vps:
        assert np.array_equal(kd.order(vp), okd.order(ot, vp))
    # the mirrored Python traversal gives the same order
    idx = {id(n): i for i, n in enumerate(kd.leaves())}
    assert np.array_equal(kd.order(vps[0]),
                          [idx[id(n)] for n in kd.trunk.kd_traverse(viewpoint=vps[0])])


def test_oracle_kdtree_invariants():
    """oracle/kdtree.py is pinned by what a kd-tree of the hierarchy must
    satisfy: the leaves tile the domain; each leaf's grid covers it and no
    finer grid overlaps it; painter's order along lines of sight."""
    from oracle import kdtree as okd

    ds = _small_amr()
    ot = okd.build([(g.LeftEdge, g.RightEdge, g.id, g.Level) for g, _ in ds.blocks()])
    leaves = list(ot.kd_traverse())
    vol = sum(float(np.prod(n.right_edge - n.left_edge)) for n in leaves)
    assert abs(vol - 1.0) < 1e-12
    by_id = {g.id: g for g in ds.grids}
    for node in leaves:
        g = by_id[node.grid]
        assert np.all(g.LeftEdge <= node.left_edge) and np.all(g.RightEdge >= node.right_edge)
        for h in ds.grids:
            if h.Level > g.Level:
                lo = np.maximum(h.LeftEdge, node.left_edge)
                hi = np.minimum(h.RightEdge, node.right_edge)
                assert np.any(hi <= lo)
    rng = np.random.default_rng(9)
    for _ in range(10):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        vp = 0.5 + n * 1e6
        pos = rng.uniform(0.2, 0.8, 3) - n
        tin = []
        for node in ot.kd_traverse(viewpoint=vp):
            t0, t1 = -np.inf, np.inf
            for a in range(3):
                ta = (node.left_edge[a] - pos[a]) / n[a]
                tb = (node.right_edge[a] - pos[a]) / n[a]
                t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
            if t1 > t0:
                tin.append(t0)
        assert len(tin) > 1
        assert all(a <= b + 1e-12 for a, b in zip(tin, tin[1:]))


def test_native_kdtree_errors():
    with pytest.raises(ValueError):
        _lib.KdTree([[0, 0, 0]], [[1, 0, 1]], [0], [0])   # empty extent
    with pytest.raises(ValueError):
        _lib.KdTree([[0, 0, 0]], [[1, 1, 1]], [-1], [0])  # -1 means "no grid"
    t = _lib.KdTree(np.zeros((0, 3)), np.zeros((0, 3)), [], [])
    assert t.n_leaves == 0 and t.n_nodes == 1


def test_tf_from_ctf_object():
    from yt_amd.transfer_functions import gaussian_ctf

    tf = gaussian_ctf((-1.0, 2.0), n_gauss=8)
    t = _lib.TF.from_tf_obj(tf)
    assert t.desc.n_fits == 4
    assert list(t.desc.weight_table_ids)[:4] == [3, 3, 3, -1]
    assert t.desc.nbins[0] == 256


def _declared():
    src = open(os.path.join(ROOT, "include", "ytrc.h")).read()
    return sorted(set(re.findall(r"\b(ytrc_[a-z_]+)\s*\(", src)))


def test_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = _declared()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(decl) == sorted(_lib.EXPORTED)


def test_library_missing_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libytrc.so")
    with pytest.raises(_lib.YtrcLibraryError):
        _lib.lib()


def test_autotune_signature_and_cache(tmp_path):
    """The autotune cache key separates kernel variants and image-size
    classes, and the JSON cache round-trips."""
    from yt_amd import _lib, autotune

    p = {"nv": (64, 64)}
    a = _lib.RequestSpec(_lib.PROJECTION, p, 0, [0])
    b = _lib.RequestSpec(_lib.VOLUME_RENDER, p, 0, [0])
    big = _lib.RequestSpec(_lib.PROJECTION, {"nv": (1024, 1024)}, 0, [0])
    s1 = autotune.signature([a, a], "dev")
    assert s1 != autotune.signature([a, b], "dev")
    assert s1 != autotune.signature([big, big], "dev")
    assert s1 == autotune.signature([a, a], "dev")
    path = str(tmp_path / "sub" / "t.json")
    autotune._store(path, {"k": {"cfg": autotune.DEFAULT}})
    assert autotune._load(path)["k"]["cfg"] == autotune.DEFAULT
    assert autotune._load(str(tmp_path / "missing.json")) == {}


def test_leaf_slice_matches_brick_data():
    """The device-prep inputs (leaf_slice) are the slices get_brick_data
    takes (amr_kdtree.py:325-374)."""
    from tests.helpers import small_amr
    from yt_amd.amr_kdtree import AMRKDTree

    ds = small_amr(root_n=16, levels=2, seed=5)
    kd = AMRKDTree(ds)
    kd.set_fields(["density"], [True])
    for node in kd.leaves():
        grid, li, ri, nle, nre, mask = kd.leaf_slice(node)
        pg = kd.get_brick_data(node)
        assert np.array_equal(ri - li, pg.dims)
        assert np.array_equal(nle, pg.LeftEdge) and np.array_equal(nre, pg.RightEdge)
        assert np.array_equal(mask, pg.source_mask)
        node.data = None


def test_compute_stddev_image():
    """math_utils.py:1546-1583: sqrt(<v^2> - <v>^2); round-off negatives
    (a constant field) become 0, significant ones raise."""
    from yt_amd import compute_stddev_image

    rng = np.random.default_rng(12)
    # two values a, b with weights w, 1-w: variance w(1-w)(a-b)^2
    a, b, w = rng.uniform(1, 3, (3, 5, 4)), rng.uniform(1, 3, (3, 5, 4)), rng.uniform(0, 1, (5, 4))
    a, b = a[0], b[0]
    mean = w * a + (1 - w) * b
    mean2 = w * a * a + (1 - w) * b * b
    sd = compute_stddev_image(mean2, mean)
    assert np.allclose(sd, np.sqrt(w * (1 - w)) * np.abs(a - b), rtol=1e-12, atol=1e-12)
    # constant field: <v^2> and <v>^2 differ by round-off only
    c = np.full((6, 6), 0.1) * 3.0
    c2 = np.full((6, 6), 0.3 * 0.3)
    c2[0, 0] = np.nextafter(c[0, 0] * c[0, 0], 0.0)
    out = compute_stddev_image(c2, c)
    assert out[0, 0] == 0.0 and np.all(np.isfinite(out))
    with pytest.raises(ValueError):
        compute_stddev_image(np.full((2, 2), 1.0), np.full((2, 2), 2.0))


def test_product_modules_do_not_import_torch():
    """The package, its autotuner and server driver load no second HIP
    runtime: torch is never imported by the product path."""
    import subprocess
    import sys

    code = ("import sys; import yt_amd, yt_amd.autotune, yt_amd.serve, yt_amd.scene, "
            "yt_amd.off_axis_projection, yt_amd.image_samplers; "
            "assert 'torch' not in sys.modules, 'torch imported'")
    subprocess.run([sys.executable, "-c", code], check=True, cwd=ROOT)


def test_serving_summary_and_p99_interval():
    """serve.summarize: percentiles over all requests and the side <= 1024
    ones; the bootstrap interval of the p99 brackets the sample p99 of a
    large sample and narrows as the sample grows."""
    from yt_amd import serve

    rng = np.random.default_rng(4)
    lat = rng.exponential(10.0, size=4000)
    lo, hi = serve.p99_interval(lat)
    p99 = float(np.percentile(lat, 99))
    assert lo <= p99 <= hi
    lo2, hi2 = serve.p99_interval(lat[:200])
    assert hi2 - lo2 > hi - lo

    class S:
        def __init__(self, side):
            self.nv = (side, side)

    specs = [S(256), S(2048), S(512), S(1024)]
    res = dict(latency_s=np.array([0.001, 0.1, 0.002, 0.003]), wall_s=2.0, first_arrival=0.0,
               last_arrival=1.5, admission_waits=3)
    out = serve.summarize(res, specs)
    assert out["n"] == 4 and out["n_side_le_1024"] == 3
    assert out["p99_ms_side_le_1024"] < 3.0 + 1e-9 and out["p99_ms"] > 90.0
    assert abs(out["offered_images_per_s"] - 2.0) < 1e-12 and out["images_per_s"] == 2.0


def test_autotune_descent_finds_a_separable_optimum(monkeypatch, tmp_path):
    """Coordinate descent over the six launch knobs reaches the minimum of a
    separable cost (what the knobs' effects roughly are) with far fewer
    measurements than the grid, and the grid finds the same point."""
    from yt_amd import autotune

    target = {"tile": [4, 16], "item_tiles": 4, "tf_lds": "16k", "vertex_lds": True, "waves_per_group": 2,
              "groups_per_cu": 2}

    class FakeTimer:
        def __init__(self, ctx, specs, flags, reps, log):
            self.seen = {}

        def __call__(self, cfg):
            import json
            key = json.dumps(cfg, sort_keys=True)
            cost = 10.0 + sum(0.0 if cfg[k] == target[k] else 1.0 + 0.1 * i
                              for i, k in enumerate(autotune.KNOBS))
            self.seen[key] = cost
            return cost

        def close(self):
            pass

    class Ctx:
        device_name = "fake"

    monkeypatch.setattr(autotune, "_Timer", FakeTimer)
    best, table = autotune.tune(Ctx(), [], use_cache=False)
    assert best == target
    assert len(table) < 3 * 4 * 4 * 2 * 3 * 5 // 3
    gbest, gtable = autotune.tune(Ctx(), [], use_cache=False, search="grid")
    assert gbest == target and len(gtable) == 3 * 4 * 4 * 2 * 3 * 5
    with pytest.raises(ValueError):
        autotune.tune(Ctx(), [], use_cache=False, search="random")


def test_autotune_shares_timings_of_identical_launch_plans(monkeypatch):
    """Table-size values that give the same launches (the batch's launch plan
    has as many launches) share one batch and one timing; values that split
    a launch, and the vertex-block knob, are timed on batches of their own."""
    from yt_amd import _lib, autotune

    made, timed = [], []

    class FakeBatch:
        def __init__(self, ctx, specs, flags, tile=(8, 8), item_tiles=1, tf_lds="all", vertex_lds=False):
            self.key = (tuple(tile), item_tiles, tf_lds, vertex_lds)
            self.wpg = 4
            made.append(self.key)

        def launch_plan(self):
            # two distinct tables of 12 KB: split by "16k" and "one", not by "32k"
            return (3 if self.key[2] in ("16k", "one") else 2), 24 << 10

        def set_launch(self, wpg, gpc):
            self.wpg = wpg

        def run(self):
            timed.append((self.key, self.wpg))
            return 10.0 + (0.5 if self.key[2] in ("16k", "one") else 0.0) + 0.1 * self.wpg

        def close(self):
            pass

    class Ctx:
        device_name = "fake"

    monkeypatch.setattr(_lib, "Batch", FakeBatch)
    best, table = autotune.tune(Ctx(), [], tiles=((8, 8),), item_tiles=(1,), waves_per_group=(1, 4),
                                groups_per_cu=(0,), use_cache=False, search="grid", reps=1)
    assert len(table) == 4 * 2 * 2
    by = {(c["tf_lds"], c["vertex_lds"], c["waves_per_group"]): ms for c, ms in table}
    for v in (False, True):
        for w in (1, 4):
            assert by[("32k", v, w)] == by[("all", v, w)]
            assert by[("one", v, w)] == by[("16k", v, w)]
            assert by[("16k", v, w)] != by[("all", v, w)]
    # batches: one per (tf_lds, vertex_lds) built, the aliases closed; timings:
    # one per distinct plan x vertex knob x waves, each a warm run plus one
    assert len(made) == 8
    assert len(set(timed)) == 2 * 2 * 2 and len(timed) == 2 * len(set(timed))
    assert best["tf_lds"] == "all" and best["waves_per_group"] == 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    so, before anything touches a device."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120, cwd=root)
    assert res.returncode == 2
    assert "one GPU" in res.stderr and res.stdout == ""
