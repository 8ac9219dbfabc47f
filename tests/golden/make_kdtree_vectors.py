"""Writes golden vectors for the two brick-producer functions of the
reference that no closed form pins:

  kd-tree build and viewpoint order   amr_kdtools.pyx:201-262 (Node.add_grids),
                                       :760-823 (kdtree_get_choices),
                                       :565-622, 658-755 (depth / viewpoint traversal)
  no-ghost vertex centring            interpolators.pyx:177-226
                                       (ghost_zone_interpolate, rf=1,
                                       input_left 0.5, output_left 0 as
                                       grid_patch.py:284-296 calls it)

Both Cython modules need only numpy and libc.  This script cythonizes them
from the reference tree into tests/golden/_build/ (git-ignored, never
shipped), compiles them with gcc (-O2, no FMA contraction: yt's wheels are
built for baseline x86-64), imports them from there and writes
tests/golden/kdtree_vectors.npz and tests/golden/vertex_centre_vectors.npz.
Only the .npz files are committed.

Grid hierarchies are inserted level by level, each level's grids in blocks
order, as Tree.build does (amr_kdtree.py:100-118), with rank 0 of 1.

Run: python tests/golden/make_kdtree_vectors.py  (needs /root/reference)
"""

import importlib.util
import os
import subprocess
import sys
import sysconfig

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.environ.get("YT_REFERENCE", "/root/reference")
LIBDIR = os.path.join(REF, "yt", "utilities", "lib")
BUILD = os.path.join(HERE, "_build")


def build_module(name):
    """Cythonize LIBDIR/name.pyx into BUILD and compile it; returns the
    imported module."""
    os.makedirs(BUILD, exist_ok=True)
    pyx = os.path.join(LIBDIR, name + ".pyx")
    c = os.path.join(BUILD, name + ".c")
    so = os.path.join(BUILD, name + sysconfig.get_config_var("EXT_SUFFIX"))
    subprocess.run([sys.executable, "-m", "cython", "-3", "-I", REF, "-o", c, pyx], check=True)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-w",
                    "-I" + sysconfig.get_paths()["include"], "-I" + np.get_include(),
                    "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION", c, "-o", so], check=True)
    spec = importlib.util.spec_from_file_location(name, so)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def hierarchies():
    from tests.helpers import small_amr, uniform_lognormal
    from yt_amd import data

    yield "amr16", small_amr(root_n=16, levels=2, seed=11)
    yield "amr32", small_amr(root_n=32, levels=2, seed=7)
    yield "amr64x3", small_amr(root_n=64, levels=3, seed=3)
    yield "uniform8", uniform_lognormal(16, nprocs=8)
    yield "uniform64", uniform_lognormal(16, psize=(4, 4, 4))
    yield "psize343", uniform_lognormal(21, psize=(3, 4, 3))
    gd = data.make_amr_hierarchy(root_n=128, levels=3)
    yield "c2", data.load_amr_grids(gd, [128] * 3)


def viewpoints(splits, rng):
    """20 random points in and around the unit domain, 4 fixed ones, and 8
    lying exactly on split planes (the <= tie of step_viewpoint)."""
    vps = list(rng.uniform(-1.0, 2.0, size=(20, 3)))
    vps += [np.array([0.5, 0.5, 0.5]), np.array([0.25, 0.5, 0.75]),
            np.array([1e6, -1e6, 3.0]), np.array([0.0, 0.0, 0.0])]
    inner = [(d, p) for d, p in splits if d >= 0 and np.isfinite(p)]
    for _ in range(8):
        v = rng.uniform(0.0, 1.0, 3)
        for _ in range(2):
            d, p = inner[rng.integers(len(inner))]
            v[d] = p
        vps.append(v)
    return np.array(vps, dtype=np.float64)


def kd_vectors(kdt):
    out, cases = {}, []
    rng = np.random.default_rng(2024)
    for case, ds in hierarchies():
        cases.append(case)
        blocks = [g for g, _ in ds.blocks()]
        gle = np.array([g.LeftEdge for g in blocks], dtype=np.float64).reshape(-1, 3)
        gre = np.array([g.RightEdge for g in blocks], dtype=np.float64).reshape(-1, 3)
        gid = np.array([g.id for g in blocks], dtype=np.int64)
        lvl = np.array([g.Level for g in blocks], dtype=np.int32)
        trunk = kdt.Node(None, None, None, np.array([-np.inf] * 3), np.array([np.inf] * 3), -1, 1)
        for level in sorted(set(lvl.tolist())):
            sel = lvl == level
            trunk.add_grids(int(sel.sum()), np.ascontiguousarray(gle[sel]),
                            np.ascontiguousarray(gre[sel]), np.ascontiguousarray(gid[sel]), 0, 1)
        # depth_first_touch: each node once, in depth-first pre-order
        # (depth_traverse yields an internal node again on each return to it)
        nodes = list(trunk.depth_first_touch())
        node_id = np.array([n.node_id for n in nodes], dtype=np.int64)
        grid = np.array([n.grid for n in nodes], dtype=np.int64)
        sdim = np.array([n.get_split_dim() for n in nodes], dtype=np.int32)
        spos = np.array([n.get_split_pos() if n.get_split_dim() >= 0 else np.nan for n in nodes],
                        dtype=np.float64)
        nle = np.array([n.get_left_edge() for n in nodes], dtype=np.float64)
        nre = np.array([n.get_right_edge() for n in nodes], dtype=np.float64)
        leaf = np.array([n.kd_is_leaf() for n in nodes], dtype=np.int8)
        vps = viewpoints(list(zip(sdim.tolist(), spos.tolist())), rng)
        orders = np.array([[n.node_id for n in trunk.kd_traverse(viewpoint=vp)] for vp in vps],
                          dtype=np.int64)
        depth_leaves = np.array([n.node_id for n in trunk.kd_traverse()], dtype=np.int64)
        out.update({f"{case}/gle": gle, f"{case}/gre": gre, f"{case}/gid": gid,
                    f"{case}/level": lvl, f"{case}/node_id": node_id, f"{case}/grid": grid,
                    f"{case}/split_dim": sdim, f"{case}/split_pos": spos,
                    f"{case}/left_edge": nle, f"{case}/right_edge": nre, f"{case}/leaf": leaf,
                    f"{case}/viewpoints": vps, f"{case}/orders": orders,
                    f"{case}/leaves": depth_leaves})
        print(f"{case}: {len(blocks)} grids, {len(nodes)} nodes, {len(depth_leaves)} leaves",
              flush=True)
    out["cases"] = np.array(cases)
    return out


def vertex_vectors(interp):
    """ghost_zone_interpolate as get_vertex_centered_data(no_ghost=True)
    calls it, on seeded lognormal, signed and thin arrays."""
    rng = np.random.default_rng(77)
    shapes = [(2, 2, 2), (2, 3, 5), (3, 5, 4), (8, 8, 8), (16, 16, 32), (7, 2, 9)]
    out = {}
    for i, shp in enumerate(shapes):
        if i % 3 == 1:
            a = rng.normal(size=shp)
        else:
            a = np.exp(rng.normal(size=shp) * 1.5)
        a = np.ascontiguousarray(a, dtype=np.float64)
        o = np.zeros(tuple(s + 1 for s in shp), dtype=np.float64)
        interp.ghost_zone_interpolate(1, a, np.array([0.5, 0.5, 0.5]), o, np.array([0.0, 0.0, 0.0]))
        out[f"in{i}"] = a
        out[f"out{i}"] = o
    out["n"] = np.array(len(shapes))
    return out


def main():
    if not os.path.isdir(LIBDIR):
        raise SystemExit(f"reference tree not found at {REF}")
    kdt = build_module("amr_kdtools")
    interp = build_module("interpolators")
    np.savez_compressed(os.path.join(HERE, "kdtree_vectors.npz"), **kd_vectors(kdt))
    np.savez_compressed(os.path.join(HERE, "vertex_centre_vectors.npz"), **vertex_vectors(interp))
    print("wrote kdtree_vectors.npz, vertex_centre_vectors.npz", flush=True)


if __name__ == "__main__":
    main()
