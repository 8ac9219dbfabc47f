"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the FP64 CPU oracle.

oracle/ytrc_oracle.c restates yt's brick ray-caster (walk_volume,
ProjectionSampler, VolumeRenderSampler, FIT_*, offset_interpolate, the
plane-parallel lens and cast_through_kdtree) in plain C.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker or the CPU baseline.  The product package
(yt_amd) never imports it.

Parity pinning of each oracle function (see tests/test_oracle_*.py):
  walk_volume            exact-rational slab brute force, invariants
  ProjectionSampler      closed forms (sqrt(3) chord, constant field), linearity
  VolumeRenderSampler    closed form 1 - prod(1 - dt*e) for constant fields
  FIT_get_value          exact at nodes, linear between, 0 outside/NaN/0
  offset_interpolate     exact on linear fields, corner values, and bit-equal to
                         yt's own fixed_interpolator.cpp compiled into _ref/
  lens extent            conservativeness against the ray lattice
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libytrc_oracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libyt_fixed_interpolator.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        D = C.POINTER(C.c_double)
        I32 = C.POINTER(C.c_int32)
        L.orc_set_rows.restype = None
        L.orc_set_rows.argtypes = [C.c_int, C.c_int]
        L.orc_cast.restype = C.c_int
        L.orc_cast.argtypes = [
            C.c_int, D, D, D, D, D, D, D, C.c_int, C.c_int, C.c_int, D, D, I32, C.c_int,
            C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int, I32, C.c_void_p, C.c_int,
            D, D, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]
        L.orc_cast_lens.restype = C.c_int
        L.orc_cast_lens.argtypes = [C.c_int, C.c_int, D, D, D] + L.orc_cast.argtypes[1:]
        L.orc_extents_perspective.restype = C.c_int
        L.orc_extents_perspective.argtypes = [D, C.c_int, C.c_int, C.c_int, D, D, I32,
                                              C.POINTER(C.c_int64)]
        L.orc_extents.restype = C.c_int
        L.orc_extents.argtypes = [D, D, D, D, C.c_int, C.c_int, C.c_int, D, D, I32,
                                  C.POINTER(C.c_int64)]
        L.orc_walk_log.restype = C.c_int
        L.orc_walk_log.argtypes = [D, D, I32, D, D, C.c_double, I32, D, C.c_int]
        L.orc_fit_values.restype = C.c_int
        L.orc_fit_values.argtypes = [C.c_int, D, C.c_double, C.c_double, C.c_int, D, D]
        L.orc_trilinear.restype = C.c_int
        L.orc_trilinear.argtypes = [I32, D, C.c_int, I32, D, D]
        L.orc_max_threads.restype = C.c_int
        L.orc_gradients.restype = C.c_int
        L.orc_gradients.argtypes = [I32, D, C.c_int, I32, D, D]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _v3(x, n):
    a = np.ascontiguousarray(np.ravel(np.asarray(x, dtype=np.float64)))
    assert a.size == n, (a.size, n)
    return a


class _TFDesc(C.Structure):
    _fields_ = [
        ("n_fits", C.c_int32),
        ("nbins", C.c_int32 * 6),
        ("bounds", C.c_double * 12),
        ("y", C.POINTER(C.c_double) * 6),
        ("field_ids", C.c_int32 * 6),
        ("weight_field_ids", C.c_int32 * 6),
        ("weight_table_ids", C.c_int32 * 6),
        ("field_table_ids", C.c_int32 * 6),
        ("grey_opacity", C.c_int32),
        ("light_dir", C.c_double * 3),
        ("light_rgba", C.c_double * 4),
    ]


def cast(kind, params, bricks, order=None, tf=None, n_samples=10, image=None, zbuffer=None,
         num_threads=0, record=False, with_samples=False, rows=None):
    """Runs the oracle for one image.

    params: sampler parameters (vp_pos, vp_dir, center, bounds, x_vec, y_vec,
    width, nv).  bricks: list of dicts with le, re, dims, data (list of
    float64 arrays), mask (uint8 array or None).  tf: dict with tables
    [(y, b0, b1)], field_ids, weight_field_ids, weight_table_ids,
    field_table_ids, grey_opacity.  Returns image (nv0, nv1, 4) [, visits,
    hashes].  rows=(r0, r1) casts only the pixels of image rows r0 <= vi < r1
    (a bounded timing sample of a large image); the others keep their
    input values."""
    L = lib()
    nv0, nv1 = int(params["nv"][0]), int(params["nv"][1])
    if image is None:
        image = np.zeros((nv0, nv1, 4), dtype=np.float64)
    image = np.ascontiguousarray(image, dtype=np.float64)
    n = len(bricks)
    nf = len(bricks[0]["data"])
    les = np.ascontiguousarray([b["le"] for b in bricks], dtype=np.float64)
    res = np.ascontiguousarray([b["re"] for b in bricks], dtype=np.float64)
    dims = np.ascontiguousarray([b["dims"] for b in bricks], dtype=np.int32)
    keep = []
    fptr = (C.c_void_p * (n * nf))()
    mptr = (C.c_void_p * n)()
    for bi, b in enumerate(bricks):
        for f in range(nf):
            a = np.ascontiguousarray(b["data"][f], dtype=np.float64)
            keep.append(a)
            fptr[bi * nf + f] = a.ctypes.data
        m = b.get("mask")
        if m is not None:
            m = np.ascontiguousarray(m, dtype=np.uint8)
            keep.append(m)
            mptr[bi] = m.ctypes.data
        else:
            mptr[bi] = None
    if order is None:
        order = np.arange(n, dtype=np.int32)
    order = np.ascontiguousarray(order, dtype=np.int32)
    tfd = None
    if kind in (1, 3):
        tfd = _TFDesc()
        tables = tf["tables"]
        tfd.n_fits = len(tables)
        for i, (y, b0, b1) in enumerate(tables):
            yy = np.ascontiguousarray(y, dtype=np.float64)
            keep.append(yy)
            tfd.nbins[i] = len(yy)
            tfd.bounds[2 * i] = b0
            tfd.bounds[2 * i + 1] = b1
            tfd.y[i] = _d(yy)
        for i in range(6):
            tfd.field_ids[i] = tf["field_ids"][i]
            tfd.weight_field_ids[i] = tf["weight_field_ids"][i]
            tfd.weight_table_ids[i] = tf["weight_table_ids"][i]
            tfd.field_table_ids[i] = tf["field_table_ids"][i]
        tfd.grey_opacity = int(tf.get("grey_opacity", 0))
        tfd.light_dir[:] = [float(x) for x in tf.get("light_dir", (0.0, 0.0, 0.0))]
        tfd.light_rgba[:] = [float(x) for x in tf.get("light_rgba", (0.0, 0.0, 0.0, 0.0))]
    zb = None
    if zbuffer is not None:
        zb = np.ascontiguousarray(zbuffer, dtype=np.float64)
    visits = hashes = samples = None
    if record:
        visits = np.zeros((nv0, nv1), dtype=np.int64)
        hashes = np.zeros((nv0, nv1), dtype=np.uint64)
        samples = np.zeros((nv0, nv1), dtype=np.int64)
    lens = 0 if params.get("lens_type", "plane-parallel") == "plane-parallel" else 1
    pix_pos = pix_dir = cam = None
    if lens:
        pix_pos = np.ascontiguousarray(params["vp_pos"], dtype=np.float64).reshape(nv0, nv1, 3)
        pix_dir = np.ascontiguousarray(params["vp_dir"], dtype=np.float64).reshape(nv0, nv1, 3)
        if params.get("lens_type") == "perspective" and params.get("camera_data") is not None:
            cam = _v3(params["camera_data"], 15)
        vp_pos = np.zeros(12)
        vp_dir = np.zeros(3)
    else:
        vp_pos = _v3(params["vp_pos"], 12)
        vp_dir = _v3(params["vp_dir"], 3)
    center = _v3(params["center"], 3)
    bounds = _v3(params["bounds"], 4)
    x_vec = _v3(params["x_vec"], 3)
    y_vec = _v3(params["y_vec"], 3)
    width = _v3(params["width"], 3)
    L.orc_set_rows(*((0, 0x7fffffff) if rows is None else (int(rows[0]), int(rows[1]))))
    rc = L.orc_cast_lens(
        int(kind), lens, _d(pix_pos) if lens else None, _d(pix_dir) if lens else None,
        _d(cam) if cam is not None else None,
        _d(vp_pos), _d(vp_dir), _d(center), _d(bounds), _d(x_vec), _d(y_vec),
        _d(width), nv0, nv1, n, _d(les), _d(res), _i(dims), nf, fptr, mptr, len(order),
        _i(order), C.cast(C.pointer(tfd), C.c_void_p) if tfd is not None else None,
        int(n_samples), _d(image), _d(zb) if zb is not None else None, int(num_threads),
        visits.ctypes.data_as(C.POINTER(C.c_int64)) if record else None,
        hashes.ctypes.data_as(C.POINTER(C.c_uint64)) if record else None,
        samples.ctypes.data_as(C.POINTER(C.c_int64)) if record else None)
    if rows is not None:
        L.orc_set_rows(0, 0x7fffffff)
    if rc != 0:
        raise ValueError(f"orc_cast failed ({rc})")
    if record and with_samples:
        return image, visits, hashes, samples
    if record:
        return image, visits, hashes
    return image


def extents(params, bricks):
    L = lib()
    n = len(bricks)
    les = np.ascontiguousarray([b["le"] for b in bricks], dtype=np.float64)
    res = np.ascontiguousarray([b["re"] for b in bricks], dtype=np.float64)
    dims = np.ascontiguousarray([b["dims"] for b in bricks], dtype=np.int32)
    out = np.zeros((n, 4), dtype=np.int64)
    center = _v3(params["center"], 3)
    bounds = _v3(params["bounds"], 4)
    x_vec = _v3(params["x_vec"], 3)
    y_vec = _v3(params["y_vec"], 3)
    L.orc_extents(_d(center), _d(bounds), _d(x_vec), _d(y_vec), int(params["nv"][0]),
                  int(params["nv"][1]), n, _d(les), _d(res), _i(dims),
                  out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


def extents_perspective(camera_data, nv, bricks):
    """Clipped calculate_extent_perspective rectangles, n_bricks x 4."""
    L = lib()
    n = len(bricks)
    les = np.ascontiguousarray([b["le"] for b in bricks], dtype=np.float64)
    res = np.ascontiguousarray([b["re"] for b in bricks], dtype=np.float64)
    dims = np.ascontiguousarray([b["dims"] for b in bricks], dtype=np.int32)
    out = np.zeros((n, 4), dtype=np.int64)
    cam = _v3(camera_data, 15)
    L.orc_extents_perspective(_d(cam), int(nv[0]), int(nv[1]), n, _d(les), _d(res), _i(dims),
                              out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


def walk_log(le, re, dims, v_pos, v_dir, max_t=1.0, cap=4096):
    """(hit_count, idx (n,3) int32, t (n,2) float64) of one ray through one brick."""
    L = lib()
    le = _v3(le, 3)
    re = _v3(re, 3)
    d = np.ascontiguousarray(dims, dtype=np.int32)
    p = _v3(v_pos, 3)
    v = _v3(v_dir, 3)
    idx = np.zeros((cap, 3), dtype=np.int32)
    t = np.zeros((cap, 2), dtype=np.float64)
    hit = L.orc_walk_log(_d(le), _d(re), _i(d), _d(p), _d(v), float(max_t), _i(idx), _d(t), cap)
    k = min(hit, cap)
    return hit, idx[:k].copy(), t[:k].copy()


def fit_values(y, b0, b1, v):
    L = lib()
    y = np.ascontiguousarray(y, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(v)
    L.orc_fit_values(len(y), _d(y), float(b0), float(b1), v.size, _d(v), _d(out))
    return out


def trilinear(dims, vdata, cells, dps):
    L = lib()
    d = np.ascontiguousarray(dims, dtype=np.int32)
    vd = np.ascontiguousarray(vdata, dtype=np.float64)
    c = np.ascontiguousarray(cells, dtype=np.int32)
    p = np.ascontiguousarray(dps, dtype=np.float64)
    out = np.zeros(len(c), dtype=np.float64)
    L.orc_trilinear(_i(d), _d(vd), len(c), _i(c), _d(p), _d(out))
    return out


def gradients(dims, vdata, cells, dps):
    """eval_gradient (fixed_interpolator.cpp:116-148) at each (cell, dp)."""
    L = lib()
    d = np.ascontiguousarray(dims, dtype=np.int32)
    vd = np.ascontiguousarray(vdata, dtype=np.float64)
    c = np.ascontiguousarray(cells, dtype=np.int32)
    p = np.ascontiguousarray(dps, dtype=np.float64)
    out = np.zeros((len(c), 3), dtype=np.float64)
    L.orc_gradients(_i(d), _d(vd), len(c), _i(c), _d(p), _d(out))
    return out


def ref_eval_gradient(dims, vdata, cells, dps):
    """yt's eval_gradient (fixed_interpolator.cpp:116-148) from oracle/_ref."""
    L = ref_lib()
    ds = (C.c_int * 3)(*[int(x) for x in dims])
    vd = np.ascontiguousarray(vdata, dtype=np.float64).ravel()
    out = np.zeros((len(cells), 3))
    for n, (c, dp) in enumerate(zip(cells, dps)):
        off = (int(c[0]) * (dims[1] + 1) + int(c[1])) * (dims[2] + 1) + int(c[2])
        p = (C.c_double * 3)(*[float(x) for x in dp])
        g = (C.c_double * 3)()
        L._Z13eval_gradientPiPdS0_S0_(ds, p, _d(vd[off:]), g)
        out[n] = list(g)
    return out


def max_threads():
    return lib().orc_max_threads()


_ref = None


def ref_lib():
    """yt's fixed_interpolator.cpp built from the reference tree (oracle/Makefile
    target ``ref``), or None where it was not built.  A checker only."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        L = C.CDLL(REF_LIB_PATH)
        # C++ symbols of fixed_interpolator.hpp:17-24 (g++ name mangling)
        f = L._Z18offset_interpolatePiPdS0_
        f.restype = C.c_double
        f.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        g = L._Z13eval_gradientPiPdS0_S0_
        g.restype = None
        g.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                      C.POINTER(C.c_double)]
        _ref = L
    return _ref


def ref_offset_interpolate(dims, vdata, cells, dps):
    """yt's offset_interpolate (fixed_interpolator.cpp:34-52) at each cell's
    offset (i*(d1+1)+j)*(d2+1)+k, as image_samplers.pyx:465-466 computes it."""
    L = ref_lib()
    ds = (C.c_int * 3)(*[int(x) for x in dims])
    vd = np.ascontiguousarray(vdata, dtype=np.float64).ravel()
    out = np.empty(len(cells))
    for n, (c, dp) in enumerate(zip(cells, dps)):
        off = (int(c[0]) * (dims[1] + 1) + int(c[1])) * (dims[2] + 1) + int(c[2])
        p = (C.c_double * 3)(*[float(x) for x in dp])
        out[n] = L._Z18offset_interpolatePiPdS0_(ds, p, _d(vd[off:]))
    return out
