"""ctypes binding of libytrc.so (include/ytrc.h).

The product path has no CPU fallback: importing this module succeeds without
the library, but every entry point raises YtrcLibraryError when libytrc.so
is absent or cannot be loaded.
"""

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("YTRC_LIB", os.path.join(HERE, "libytrc.so"))

YTRC_OK = 0
YTRC_EINVAL = -1
YTRC_EDEVICE = -2
YTRC_EBUDGET = -3
YTRC_ENOTIMPL = -4
YTRC_ECANCELLED = -5
YTRC_EBUSY = -6
YTRC_ETIMEDOUT = -7

PROJECTION = 0
VOLUME_RENDER = 1
INTERPOLATED_PROJECTION = 2
LIGHT_SOURCE = 3
LENS_PLANE_PARALLEL = 0
LENS_PER_PIXEL = 1

FLAG_RECORD = 1
FLAG_ERT = 2
FLAG_PARTIAL = 4

BATCH_PLAIN = 0
BATCH_PERSISTENT = 1
BATCH_GRAPH = 2
# rays per 64-lane tile, rows x columns
BATCH_TILE = {(8, 8): 0, (4, 16): 16, (16, 4): 32}
BATCH_ITEM_TILES = {1: 0, 2: 64, 4: 128, 8: 192}
# transfer-function table bytes one persistent launch stages in LDS:
# "all" (one launch per kernel variant), at most 32 or 16 KiB, one table
BATCH_TF_LDS = {"all": 0, "32k": 256, "16k": 512, "one": 768}
# persistent volume renders (ColorTransferFunction, one field, 10 sub-samples)
# stage vertex blocks in LDS where a wave's lanes share one
BATCH_VERTEX_LDS = 1024

EXPORTED = [
    "ytrc_ctx_create",
    "ytrc_ctx_destroy",
    "ytrc_ctx_used_bytes",
    "ytrc_last_error",
    "ytrc_dataset_upload",
    "ytrc_dataset_free",
    "ytrc_dataset_bytes",
    "ytrc_batch_create",
    "ytrc_batch_launch",
    "ytrc_batch_sync",
    "ytrc_batch_read_image",
    "ytrc_batch_read_record",
    "ytrc_batch_read_partial",
    "ytrc_batch_set_launch",
    "ytrc_batch_bytes",
    "ytrc_batch_launch_plan",
    "ytrc_batch_destroy",
    "ytrc_batch_kernel_ms",
    "ytrc_batch_request_ms",
    "ytrc_hbm_probe",
    "ytrc_kdtree_build",
    "ytrc_kdtree_size",
    "ytrc_kdtree_read_nodes",
    "ytrc_kdtree_order",
    "ytrc_kdtree_destroy",
    "ytrc_server_create",
    "ytrc_server_submit",
    "ytrc_server_progress",
    "ytrc_server_wait",
    "ytrc_server_read",
    "ytrc_server_read_planar",
    "ytrc_server_cancel",
    "ytrc_server_set_policy",
    "ytrc_ctx_max_lds",
    "ytrc_ctx_device_name",
    "ytrc_render_noalloc",
    "ytrc_batch_sync_for",
    "ytrc_batch_cancel",
    "ytrc_server_wait_for",
    "ytrc_server_stats",
    "ytrc_server_destroy",
    "ytrc_dataset_build_vr",
    "ytrc_dataset_read_points",
]


class YtrcLibraryError(RuntimeError):
    """libytrc.so is missing or failed to load: there is no CPU fallback."""


class YtrcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ytrc error {code}: {msg}")
        self.code = code


class YtrcBudgetError(YtrcError):
    """The request does not fit the context's HBM budget."""


class YtrcCancelled(YtrcError):
    """The launch was stopped (Batch.cancel); its images are incomplete."""


class YtrcTimeout(YtrcError, TimeoutError):
    """A wait's deadline passed before the work finished."""


class BrickDesc(C.Structure):
    _fields_ = [
        ("left_edge", C.c_double * 3),
        ("right_edge", C.c_double * 3),
        ("dims", C.c_int32 * 3),
        ("vertex_centred", C.c_int32),
    ]


class GridDesc(C.Structure):
    _fields_ = [
        ("left_edge", C.c_double * 3),
        ("right_edge", C.c_double * 3),
        ("dims", C.c_int32 * 3),
        ("pad", C.c_int32),
    ]


class LeafDesc(C.Structure):
    _fields_ = [
        ("grid", C.c_int32),
        ("li", C.c_int32 * 3),
        ("ri", C.c_int32 * 3),
        ("pad", C.c_int32),
        ("left_edge", C.c_double * 3),
        ("right_edge", C.c_double * 3),
    ]


class KdNode(C.Structure):
    """ytrc_kd_node."""
    _fields_ = [
        ("left_edge", C.c_double * 3),
        ("right_edge", C.c_double * 3),
        ("split_pos", C.c_double),
        ("grid", C.c_int64),
        ("node_id", C.c_int64),
        ("parent", C.c_int32),
        ("left", C.c_int32),
        ("right", C.c_int32),
        ("split_dim", C.c_int32),
    ]


class TFDesc(C.Structure):
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
    ]


class Request(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("lens", C.c_int32),
        ("nv", C.c_int32 * 2),
        ("vp_pos", C.c_double * 12),
        ("vp_dir", C.c_double * 3),
        ("center", C.c_double * 3),
        ("bounds", C.c_double * 4),
        ("x_vec", C.c_double * 3),
        ("y_vec", C.c_double * 3),
        ("width", C.c_double * 3),
        ("n_samples", C.c_int32),
        ("ds_id", C.c_int32),
        ("n_order", C.c_int32),
        ("order", C.POINTER(C.c_int32)),
        ("zbuffer", C.POINTER(C.c_double)),
        ("image_in", C.POINTER(C.c_double)),
        ("tf", C.POINTER(TFDesc)),
        ("flags", C.c_uint32),
        ("pix_pos", C.POINTER(C.c_double)),
        ("pix_dir", C.POINTER(C.c_double)),
        ("camera_data", C.POINTER(C.c_double)),
        ("light_dir", C.c_double * 3),
        ("light_rgba", C.c_double * 4),
    ]


_lib = None
_lib_lock = threading.Lock()


def _declare(lib):
    P = C.c_void_p
    lib.ytrc_ctx_create.argtypes = [C.c_int32, C.c_uint64, C.POINTER(P)]
    lib.ytrc_ctx_destroy.argtypes = [P]
    lib.ytrc_ctx_destroy.restype = None
    lib.ytrc_ctx_used_bytes.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.ytrc_last_error.restype = C.c_char_p
    lib.ytrc_last_error.argtypes = []
    lib.ytrc_dataset_upload.argtypes = [
        P, C.c_int32, C.POINTER(BrickDesc), C.c_int32, C.c_int32,
        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_int32)]
    lib.ytrc_dataset_build_vr.argtypes = [
        P, C.c_int32, C.POINTER(GridDesc), C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
        C.POINTER(C.c_int32), C.c_int32, C.POINTER(LeafDesc), C.POINTER(C.c_void_p),
        C.POINTER(C.c_int32)]
    lib.ytrc_dataset_read_points.argtypes = [
        P, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_int64)]
    lib.ytrc_dataset_free.argtypes = [P, C.c_int32]
    lib.ytrc_dataset_bytes.argtypes = [P, C.c_int32, C.POINTER(C.c_uint64)]
    lib.ytrc_batch_create.argtypes = [P, C.c_int32, C.POINTER(Request), C.c_uint32, C.POINTER(P)]
    lib.ytrc_batch_launch.argtypes = [P]
    lib.ytrc_batch_sync.argtypes = [P, C.POINTER(C.c_float)]
    lib.ytrc_batch_sync_for.argtypes = [P, C.c_double, C.POINTER(C.c_float)]
    lib.ytrc_batch_cancel.argtypes = [P]
    lib.ytrc_batch_read_image.argtypes = [P, C.c_int32, C.POINTER(C.c_double)]
    lib.ytrc_batch_read_partial.argtypes = [P, C.c_int32, C.POINTER(C.c_double)]
    lib.ytrc_kdtree_build.argtypes = [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(P)]
    lib.ytrc_kdtree_size.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.ytrc_kdtree_read_nodes.argtypes = [P, C.POINTER(KdNode)]
    lib.ytrc_kdtree_order.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    lib.ytrc_kdtree_destroy.argtypes = [P]
    lib.ytrc_kdtree_destroy.restype = None
    lib.ytrc_batch_read_record.argtypes = [
        P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]
    lib.ytrc_batch_set_launch.argtypes = [P, C.c_int32, C.c_int32]
    lib.ytrc_batch_bytes.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.ytrc_batch_launch_plan.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
    lib.ytrc_batch_destroy.argtypes = [P]
    lib.ytrc_batch_destroy.restype = None
    lib.ytrc_batch_kernel_ms.argtypes = [P, C.POINTER(C.c_float), C.POINTER(C.c_int32)]
    lib.ytrc_batch_request_ms.argtypes = [P, C.c_int32, C.POINTER(C.c_float)]
    lib.ytrc_hbm_probe.argtypes = [P, C.c_uint64, C.c_int32, C.POINTER(C.c_float),
                                   C.POINTER(C.c_float)]
    I64P = C.POINTER(C.c_int64)
    lib.ytrc_server_create.argtypes = [P, C.c_int32, C.c_uint64, C.c_int32, C.POINTER(P)]
    lib.ytrc_server_submit.argtypes = [P, C.POINTER(Request), I64P]
    lib.ytrc_server_progress.argtypes = [P, I64P, C.c_int32, C.POINTER(C.c_int32)]
    lib.ytrc_server_wait.argtypes = [P, I64P, C.c_int32, C.POINTER(C.c_int32)]
    lib.ytrc_server_wait_for.argtypes = [P, C.c_double, I64P, C.c_int32, C.POINTER(C.c_int32)]
    lib.ytrc_server_read.argtypes = [P, C.c_int64, C.POINTER(C.c_double)]
    lib.ytrc_server_read_planar.argtypes = [P, C.c_int64, C.POINTER(C.c_float), C.c_int32,
                                            C.POINTER(C.c_int32)]
    lib.ytrc_server_cancel.argtypes = [P, C.c_int64]
    lib.ytrc_server_set_policy.argtypes = [P, C.c_int32, C.c_double]
    lib.ytrc_ctx_max_lds.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.ytrc_ctx_device_name.argtypes = [P, C.c_char_p, C.c_int32]
    lib.ytrc_render_noalloc.argtypes = [P, C.POINTER(Request), C.POINTER(C.c_double)]
    lib.ytrc_server_stats.argtypes = [P, I64P]
    lib.ytrc_server_destroy.argtypes = [P]
    lib.ytrc_server_destroy.restype = None
    for name in EXPORTED:
        f = getattr(lib, name)
        if name not in ("ytrc_ctx_destroy", "ytrc_batch_destroy", "ytrc_last_error",
                        "ytrc_server_destroy"):
            f.restype = C.c_int


def lib():
    """Loads libytrc.so once; raises YtrcLibraryError if it is unavailable."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise YtrcLibraryError(
                    f"{LIB_PATH} not found: build it with yt_amd.build.build() "
                    "(there is no CPU fallback for the ray-caster)")
            try:
                handle = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
            except OSError as e:
                raise YtrcLibraryError(f"cannot load {LIB_PATH}: {e}") from e
            _declare(handle)
            _lib = handle
        return _lib


def _check(rc):
    if rc != YTRC_OK:
        msg = lib().ytrc_last_error().decode(errors="replace")
        if rc == YTRC_EBUDGET:
            raise YtrcBudgetError(rc, msg)
        if rc == YTRC_ENOTIMPL:
            raise NotImplementedError(msg)
        if rc == YTRC_EINVAL:
            raise ValueError(msg)
        if rc == YTRC_ECANCELLED:
            raise YtrcCancelled(rc, msg)
        if rc == YTRC_ETIMEDOUT:
            raise YtrcTimeout(rc, msg)
        raise YtrcError(rc, msg)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class KdTree:
    """The native kd-tree of a grid hierarchy (ytrc_kdtree_*): host code,
    no device needed.  left_edges / right_edges: (n, 3); ids: grid ids
    (>= 0); levels: grid levels."""

    def __init__(self, left_edges, right_edges, ids, levels):
        le = np.ascontiguousarray(left_edges, dtype=np.float64).reshape(-1, 3)
        re = np.ascontiguousarray(right_edges, dtype=np.float64).reshape(-1, 3)
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        levels = np.ascontiguousarray(levels, dtype=np.int32)
        n = len(ids)
        if not (len(le) == len(re) == len(levels) == n):
            raise ValueError("KdTree: one edge pair, id and level per grid")
        self._h = C.c_void_p()
        _check(lib().ytrc_kdtree_build(
            n, _dptr(le), _dptr(re), ids.ctypes.data_as(C.POINTER(C.c_int64)),
            levels.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(self._h)))
        nn, nl = C.c_int32(), C.c_int32()
        _check(lib().ytrc_kdtree_size(self._h, C.byref(nn), C.byref(nl)))
        self.n_nodes, self.n_leaves = nn.value, nl.value

    def nodes(self):
        out = (KdNode * self.n_nodes)()
        _check(lib().ytrc_kdtree_read_nodes(self._h, out))
        return list(out)

    def order(self, viewpoint):
        vp = np.ascontiguousarray(viewpoint, dtype=np.float64).ravel()[:3]
        out = np.empty(self.n_leaves, dtype=np.int32)
        _check(lib().ytrc_kdtree_order(self._h, _dptr(vp),
                                       out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def close(self):
        if self._h:
            lib().ytrc_kdtree_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device context: a stream, resident datasets and an HBM budget."""

    def __init__(self, device=0, budget_bytes=0):
        self._h = C.c_void_p()
        _check(lib().ytrc_ctx_create(int(device), int(budget_bytes), C.byref(self._h)))
        self.device = device

    def close(self):
        if self._h:
            lib().ytrc_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def used_bytes(self):
        u, b = C.c_uint64(), C.c_uint64()
        _check(lib().ytrc_ctx_used_bytes(self._h, C.byref(u), C.byref(b)))
        return u.value, b.value

    @property
    def max_lds(self):
        """LDS bytes one workgroup may allocate: the bound on the
        transfer-function tables of one launch."""
        b = C.c_uint64()
        _check(lib().ytrc_ctx_max_lds(self._h, C.byref(b)))
        return b.value

    @property
    def device_name(self):
        """The device's name, ISA and CU count (labels tuning caches)."""
        buf = C.create_string_buffer(256)
        _check(lib().ytrc_ctx_device_name(self._h, buf, len(buf)))
        return buf.value.decode(errors="replace")

    def render_noalloc(self, spec):
        """One request through ytrc_render_noalloc (no device allocation:
        the library's fixed per-device array); returns its (nv0, nv1, 4)
        float64 image."""
        out = np.empty((spec.nv[0], spec.nv[1], 4), dtype=np.float64)
        r = spec.to_c()
        _check(lib().ytrc_render_noalloc(self._h, C.byref(r), _dptr(out)))
        return out

    def hbm_probe(self, nbytes=1 << 30, iters=20):
        """(GB/s, ms per copy) of a streaming device copy of nbytes."""
        g, ms = C.c_float(), C.c_float()
        _check(lib().ytrc_hbm_probe(self._h, int(nbytes), int(iters), C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def upload(self, bricks, fields, masks=None):
        """bricks: list of (left_edge, right_edge, dims, vertex_centred);
        fields: list (per brick) of lists of arrays (float64 or float32, all
        the same dtype); masks: None or list of uint8 arrays / None."""
        n = len(bricks)
        if n == 0:
            raise ValueError("upload: no bricks")
        nf = len(fields[0])
        descs = (BrickDesc * n)()
        keep = []
        ptrs = (C.c_void_p * (n * nf))()
        dt = None
        for b, (le, re, dims, vc) in enumerate(bricks):
            descs[b].left_edge[:] = [float(x) for x in le]
            descs[b].right_edge[:] = [float(x) for x in re]
            descs[b].dims[:] = [int(x) for x in dims]
            descs[b].vertex_centred = int(bool(vc))
            if len(fields[b]) != nf:
                raise ValueError("upload: every brick needs the same number of fields")
            shape = tuple(int(d) + int(bool(vc)) for d in dims)
            for f, arr in enumerate(fields[b]):
                a = np.asarray(arr)
                if a.dtype not in (np.float64, np.float32):
                    a = a.astype(np.float64)
                if dt is None:
                    dt = a.dtype
                a = np.ascontiguousarray(a, dtype=dt)
                if a.shape != shape:
                    raise ValueError(f"upload: brick {b} field {f} has shape {a.shape}, expected {shape}")
                keep.append(a)
                ptrs[b * nf + f] = a.ctypes.data
        mptr = None
        if masks is not None:
            mptr = (C.c_void_p * n)()
            for b, m in enumerate(masks):
                if m is None:
                    mptr[b] = None
                    continue
                mm = np.ascontiguousarray(m, dtype=np.uint8)
                if mm.shape != tuple(int(d) for d in bricks[b][2]):
                    raise ValueError(f"upload: mask {b} has shape {mm.shape}")
                keep.append(mm)
                mptr[b] = mm.ctypes.data
        ds = C.c_int32()
        _check(lib().ytrc_dataset_upload(
            self._h, n, descs, nf, 0 if dt == np.float64 else 1, ptrs,
            C.cast(mptr, C.POINTER(C.c_void_p)) if mptr is not None else None, C.byref(ds)))
        return ds.value

    def build_vr(self, grids, cells, log_fields, leaves, masks=None):
        """Device brick preparation (ytrc_dataset_build_vr).  grids: list of
        (left_edge, right_edge, dims); cells: per grid, a list of C-order
        dims-shaped cell arrays (one per field, all float64 or all float32);
        log_fields: per field, log10 or not; leaves: list of (grid index,
        li, ri, left_edge, right_edge); masks: None or per-leaf uint8 arrays
        of shape ri - li (None = all ones)."""
        ng, nl = len(grids), len(leaves)
        if ng == 0 or nl == 0:
            raise ValueError("build_vr: no grids or no leaves")
        nf = len(cells[0])
        gd = (GridDesc * ng)()
        keep = []
        ptrs = (C.c_void_p * (ng * nf))()
        dt = None
        for g, (le, re, dims) in enumerate(grids):
            gd[g].left_edge[:] = [float(x) for x in le]
            gd[g].right_edge[:] = [float(x) for x in re]
            gd[g].dims[:] = [int(x) for x in dims]
            if len(cells[g]) != nf:
                raise ValueError("build_vr: every grid needs the same number of fields")
            for f, arr in enumerate(cells[g]):
                a = np.asarray(arr)
                if dt is None:
                    dt = np.float32 if a.dtype == np.float32 else np.float64
                a = np.ascontiguousarray(a, dtype=dt)
                if a.shape != tuple(int(d) for d in dims):
                    raise ValueError(f"build_vr: grid {g} field {f} has shape {a.shape}")
                keep.append(a)
                ptrs[g * nf + f] = a.ctypes.data
        ld = (LeafDesc * nl)()
        for l, (g, li, ri, le, re) in enumerate(leaves):
            ld[l].grid = int(g)
            ld[l].li[:] = [int(x) for x in li]
            ld[l].ri[:] = [int(x) for x in ri]
            ld[l].left_edge[:] = [float(x) for x in le]
            ld[l].right_edge[:] = [float(x) for x in re]
        mptr = None
        if masks is not None:
            mptr = (C.c_void_p * nl)()
            for l, m in enumerate(masks):
                if m is None:
                    continue
                mm = np.ascontiguousarray(m, dtype=np.uint8)
                want = tuple(int(r) - int(i) for i, r in zip(leaves[l][1], leaves[l][2]))
                if mm.shape != want:
                    raise ValueError(f"build_vr: mask {l} has shape {mm.shape}, expected {want}")
                keep.append(mm)
                mptr[l] = mm.ctypes.data
        logs = (C.c_int32 * nf)(*[int(bool(x)) for x in log_fields])
        ds = C.c_int32()
        _check(lib().ytrc_dataset_build_vr(
            self._h, ng, gd, nf, 0 if dt == np.float64 else 1, ptrs, logs, nl, ld,
            C.cast(mptr, C.POINTER(C.c_void_p)) if mptr is not None else None, C.byref(ds)))
        return ds.value

    def read_points(self, ds_id, brick):
        """A brick's stored float32 points, fields interleaved, storage order."""
        n = C.c_int64()
        _check(lib().ytrc_dataset_read_points(self._h, int(ds_id), int(brick), None, 0, C.byref(n)))
        out = np.empty(n.value, dtype=np.float32)
        _check(lib().ytrc_dataset_read_points(
            self._h, int(ds_id), int(brick), out.ctypes.data_as(C.POINTER(C.c_float)), n.value,
            C.byref(n)))
        return out

    def free(self, ds_id):
        _check(lib().ytrc_dataset_free(self._h, int(ds_id)))

    def dataset_bytes(self, ds_id):
        v = C.c_uint64()
        _check(lib().ytrc_dataset_bytes(self._h, int(ds_id), C.byref(v)))
        return v.value


class TF:
    """Host copy of a transfer function in the C-ABI layout."""

    def __init__(self, tables, field_ids, weight_field_ids, weight_table_ids,
                 field_table_ids, grey_opacity):
        n = len(tables)
        if n > 6:
            raise ValueError("at most 6 field tables (image_samplers.pyx:430)")
        self._ys = [np.ascontiguousarray(y, dtype=np.float64) for (y, _, _) in tables]
        d = TFDesc()
        d.n_fits = n
        for i, (y, b0, b1) in enumerate(tables):
            d.nbins[i] = len(self._ys[i])
            d.bounds[2 * i] = float(b0)
            d.bounds[2 * i + 1] = float(b1)
            d.y[i] = _dptr(self._ys[i])
        for i in range(6):
            d.field_ids[i] = int(field_ids[i])
            d.weight_field_ids[i] = int(weight_field_ids[i])
            d.weight_table_ids[i] = int(weight_table_ids[i])
            d.field_table_ids[i] = int(field_table_ids[i])
        d.grey_opacity = int(bool(grey_opacity))
        self.desc = d

    @classmethod
    def from_tf_obj(cls, tf_obj):
        """Takes the attributes VolumeRenderSampler reads from tf_obj
        (image_samplers.pyx:429-446)."""
        n = tf_obj.n_field_tables
        tables = [(tf_obj.tables[i].y, tf_obj.tables[i].x_bounds[0],
                   tf_obj.tables[i].x_bounds[1]) for i in range(n)]
        return cls(tables, tf_obj.field_ids, tf_obj.weight_field_ids, tf_obj.weight_table_ids,
                   tf_obj.field_table_ids, getattr(tf_obj, "grey_opacity", 0))


class RequestSpec:
    """One image: sampler constructor parameters + brick order."""

    def __init__(self, kind, params, ds_id, order, n_samples=10, tf=None, zbuffer=None,
                 image_in=None, record=False, ert=False, light_dir=(1.0, 1.0, 1.0),
                 light_rgba=(1.0, 1.0, 1.0, 1.0), partial=False):
        self.kind = kind
        # LightSourceRenderSampler normalises the light direction
        # (image_samplers.pyx:526-528).
        ld = np.asarray(light_dir, dtype=np.float64).ravel()[:3]
        self.light_dir = ld / np.sqrt(ld[0] * ld[0] + ld[1] * ld[1] + ld[2] * ld[2])
        self.light_rgba = np.asarray(light_rgba, dtype=np.float64).ravel()[:4]
        self.params = params
        self.ds_id = ds_id
        self.order = np.ascontiguousarray(order, dtype=np.int32)
        self.n_samples = int(n_samples)
        self.tf = tf
        self.nv = (int(params["nv"][0]), int(params["nv"][1]))
        self.zbuffer = None if zbuffer is None else np.ascontiguousarray(zbuffer, dtype=np.float64)
        self.image_in = None if image_in is None else np.ascontiguousarray(image_in, dtype=np.float64)
        self.record = record
        self.ert = ert
        self.partial = partial
        # Per-pixel lenses (perspective, fisheye, ...): vp_pos / vp_dir are
        # (nv0, nv1, 3) arrays (image_samplers.pyx:89-96).
        self.lens_type = params.get("lens_type", "plane-parallel")
        self.pix_pos = self.pix_dir = self.camera_data = None
        if self.lens_type != "plane-parallel":
            pos = np.asarray(params["vp_pos"], dtype=np.float64)
            dirs = np.asarray(params["vp_dir"], dtype=np.float64)
            if pos.shape[:2] != self.nv or dirs.shape[:2] != self.nv:
                raise RuntimeError(
                    f"Bad lens shape / direction for {self.lens_type}\n"
                    f"Shapes: ({pos.shape[0]} - {dirs.shape[0]} - {self.nv[0]}) and "
                    f"({pos.shape[1]} - {dirs.shape[1]} - {self.nv[1]})")
            self.pix_pos = np.ascontiguousarray(pos.reshape(self.nv[0], self.nv[1], 3))
            self.pix_dir = np.ascontiguousarray(dirs.reshape(self.nv[0], self.nv[1], 3))
            cd = params.get("camera_data")
            if self.lens_type == "perspective" and cd is not None:
                self.camera_data = np.ascontiguousarray(cd, dtype=np.float64).reshape(15)

    def to_c(self):
        r = Request()
        p = self.params
        r.kind = self.kind
        r.nv[:] = list(self.nv)
        if self.pix_pos is not None:
            r.lens = LENS_PER_PIXEL
            r.pix_pos = _dptr(self.pix_pos)
            r.pix_dir = _dptr(self.pix_dir)
            if self.camera_data is not None:
                r.camera_data = _dptr(self.camera_data)
        else:
            r.lens = LENS_PLANE_PARALLEL
            r.vp_pos[:] = [float(x) for x in np.ravel(p["vp_pos"])]
            r.vp_dir[:] = [float(x) for x in np.ravel(p["vp_dir"])]
        r.center[:] = [float(x) for x in np.ravel(p["center"])]
        r.bounds[:] = [float(x) for x in p["bounds"]]
        r.x_vec[:] = [float(x) for x in np.ravel(p["x_vec"])]
        r.y_vec[:] = [float(x) for x in np.ravel(p["y_vec"])]
        r.width[:] = [float(x) for x in np.ravel(p["width"])]
        r.n_samples = self.n_samples
        r.ds_id = self.ds_id
        r.n_order = len(self.order)
        r.order = self.order.ctypes.data_as(C.POINTER(C.c_int32))
        if self.zbuffer is not None:
            if self.zbuffer.size != self.nv[0] * self.nv[1]:
                raise ValueError("zbuffer shape does not match the image")
            r.zbuffer = _dptr(self.zbuffer)
        if self.image_in is not None:
            if self.image_in.size != self.nv[0] * self.nv[1] * 4:
                raise ValueError("image shape does not match (nv0, nv1, 4)")
            r.image_in = _dptr(self.image_in)
        if self.tf is not None:
            r.tf = C.pointer(self.tf.desc)
        r.flags = ((FLAG_RECORD if self.record else 0) | (FLAG_ERT if self.ert else 0) |
                   (FLAG_PARTIAL if self.partial else 0))
        r.light_dir[:] = [float(x) for x in self.light_dir]
        r.light_rgba[:] = [float(x) for x in self.light_rgba]
        return r


class Batch:
    """A set of requests staged on the device; launch() runs them."""

    def __init__(self, ctx, specs, flags=BATCH_PLAIN, tile=(8, 8), item_tiles=1, tf_lds="all",
                 vertex_lds=False):
        self.ctx = ctx
        self.specs = list(specs)
        if tuple(tile) not in BATCH_TILE:
            raise ValueError(f"tile must be one of {sorted(BATCH_TILE)}")
        if item_tiles not in BATCH_ITEM_TILES:
            raise ValueError(f"item_tiles must be one of {sorted(BATCH_ITEM_TILES)}")
        if tf_lds not in BATCH_TF_LDS:
            raise ValueError(f"tf_lds must be one of {sorted(BATCH_TF_LDS)}")
        self.tile = tuple(tile)
        self.item_tiles = int(item_tiles)
        self.tf_lds = tf_lds
        self.vertex_lds = bool(vertex_lds)
        flags = (int(flags) | BATCH_TILE[self.tile] | BATCH_ITEM_TILES[self.item_tiles] |
                 BATCH_TF_LDS[tf_lds] | (BATCH_VERTEX_LDS if self.vertex_lds else 0))
        n = len(self.specs)
        arr = (Request * n)()
        for i, s in enumerate(self.specs):
            arr[i] = s.to_c()
        self._h = C.c_void_p()
        _check(lib().ytrc_batch_create(ctx._h, n, arr, int(flags), C.byref(self._h)))

    def set_launch(self, waves_per_group=4, groups_per_cu=0):
        _check(lib().ytrc_batch_set_launch(self._h, int(waves_per_group), int(groups_per_cu)))

    def launch(self):
        _check(lib().ytrc_batch_launch(self._h))

    def sync(self, timeout=None):
        """Waits for the last launch; returns its GPU ms.  With a timeout
        (seconds) raises YtrcTimeout if it is still running then; raises
        YtrcCancelled if it was stopped."""
        ms = C.c_float()
        if timeout is None:
            _check(lib().ytrc_batch_sync(self._h, C.byref(ms)))
        else:
            _check(lib().ytrc_batch_sync_for(self._h, float(timeout) * 1e3, C.byref(ms)))
        return ms.value

    def cancel(self):
        """Asks the running launch to stop at every wave's next tile."""
        _check(lib().ytrc_batch_cancel(self._h))

    def wait_interruptibly(self, poll=0.05):
        """sync() in slices of `poll` seconds, so that Python handles
        signals between them (the reference checks PyErr_CheckSignals every
        1000 rays, image_samplers.pyx:188-190): on KeyboardInterrupt the
        launch is stopped, waited for, and the interrupt re-raised."""
        while True:
            try:
                return self.sync(timeout=poll)
            except YtrcTimeout:
                continue
            except KeyboardInterrupt:
                self.cancel()
                try:
                    self.sync()
                except YtrcCancelled:
                    pass
                raise

    def run(self):
        self.launch()
        return self.sync()

    def kernel_ms(self):
        ms, n = C.c_float(), C.c_int32()
        _check(lib().ytrc_batch_kernel_ms(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def request_ms(self, i):
        ms = C.c_float()
        _check(lib().ytrc_batch_request_ms(self._h, int(i), C.byref(ms)))
        return ms.value

    def nbytes(self):
        v = C.c_uint64()
        _check(lib().ytrc_batch_bytes(self._h, C.byref(v)))
        return v.value

    def launch_plan(self):
        """(kernel launches per launch of a persistent batch, largest LDS
        bytes of transfer-function tables one of them stages)."""
        n, b = C.c_int32(), C.c_uint64()
        _check(lib().ytrc_batch_launch_plan(self._h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def read_image(self, i):
        nv = self.specs[i].nv
        out = np.empty((nv[0], nv[1], 4), dtype=np.float64)
        _check(lib().ytrc_batch_read_image(self._h, int(i), _dptr(out)))
        return out

    def read_partial(self, i):
        """(nv0, nv1, 8) float64: C[0..3], T[0..3] of a partial-map request."""
        nv = self.specs[i].nv
        out = np.empty((nv[0], nv[1], 8), dtype=np.float64)
        _check(lib().ytrc_batch_read_partial(self._h, int(i), _dptr(out)))
        return out

    def read_record(self, i):
        nv = self.specs[i].nv
        v = np.empty(nv, dtype=np.int64)
        h = np.empty(nv, dtype=np.uint64)
        s = np.empty(nv, dtype=np.int64)
        _check(lib().ytrc_batch_read_record(
            self._h, int(i), v.ctypes.data_as(C.POINTER(C.c_int64)),
            h.ctypes.data_as(C.POINTER(C.c_uint64)), s.ctypes.data_as(C.POINTER(C.c_int64))))
        return v, h, s

    def close(self):
        if self._h:
            lib().ytrc_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server (ytrc_server_*): submit requests as they
    arrive, call progress() to advance, read() finished images."""

    def __init__(self, ctx, max_requests=256, budget_bytes=8 << 30, chunk_tiles=16384):
        self.ctx = ctx
        self._h = C.c_void_p()
        _check(lib().ytrc_server_create(ctx._h, int(max_requests), int(budget_bytes),
                                        int(chunk_tiles), C.byref(self._h)))
        self._specs = {}
        self._ids = (C.c_int64 * 1024)()

    def submit(self, spec):
        """Returns the request id, or None when it is not admitted now."""
        rid = C.c_int64()
        r = spec.to_c()
        rc = lib().ytrc_server_submit(self._h, C.byref(r), C.byref(rid))
        if rc == YTRC_EBUDGET:
            return None
        _check(rc)
        self._specs[rid.value] = spec
        return rid.value

    def _done(self, fn):
        n = C.c_int32()
        _check(fn(self._h, self._ids, 1024, C.byref(n)))
        return [self._ids[i] for i in range(n.value)]

    def progress(self):
        return self._done(lib().ytrc_server_progress)

    def wait(self, timeout=None):
        """Ids finished since the last call, waiting for at least one; with
        a timeout (seconds) raises YtrcTimeout when none finished by then.
        Returns [] when nothing is queued or running."""
        if timeout is None:
            return self._done(lib().ytrc_server_wait)
        n = C.c_int32()
        _check(lib().ytrc_server_wait_for(self._h, float(timeout) * 1e3, self._ids, 1024, C.byref(n)))
        return [self._ids[i] for i in range(n.value)]

    def read(self, rid):
        spec = self._specs.pop(rid)
        out = np.empty((spec.nv[0], spec.nv[1], 4), dtype=np.float64)
        _check(lib().ytrc_server_read(self._h, int(rid), _dptr(out)))
        return out

    def read_planar(self, rid):
        """(nch, nv0, nv1) float32: the channels the kernel wrote."""
        spec = self._specs.pop(rid)
        out = np.empty((4, spec.nv[0], spec.nv[1]), dtype=np.float32)
        n = C.c_int32()
        _check(lib().ytrc_server_read_planar(self._h, int(rid),
                                             out.ctypes.data_as(C.POINTER(C.c_float)), 4,
                                             C.byref(n)))
        return out[:n.value]

    POLICIES = {"fifo": 0, "shortest_first": 1}

    def set_policy(self, policy, age_limit_ms=100.0):
        """Chunk fill order: "fifo" (arrival order) or "shortest_first"
        (least estimated work left first; requests older than
        age_limit_ms go first, in arrival order)."""
        if policy not in self.POLICIES:
            raise ValueError(f"unknown server policy {policy!r}")
        _check(lib().ytrc_server_set_policy(self._h, self.POLICIES[policy], float(age_limit_ms)))

    def cancel(self, rid):
        _check(lib().ytrc_server_cancel(self._h, int(rid)))
        self._specs.pop(rid, None)

    def stats(self):
        s = (C.c_int64 * 8)()
        _check(lib().ytrc_server_stats(self._h, s))
        keys = ("held", "queued", "chunks_inflight", "used", "budget", "tiles", "chunks",
                "free_slots")
        return dict(zip(keys, list(s)))

    def close(self):
        if self._h:
            lib().ytrc_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    """The process-wide context on device 0 used by the drop-in samplers."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("YTRC_DEVICE", "0")),
                               int(os.environ.get("YTRC_HBM_BUDGET", "0")))
    return _default_ctx
