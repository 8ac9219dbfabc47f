"""Launch-shape autotuner for persistent ray-casting batches.

The reference has no launch parameters: each sampler call loops over pixels
with an OpenMP ``prange(schedule="static", chunk=100)``
(image_samplers.pyx:167-197).  On the GPU the same work is cut into 64-ray
tiles, one per wavefront, and these choices set how well the tiles fill the
chip:

  * tile shape: 8x8, 4x16 or 16x4 rays (rows along image axis 0);
  * rays per work item: 1, 2, 4 or 8 consecutive tiles (64-512 rays) that
    one wave renders in a row per item of the persistent queue (the
    analog of the prange chunk);
  * transfer-function table bytes per launch in LDS: every workgroup of a
    persistent launch stages the distinct tables of the launch's requests;
    "all" keeps one launch per kernel variant, "32k"/"16k" split a variant's
    requests over launches so that each stages at most that many bytes,
    "one" gives every distinct table its own launch (fewer LDS bytes per
    workgroup, more workgroups per CU, more launches).  Values that give
    the same launches (a batch whose variants each have one table) are
    measured once;
  * vertex blocks staged in LDS (volume renders with a ColorTransferFunction,
    one field, 10 sub-samples): off, or each wave keeps one 512-byte vertex
    block in LDS and reads a visit's corners from it when all its lanes are
    in that block;
  * waves per workgroup: 1, 2 or 4 tiles launched together, which share
    one copy of the launch's tables in LDS (LDS per wave: table bytes /
    waves per workgroup);
  * workgroups per CU for the persistent grid (0 = the occupancy limit).

``tune`` measures candidates on the caller's own batch (HIP events around
each launch, best of ``reps`` after one warm launch that also captures the
graph) and keeps the fastest in a JSON cache keyed by the device and a
signature of the batch, so later batches of the same shape skip the search.
The default search is coordinate descent: starting from DEFAULT, each knob
in turn takes every value with the others held, the best is kept, and the
sweep repeats until no knob changes (at most ``rounds`` sweeps); ``search =
"grid"`` measures every combination.

``tune_bricks`` picks the brick edge of a uniform dataset (the brick chunk
the kd-tree and the projection blocks are cut into) by the same measurement
over the candidate splits.
"""

import json
import os
import time

from . import _lib

TILES = ((8, 8), (4, 16), (16, 4))
ITEM_TILES = (1, 2, 4, 8)
WAVES_PER_GROUP = (1, 2, 4)
GROUPS_PER_CU = (0, 1, 2, 4, 8)
TF_LDS = ("all", "32k", "16k", "one")
VERTEX_LDS = (False, True)
DEFAULT = {"tile": [8, 8], "item_tiles": 1, "tf_lds": "all", "vertex_lds": False, "waves_per_group": 4,
           "groups_per_cu": 0}
KNOBS = ("tile", "item_tiles", "tf_lds", "vertex_lds", "waves_per_group", "groups_per_cu")


def cache_path():
    return os.environ.get("YT_AMD_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "yt_amd", "autotune.json"))


def _load(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path, d):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def signature(specs, device_name=""):
    """Cache key: device, then per kernel variant the request count, the total
    pixels (to a power of two) and the sampling settings."""
    groups = {}
    for s in specs:
        k = (s.kind, bool(s.record), bool(s.ert), int(s.n_samples) if s.kind == _lib.VOLUME_RENDER else 0)
        n, px = groups.get(k, (0, 0))
        groups[k] = (n + 1, px + s.nv[0] * s.nv[1])
    parts = [device_name]
    for k in sorted(groups):
        n, px = groups[k]
        parts.append(f"k{k[0]}r{int(k[1])}e{int(k[2])}s{k[3]}:n{n}:p{max(px, 1).bit_length()}")
    return "|".join(parts)


def _full(cfg):
    out = dict(DEFAULT)
    out.update(cfg or {})
    out["tile"] = list(out["tile"])
    return out


def make_batch(ctx, specs, cfg, flags=_lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH):
    """A batch built and launched with the launch shape ``cfg``."""
    cfg = _full(cfg)
    b = _lib.Batch(ctx, specs, flags, tile=tuple(cfg["tile"]), item_tiles=int(cfg["item_tiles"]),
                   tf_lds=cfg["tf_lds"], vertex_lds=bool(cfg["vertex_lds"]))
    b.set_launch(cfg["waves_per_group"], cfg["groups_per_cu"])
    return b


class _Timer:
    """Times launch shapes on one set of specs, reusing one batch per
    (tile shape, rays per item, launch plan) and each measurement per
    configuration; table-size values that give the same launches share a
    batch and a measurement."""

    def __init__(self, ctx, specs, flags, reps, log):
        self.ctx, self.specs, self.flags, self.reps, self.log = ctx, specs, flags, reps, log
        self.batches = {}  # (tile, item_tiles, vertex_lds, n_launches) -> batch
        self.alias = {}    # (tile, item_tiles, vertex_lds, tf_lds) -> that key
        self.ms = {}       # (batch key, waves, groups) -> ms
        self.seen = {}

    def _batch(self, cfg):
        ak = (tuple(cfg["tile"]), cfg["item_tiles"], bool(cfg["vertex_lds"]), cfg["tf_lds"])
        if ak not in self.alias:
            b = _lib.Batch(self.ctx, self.specs, self.flags, tile=ak[0], item_tiles=ak[1],
                           vertex_lds=ak[2], tf_lds=ak[3])
            bk = (ak[0], ak[1], ak[2], b.launch_plan()[0])
            if bk in self.batches:
                b.close()  # the same launches as a batch already built
            else:
                self.batches[bk] = b
            self.alias[ak] = bk
        bk = self.alias[ak]
        return bk, self.batches[bk]

    def __call__(self, cfg):
        key = json.dumps(cfg, sort_keys=True)
        if key in self.seen:
            return self.seen[key]
        bk, b = self._batch(cfg)
        mk = (bk, cfg["waves_per_group"], cfg["groups_per_cu"])
        if mk not in self.ms:
            b.set_launch(cfg["waves_per_group"], cfg["groups_per_cu"])
            b.run()
            self.ms[mk] = min(b.run() for _ in range(max(1, self.reps)))
            if self.log:
                self.log(f"autotune {describe(cfg)}: {self.ms[mk]:.3f} ms")
        self.seen[key] = self.ms[mk]
        return self.ms[mk]

    def measured(self):
        """Distinct launch configurations actually timed."""
        return len(self.ms)

    def close(self):
        for b in self.batches.values():
            b.close()
        self.batches = {}


def describe(cfg):
    cfg = _full(cfg)
    return (f"{cfg['tile'][0]}x{cfg['tile'][1]}/i{cfg['item_tiles']}/t{cfg['tf_lds']}"
            f"/v{int(bool(cfg['vertex_lds']))}/w{cfg['waves_per_group']}/g{cfg['groups_per_cu']}")


def tune(ctx, specs, flags=_lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH, tiles=TILES,
         item_tiles=ITEM_TILES, waves_per_group=WAVES_PER_GROUP, groups_per_cu=GROUPS_PER_CU,
         reps=2, use_cache=True, path=None, log=None, search="descent", rounds=3, tf_lds=TF_LDS,
         vertex_lds=VERTEX_LDS):
    """Returns (best config dict, list of (config, ms) measured).  With
    ``use_cache`` a cached entry for this signature is returned unmeasured."""
    path = path or cache_path()
    key = signature(specs, ctx.device_name)
    if use_cache:
        hit = _load(path).get(key)
        if hit is not None:
            return _full(hit["cfg"]), []
    values = {"tile": [list(t) for t in tiles], "item_tiles": list(item_tiles), "tf_lds": list(tf_lds),
              "vertex_lds": list(vertex_lds), "waves_per_group": list(waves_per_group),
              "groups_per_cu": list(groups_per_cu)}
    t0 = time.time()
    timer = _Timer(ctx, specs, flags, reps, log)
    try:
        if search == "grid":
            for tile in values["tile"]:
                for it in values["item_tiles"]:
                    for tl in values["tf_lds"]:
                        for vl in values["vertex_lds"]:
                            for wpg in values["waves_per_group"]:
                                for gpc in values["groups_per_cu"]:
                                    timer({"tile": tile, "item_tiles": it, "tf_lds": tl, "vertex_lds": vl,
                                           "waves_per_group": wpg, "groups_per_cu": gpc})
            best = min(timer.seen.items(), key=lambda kv: kv[1])
            best = json.loads(best[0])
        elif search == "descent":
            best = _full(None)
            for v in values.values():
                if not v:
                    raise ValueError("tune: every knob needs at least one value")
            for knob in KNOBS:  # start from a point inside the given ranges
                if best[knob] not in values[knob]:
                    best[knob] = values[knob][0]
            best_ms = timer(best)
            for _ in range(max(1, rounds)):
                changed = False
                for knob in KNOBS:
                    for v in values[knob]:
                        cand = dict(best, **{knob: v})
                        ms = timer(cand)
                        if ms < best_ms:
                            best, best_ms, changed = cand, ms, True
                if not changed:
                    break
        else:
            raise ValueError(f"tune: unknown search {search!r}")
    finally:
        timer.close()
    table = [(json.loads(k), ms) for k, ms in timer.seen.items()]
    if use_cache:
        d = _load(path)
        d[key] = {"cfg": best, "ms": min(t[1] for t in table), "searched": len(table),
                  "timed": timer.measured(),
                  "seconds": round(time.time() - t0, 2)}
        _store(path, d)
    return _full(best), table


def tune_bricks(build, candidates, cfg=None, flags=_lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH, reps=3,
                log=None):
    """Picks the brick split of a dataset by measurement: for each candidate
    (e.g. bricks per axis of load_uniform_grid), ``build(candidate)`` returns
    (context, specs, release) -- the resident dataset's requests for the same
    cameras and a function that frees it; the batch runs with launch shape
    ``cfg``.  Returns (best candidate, {candidate: ms per launch})."""
    out = {}
    for c in candidates:
        ctx, specs, release = build(c)
        try:
            b = make_batch(ctx, specs, cfg or DEFAULT, flags)
            try:
                b.run()
                ms = min(b.run() for _ in range(max(1, reps)))
            finally:
                b.close()
        finally:
            release()
        out[c] = ms
        if log:
            log(f"brick split {c}: {ms:.3f} ms")
    best = min(out, key=lambda c: out[c])
    return best, out
