"""Launch-shape autotuner for persistent ray-casting batches.

The reference has no launch parameters: each sampler call loops over pixels
with an OpenMP ``prange`` (image_samplers.pyx:167-197).  On the GPU the same
work is cut into 64-ray tiles, one per wavefront, and three choices set how
well the tiles fill the chip:

  * tile shape: 8x8, 4x16 or 16x4 rays (rows along image axis 0);
  * waves per workgroup: 1, 2 or 4 tiles launched together;
  * workgroups per CU for the persistent grid (0 = the occupancy limit).

``tune`` times every candidate on the caller's own batch (HIP events around
each launch, best of ``reps`` after one warm launch that also captures the
graph) and keeps the fastest in a JSON cache keyed by the device and a
signature of the batch, so later batches of the same shape skip the search.
"""

import json
import os
import time

from . import _lib

TILES = ((8, 8), (4, 16), (16, 4))
WAVES_PER_GROUP = (1, 2, 4)
GROUPS_PER_CU = (0, 1, 2, 4, 8)
DEFAULT = {"tile": [8, 8], "waves_per_group": 4, "groups_per_cu": 0}


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


def make_batch(ctx, specs, cfg, flags=_lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH):
    """A batch built and launched with the launch shape ``cfg``."""
    b = _lib.Batch(ctx, specs, flags, tile=tuple(cfg["tile"]))
    b.set_launch(cfg["waves_per_group"], cfg["groups_per_cu"])
    return b


def tune(ctx, specs, flags=_lib.BATCH_PERSISTENT | _lib.BATCH_GRAPH, tiles=TILES,
         waves_per_group=WAVES_PER_GROUP, groups_per_cu=GROUPS_PER_CU, reps=2,
         use_cache=True, path=None, log=None):
    """Returns (best config dict, list of (config, ms) measured).  With
    ``use_cache`` a cached entry for this signature is returned unmeasured."""
    path = path or cache_path()
    key = signature(specs, ctx.device_name)
    if use_cache:
        hit = _load(path).get(key)
        if hit is not None:
            return dict(hit["cfg"]), []
    table = []
    t0 = time.time()
    for tile in tiles:
        b = _lib.Batch(ctx, specs, flags, tile=tile)
        try:
            for wpg in waves_per_group:
                for gpc in groups_per_cu:
                    b.set_launch(wpg, gpc)
                    b.run()
                    ms = min(b.run() for _ in range(max(1, reps)))
                    cfg = {"tile": list(tile), "waves_per_group": wpg, "groups_per_cu": gpc}
                    table.append((cfg, ms))
                    if log:
                        log(f"autotune {tile[0]}x{tile[1]} wpg={wpg} gpc={gpc}: {ms:.3f} ms")
        finally:
            b.close()
    best = min(table, key=lambda t: t[1])[0]
    if use_cache:
        d = _load(path)
        d[key] = {"cfg": best, "ms": min(t[1] for t in table), "searched": len(table),
                  "seconds": round(time.time() - t0, 2)}
        _store(path, d)
    return best, table
