"""Serving loop: Poisson request arrivals against resident volumes.

Requests (a camera, an image size, projection or volume render) arrive at
seeded Poisson times.  Each is offered to the continuous-batching server
(ytrc_server_*) when it arrives; a request the server cannot admit yet (no
slot or HBM budget) waits in arrival order and is offered again as finished
requests are read back and release their memory.  Latency is measured from
arrival to the image being in host memory; with deferred requests it
includes the host's preparation of the request at arrival (camera
parameters, the kd-tree brick order for the viewpoint).
"""

import time

import numpy as np

from . import _lib
from .render import camera_params


def make_requests(volumes, n, seed=1006, side_range=(256, 2048), vr_fraction=0.5,
                  tfs=None, rng_cam_seed=1004, deferred=False):
    """n requests spread over the resident volumes: random view direction
    and north, side uniform in side_range, projection or VR.  Returns
    request specs, or with deferred=True callables that build them (so that
    run_poisson prepares each request when it arrives)."""
    rng = np.random.default_rng(rng_cam_seed)
    rs = np.random.default_rng(seed)
    tfc = {k: (v if isinstance(v, _lib.TF) else _lib.TF.from_tf_obj(v))
           for k, v in (tfs or {}).items()}
    out = []
    for i in range(n):
        vol = volumes[i % len(volumes)]
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        side = int(rs.integers(side_range[0], side_range[1] + 1))
        north = rng.normal(size=3)
        vr = (vol.vr_ds_id is not None) and (vol.proj_ds_id is None or rs.random() < vr_fraction)

        def build(vol=vol, nrm=nrm, side=side, north=north, vr=vr):
            p = camera_params(vol.ds.domain_center, nrm, float(vol.ds.domain_width.max()), side,
                              north)
            if vr:
                return vol.vr_request(p, tfc.get(id(vol)))
            return vol.projection_request(p)

        out.append(build if deferred else build())
    return out


def run_poisson(server, specs, rate, seed=1005, readback="planar", idle_sleep=2e-4):
    """Offers specs[i] at Poisson arrival times of the given rate (req/s).
    Returns latencies (s), wall time and the number of admission waits."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(specs)))
    lat = np.full(len(specs), np.nan)
    id_to_idx = {}
    waiting = []
    nxt = 0
    done = 0
    waits = 0
    t0 = time.perf_counter()
    while done < len(specs):
        now = time.perf_counter() - t0
        while nxt < len(specs) and arrivals[nxt] <= now:
            waiting.append(nxt)
            nxt += 1
        while waiting:
            i = waiting[0]
            if callable(specs[i]):
                specs[i] = specs[i]()
            rid = server.submit(specs[i])
            if rid is None:
                waits += 1
                break
            id_to_idx[rid] = i
            waiting.pop(0)
        fin = server.progress()
        for rid in fin:
            i = id_to_idx.pop(rid)
            if readback == "planar":
                server.read_planar(rid)
            else:
                server.read(rid)
            lat[i] = time.perf_counter() - t0 - arrivals[i]
            done += 1
        if not fin and (not waiting or nxt < len(specs)):
            if server.stats()["chunks_inflight"] == 0 and not waiting:
                time.sleep(min(idle_sleep, max(0.0, arrivals[nxt] - now) if nxt < len(specs) else idle_sleep))
            else:
                time.sleep(idle_sleep / 4)
    wall = time.perf_counter() - t0
    return dict(latency_s=lat, wall_s=wall, admission_waits=waits,
                first_arrival=float(arrivals[0]), last_arrival=float(arrivals[-1]))


def p99_interval(lat_ms, n_boot=2000, seed=0):
    """Bootstrap 95 % interval of the 99th percentile of lat_ms: the 2.5th
    and 97.5th percentiles of the p99 of n_boot resamples (with
    replacement) of the same size."""
    lat_ms = np.asarray(lat_ms, dtype=np.float64)
    if lat_ms.size < 2:
        return None
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, lat_ms.size, size=(n_boot, lat_ms.size))
    boots = np.percentile(lat_ms[idx], 99, axis=1)
    return [round(float(np.percentile(boots, 2.5)), 3), round(float(np.percentile(boots, 97.5)), 3)]


def summarize(res, specs):
    """Latency percentiles over all requests and over those of side <= 1024
    (the 1024^2-and-smaller requests), with bootstrap 95 % intervals of the
    p99s; images/s offered (arrivals) and served; admission waits."""
    lat = res["latency_s"] * 1e3
    span = res["wall_s"] - res["first_arrival"]
    sides = np.array([max(s.nv) if hasattr(s, "nv") else 0 for s in specs])
    small = lat[sides <= 1024]
    arr_span = res["last_arrival"] - res["first_arrival"]
    return dict(
        n=len(specs),
        n_side_le_1024=int(small.size),
        p50_ms=float(np.percentile(lat, 50)),
        p99_ms=float(np.percentile(lat, 99)),
        p99_ms_ci95=p99_interval(lat),
        mean_ms=float(np.mean(lat)),
        p99_ms_side_le_1024=float(np.percentile(small, 99)) if small.size else None,
        p99_ms_side_le_1024_ci95=p99_interval(small) if small.size else None,
        offered_images_per_s=(len(specs) - 1) / arr_span if arr_span > 0 else None,
        images_per_s=len(specs) / span,
        admission_waits=int(res["admission_waits"]),
    )
