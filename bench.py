#!/usr/bin/env python
"""Headline benchmark: batched off-axis projection + volume rendering of a
512^3 float32 lognormal field on one MI355X.

Workload (one "step"): 64 plane-parallel cameras on seeded directions
(seed 1004) with random north vectors, image sides drawn from
{128, 256, 512, 1024} (seed 1006), alternating unweighted projection and
volume rendering (8-Gaussian ColorTransferFunction over log10 density,
10 sub-samples per cell).  The field (seed 1001, P(k) ~ k^-3, sigma_ln = 1)
is split into 2x2x2 grids of 256^3 (load_uniform_grid), which are the
projection blocks and the AMRKDTree bricks (vertex-centred, log10).  All
bricks are resident in HBM before timing starts; a step is one launch of the
whole batch (persistent tile-queue kernels captured in a hipGraph).

Prints one JSON line.  Progress goes to stderr.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gsamples/s off-axis proj+VR at 512^3 (1 GPU); p50/p99 render latency"
HBM_PEAK_GBS = 8000.0
# MI355X vector FP64 peak (TFLOP/s), for the DDA's arithmetic floor.
FP64_VECTOR_TFLOPS = 78.6
# FP64 operations of one DDA step: the min-tmax selection (2 compares),
# exit_t = fmin(tmax, max_t), enter_t - exit_t, tmax += tdelta and the
# enter_t >= max_t test, plus the per-visit sample position (2).
FP64_OPS_PER_VISIT = 8


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(n, psize, n_cams, sizes=(128, 256, 512, 1024)):
    from yt_amd import data, render
    from yt_amd.transfer_functions import gaussian_ctf

    t = time.time()
    rho = data.lognormal_density(n, seed=1001, dtype=np.float32)
    log(f"field {n}^3 in {time.time() - t:.1f}s")
    ds = data.load_uniform_grid({"density": rho}, psize=(psize,) * 3)
    lo, hi = float(np.log10(rho.min())), float(np.log10(rho.max()))
    tf = gaussian_ctf((lo, hi), n_gauss=8, nbins=256)
    rng = np.random.default_rng(1004)
    rs = np.random.default_rng(1006)
    cams = []
    for i in range(n_cams):
        nv = rng.normal(size=3)
        nv /= np.linalg.norm(nv)
        north = rng.normal(size=3)
        res = int(rs.choice(sizes))
        p = render.camera_params((0.5, 0.5, 0.5), nv, 1.0, res, north)
        cams.append((i % 2, p))
    return rho, ds, tf, cams


def make_specs(vol, tf, cams, record=False):
    specs = []
    for kind, p in cams:
        if kind == 0:
            specs.append(vol.projection_request(p, record=record))
        else:
            specs.append(vol.vr_request(p, tf, n_samples=10, record=record))
    return specs


def alg_bytes(kind, samples, npix, nf=1, n_samples=10):
    """Algorithmic HBM bytes: projection n_fields*4 B per sample, VR
    (8*n_fields*4 + 1)/n_samples B per sub-sample, plus the image channels
    written (nf for projection, 3 for VR) at 4 B."""
    if kind == 0:
        return samples * nf * 4.0 + npix * nf * 4.0
    return samples * (8.0 * nf * 4.0 + 1.0) / n_samples + npix * 3 * 4.0


# MI355X FP32 vector peak 157.3 TFLOP/s = 7.86e13 lane FMAs per second: the
# rate at which the VALUs retire lane operations when every instruction is a
# full-rate FP32 one (FP64 operations issue at half that rate).
VALU_LANE_OPS_PEAK = 157.3e12 / 2.0


def measured_valu(path, kind, args):
    """VALU instructions per dispatch of the kernel of `kind` and its
    VALUUtilization from the committed rocprofv3 --pmc summary, when taken
    on this same workload; None otherwise."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    want = {"n": args.n, "psize": args.psize, "cams": args.cams}
    if any(t.get("workload", {}).get(k) != v for k, v in want.items()):
        return None
    return t.get("valu", {}).get(str(kind))


def measured_traffic(path, kind, args):
    """HBM bytes per dispatch of the dominant kernel from a committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE in KiB per dispatch,
    each multiplied by its copy-kernel calibration factor), when that summary
    was taken on this same workload; None otherwise."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    want = {"n": args.n, "psize": args.psize, "cams": args.cams, "kind": kind}
    if any(t.get("workload", {}).get(k) != v for k, v in want.items()):
        return None
    return int(round(1024 * (t["fetch_kib"] * t["fetch_factor"] + t["write_kib"] * t["write_factor"])))


def host_threads():
    """Host threads this process may use: its CPU affinity mask, bounded by
    the cgroup CPU quota when one is set (cpu.max)."""
    n = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    return (n if quota is None else min(n, quota)), n, quota


# Rows of each image side the CPU sample casts (a central strip of about
# 16k pixels per image).
SAMPLE_ROWS = {128: 128, 256: 64, 512: 32, 1024: 16}


def sample_strata(cams):
    """One camera per (kind, side): the first of the 64 with that kind and
    side, and the central rows of its image the CPU sample casts."""
    out = []
    for kind in (0, 1):
        for side in sorted(SAMPLE_ROWS):
            idx = [i for i in range(len(cams)) if cams[i][0] == kind and cams[i][1]["nv"][0] == side]
            if idx:
                n = min(side, SAMPLE_ROWS[side])
                r0 = (side - n) // 2
                out.append((idx[0], (r0, r0 + n)))
    return out


def cpu_baseline(ds, vol, tf, cams, strata, window_samples, gpu_images, runs=5):
    """The FP64 oracle (OpenMP static/100) on a bounded, stratified sample of
    the workload: for each kind and image side one camera, cast over a
    central strip of its image (sample_strata).  Timed as the median of
    `runs` passes over the whole sample after one warm-up pass, on every
    host thread available, and on one thread (as off_axis_projection runs
    by default, off_axis_projection.py:34) as the median of 3 passes.  The
    oracle's images of the sample are also compared with the GPU's images
    of the timed batch (per-image rel-L2 and max-abs / max)."""
    from oracle import oracle
    from tests.helpers import oracle_proj_bricks, tf_dict

    threads, affinity, quota = host_threads()
    pbricks = oracle_proj_bricks(ds, ["density"])
    vbricks = []
    for node in vol.leaves:
        pg = vol.kd.get_brick_data(node)
        vbricks.append(dict(le=pg.LeftEdge, re=pg.RightEdge, dims=pg.dims, data=pg.my_data,
                            mask=pg.source_mask))
        node.data = None
    tfd = tf_dict(tf)
    orders = {i: vol.vr_order(cams[i][1]["viewpoint"]) for i, _ in strata if cams[i][0] == 1}

    def one_pass(nt, keep=None):
        t = time.perf_counter()
        for i, rows in strata:
            kind, p = cams[i]
            if kind == 0:
                img = oracle.cast(0, p, pbricks, num_threads=nt, rows=rows)
            else:
                img = oracle.cast(1, p, vbricks, order=orders[i], tf=tfd, n_samples=10,
                                  num_threads=nt, rows=rows)
            if keep is not None:
                keep[i] = img[rows[0]:rows[1]]
        return time.perf_counter() - t

    total = int(sum(window_samples[i] for i, _ in strata))
    oimgs = {}
    one_pass(threads, oimgs)
    times = [one_pass(threads) for _ in range(runs)]
    t_med = float(np.median(times))
    single = [one_pass(1) for _ in range(3)]
    t_one = float(np.median(single))
    names = [f"{'proj' if cams[i][0] == 0 else 'vr'}{cams[i][1]['nv'][0]} rows {r[0]}-{r[1]}"
             for i, r in strata]
    log(f"cpu oracle: {len(strata)} images ({', '.join(names)}), {total} samples, "
        f"median {t_med:.2f}s of {runs} on {threads} threads, {t_one:.2f}s on 1 thread")
    parity = []
    for i, rows in strata:
        kind = cams[i][0]
        nch = 1 if kind == 0 else 3
        g = gpu_images[i][rows[0]:rows[1]]
        o = oimgs[i]
        errs = []
        for c in range(nch):
            d = g[..., c].astype(np.float64) - o[..., c]
            den = float(np.sqrt(np.sum(o[..., c] ** 2)))
            mx = float(np.max(np.abs(o[..., c])))
            errs.append({"rel_l2": float(np.sqrt(np.sum(d * d)) / den) if den else 0.0,
                         "max_abs_over_max": float(np.max(np.abs(d)) / mx) if mx else 0.0})
        parity.append({"image": i, "kind": "projection" if kind == 0 else "vr",
                       "side": int(cams[i][1]["nv"][0]), "rows": list(rows), "channels": errs})
    worst = max(max(c["rel_l2"] for c in e["channels"]) for e in parity)
    log(f"sampled images vs the FP64 oracle: largest rel-L2 {worst:.2e}")
    # The 64^3 float64 field at 256^2 (the off_axis_projection case of the
    # drop-in leg), per call, on 1 and on `threads` threads.
    from yt_amd import data, render

    ds0 = data.load_uniform_grid({"density": data.lognormal_density(64, seed=1001)})
    p0 = render.camera_params((0.5, 0.5, 0.5), (0.3, 0.5, 0.81), 1.0, 256)
    b0 = oracle_proj_bricks(ds0, ["density"])
    c0 = {}
    for nt in (1, threads):
        xs = []
        for _ in range(6):
            t = time.perf_counter()
            oracle.cast(0, p0, b0, num_threads=nt)
            xs.append((time.perf_counter() - t) * 1e3)
        c0[f"{nt}_threads"] = round(float(np.median(xs[1:])), 2)
    log(f"cpu oracle, 64^3 projection at 256^2 (ms): {c0}")
    return {
        "value": total / t_med / 1e9,
        "unit": "Gsamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one camera per kind and side of the 64 ({', '.join(names)}): "
                  f"{total} samples; FP64 C oracle, OpenMP static/100; median of {runs} passes "
                  f"after one warm-up",
        "affinity_cpus": affinity,
        "cgroup_quota_cpus": quota,
        "seconds": [round(x, 3) for x in times],
        "single_thread": {"value": total / t_one / 1e9, "seconds": [round(x, 3) for x in single]},
        "c0_projection_64cubed_256_ms": c0,
        "parity_sample": parity,
    }


def configs_leg(ctx, cpu=True, reps=5):
    """The single-image configurations beside the batched workload: 64^3
    projection at 256^2, the 256^3 density-weighted projection of
    temperature at 1024^2 and the 643-grid AMR volume render at 1024^2.
    Per configuration: GPU time of one plain launch (one kernel per image)
    and of one persistent launch (HIP events, median of `reps`), the sample
    count (record pass), Gsamples/s, and -- with `cpu` -- the FP64 oracle's
    wall time on all host threads (and on one thread for the 64^3 case,
    off_axis_projection's default) with the GPU image's rel-L2 against it."""
    from oracle import oracle
    from tests.helpers import oracle_proj_bricks, oracle_vr_bricks, tf_dict, vr_tf_for
    from yt_amd import _lib, data, render

    threads = host_threads()[0]
    out = {}

    def gpu_times(spec):
        res = {}
        for name, flags in (("plain", _lib.BATCH_PLAIN), ("persistent", _lib.BATCH_PERSISTENT)):
            b = _lib.Batch(ctx, [spec], flags)
            try:
                b.run()
                res[name] = float(np.median([b.run() for _ in range(reps)]))
                img = b.read_image(0)
            finally:
                b.close()
        spec.record = True
        b = _lib.Batch(ctx, [spec], _lib.BATCH_PERSISTENT)
        try:
            b.run()
            samples = int(b.read_record(0)[2].sum())
        finally:
            b.close()
            spec.record = False
        return res, samples, img

    def cpu_times(fn, threads_list):
        res = {}
        for nt in threads_list:
            fn(nt)
            xs = []
            for _ in range(3):
                t = time.perf_counter()
                o = fn(nt)
                xs.append((time.perf_counter() - t) * 1e3)
            res[f"{nt}_threads"] = round(float(np.median(xs)), 2)
        return res, o

    cases = []
    ds0 = data.load_uniform_grid({"density": data.lognormal_density(64, seed=1001)})
    cases.append(("c0_64cubed_projection_256", ds0, "proj", "density", None,
                  render.camera_params((0.5,) * 3, (0.3, 0.5, 0.81), 1.0, 256), (1, threads)))
    rho = data.lognormal_density(256, seed=1001)
    ds1 = data.load_uniform_grid({"density": rho,
                                  "temperature": data.temperature_from_density(rho, seed=1002)})
    cases.append(("c1_256cubed_weighted_projection_1024", ds1, "proj", "temperature", "density",
                  render.camera_params((0.5,) * 3, (0.3, 0.5, 0.81), 1.0, 1024), (threads,)))
    amr = data.load_amr_grids(data.make_amr_hierarchy(128, 3), [128] * 3)
    cases.append(("c2_amr_vr_1024", amr, "vr", "density", None,
                  render.camera_params((0.5,) * 3, (0.3, 0.5, 0.81), 1.0, 1024), (threads,)))
    for name, ds, kind, field, weight, p, tl in cases:
        if kind == "proj":
            vol = render.ResidentVolume(ctx, ds, proj_fields=field, weight=weight)
            spec = vol.projection_request(p)
        else:
            vol = render.ResidentVolume(ctx, ds, vr_field=field)
            tf = vr_tf_for(ds)
            spec = vol.vr_request(p, tf)
        try:
            g, samples, img = gpu_times(spec)
            r = {"gpu_plain_ms": round(g["plain"], 3), "gpu_persistent_ms": round(g["persistent"], 3),
                 "samples": samples,
                 "gsamples_per_s": round(samples / (g["persistent"] * 1e-3) / 1e9, 2)}
            if cpu:
                if kind == "proj":
                    bricks = oracle_proj_bricks(ds, [field], weight=weight)
                    c, oimg = cpu_times(lambda nt: oracle.cast(0, p, bricks, num_threads=nt), tl)
                    nch = 1
                else:
                    bricks = oracle_vr_bricks(vol)
                    tfd = tf_dict(tf)
                    c, oimg = cpu_times(lambda nt: oracle.cast(1, p, bricks, order=spec.order, tf=tfd,
                                                               num_threads=nt), tl)
                    nch = 3
                r["cpu_oracle_ms"] = c
                r["rel_l2"] = max(float(np.linalg.norm(img[..., k] - oimg[..., k]) /
                                        np.linalg.norm(oimg[..., k])) for k in range(nch))
        finally:
            vol.free()
        out[name] = r
        log(f"config {name}: {r}")
    return out


def drop_in_leg(reps=3):
    """End to end through the drop-in entry points, host arrays in, image
    out (upload, brick preparation and cast included): off_axis_projection
    of the 64^3 field at 256^2 and of the 256^3 field, density-weighted
    temperature, at 1024^2; create_scene(...).render() of the AMR hierarchy
    (8-Gaussian CTF) at 1024^2.  Median wall ms of `reps` calls after one."""
    from yt_amd import data, off_axis_projection, scene

    def timed(fn):
        fn()
        xs = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            xs.append((time.perf_counter() - t0) * 1e3)
        return round(float(np.median(xs)), 2)

    out = {}
    ds0 = data.load_uniform_grid({"density": data.lognormal_density(64, seed=1001)})
    out["c0_projection_64cubed_256"] = timed(
        lambda: off_axis_projection(ds0, (0.5,) * 3, (0.3, 0.5, 0.81), 1.0, 256, "density"))
    rho = data.lognormal_density(256, seed=1001)
    ds1 = data.load_uniform_grid({"density": rho,
                                  "temperature": data.temperature_from_density(rho, seed=1002)})
    out["c1_weighted_projection_256cubed_1024"] = timed(
        lambda: off_axis_projection(ds1, (0.5,) * 3, (0.3, 0.5, 0.81), 1.0, 1024, "temperature",
                                    weight="density"))
    amr = data.load_amr_grids(data.make_amr_hierarchy(128, 3), [128] * 3)

    def render_amr():
        sc = scene.create_scene(amr, "density")
        sc.camera.resolution = (1024, 1024)
        sc.render()
        sc.sources["source_00"].free()

    out["c2_amr_scene_render_1024"] = timed(render_amr)
    log(f"drop-in end to end (ms): {out}")
    return out


def serve_leg(ctx, vol512, tf512, rates, n, budgets=(), policy_rates=(), age_limit_ms=100.0):
    """Continuous-batching server over three resident datasets: 256^3 and
    512^3 uniform lognormal fields and the seeded AMR hierarchy (128^3 root,
    3 levels); requests of side uniform in [256, 2048], half projection and
    half VR (8-Gaussian CTF per dataset), Poisson arrivals (seed 1005).
    The server fills its chunks shortest first (requests older than
    age_limit_ms first); policy_rates repeat the same request sets and
    arrivals with chunks filled in arrival order, for comparison."""
    from yt_amd import _lib, data, render, serve
    from yt_amd.transfer_functions import gaussian_ctf

    t = time.time()
    rho = data.lognormal_density(256, seed=1001, dtype=np.float32)
    v256 = render.ResidentVolume(ctx, data.load_uniform_grid({"density": rho}),
                                 proj_fields="density", vr_field="density")
    amr = data.load_amr_grids(data.make_amr_hierarchy(128, 3), [128] * 3)
    vamr = render.ResidentVolume(ctx, amr, proj_fields="density", vr_field="density")
    log(f"serving datasets resident in {time.time() - t:.1f}s "
        f"(AMR: {len(amr.grids)} grids, {vamr.n_vr_bricks} kd bricks)")

    def tf_for(ds):
        lo = min(float(np.log10(np.min(g["density"]))) for g in ds.grids)
        hi = max(float(np.log10(np.max(g["density"]))) for g in ds.grids)
        return gaussian_ctf((lo, hi), n_gauss=8, nbins=256)

    vols = [v256, vol512, vamr]
    tfs = {id(v256): tf_for(v256.ds), id(vol512): tf512, id(vamr): tf_for(amr)}
    def server(budget):
        srv = _lib.Server(ctx, max_requests=256, budget_bytes=budget, chunk_tiles=16384)
        srv.set_policy("shortest_first", age_limit_ms=age_limit_ms)
        return srv

    srv = server(8 << 30)
    out = {}
    for k, lam in enumerate(rates):
        specs = serve.make_requests(vols, n, seed=1006 + k, tfs=tfs, rng_cam_seed=1004 + k,
                                     deferred=True)
        res = serve.run_poisson(srv, specs, lam, seed=1005 + k)
        out[f"{lam:g}"] = {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                           for kk, vv in serve.summarize(res, specs).items()}
        log(f"serving rate {lam:g}/s: {out[f'{lam:g}']}")
    srv.close()
    by_policy = {}
    srv = _lib.Server(ctx, max_requests=256, budget_bytes=8 << 30, chunk_tiles=16384)
    srv.set_policy("fifo")
    for k, lam in enumerate(rates):
        if lam not in policy_rates:
            continue
        specs = serve.make_requests(vols, n, seed=1006 + k, tfs=tfs, rng_cam_seed=1004 + k,
                                     deferred=True)
        res = serve.run_poisson(srv, specs, lam, seed=1005 + k)
        by_policy[f"{lam:g}"] = {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                                 for kk, vv in serve.summarize(res, specs).items()}
        log(f"serving rate {lam:g}/s, arrival order: {by_policy[f'{lam:g}']}")
    srv.close()
    # Smaller HBM budgets at the highest rate: admission control holds
    # requests back until earlier ones are read.
    by_budget = {}
    lam = rates[-1] if rates else 200.0
    for k, gib in enumerate(budgets):
        srv = server(int(gib * 2**30))
        specs = serve.make_requests(vols, n, seed=1106 + k, tfs=tfs, rng_cam_seed=1104 + k,
                                     deferred=True)
        res = serve.run_poisson(srv, specs, lam, seed=1105 + k)
        by_budget[f"{gib:g}"] = {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                                 for kk, vv in serve.summarize(res, specs).items()}
        log(f"serving rate {lam:g}/s, budget {gib:g} GiB: {by_budget[f'{gib:g}']}")
        srv.close()
    v256.free()
    vamr.free()
    return {"datasets": ["256^3 uniform", "512^3 uniform (8 bricks)",
                         f"AMR 128^3+3 levels ({len(amr.grids)} grids)"],
            "sides": [256, 2048], "budget_gib": 8,
            "policy": {"fill": "shortest_first", "age_limit_ms": age_limit_ms}, "rates": out,
            "budget_sweep": {"rate": lam, "budgets_gib": by_budget},
            "policy_sweep": {"policy": "fifo", "rates": by_policy}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--psize", type=int, default=2, help="grids (bricks) per axis")
    ap.add_argument("--cams", type=int, default=64)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="timed passes of the CPU baseline's sample (median reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--lat-reps", type=int, default=20)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the end-to-end drop-in calls (off_axis_projection, Scene.render)")
    ap.add_argument("--serve-rates", default="5,10,20,50,100,200",
                    help="Poisson arrival rates (req/s) of the serving leg")
    ap.add_argument("--serve-budgets", default="1,0.25",
                    help="smaller server budgets (GiB) swept at the highest rate")
    ap.add_argument("--serve-n", type=int, default=200, help="requests per rate")
    ap.add_argument("--serve-age-ms", type=float, default=100.0,
                    help="age limit of the shortest-first chunk fill order (ms)")
    ap.add_argument("--serve-policy-rates", default="50,100,200",
                    help="rates repeated with chunks filled in arrival order")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the single-image configurations (64^3, 256^3 weighted, AMR VR)")
    ap.add_argument("--no-alt", action="store_true",
                    help="skip the 64-brick (128^3 bricks) variant of the batched step")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape autotuner (8x8 tiles, 4 waves/group, occupancy grid)")
    ap.add_argument("--launch", default=None,
                    help="JSON object of launch knobs (autotune.KNOBS) applied after tuning, "
                         'e.g. \'{"vertex_lds": true}\'')
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_k_persist.json"),
                    help="measured HBM bytes per dispatch (rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run the HBM copy probe on this many GiB and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the renderer has no multi-GPU path "
              "(how it would shard is described, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from yt_amd import _lib, autotune, render

    if args.copy_probe > 0:
        ctx = _lib.Context(0)
        g, ms = ctx.hbm_probe(int(args.copy_probe * 2**30), 20)
        print(json.dumps({"copy_bytes": int(args.copy_probe * 2**30), "gbps": g, "ms": ms}))
        ctx.close()
        return

    rho, ds, tf, cams = build_workload(args.n, args.psize, args.cams)
    ctx = _lib.Context(0)
    copy_gbps, _ = ctx.hbm_probe(1 << 30, 10)
    log(f"HBM copy probe: {copy_gbps:.0f} GB/s")
    t = time.time()
    vol = render.ResidentVolume(ctx, ds, proj_fields="density", vr_field="density")
    log(f"resident: {vol.n_proj_bricks} projection blocks, {vol.n_vr_bricks} VR bricks, "
        f"{vol.device_bytes() / 2**30:.2f} GiB in {time.time() - t:.1f}s")
    dataset_bytes = vol.device_bytes()

    # Sample counts (untimed record pass): visits with dt > 0 for projection,
    # mask-1 visits x n_samples for VR.
    strata = sample_strata(cams)
    strata_rows = dict(strata)
    window_samples = {}
    rec = _lib.Batch(ctx, make_specs(vol, tf, cams, record=True), _lib.BATCH_PERSISTENT)
    rec.run()
    samples, npix, visits = [], [], 0
    for i in range(len(cams)):
        v, _, s = rec.read_record(i)
        if i in strata_rows:
            r0, r1 = strata_rows[i]
            window_samples[i] = int(s[r0:r1].sum())
        samples.append(int(s.sum()))
        npix.append(int(v.size))
        visits += int(v.sum())
    rec.close()
    step_samples = sum(samples)
    log(f"samples/step {step_samples:.4e}, visits/step {visits:.4e}, pixels/step {sum(npix)}")

    flags = _lib.BATCH_PERSISTENT | (0 if args.no_graph else _lib.BATCH_GRAPH)
    specs = make_specs(vol, tf, cams)
    cfg = dict(autotune.DEFAULT)
    tune_info = None
    if not args.no_tune:
        t = time.time()
        cfg, table = autotune.tune(ctx, specs, flags, use_cache=False)
        ranked, shown = [], set()
        for c, ms in sorted(table, key=lambda x: x[1]):
            if ms not in shown:  # one line per timed launch configuration
                shown.add(ms)
                ranked.append((c, ms))
        by_key = {json.dumps(c, sort_keys=True): ms for c, ms in table}
        default_ms = by_key.get(json.dumps(autotune.DEFAULT, sort_keys=True), float("nan"))
        timed = len({ms for _, ms in table})  # configurations with the same launches share a timing
        tune_info = {"search": "coordinate descent over tile shape, rays per item, transfer-function "
                               "table bytes per launch, vertex blocks in LDS, waves per workgroup and "
                               "workgroups per CU",
                     "configurations": len(table), "timed": timed, "seconds": round(time.time() - t, 1),
                     "best": autotune.describe(cfg), "best_ms": round(ranked[0][1], 3),
                     "default_ms": round(default_ms, 3),
                     "worst_ms": round(ranked[-1][1], 3)}
        log(f"autotune: {len(table)} configurations ({timed} timed) in {time.time() - t:.1f}s; best "
            + ", ".join(f"{autotune.describe(c)}={ms:.2f}ms" for c, ms in ranked[:4])
            + f"; default {default_ms:.2f}ms; worst {ranked[-1][1]:.2f}ms")
    if args.launch:
        extra = json.loads(args.launch)
        unknown = set(extra) - set(autotune.KNOBS)
        if unknown:
            raise SystemExit(f"--launch: unknown knobs {sorted(unknown)}")
        cfg = dict(autotune._full(cfg), **extra)
        log(f"launch shape fixed to {autotune.describe(cfg)}")
    batch = autotune.make_batch(ctx, specs, cfg, flags)
    for _ in range(args.warmup):
        batch.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.launch()
    batch.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    gsps = step_samples / (wall / args.steps) / 1e9
    # Compulsory bytes of a step: every resident brick is in view of every
    # camera here, so each is read at least once, and every image channel
    # is written once (1 for projection, 3 for VR).
    image_bytes = sum(4 * npix[i] * (1 if cams[i][0] == 0 else 3) for i in range(len(cams)))
    compulsory = dataset_bytes + image_bytes
    fp64_ms = visits * FP64_OPS_PER_VISIT / (FP64_VECTOR_TFLOPS * 1e12) * 1e3
    imgs_s = len(cams) / (wall / args.steps)
    log(f"batched: {ms_step:.2f} ms/step, {gsps:.1f} Gsamples/s, {imgs_s:.1f} images/s")
    # the timed batch's images of the CPU-sampled cameras (compared with the
    # oracle in the cpu_baseline leg)
    gpu_images = {i: batch.read_image(i) for i, _ in strata}
    batch.close()

    # Per-kind persistent kernels (dominant-kernel roofline).
    kinds = {}
    for kind in (0, 1):
        idx = [i for i in range(len(cams)) if cams[i][0] == kind]
        b = autotune.make_batch(ctx, [specs[i] for i in idx], cfg, _lib.BATCH_PERSISTENT)
        b.run()
        tms = []
        for _ in range(max(3, args.steps)):
            tms.append(b.run())
        b.close()
        ms = float(np.mean(tms))
        byt = sum(alg_bytes(kind, samples[i], npix[i]) for i in idx)
        kinds[kind] = dict(ms=ms, bytes=byt, samples=sum(samples[i] for i in idx))
        log(f"{'projection' if kind == 0 else 'VR'} persistent kernel: {ms:.3f} ms, "
            f"{kinds[kind]['samples']:.4e} samples, {byt / ms / 1e6:.1f} GB/s algorithmic")
    dom = max(kinds, key=lambda k: kinds[k]["ms"])
    ach = kinds[dom]["bytes"] / kinds[dom]["ms"] / 1e6
    traffic = measured_traffic(args.traffic_file, dom, args)
    # The binding roof of both kernels is instruction issue: lane-VALU
    # operations per sample and their rate against the FP32 lane peak.
    valu = {}
    for kind in kinds:
        v = measured_valu(args.traffic_file, kind, args)
        if v is None:
            continue
        lane_ops = v["sq_insts_valu"] * 64.0
        valu["vr" if kind == 1 else "projection"] = {
            "lane_valu_per_sample": round(lane_ops / kinds[kind]["samples"], 2),
            "lane_ops_per_s": round(lane_ops / (kinds[kind]["ms"] * 1e-3), 1),
            "frac_of_fp32_lane_peak": round(lane_ops / (kinds[kind]["ms"] * 1e-3) / VALU_LANE_OPS_PEAK, 4),
            "valu_utilization_pct": v["valu_utilization_pct"],
            "insts_per_launch": v["sq_insts_valu"],
        }
    for k, v in valu.items():
        log(f"{k} kernel VALU: {v}")

    # Plain: one launch + sync per image.
    plain = None
    if not args.no_plain:
        singles = [_lib.Batch(ctx, [s], _lib.BATCH_PLAIN) for s in specs]
        for b in singles:
            b.run()
        t0 = time.perf_counter()
        for b in singles:
            b.launch()
            b.sync()
        pw = time.perf_counter() - t0
        for b in singles:
            b.close()
        plain = dict(images_per_s=len(cams) / pw, ms_per_step=pw * 1e3)
        log(f"plain one-call-per-image: {pw * 1e3:.2f} ms for {len(cams)} images "
            f"({len(cams) / pw:.1f} images/s); batched speed-up {imgs_s / plain['images_per_s']:.2f}x")
        # The same comparison per image side: the cameras of one side as a
        # batch (persistent launch in a hipGraph) and one call per image.
        by_side = {}
        for side in sorted({int(p["nv"][0]) for _, p in cams}):
            idx = [i for i in range(len(cams)) if int(cams[i][1]["nv"][0]) == side]
            sub = [specs[i] for i in idx]
            bb = autotune.make_batch(ctx, sub, cfg, flags)
            bb.run()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                bb.launch()
            bb.sync()
            tb = (time.perf_counter() - t0) / args.steps
            bb.close()
            singles = [_lib.Batch(ctx, [s], _lib.BATCH_PLAIN) for s in sub]
            for b in singles:
                b.run()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                for b in singles:
                    b.launch()
                    b.sync()
            tp = (time.perf_counter() - t0) / args.steps
            for b in singles:
                b.close()
            by_side[str(side)] = dict(images=len(sub), batched_images_per_s=round(len(sub) / tb, 1),
                                      plain_images_per_s=round(len(sub) / tp, 1),
                                      speedup=round(tp / tb, 2))
            log(f"side {side}: {by_side[str(side)]}")
        plain["by_side"] = by_side

    # Latency of single 1024^2 requests (launch -> sync on the host clock).
    lat = {}
    for kind, name in (((1, "vr_1024"), (0, "proj_1024")) if args.lat_reps > 0 else ()):
        p = render.camera_params((0.5, 0.5, 0.5), (0.3, 0.5, 0.81), 1.0, 1024)
        s = vol.projection_request(p) if kind == 0 else vol.vr_request(p, tf)
        b = _lib.Batch(ctx, [s], _lib.BATCH_PERSISTENT)
        b.run()
        xs = []
        for _ in range(args.lat_reps):
            t0 = time.perf_counter()
            b.launch()
            b.sync()
            xs.append((time.perf_counter() - t0) * 1e3)
        b.close()
        lat[name] = dict(p50_ms=float(np.percentile(xs, 50)), p99_ms=float(np.percentile(xs, 99)))
    log(f"latency {lat}")

    # The same field in 4x4x4 bricks of 128^3 (64 projection blocks and kd
    # bricks): the batched step again, with its own sample count.
    alt = None
    if not args.no_alt and args.psize != 4:
        from yt_amd import data as ydata

        ds4 = ydata.load_uniform_grid({"density": rho}, psize=(4, 4, 4))
        vol4 = render.ResidentVolume(ctx, ds4, proj_fields="density", vr_field="density")
        rec4 = _lib.Batch(ctx, make_specs(vol4, tf, cams, record=True), _lib.BATCH_PERSISTENT)
        rec4.run()
        s4 = sum(int(rec4.read_record(i)[2].sum()) for i in range(len(cams)))
        rec4.close()
        b4 = autotune.make_batch(ctx, make_specs(vol4, tf, cams), cfg, flags)
        for _ in range(max(1, args.warmup)):
            b4.run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b4.launch()
        b4.sync()
        w4 = (time.perf_counter() - t0) / args.steps
        b4.close()
        vol4.free()
        # the brick split (brick chunk) as the tuner would choose it
        # (autotune.tune_bricks): the faster of the two measured splits
        alt = dict(bricks="4x4x4 of 128^3", samples_per_step=s4, ms_per_step=round(w4 * 1e3, 3),
                   gsamples_per_s=round(s4 / w4 / 1e9, 3),
                   chosen_split=f"{args.psize}^3" if ms_step <= w4 * 1e3 else "4^3")
        log(f"64 bricks of 128^3: {w4 * 1e3:.2f} ms/step, {s4 / w4 / 1e9:.1f} Gsamples/s")
    del rho

    serving = None
    if not args.no_serve:
        serving = serve_leg(ctx, vol, tf, [float(x) for x in args.serve_rates.split(",")],
                            args.serve_n,
                            [float(x) for x in args.serve_budgets.split(",") if x.strip()],
                            [float(x) for x in args.serve_policy_rates.split(",") if x.strip()],
                            args.serve_age_ms)

    dropin = None if args.no_dropin else drop_in_leg()
    configs = None if args.no_configs else configs_leg(ctx, cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ds, vol, tf, cams, strata, window_samples, gpu_images,
                           runs=args.cpu_runs)
        log(f"cpu baseline {cpu['value']:.4f} Gsamples/s on {cpu['cores']} threads")

    vol.free()
    ctx.close()
    out = {
        "metric": METRIC,
        "value": round(gsps, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.n}^3 float32 lognormal (k^-3, sigma_ln=1), {args.psize}^3 bricks, "
                        f"{args.cams} cameras (sides 128-1024), alternating projection / VR "
                        "(8-Gaussian CTF, 10 sub-samples), persistent tile-queue kernels"
                        + ("" if args.no_graph else " in a hipGraph"),
            "images": args.cams,
            "pixels": int(sum(npix)),
            "samples_per_step": int(step_samples),
            "visits_per_step": int(visits),
            "dda": "f64",
            "launch": cfg,
            "parallelism": "1 GPU",
        },
        "autotune": tune_info,
        "images_per_s": round(imgs_s, 2),
        "compulsory": {"bytes": int(compulsory), "ms_at_peak": round(compulsory / HBM_PEAK_GBS / 1e6, 4),
                       "note": "resident bricks read once + image written once, at 8 TB/s"},
        "fp64_dda_ceiling_ms": round(fp64_ms, 3),
        "bricks_64": alt,
        "plain_images_per_s": None if plain is None else round(plain["images_per_s"], 2),
        "batched_vs_plain_by_side": None if plain is None else plain["by_side"],
        "latency_ms": lat,
        "serving": serving,
        "drop_in_ms": dropin,
        "configs": configs,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_persist<VR>" if dom == 1 else "k_persist<projection>",
            "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_ms": round(kinds[dom]["ms"], 4),
            "algorithmic_bytes": int(kinds[dom]["bytes"]),
            "samples": int(kinds[dom]["samples"]),
            "copy_probe_gbps": round(copy_gbps, 1),
            "valu": valu or None,
            "projection_kernel": {"kernel_ms": round(kinds[0]["ms"], 4),
                                  "achieved": round(kinds[0]["bytes"] / kinds[0]["ms"] / 1e6, 1),
                                  "frac": round(kinds[0]["bytes"] / kinds[0]["ms"] / 1e6 / HBM_PEAK_GBS, 4)},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
