#!/usr/bin/env python3
"""Throughput of the full octree build on one MI355X.

One step = one ocpi_generate_octree of the whole cloud (counting, merging, distribution,
subsampling with colour filtering, final placement) with the cloud already resident in HBM
(uploaded once by ocpi_init_octree, outside the timed region). The default workload is the
morrowbay-sized cloud: 119,701,547 points, 27-byte records (3 x float64 + RGB, coordinates offset
by 1e6), grids 512/128, merge threshold 10000, implicit subsampling with distance-weighted
inter-cell colour filtering, output factor 2.2.

Prints one JSON line. `roofline` is for the dominant kernel class, timed with HIP events on the
session stream by the library; `cpu_baseline` times the CPU oracle on the same cloud on this host.
`host_to_host` is one cloud from host memory to results in host memory (never `value`);
`other_workloads` the device-resident rate of the other named single-cloud shapes. `batched` is the
batched-build workload (256 clouds of 10K-2M points) through the batched pipeline (one launch per
stage for all clouds, launch shapes from its autotuner, HIP graphs) device-resident and from host
memory, beside the slot service (slot count from its tuner) and the same clouds built one after
another through the plain session API; `serving` is a sweep of Poisson arrival rates through the
batch service.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "octree-cuda_amd"))

import numpy as np  # noqa: E402

from octree_amd import api  # noqa: E402
from octree_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kind, n, stride, grid, summ):
    """Compulsory HBM bytes of one build per stage (one input point = one unit):
    count       N*S + 4*G^3                   read every record, write the dense histogram
    distribute  N*(S + 4 + 4 + 18)            record, cell -> leaf lookup, index write, 18-byte record write
    subsample   C*(4 + S + 6) + I*(4 + 18)    C = N + I - I_root candidate points (index, record, colour),
                                              I subsampled points (index + record written)"""
    internal_points = summ["used"] - n
    if kind == "count":
        return n * stride + 4 * grid ** 3
    if kind == "distribute":
        return n * (stride + 4 + 4 + 18)
    if kind == "subsample":
        candidates = n + internal_points - summ["root_points"]
        return candidates * (4 + stride + 6) + internal_points * (4 + 18)
    raise ValueError(kind)


def profiled_traffic(workload):
    """HBM bytes per build and stage from the committed FETCH_SIZE / WRITE_SIZE summary of this
    workload (calibrated against a plain copy), or None when there is none."""
    path = os.path.join(REPO, "profiles", f"{workload.lower()}_traffic_fetch_write.md")
    if not os.path.exists(path):
        return None
    for text in open(path):
        if text.startswith("Machine-readable (bytes per build"):
            pairs = text.split(":", 1)[1].split(",")
            return {k.strip(): int(v) for k, v in (p.split("=") for p in pairs)}
    return None


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size (for scaling the FETCH_SIZE / WRITE_SIZE counters)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"bench: copied {nbytes} bytes device to device 3 times (counter calibration)")


def cpu_baseline(cloud, wl, cubic, sample_points, threads):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_binding as O

    import dataclasses
    s = wl.cloud
    sample = cloud[: sample_points * s.stride]
    spec = wl.build
    if sample_points < s.points:
        # a prefix subsamples relatively more than the whole cloud, so give it room (the output
        # capacity only sizes buffers; the work is the same)
        spec = dataclasses.replace(wl.build, output_factor=max(wl.build.output_factor, 3.0))
    t0 = time.perf_counter()
    r = O.build(sample, sample_points, s.stride, s.cloud_type, cubic, [s.scale] * 3, spec, threads=threads,
                want_data=False)
    dt = time.perf_counter() - t0
    if r.status != 0:
        log(f"bench: CPU oracle failed with status {r.status}")
        return None
    what = "the whole cloud" if sample_points == s.points else f"the first {sample_points:,} points of the same cloud"
    return {"value": sample_points / dt, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (std::thread x {threads}) building {what} with the same bounding box and "
                      f"configuration once, {dt:.1f} s"}


def run_batched(args):
    from octree_amd import batch
    log("bench: batched builds: generating 256 clouds")
    reqs = batch.batch_requests()
    pipeline = run_batch_pipeline(args, batch, reqs)
    slots = 16
    svc = api.Service(args.device, slots=slots, memory_budget=200 << 30)
    try:
        # steady state: every slot's arena grown for this batch (growing an arena frees and
        # reallocates device memory, which synchronises the whole device), then one more pass
        batch.run_batch(svc, reqs, slots)
        batch.run_batch(svc, reqs, slots)
        best, tuned = batch.tune_slots(svc, reqs, candidates=(1, 2, 4, 8, 12, 16), sample=128, repeats=3)
        log(f"bench: slot tuner {dict((k, round(v / 1e6, 1)) for k, v in tuned.items())} Mpts/s -> {best} slots")
        seq = batch.run_batch(svc, reqs, 1)
        par = batch.run_batch(svc, reqs, best)
        log(f"bench: one slot {seq}")
        log(f"bench: batched {par}")
    finally:
        svc.close()
    batch.run_plain_sequential(reqs[:16])  # warm-up of the session path
    plain = batch.run_plain_sequential(reqs)
    log(f"bench: plain sequential {plain}")
    e2e = pipeline["host_clouds"]
    return {"workload": "256 clouds, 10K-2M points log-uniform (seed 4), f32/f64 mixed, grids 256/64, T 5000, "
                        "implicit + inter, output factor 3.0",
            "pipeline": pipeline,
            "points_per_s": e2e["points_per_s"], "p50_ms": e2e["p50_ms"], "p99_ms": e2e["p99_ms"],
            "slot_service": {"points_per_s": par["points_per_s"], "clouds_per_s": par["clouds_per_s"],
                             "p50_ms": par["p50_ms"], "p99_ms": par["p99_ms"], "slots": best,
                             "tuner_points_per_s": {str(k): v for k, v in tuned.items()},
                             "one_slot_points_per_s": seq["points_per_s"], "one_slot_p50_ms": seq["p50_ms"],
                             "one_slot_p99_ms": seq["p99_ms"]},
            "plain_sequential_points_per_s": plain["points_per_s"], "plain_sequential_p50_ms": plain["p50_ms"],
            "plain_sequential_p99_ms": plain["p99_ms"],
            "speedup_vs_plain_sequential": e2e["points_per_s"] / plain["points_per_s"],
            "speedup_vs_one_slot": e2e["points_per_s"] / seq["points_per_s"],
            "failed": par["failed"] + seq["failed"] + e2e["failed"]}


def run_batch_pipeline(args, batch, reqs):
    """The 256 clouds through the batched pipeline (ocpi_batch_build): launch shapes chosen by the
    autotuner on the device-resident batch, then the device-resident rate (clouds already in HBM,
    results copied to host memory) and the end-to-end rate (clouds in host memory) over `repeats`."""
    arr, addrs = batch.upload_clouds(reqs)
    try:
        with api.BatchBuilder(args.device) as bb:
            batch.run_batch_builder(bb, reqs, addrs)  # warm-up: arena, pinned block, graphs
            best, history = batch.tune_batch_launch(bb, reqs, addrs, repeats=2)
            log(f"bench: batch launch tuner -> {best}")
            dev = [batch.run_batch_builder(bb, reqs, addrs) for _ in range(3)]
            host = [batch.run_batch_builder(bb, reqs, None) for _ in range(3)]
            timings = bb.timings()
    finally:
        arr.close()
    d = min(dev, key=lambda r: r["device_ms"])
    h = min(host, key=lambda r: r["seconds"])
    log(f"bench: batched pipeline device-resident {d}")
    log(f"bench: batched pipeline host clouds {h}")
    return {"launch": best, "tuner": history, "graphs": True,
            "device_ms": d["device_ms"], "device_points_per_s": d["device_points_per_s"],
            "device_resident": {k: d[k] for k in ("points_per_s", "seconds", "p50_ms", "p99_ms", "copy_ms", "failed")},
            "host_clouds": {k: h[k] for k in ("points_per_s", "seconds", "p50_ms", "p99_ms", "copy_ms", "failed")},
            "last_timings_ms": timings}


def run_serving(args):
    from octree_amd import batch
    rates = [float(r) for r in args.serving_rates.split(",") if r]
    log("bench: serving: generating the request clouds")
    pool = batch.serving_pool()
    svc = api.Service(args.device, slots=args.serving_workers, memory_budget=200 << 30,
                      batch_size=args.serving_batch)
    out = []
    try:
        warm = batch.serving_requests(pool, 20.0, 1.0, seed=batch.SERVE_SEED + 100)
        batch.run_serving(svc, warm, 1.0)  # arenas and pinned blocks
        for rate in rates:
            reqs = batch.serving_requests(pool, rate, args.serving_seconds)
            r = batch.run_serving(svc, reqs, args.serving_seconds)
            log(f"bench: serving {rate} req/s: {r}")
            out.append(r)
    finally:
        svc.close()
    return {"workload": "Poisson arrivals, 50K-20M points log-uniform (seed 5), f32/f64 50/50, method and filter "
                        "uniform, grids 512/128, T 10000, output factor 3.0 (these clouds need N + I = 2.45 N, "
                        "so the 2.2 estimate would fail them), admission by footprint against a 200 GB budget",
            "service": f"batch service, {args.serving_workers} workers, up to {args.serving_batch} requests per batch",
            "seconds_per_rate": args.serving_seconds, "rates": out}


def run_other_workloads(args):
    """Device-resident build rate and the stage fractions of the other named single-cloud shapes."""
    out = {}
    for name in [w for w in args.other_workloads.split(",") if w]:
        wl = W.WORKLOADS[name]
        s = wl.cloud
        log(f"bench: {name}: generating {wl.name}")
        cloud = W.generate(s)
        _, cubic = W.bounding_box(cloud, s.points, s.stride, s.cloud_type)
        with api.Session(args.device) as sess:
            sess.set_cloud(cloud, s.points, s.stride, s.cloud_type, cubic, [s.scale] * 3)
            sess.configure(wl.build)
            sess.init()
            for _ in range(2):
                sess.generate()
            ms, tims = [], []
            for _ in range(5):
                t0 = time.perf_counter()
                sess.generate()
                ms.append(1e3 * (time.perf_counter() - t0))
                tims.append(sess.timings())
            summ = sess.summary()
            nodes = sess.nodes()
        summ["root_points"] = int(nodes.point_count[-1]) if nodes.internal[-1] else 0
        avg = {k: float(np.mean([t[k] for t in tims])) for k in tims[0]}
        stages = {}
        for kind, key in (("count", "count_ms"), ("distribute", "distribute_ms"), ("subsample", "subsample_ms")):
            b = algorithmic_bytes(kind, s.points, s.stride, wl.build.chunking_grid, summ)
            gbs = b / (avg[key] * 1e-3) / 1e9
            stages[kind] = {"ms": round(avg[key], 4), "bytes": b, "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        out[name] = {"workload": wl.name, "points_per_s": s.points / (float(np.mean(ms)) * 1e-3),
                     "ms_per_build": round(float(np.mean(ms)), 3), "stages": stages}
        log(f"bench: {name}: {out[name]}")
    return out


def host_to_host(args, wl, cloud, cubic):
    """One cloud from host memory to its results in host memory: upload (ocpi_init_octree), build,
    and the copies of the nodes, LUT and output records (what an export reads)."""
    s = wl.cloud
    best = None
    with api.Session(args.device) as sess:
        sess.set_cloud(cloud, s.points, s.stride, s.cloud_type, cubic, [s.scale] * 3)
        sess.configure(wl.build)
        for _ in range(3):
            t0 = time.perf_counter()
            sess.init()
            t1 = time.perf_counter()
            sess.generate()
            t2 = time.perf_counter()
            used = sess.summary()["used"]
            sess.nodes()
            sess.point_lut(used)
            sess.output(used)
            t3 = time.perf_counter()
            r = {"points_per_s": s.points / (t3 - t0), "ms": 1e3 * (t3 - t0), "upload_ms": 1e3 * (t1 - t0),
                 "build_ms": 1e3 * (t2 - t1), "download_ms": 1e3 * (t3 - t2)}
            if best is None or r["ms"] < best["ms"]:
                best = r
    log(f"bench: host to host {best}")
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="C3", choices=sorted(W.WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="points of the cloud the CPU oracle baseline builds with all cores (-1 = all, 0 = skip)")
    ap.add_argument("--cpu-sample-1t", type=int, default=4_000_000,
                    help="points of the cloud the single-thread CPU oracle baseline builds (0 = skip)")
    ap.add_argument("--batched", type=int, default=1, help="also run the batched-build workload (0 = skip)")
    ap.add_argument("--serving-rates", default="5,20,50,100,200",
                    help="Poisson arrival rates (req/s) of the serving sweep, '' to skip")
    ap.add_argument("--serving-seconds", type=float, default=6.0,
                    help="seconds of arrivals per rate (the configured sweep is 60; the default keeps the run short)")
    ap.add_argument("--serving-workers", type=int, default=2)
    ap.add_argument("--serving-batch", type=int, default=64)
    ap.add_argument("--other-workloads", default="C1,C2", help="single-cloud shapes also measured ('' to skip)")
    ap.add_argument("--host-to-host", type=int, default=1, help="also time upload + build + result copies")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, help="also run a plain D2D copy of this many bytes")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the octree build of one cloud "
              f"runs on a single MI355X; building one cloud across several GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    wl = W.WORKLOADS[args.workload]
    s = wl.cloud
    log(f"bench: generating {wl.name} ({s.points:,} points)")
    t0 = time.perf_counter()
    cloud = W.generate(s)
    _, cubic = W.bounding_box(cloud, s.points, s.stride, s.cloud_type)
    log(f"bench: cloud ready in {time.perf_counter() - t0:.1f} s, cubic bb {cubic}")

    sess = api.Session(args.device)
    sess.set_cloud(cloud, s.points, s.stride, s.cloud_type, cubic, [s.scale] * 3)
    sess.configure(wl.build)
    t0 = time.perf_counter()
    sess.init()
    log(f"bench: init (allocation + upload) {time.perf_counter() - t0:.2f} s")

    for i in range(args.warmup):
        t0 = time.perf_counter()
        sess.generate()
        log(f"bench: warmup {i} {1e3 * (time.perf_counter() - t0):.2f} ms")

    step_ms, timings = [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        sess.generate()  # synchronous: returns after the stream has drained
        step_ms.append(1e3 * (time.perf_counter() - t0))
        timings.append(sess.timings())
        log(f"bench: step {i} {step_ms[-1]:.2f} ms {timings[-1]}")
    summ = sess.summary()
    nodes = sess.nodes()
    summ["root_points"] = int(nodes.point_count[-1]) if nodes.internal[-1] else 0
    ms = float(np.mean(step_ms))
    avg = {k: float(np.mean([t[k] for t in timings])) for k in timings[0]}

    # roofline of the dominant stage (device time from HIP events on the session stream)
    stage_ms = {"count": avg["count_ms"], "distribute": avg["distribute_ms"], "subsample": avg["subsample_ms"]}
    grid = wl.build.chunking_grid

    def rl(kind):
        b = algorithmic_bytes(kind, s.points, s.stride, grid, summ)
        gbs = b / (stage_ms[kind] * 1e-3) / 1e9
        return b, gbs

    dominant = max(stage_ms, key=stage_ms.get)
    rl_bytes, achieved = rl(dominant)
    traffic = profiled_traffic(args.workload)
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic.get(dominant) if traffic else None,
                "traffic_stages": traffic,
                "traffic_source": f"profiles/{args.workload.lower()}_traffic_fetch_write.md" if traffic else None,
                "bytes_per_launch": rl_bytes, "avg_launch_ms": round(stage_ms[dominant], 4),
                "stages": {k: {"ms": round(v, 4), "bytes": rl(k)[0], "GBs": round(rl(k)[1], 1),
                               "frac": round(rl(k)[1] / HBM_PEAK_GBS, 4)} for k, v in stage_ms.items()},
                "whole_build": {"bytes": sum(rl(k)[0] for k in stage_ms) + 4 * summ["dense_nodes"],
                                "GBs": round((sum(rl(k)[0] for k in stage_ms) + 4 * summ["dense_nodes"])
                                             / (avg["build_ms"] * 1e-3) / 1e9, 1)}}
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)

    sess.close()
    h2h = host_to_host(args, wl, cloud, cubic) if args.host_to_host else None
    others = run_other_workloads(args) if args.other_workloads else None
    batched = run_batched(args) if args.batched else None
    serving = run_serving(args) if args.serving_rates else None

    cpu = None
    if args.cpu_sample != 0:
        # every core this process may run on, but no more worker threads than the host grants
        # one GPU's job (OMP_NUM_THREADS, when set)
        affinity = len(os.sched_getaffinity(0))
        threads = min(affinity, int(os.environ.get("OMP_NUM_THREADS", affinity)))
        sample = s.points if args.cpu_sample < 0 else min(args.cpu_sample, s.points)
        log(f"bench: CPU oracle baseline on {sample:,} points with {threads} threads")
        cpu = cpu_baseline(cloud, wl, cubic, sample, threads)
        if cpu is not None and args.cpu_sample_1t > 0:
            sample1 = min(args.cpu_sample_1t, s.points)
            log(f"bench: CPU oracle baseline on {sample1:,} points with 1 thread")
            one = cpu_baseline(cloud, wl, cubic, sample1, 1)
            if one is not None:
                cpu["single_thread"] = {"value": one["value"], "unit": one["unit"], "sample": one["sample"]}
        if cpu is not None:
            cpu["hardware_concurrency"] = os.cpu_count()
            cpu["affinity_cores"] = affinity
            cpu["note"] = ("threads = the cores this process may use, capped by OMP_NUM_THREADS (the host's share "
                           "for one GPU); the oracle distributes the points into the leaves serially")

    line = {
        "metric": "points/sec full octree build (1 MI355X, % HBM roofline)",
        "value": s.points / (ms * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if s.cloud_type == 1 else "f32",
        "data": "synthetic (seeded heightfield + Gaussian clusters, generated on the host)",
        "config": {"workload": wl.name, "points": s.points, "record_bytes": s.stride,
                   "chunking_grid": wl.build.chunking_grid, "subsampling_grid": wl.build.subsampling_grid,
                   "merging_threshold": wl.build.merging_threshold, "output_factor": wl.build.output_factor,
                   "intra": wl.build.intra, "inter": wl.build.inter, "weighted": wl.build.weighted,
                   "explicit_random": wl.build.random, "sparse_nodes": summ["sparse_nodes"],
                   "output_records": summ["used"], "stage_ms": {k: round(v, 4) for k, v in avg.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_to_host": h2h,
        "other_workloads": others,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
