#!/usr/bin/env python3
"""Throughput of the full octree build on one MI355X.

One step = one ocpi_generate_octree of the whole cloud (counting, merging, distribution,
subsampling with colour filtering, final placement) with the cloud already resident in HBM
(uploaded once by ocpi_init_octree, outside the timed region). The default workload is the
morrowbay-sized cloud: 119,701,547 points, 27-byte records (3 x float64 + RGB, coordinates offset
by 1e6), grids 512/128, merge threshold 10000, implicit subsampling with distance-weighted
inter-cell colour filtering, output factor 2.2.

Prints one JSON line. `roofline` is for the dominant kernel class, timed with HIP events on the
session stream by the library; `cpu_baseline` times the CPU oracle on a bounded prefix of the same
cloud on this host. `batched` is the batched-build workload (256 clouds of 10K-2M points through
the build service, slot count chosen by the tuner) with its p50/p99 latency (every cloud
submitted at once), against the same clouds built one after another through the plain session
API and through one slot of the service; `serving` is a short sweep of Poisson arrival rates through the service.
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
    return {"workload": "256 clouds, 10K-2M points log-uniform (seed 4), f32/f64 mixed, grids 256/64, T 5000, "
                        "implicit + inter, output factor 3.0",
            "points_per_s": par["points_per_s"], "clouds_per_s": par["clouds_per_s"], "p50_ms": par["p50_ms"],
            "p99_ms": par["p99_ms"], "slots": best, "tuner_points_per_s": {str(k): v for k, v in tuned.items()},
            "plain_sequential_points_per_s": plain["points_per_s"], "plain_sequential_p50_ms": plain["p50_ms"],
            "plain_sequential_p99_ms": plain["p99_ms"],
            "speedup_vs_plain_sequential": par["points_per_s"] / plain["points_per_s"],
            "one_slot_points_per_s": seq["points_per_s"], "one_slot_p50_ms": seq["p50_ms"],
            "one_slot_p99_ms": seq["p99_ms"], "speedup_vs_one_slot": par["points_per_s"] / seq["points_per_s"],
            "failed": par["failed"] + seq["failed"]}


def run_serving(args):
    from octree_amd import batch
    rates = [float(r) for r in args.serving_rates.split(",") if r]
    log("bench: serving: generating the request clouds")
    pool = batch.serving_pool()
    svc = api.Service(args.device, slots=4, memory_budget=200 << 30)
    out = []
    try:
        for rate in rates:
            reqs = batch.serving_requests(pool, rate, args.serving_seconds)
            r = batch.run_serving(svc, reqs, args.serving_seconds)
            log(f"bench: serving {rate} req/s: {r}")
            out.append(r)
    finally:
        svc.close()
    return {"workload": "Poisson arrivals, 50K-20M points log-uniform (seed 5), f32/f64 50/50, method and filter "
                        "uniform, grids 512/128, 200 GB budget, 4 slots",
            "seconds_per_rate": args.serving_seconds, "rates": out}


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
    ap.add_argument("--serving-rates", default="10,20,40,80",
                    help="Poisson arrival rates (req/s) of the serving sweep, '' to skip")
    ap.add_argument("--serving-seconds", type=float, default=4.0)
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
    batched = run_batched(args) if args.batched else None
    serving = run_serving(args) if args.serving_rates else None

    cpu = None
    if args.cpu_sample != 0:
        threads = min(os.cpu_count() or 1, 16)
        sample = s.points if args.cpu_sample < 0 else min(args.cpu_sample, s.points)
        log(f"bench: CPU oracle baseline on {sample:,} points with {threads} threads")
        cpu = cpu_baseline(cloud, wl, cubic, sample, threads)
        if cpu is not None and args.cpu_sample_1t > 0:
            sample1 = min(args.cpu_sample_1t, s.points)
            log(f"bench: CPU oracle baseline on {sample1:,} points with 1 thread")
            one = cpu_baseline(cloud, wl, cubic, sample1, 1)
            if one is not None:
                cpu["single_thread"] = {"value": one["value"], "unit": one["unit"], "sample": one["sample"]}

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
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
