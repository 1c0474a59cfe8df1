"""Batched builds and the serving loop on the build service (ocpi_create_service).

Batched build: 256 independent clouds, sizes log-uniform in [10K, 2M] points (seed 4), float32
and float64 records mixed, boxes log-uniform in [1, 1000] units, chunking / subsampling grids
256 / 64, merge threshold 5000, implicit subsampling with inter-cell averaging. All clouds are
submitted at once; the service builds them on its slots with continuous admission. The slot
count is chosen by `tune_slots`, which measures a sample of the batch at each candidate count.
The same clouds through one slot are the sequential reference.

Serving loop: Poisson arrivals at a given rate; sizes log-uniform in [50K, 20M] points, float32
or float64 50/50, subsampling method and colour filter uniform per request (seed 5); grids
512 / 128, threshold 10000; admission against a 200 GB device budget by the footprint estimate
(E = N * output factor). Latency is submit to result attached to the session; the queue depth is
sampled while the run lasts.

Output factor: these synthetic clouds need N + I of about 2.45 N output records (a build with
N + I > E is refused), so requests use 3.0 where the reference's default is 2.2.
"""
from __future__ import annotations

import dataclasses
import math
import queue
import threading
import time

import numpy as np

from . import api
from . import workloads as W

BATCH_SEED = 4
SERVE_SEED = 5
OUTPUT_FACTOR = 3.0

# (intra, inter, weighted, random): implicit / explicit method x none / intra / inter / weighted filter
MODES = [(False, False, False, False), (True, False, False, False), (False, True, False, False),
         (False, True, True, False), (False, False, False, True), (True, False, False, True),
         (False, True, False, True), (False, True, True, True)]


@dataclasses.dataclass
class Request:
    cloud: np.ndarray  # host records (possibly a prefix view of a larger cloud)
    points: int
    cloud_type: int
    cubic: list
    scale: float
    build: W.BuildSpec
    arrival: float = 0.0  # seconds from the start of a serving run

    @property
    def stride(self):
        return 15 if self.cloud_type == 0 else 27

    def session(self, device: int = 0, device_pointer: int | None = None) -> api.Session:
        """A configured session; with device_pointer, the cloud is read in place from HBM."""
        s = api.Session(device)
        s.set_cloud(self.cloud, self.points, self.stride, self.cloud_type, self.cubic, [self.scale] * 3,
                    device_pointer=device_pointer)
        s.configure(self.build)
        return s


def _log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def batch_requests(count: int = 256, seed: int = BATCH_SEED, lo: int = 10_000, hi: int = 2_000_000):
    """The batched-build workload: `count` independent clouds."""
    rng = np.random.default_rng(seed)
    sizes = _log_uniform(rng, lo, hi, count).astype(np.int64)
    sides = _log_uniform(rng, 1.0, 1000.0, count)
    types = rng.integers(0, 2, count)
    out = []
    for i in range(count):
        side = float(sides[i])
        spec = W.CloudSpec(int(sizes[i]), int(types[i]), (side, side * float(rng.uniform(0.5, 1.0)), side * 0.2),
                           int(rng.integers(2, 12)), seed * 1000 + i, sigma=side * 0.01,
                           offset=0.0 if types[i] == 0 else 1e5)
        cloud = W.generate(spec)
        _, cubic = W.bounding_box(cloud, spec.points, spec.stride, spec.cloud_type)
        build = W.BuildSpec(256, 5000, OUTPUT_FACTOR, 64, inter=True)
        out.append(Request(cloud, spec.points, spec.cloud_type, cubic, spec.scale, build))
    return out


def run_batch(service: api.Service, requests, slots: int | None = None):
    """Submits every request at once and waits for all; returns throughput and latency figures."""
    if slots is not None:
        service.set_active_slots(slots)
    service.reset_stats()
    sessions = [r.session() for r in requests]
    t0 = time.perf_counter()
    for s in sessions:
        service.submit(s)
    for s in sessions:
        service.wait(s)
    wall = time.perf_counter() - t0
    st = service.stats()
    for s in sessions:
        s.close()
    points = sum(r.points for r in requests)
    return {"clouds": len(requests), "points": points, "seconds": wall, "points_per_s": points / wall,
            "clouds_per_s": len(requests) / wall, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
            "max_queue_depth": st["max_queue_depth"], "slots": int(st["active_slots"]),
            "peak_reserved_bytes": st["peak_reserved_bytes"], "failed": int(st["failed"]),
            "setup_ms": st["setup_ms"], "generate_ms": st["generate_ms"], "capture_ms": st["capture_ms"]}


def run_plain_sequential(requests, device: int = 0):
    """The same clouds built one after another through the plain session API, as a caller of the
    reference API would: init (upload), generate, copy the nodes, LUT and output to host memory."""
    lat = []
    t0 = time.perf_counter()
    for r in requests:
        t = time.perf_counter()
        s = r.session(device)
        s.init()
        s.generate()
        used = s.summary()["used"]
        s.nodes()
        s.point_lut(used)
        s.output(used)
        s.close()
        lat.append(time.perf_counter() - t0)
    wall = time.perf_counter() - t0
    points = sum(r.points for r in requests)
    lat_ms = np.array(lat) * 1e3  # completion times with every request submitted at t0
    return {"clouds": len(requests), "points": points, "seconds": wall, "points_per_s": points / wall,
            "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99))}


def tune_slots(service: api.Service, requests, candidates=(1, 2, 3, 4, 6, 8), sample: int = 64, seed: int = 0,
               repeats: int = 2):
    """Slot-count tuner: builds a fixed sample of the batch at each candidate count (after one
    warm-up pass that grows the arenas), `repeats` times, and keeps the count with the highest
    best-of-repeats throughput."""
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(requests), size=min(sample, len(requests)), replace=False)
    subset = [requests[i] for i in sorted(pick)]
    run_batch(service, subset, max(candidates))  # warm-up: code objects and, for a cold service, arenas
    results = {}
    for k in candidates:
        results[k] = max(run_batch(service, subset, k)["points_per_s"] for _ in range(repeats))
    best = max(results, key=results.get)
    return best, results


def upload_clouds(requests):
    """All request clouds in one device array (256-byte aligned); returns (array, addresses)."""
    from .device import DeviceArray
    offsets, total = [], 0
    for r in requests:
        offsets.append(total)
        total += (r.points * r.stride + 255) & ~255
    arr = DeviceArray(total)
    for r, off in zip(requests, offsets):
        arr.upload(r.cloud[: r.points * r.stride], off)
    return arr, [arr.address(off) for off in offsets]


def run_batch_builder(builder: api.BatchBuilder, requests, addresses=None):
    """All requests built in one batched call (ocpi_batch_build); with addresses (clouds already in
    HBM, upload_clouds) the clouds are read in place, otherwise uploaded from host memory. Every
    request is submitted at t0 and finishes when the call returns, so p50 = p99 = the call's time.
    Returns throughput figures: end to end (results in host memory) and the device pipeline alone."""
    sessions = [r.session(device_pointer=None if addresses is None else addresses[i])
                for i, r in enumerate(requests)]
    t0 = time.perf_counter()
    status = builder.build(sessions, raise_on_failure=False)
    wall = time.perf_counter() - t0
    t = builder.timings()
    for s in sessions:
        s.close()
    points = sum(r.points for r in requests)
    return {"clouds": len(requests), "points": points, "seconds": wall, "points_per_s": points / wall,
            "device_ms": t["device_ms"], "device_points_per_s": points / (t["device_ms"] * 1e-3),
            "copy_ms": t["copy_ms"], "p50_ms": wall * 1e3, "p99_ms": wall * 1e3,
            "failed": int(np.count_nonzero(status))}


def tune_batch_launch(builder: api.BatchBuilder, requests, addresses=None, repeats: int = 2,
                      candidates=None):
    """Launch-shape autotuner of the batched pipeline: coordinate descent over the persistent
    workgroups per launch of its three kernel classes (sort chunks of 1024 threads, subsampling
    tiles, one item per thread), each shape timed by the device time of the whole batch, best of
    `repeats` builds (the first build of a shape also captures its graphs). Returns the chosen
    (thread, tile, sort) grid and every measurement."""
    if candidates is None:
        candidates = {"sort": (256, 512, 1024, 2048), "tile": (1024, 2048, 4096, 8192, 16384),
                      "thread": (512, 1024, 2048, 4096, 8192)}
    best = {"thread": 2048, "tile": 4096, "sort": 1024}
    history = []

    def measure(shape):
        builder.set_launch(shape["thread"], shape["tile"], shape["sort"])
        times = []
        for _ in range(repeats):
            r = run_batch_builder(builder, requests, addresses)
            if r["failed"]:
                raise RuntimeError(f"{r['failed']} requests failed while tuning")
            times.append(r["device_ms"])
        history.append({**shape, "device_ms": min(times)})
        return min(times)

    for knob in ("sort", "tile", "thread"):
        results = {c: measure({**best, knob: c}) for c in candidates[knob]}
        best[knob] = min(results, key=results.get)
    builder.set_launch(best["thread"], best["tile"], best["sort"])
    return best, history


PREFIX_BLOCK = 1 << 16


class PrefixCloud:
    """A large cloud whose prefixes serve as request clouds (the generator's order is random, so
    the first n points are a smaller cloud of the same shape), with running per-block minima and
    maxima so that the bounding box of any prefix costs one partial block."""

    def __init__(self, spec: W.CloudSpec):
        self.spec = spec
        self.records = W.generate(spec)
        blocks = spec.points // PREFIX_BLOCK
        lo = np.empty((blocks, 3))
        hi = np.empty((blocks, 3))
        span = PREFIX_BLOCK * spec.stride
        for b in range(blocks):
            real, _ = W.bounding_box(self.records[b * span:(b + 1) * span], PREFIX_BLOCK, spec.stride,
                                     spec.cloud_type, threads=1)
            lo[b], hi[b] = real[:3], real[3:]
        self.cummin = np.minimum.accumulate(lo, axis=0)
        self.cummax = np.maximum.accumulate(hi, axis=0)

    def prefix(self, n: int):
        """(records of the first n points, their cubic bounding box as src/boundingbox.h:10-52)."""
        spec = self.spec
        k = n // PREFIX_BLOCK
        rest = n - k * PREFIX_BLOCK
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        if k:
            lo, hi = self.cummin[k - 1].copy(), self.cummax[k - 1].copy()
        if rest:
            part = self.records[k * PREFIX_BLOCK * spec.stride: n * spec.stride]
            real, _ = W.bounding_box(part, rest, spec.stride, spec.cloud_type)
            lo = np.minimum(lo, real[:3])
            hi = np.maximum(hi, real[3:])
        dims = hi - lo
        side = dims.max()
        mn = lo - (side - dims) / 2.0
        return self.records[: n * spec.stride], [float(v) for v in mn] + [float(v) for v in mn + side]


def serving_pool(seed: int = SERVE_SEED, max_points: int = 20_000_000):
    """One large cloud per record type for the serving loop's requests."""
    return {t: PrefixCloud(W.CloudSpec(max_points, t, (400.0, 400.0, 60.0), 24, seed * 10 + t, sigma=2.0,
                                       offset=0.0 if t == 0 else 1e6)) for t in (0, 1)}


def serving_requests(pool, rate: float, duration: float, seed: int = SERVE_SEED, lo: int = 50_000,
                     hi: int = 20_000_000):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        n = int(_log_uniform(rng, lo, hi))
        ctype = int(rng.integers(0, 2))
        intra, inter, weighted, rnd = MODES[int(rng.integers(0, len(MODES)))]
        spec = pool[ctype].spec
        cloud, cubic = pool[ctype].prefix(n)
        build = W.BuildSpec(512, 10000, OUTPUT_FACTOR, 128, intra=intra, inter=inter, weighted=weighted,
                            random=rnd, seed=int(rng.integers(0, 2**31)))
        out.append(Request(cloud, n, ctype, cubic, spec.scale, build, arrival=t))
    return out


def run_serving(service: api.Service, requests, duration: float, sample_every: float = 0.01):
    """Submits each request at its arrival time, collects results as they finish and samples the
    admission-queue depth. Returns throughput, latency percentiles and queue-depth figures."""
    service.reset_stats()
    sessions = [r.session() for r in requests]
    depth, stop = [], threading.Event()

    def monitor():
        while not stop.is_set():
            depth.append(service.stats()["queued"])
            time.sleep(sample_every)

    done_points = [0]
    submitted: queue.Queue = queue.Queue()

    def collect():
        for _ in requests:
            r, s = submitted.get()
            service.wait(s)
            done_points[0] += r.points
            s.close()  # frees the host copy of the result

    mon = threading.Thread(target=monitor, daemon=True)
    col = threading.Thread(target=collect, daemon=True)
    t0 = time.perf_counter()
    mon.start()
    col.start()
    for r, s in zip(requests, sessions):
        delay = t0 + r.arrival - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        service.submit(s)
        submitted.put((r, s))
    col.join()
    wall = time.perf_counter() - t0
    stop.set()
    mon.join()
    st = service.stats()
    rate = len(requests) / duration if duration > 0 else 0.0
    return {"offered_req_per_s": rate, "requests": len(requests), "seconds": wall,
            "completed": int(st["completed"]), "failed": int(st["failed"]),
            "points_per_s": done_points[0] / wall, "req_per_s": len(requests) / wall,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "mean_queue_depth": float(np.mean(depth)) if depth else 0.0,
            "max_queue_depth": int(st["max_queue_depth"]), "peak_reserved_bytes": st["peak_reserved_bytes"]}


def main(argv=None):
    """Command line: build the batched workload, e.g. under a kernel trace:
    `python3 -m octree_amd.batch --builder batch --repeats 3` (run from octree-cuda_amd/, or with it
    on PYTHONPATH). `--builder batch` runs the batched pipeline (one launch per stage for all clouds,
    clouds resident in device memory, graph replay, the launch shape given by --launch);
    `--builder slots` the slot service with --slots slots. The first pass grows the arenas; every
    pass prints one JSON line."""
    import argparse
    import json

    ap = argparse.ArgumentParser(prog="python3 -m octree_amd.batch")
    ap.add_argument("--builder", choices=("batch", "slots"), default="batch")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--launch", default="8192,16384,2048", help="batched launch shape: thread,tile,sort grids")
    ap.add_argument("--repeats", type=int, default=2, help="passes over the batch (the first warms the arenas)")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    reqs = batch_requests()
    if args.builder == "slots":
        with api.Service(args.device, slots=args.slots, memory_budget=200 << 30) as svc:
            for rep in range(args.repeats):
                r = run_batch(svc, reqs, args.slots)
                print(json.dumps({"pass": rep, **r}), flush=True)
        return
    thread, tile, sort = (int(v) for v in args.launch.split(","))
    clouds, addresses = upload_clouds(reqs)
    try:
        with api.BatchBuilder(args.device) as bb:
            bb.set_launch(thread, tile, sort)
            bb.use_graphs(True)
            for rep in range(args.repeats):
                r = run_batch_builder(bb, reqs, addresses)
                print(json.dumps({"pass": rep, **r, "stage_ms": bb.timings()}), flush=True)
    finally:
        clouds.close()


if __name__ == "__main__":
    main()
