"""Serving loop over librtbvh.so: resident scenes, a Poisson request stream,
admission against a device-memory budget, continuous batching.

Requests (the serving configuration): a ray batch of 4k-2M rays of kind
shadow, AO or closest-hit against one of the resident scenes, or a
deform-and-rebuild of a 4k-64k-triangle grid.  Scenes stay resident for
the server's lifetime; request buffers (rays, results, rebuilt trees) are
charged against `budget_bytes` from admission to completion.

Scheduling (one host thread):
  * a request becomes visible at its arrival time and waits FIFO;
  * a tick is started only while fewer than `max_ticks` earlier ticks are
    still on the device, so requests arriving while the GPU is busy gather
    into the next tick's batch (continuous batching) instead of each going
    out alone;
  * a tick started while the device is idle takes at most `idle_batch`
    requests, so the first launch is not held back by a full tick's host-side
    setup; ticks started while others run take up to `max_batch`;
  * each tick admits from the head of the queue while the head's bytes fit
    in the budget and fewer than `max_inflight` requests are in flight
    (FIFO: a large head is never overtaken);
  * all ray requests admitted in a tick get their rays generated and are
    traced together in ONE rt_trace_batch launch (schedule and launch shape
    `trace_shape`, e.g. the tuner's choice); all rebuilds admitted in a tick
    are built together in ONE rt_gpu_build_batch launch (one workgroup per
    mesh) on the next of `rebuild_streams` streams; ticks rotate over
    `streams` streams, so work admitted while earlier work is still on the
    device overlaps with it;
  * completion is polled through events; a finished request releases its
    bytes and its latency (arrival to observed completion) is recorded.

The plain path (`run_plain`) serves the same requests one at a time with one
trace launch (or one rt_gpu_build) per request and a wait after each, with
the library's default launch shape; the results of both paths are compared
by `same_results`.  Rebuilt trees are the implicit (fast_bvh_builder) layout
on both paths.

Synthetic rays of each kind (origins uniform in the scene's bounds scaled by
1.2, directions uniform on the sphere, seeded per request): closest-hit with
unbounded tmax; shadow any-hit with tmax = 0.5 (a segment to a light at that
distance); AO any-hit with tmax = 0.05 (a short hemisphere probe).
"""
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

import rtbvh as rt
from timing import median, percentile

KINDS = ("shadow", "ao", "closest", "rebuild")
TMAX = {"shadow": 0.5, "ao": 0.05, "closest": 1e30}


# p50 = median (sorted[n // 2]) and p99 = sorted[floor(0.99 (n - 1))]: the
# statistics of common/vk_timer.cpp:96-108 (timing.py).


# --------------------------------------------------------------- scenes ----
@dataclass
class Scene:
    bvh: object
    pos: torch.Tensor
    idx: torch.Tensor
    lo: tuple
    hi: tuple
    tris: int

    @property
    def bytes(self):
        return (self.bvh.nodes.numel() + self.bvh.tris.numel() + self.pos.numel() * 4 +
                self.idx.numel() * 4)


def make_scenes(count=32, lo=10_000, hi=2_500_000, seed=5):
    """`count` resident scenes, triangle counts log-uniform in [lo, hi]
    (seeded): displaced UV spheres and random heightfields in turn, each
    built once on the device (radix builder)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        if k % 2 == 0:
            nu = max(8, int(math.sqrt(n)))
            nv = max(2, n // (2 * nu))
            pos, idx = rt.gen_sphere(nu, nv, 1.0, 0.02, 100 + k)
            b = (-1.02, -1.02, -1.02), (1.02, 1.02, 1.02)
        else:
            nx = max(4, int(math.sqrt(n / 2)))
            ny = max(2, n // (2 * nx))
            pos, idx = rt.gen_heightfield(nx, ny, 1.0, 0.05, 200 + k)
            b = (-1.0, -1.0, -0.06), (1.0, 1.0, 0.06)
        bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix")
        out.append(Scene(bvh, pos, idx, b[0], b[1], idx.shape[0]))
    torch.cuda.synchronize()
    return out


# ------------------------------------------------------------- requests ----
@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the stream starts
    kind: str
    scene: int
    n: int          # rays, or triangles for a rebuild
    seed: int
    # filled while served
    bytes: int = 0
    admitted: float = -1.0
    done: float = -1.0
    rays: object = None
    out: object = None
    bvh: object = None
    keep: list = field(default_factory=list)

    @property
    def latency(self):
        return self.done - self.arrival


def make_requests(count, qps, scenes, seed=9, rays=(4_096, 2_000_000),
                  rebuild_tris=(4_096, 65_536), kinds=KINDS):
    """Poisson arrivals at `qps` (exponential gaps, seeded); kind uniform over
    `kinds`; ray counts and rebuild sizes log-uniform in their ranges."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / qps) if qps else 0.0
        kind = kinds[int(rng.integers(len(kinds)))]
        lo, hi = rebuild_tris if kind == "rebuild" else rays
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        if kind == "rebuild":
            n -= n % 2  # the grid generator makes 2 triangles per quad
        out.append(Request(i, t, kind, int(rng.integers(len(scenes))), n,
                           int(rng.integers(1 << 30))))
    return out


def request_bytes(r):
    """Device bytes a request holds from admission to completion."""
    if r.kind == "rebuild":
        cap, scratch = rt.build_requirements(r.n, rt.build_cfg("implicit"))
        return r.n * 36 + cap * rt.NODE_BYTES + r.n * rt.TRI_BYTES + scratch + 64
    out = r.n * 8 if r.kind == "closest" else ((r.n + 31) // 32) * 4
    return r.n * 32 + out


def _rays_for(r, scene, stream=None):
    lo = tuple(1.2 * v for v in scene.lo)
    hi = tuple(1.2 * v for v in scene.hi)
    return rt.gen_random_rays(r.n, r.seed, lo, hi, TMAX[r.kind], stream=stream)


def _grid(r, stream=None):
    return rt.gen_dynamic_grid(r.n, (r.seed % 200) * 0.05, 2.0, stream=stream)


def _rebuild(r, stream=None):
    xyz = _grid(r, stream)
    bvh = rt.GpuBVH4.build(xyz, None, algorithm="implicit", stream=stream, sync=False)
    return xyz, bvh


def _trace_alone(r, scene, stream=None, shape=None):
    if r.kind == "closest":
        return scene.bvh.intersect(r.rays, stream=stream, **(shape or {}))
    return scene.bvh.occluded(r.rays, stream=stream, **(shape or {}))


# --------------------------------------------------------------- server ----
class Server:
    def __init__(self, scenes, budget_bytes=2 << 30, streams=3, rebuild_streams=4,
                 max_inflight=256, max_batch=128, keep_results=False, trace_shape=None,
                 max_ticks=None, idle_batch=16):
        self.scenes = scenes
        self.max_ticks = max_ticks or streams
        self.trace_shape = dict(trace_shape or {})
        self.keep_results = keep_results
        self.budget = budget_bytes
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        # while the device is idle a tick takes at most idle_batch requests, so
        # the first launch is not held back by a full tick's host-side setup
        self.idle_batch = min(idle_batch, max_batch) if idle_batch else max_batch
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.rebuild_streams = [torch.cuda.Stream() for _ in range(rebuild_streams)]
        self.rebuilds = 0
        self.used = 0
        self.peak = 0
        self.launches = 0

    def _admit(self, queue, inflight, now):
        took = []
        busy = sum(len(reqs) for _, _, reqs in inflight)
        limit = self.max_batch if inflight else self.idle_batch
        while queue and busy + len(took) < self.max_inflight and len(took) < limit:
            r = queue[0]
            r.bytes = request_bytes(r)
            if self.used + r.bytes > self.budget:
                if not inflight and not took:
                    raise RuntimeError(f"request {r.rid} needs {r.bytes} B > budget")
                break
            queue.popleft()
            self.used += r.bytes
            self.peak = max(self.peak, self.used)
            r.admitted = now
            took.append(r)
        return took

    def _launch(self, took, tick):
        groups = []
        ray_reqs = [r for r in took if r.kind != "rebuild"]
        if ray_reqs:
            s = self.streams[tick % len(self.streams)]
            with torch.cuda.stream(s):
                items = []
                for r in ray_reqs:
                    sc = self.scenes[r.scene]
                    r.rays = _rays_for(r, sc, stream=s)
                    items.append((sc.bvh, r.rays, "any" if r.kind != "closest" else "closest",
                                  None))
                batch = rt.TraceBatch(items, stream=s)
                outs = batch.launch(stream=s, **self.trace_shape)
                self.launches += 1
                for r, o in zip(ray_reqs, outs):
                    r.out = o
                    r.keep.append(batch)
                ev = torch.cuda.Event()
                ev.record(s)
            groups.append((ev, ray_reqs))
        # rebuilds: every mesh of the tick in one persistent batched build
        # (one workgroup each) on a stream of its own, beside the traces
        reb = [r for r in took if r.kind == "rebuild"]
        if reb:
            s = self.rebuild_streams[self.rebuilds % len(self.rebuild_streams)]
            self.rebuilds += 1
            with torch.cuda.stream(s):
                grids = [_grid(r, stream=s) for r in reb]
                bb = rt.BuildBatch([(g, None) for g in grids], stream=s)
                for r, g, t in zip(reb, grids, bb.launch(stream=s)):
                    r.bvh = t
                    r.keep.extend((g, bb))
                ev = torch.cuda.Event()
                ev.record(s)
            groups.append((ev, reb))
        return groups

    def run(self, requests):
        """Serve `requests` (sorted by arrival) in real time; returns stats."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue, inflight, finished = deque(), [], []
        t0 = time.perf_counter()
        tick = 0
        while pending or queue or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            ticks = len({k for k, _, _ in inflight})
            took = self._admit(queue, inflight, now) if queue and ticks < self.max_ticks else []
            if took:
                inflight.extend((tick, ev, reqs) for ev, reqs in self._launch(took, tick))
                tick += 1
            still = []
            for k, ev, reqs in inflight:
                if ev.query():
                    t = time.perf_counter() - t0
                    for r in reqs:
                        r.done = t
                        self.used -= r.bytes
                        r.keep.clear()
                        if not self.keep_results:
                            r.rays = r.out = r.bvh = None
                        finished.append(r)
                else:
                    still.append((k, ev, reqs))
            inflight = still
        wall = time.perf_counter() - t0
        return summarize(finished, wall, self)


def summarize(finished, wall, server=None):
    lat = [r.latency * 1e3 for r in finished]
    rays = sum(r.n for r in finished if r.kind != "rebuild")
    span = max(r.done for r in finished) - min(r.arrival for r in finished) if finished else 0.0
    out = {"requests": len(finished), "rays": rays,
           "rebuilds": sum(1 for r in finished if r.kind == "rebuild"),
           "p50_ms": round(median(lat), 3), "p99_ms": round(percentile(lat, 0.99), 3),
           "mean_ms": round(float(np.mean(lat)), 3) if lat else None,
           "span_s": round(span, 4), "wall_s": round(wall, 4),
           "mrays_per_s": round(rays / span / 1e6, 1) if span > 0 else None}
    if server is not None:
        out.update({"trace_launches": server.launches, "peak_bytes": server.peak,
                    "budget_bytes": server.budget})
    for k in KINDS:
        lk = [r.latency * 1e3 for r in finished if r.kind == k]
        if lk:
            out[f"p99_ms_{k}"] = round(percentile(lk, 0.99), 3)
    return out


def run_plain(scenes, requests, keep_results=False, trace_shape=None):
    """One request at a time, one trace launch each, waiting for each
    (arrival times ignored: the plain path's best case).  trace_shape: the
    rt_trace_ex schedule / launch shape of every trace (default: the
    library's).  Returns stats; with keep_results each request's results
    stay in r.out / r.bvh."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in requests:
        a = time.perf_counter() - t0
        if r.kind == "rebuild":
            xyz, r.bvh = _rebuild(r)
            r.keep = [xyz]
        else:
            sc = scenes[r.scene]
            r.rays = _rays_for(r, sc)
            r.out = _trace_alone(r, sc, shape=trace_shape)
        torch.cuda.synchronize()
        r.arrival, r.done = a, time.perf_counter() - t0
        if not keep_results:
            r.rays = r.out = r.bvh = None
            r.keep = []
    return summarize(requests, time.perf_counter() - t0)


def clone_requests(reqs):
    return [Request(r.rid, r.arrival, r.kind, r.scene, r.n, r.seed) for r in reqs]


def same_results(a_reqs, b_reqs):
    """Every request's results identical in the two runs (ray results bit for
    bit; rebuilt trees byte for byte).  Returns the mismatching request ids."""
    bad = []
    for a, b in zip(a_reqs, b_reqs):
        if a.kind == "rebuild":
            n = int(a.bvh.info[0]) * rt.NODE_BYTES
            ok = torch.equal(a.bvh.nodes[:n], b.bvh.nodes[:n]) and \
                torch.equal(a.bvh.tris[: a.n * rt.TRI_BYTES], b.bvh.tris[: b.n * rt.TRI_BYTES])
        elif a.kind == "closest":
            ok = torch.equal(a.out[0][: a.n], b.out[0][: b.n]) and \
                torch.equal(a.out[1][: a.n].view(torch.int32), b.out[1][: b.n].view(torch.int32))
        else:
            words = a.n // 32
            ok = torch.equal(a.out[:words], b.out[:words])
            if a.n % 32:
                mask = (1 << (a.n % 32)) - 1
                ok = ok and (int(a.out[words]) & mask) == (int(b.out[words]) & mask)
        if not ok:
            bad.append(a.rid)
    return bad
