"""Serving loop over librtbvh.so: resident scenes, a Poisson request stream,
admission against a device-memory budget, continuous batching, and
deform-and-rebuild requests that change resident scenes.

Scenes (all resident in HBM for the server's lifetime):
  * static scenes: meshes of 10k-2.5M triangles (displaced spheres and
    random heightfields), radix trees with SAH treelet restructuring built
    once;
  * grid scenes: deformable grids (include/radv_bvh.hpp:224-283
    generate_dynamic_grid_mesh) of up to GRID_MAX triangles, each with a
    ring of GRID_SETS resident buffer sets (vertices, implicit tree, build
    scratch): the current version, which requests trace, and the sets the
    next versions are written into, in turn.

Requests: a ray batch of 4k-2M rays of kind shadow, AO or closest-hit
against any scene, or a deform-and-rebuild of a grid scene (4k-64k
triangles at one of the 200 time steps t = 0.05 k): the grid is regenerated
at its new size and time into the next set of its ring and rebuilt there,
and that set becomes the scene's current version.  Requests on one scene take
effect in arrival order: a ray request traces the version left by every
rebuild of its scene that arrived before it, and no later one.

Scheduling (one host thread):
  * a request becomes visible at its arrival time and waits FIFO;
  * a tick is started only while fewer than `max_ticks` earlier ticks are on
    the device, so requests arriving while the GPU is busy gather into the
    next tick's batch (continuous batching); a tick started on an idle
    device takes at most `idle_batch` requests (its first launch is not held
    back by a large batch's setup), others up to `max_batch` requests and
    about `max_tick_rays` rays;
  * admission is FIFO for memory: it stops at the first request whose
    buffers do not fit in the budget (a first-fit arena over one device
    allocation of `budget_bytes`: rays 32 B per ray, results 8 B or 1/8 B
    per ray).  A grid scene's (GRID_TICK_REBUILDS + 1)-th rebuild in one
    tick waits for the next tick with every later request on that scene,
    while requests on other scenes behind it are admitted (each scene's
    requests stay in arrival order).  A tick's versions of one scene, the one current
    when it starts and those its rebuilds write, occupy distinct sets, and
    the ring is long enough that a rebuild seldom overwrites a set a tick
    still on the device reads;
  * a tick's rebuilds go out as ONE rt_gpu_build_batch_submit call (deform
    step and batched build, one workgroup per grid; grids above 4096
    triangles by 256-thread workgroups, RT_BUILD_BATCH_SHARED, which run
    beside the traces of other ticks instead of waiting for an idle compute
    unit), after the traces still reading the set it overwrites and the
    rebuild that wrote it last (stream waits on their events), on the tick's
    own stream ahead of its traces (rebuild_streams = 0, the default: then
    ONE rt_trace_batch_submit call per tick, rays generated on the device,
    one batched trace, after the builds) or on one of `rebuild_streams`
    streams, the tick's requests on versions whose rebuild may still run
    then going out in a second submit on that stream.  Ticks go to the next
    of `streams` streams (4: one hardware queue each).  The item tables are
    filled with numpy and staged through pinned rings allocated once: no
    device allocation, no host wait per tick;
  * completion is polled through events; a finished request releases its
    arena blocks and its latency (arrival to observed completion) is
    recorded.

The plain path (`run_plain`) serves the same requests one at a time, in
arrival order, with one trace launch (rt_trace_ex, the library's default
shape unless given one) or one grid generation and rt_gpu_build per request
and a wait after each.  Both paths start from the same scene versions
(`Scenes.reset`) and give identical results (`same_results`).

Synthetic rays of each kind (origins uniform in the scene's box, directions
uniform on the sphere, seeded per request, as rt_gen_random_rays):
closest-hit with unbounded tmax; shadow any-hit with tmax = 0.5 (a segment
to a light at that distance); AO any-hit with tmax = 0.05 (a short probe).
"""
import ctypes as C
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
GRID_MAX = 65536        # largest grid of a deform-and-rebuild request
GRID_EXTENT = 2.0       # grids span [-2, 2]^2 at z = 5 +- 0.2
GRID_STEP = 0.05        # time of step k is 0.05 k, k < 200
GRID_SETS = 8           # resident versions per grid scene (ring)
GRID_TICK_REBUILDS = 2  # rebuilds of one grid scene per tick
ITEM_DT = np.dtype(rt.rt_trace_item)
BVH_DT = np.dtype(rt.rt_bvh_dev)
GEN_DT = np.dtype(rt.rt_ray_gen)
BUILD_DT = np.dtype(rt.rt_build_item)
GRID_DT = np.dtype(rt.rt_grid_gen)
_ALIGN = 256

# p50 = median (sorted[n // 2]) and p99 = sorted[floor(0.99 (n - 1))]: the
# statistics of common/vk_timer.cpp:96-108 (timing.py).


def _al(n):
    return (n + _ALIGN - 1) // _ALIGN * _ALIGN


# --------------------------------------------------------------- scenes ----
class GridScene:
    """A deformable grid scene: a ring of GRID_SETS resident buffer sets,
    each holding the 9-float triangles of a grid of up to GRID_MAX
    triangles, its implicit (fast_bvh_builder layout) tree and the build's
    scratch; version v + 1 is written into the set after version v's."""

    def __init__(self, n0, step0, device):
        cap, need = rt.build_requirements(GRID_MAX, rt.build_cfg("implicit"))
        self.node_cap, self.scratch_bytes = cap, _al(need)
        self.sets = []
        for _ in range(GRID_SETS):
            self.sets.append({
                "pos": torch.empty(GRID_MAX * 9, dtype=torch.float32, device=device),
                "nodes": torch.empty(cap * rt.NODE_BYTES, dtype=torch.uint8, device=device),
                "tris": torch.empty(GRID_MAX * rt.TRI_BYTES, dtype=torch.uint8, device=device),
                "info": torch.empty(4, dtype=torch.int32, device=device),
                "scratch": torch.empty(self.scratch_bytes, dtype=torch.uint8, device=device)})
        self.n0, self.step0 = n0, step0
        self.readers = [[] for _ in range(GRID_SETS)]  # events of traces reading each set
        self.ready = [None] * GRID_SETS  # event of the rebuild that wrote each set
        self.cur, self.n, self.step, self.version = 0, 0, 0, 0
        self.src = [None] * GRID_SETS  # request id of the rebuild that wrote each set

    @property
    def next_set(self):
        return (self.cur + 1) % GRID_SETS

    @property
    def bytes(self):
        return sum(t.numel() * t.element_size() for s in self.sets for t in s.values())

    def build_into(self, k, n, step, stream=None):
        """Generate the grid (n triangles, time step `step`) into set k and
        build its tree there (plain path: rt_gen_dynamic_grid + rt_gpu_build)."""
        s = self.sets[k]
        rt.gen_dynamic_grid(n, GRID_STEP * step, GRID_EXTENT, stream=stream, out=s["pos"])
        cfg = rt.build_cfg("implicit")
        scratch = s["scratch"] if n > rt.FUSED_BUILD_MAX else None
        rt.call("rt_gpu_build", rt.stream_handle(stream), rt.ptr(s["pos"]), None, rt.u32(n),
                C.byref(cfg), rt.ptr(s["nodes"]), rt.u32(self.node_cap), rt.ptr(s["tris"]),
                rt.ptr(s["info"]), None, None, rt.ptr(scratch),
                C.c_size_t(scratch.numel() if scratch is not None else 0))

    def bvh(self, k=None, n=None):
        """GpuBVH4 view of set k (default: the current version)."""
        k = self.cur if k is None else k
        n = self.n if n is None else n
        s = self.sets[k]
        counts, _, tb, _ = rt.tree_layout(n)
        b = rt.GpuBVH4(s["nodes"], s["tris"], tb, n, len(counts))
        b.info, b.algorithm = s["info"], "implicit"
        return b

    def row(self, k, n):
        """The rt_bvh_dev record of set k holding a grid of n triangles."""
        counts, _, tb, _ = rt.tree_layout(n)
        s = self.sets[k]
        return (s["nodes"].data_ptr(), s["tris"].data_ptr(), tb, n, len(counts), 0)


class Scenes:
    """The resident scene set: static scenes then grid scenes.  Holds, per
    scene, the rt_bvh_dev record of its current version (a numpy table the
    server indexes by scene id) and the box synthetic rays start in."""

    def __init__(self, static, grids, device):
        self.static = static      # list of (GpuBVH4, pos, idx, lo, hi)
        self.grids = grids        # list of GridScene
        self.count = len(static) + len(grids)
        self.dev = np.zeros(self.count, BVH_DT)
        self.lo = np.zeros((self.count, 3), np.float32)
        self.hi = np.zeros((self.count, 3), np.float32)
        for i, (b, _, _, lo, hi) in enumerate(static):
            self.dev[i] = (b.nodes.data_ptr(), b.tris.data_ptr(), b.node_count, b.tri_count,
                           b.depth, 0)
            self.lo[i], self.hi[i] = [1.2 * v for v in lo], [1.2 * v for v in hi]
        for j in range(len(grids)):
            i = len(static) + j
            self.lo[i] = (-1.2 * GRID_EXTENT, -1.2 * GRID_EXTENT, 4.5)
            self.hi[i] = (1.2 * GRID_EXTENT, 1.2 * GRID_EXTENT, 5.5)
        self.tris = [b.tri_count for b, *_ in static]
        self.reset()

    def is_grid(self, sid):
        return sid >= len(self.static)

    def grid(self, sid):
        return self.grids[sid - len(self.static)]

    def bvh(self, sid):
        return self.static[sid][0] if not self.is_grid(sid) else self.grid(sid).bvh()

    def reset(self):
        """Every grid scene back to its first version (set 0: its initial size
        and time step, built by the plain path)."""
        torch.cuda.synchronize()
        for j, g in enumerate(self.grids):
            g.build_into(0, g.n0, g.step0)
            g.cur, g.n, g.step, g.version = 0, g.n0, g.step0, 0
            g.readers = [[] for _ in range(GRID_SETS)]
            g.ready, g.src = [None] * GRID_SETS, [None] * GRID_SETS
            self.dev[len(self.static) + j] = g.row(0, g.n0)
        torch.cuda.synchronize()

    def swap_in(self, sid, k, n, step, src):
        """Set k of grid scene sid, now holding (n, step), becomes current."""
        g = self.grid(sid)
        g.cur, g.n, g.step, g.version = k, n, step, g.version + 1
        g.src[k] = src
        self.dev[sid] = g.row(k, n)

    @property
    def bytes(self):
        b = sum(x.nodes.numel() + x.tris.numel() + p.numel() * 4 + i.numel() * 4
                for x, p, i, _, _ in self.static)
        return b + sum(g.bytes for g in self.grids)


def make_scenes(count=32, lo=10_000, hi=2_500_000, seed=5, grids=8, grid_tris=(4_096, 65_536),
                max_leaf_prims=1, treelet_rounds=12):
    """`count` static scenes, triangle counts log-uniform in [lo, hi]
    (seeded): displaced UV spheres and random heightfields in turn, each
    built once on the device (radix builder with `treelet_rounds` rounds of
    SAH treelet restructuring and leaves of at most `max_leaf_prims`
    triangles: a static scene stays resident and is traced many times, and on
    the displaced spheres random rays run 4-5x faster on that tree than on
    the plain radix tree with leaves of 4); plus `grids` grid scenes whose
    first version has a size log-uniform in grid_tris at a random time step."""
    rng = np.random.default_rng(seed)
    dev = rt._dev()
    static = []
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
        bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix", max_leaf_prims=max_leaf_prims,
                               treelet_rounds=treelet_rounds)
        bvh._scratch = None  # the build has finished (node count read back)
        static.append((bvh, pos, idx, b[0], b[1]))
    gl = []
    for _ in range(grids):
        n = int(math.exp(rng.uniform(math.log(grid_tris[0]), math.log(grid_tris[1]))))
        gl.append(GridScene(n - n % 2, int(rng.integers(200)), dev))
    torch.cuda.synchronize()
    return Scenes(static, gl, dev)


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
    off: int = -1
    admitted: float = -1.0
    done: float = -1.0
    out: object = None
    rays: object = None
    version: tuple = None  # (scene, set, n, source rebuild id) traced / written
    keep: list = field(default_factory=list)

    @property
    def latency(self):
        return self.done - self.arrival

    @property
    def step(self):
        """Time step of a rebuild (t = 0.05 step)."""
        return self.seed % 200


def make_requests(count, qps, scenes, seed=9, rays=(4_096, 2_000_000),
                  rebuild_tris=(4_096, GRID_MAX), kinds=KINDS):
    """Poisson arrivals at `qps` (exponential gaps, seeded); kind uniform over
    `kinds`; ray requests on a scene uniform over every scene, rebuilds on a
    grid scene uniform over the grids; ray counts and rebuild sizes
    log-uniform in their ranges."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    ng = len(scenes.grids)
    for i in range(count):
        t += rng.exponential(1.0 / qps) if qps else 0.0
        kind = kinds[int(rng.integers(len(kinds)))]
        lo, hi = rebuild_tris if kind == "rebuild" else rays
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        if kind == "rebuild":
            n = min(GRID_MAX, max(2, n - n % 2))  # 2 triangles per quad
            sid = len(scenes.static) + int(rng.integers(ng))
        else:
            sid = int(rng.integers(scenes.count))
        out.append(Request(i, t, kind, sid, n, int(rng.integers(1 << 30))))
    return out


def clone_requests(reqs):
    return [Request(r.rid, r.arrival, r.kind, r.scene, r.n, r.seed) for r in reqs]


def request_bytes(r):
    """Arena bytes a ray request holds from admission to completion: its
    rays (32 B per ray) and results (8 B per ray closest, a bit any hit).
    Rebuilds write their grid scene's resident spare set (no arena bytes)."""
    if r.kind == "rebuild":
        return 0
    out = r.n * 8 if r.kind == "closest" else ((r.n + 31) // 32) * 4
    return _al(r.n * 32) + _al(out)


class Arena:
    """First-fit allocator of 256-B aligned blocks over one device buffer."""

    def __init__(self, nbytes, device):
        self.nbytes = _al(nbytes)
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        self.base = self.buf.data_ptr()
        self.free = [(0, self.nbytes)]  # sorted (offset, size)

    def alloc(self, n):
        n = _al(max(n, 1))
        for i, (off, size) in enumerate(self.free):
            if size >= n:
                if size == n:
                    del self.free[i]
                else:
                    self.free[i] = (off + n, size - n)
                return off
        return None

    def release(self, off, n):
        n = _al(max(n, 1))
        f = self.free
        lo, hi = 0, len(f)
        while lo < hi:
            mid = (lo + hi) // 2
            if f[mid][0] < off:
                lo = mid + 1
            else:
                hi = mid
        f.insert(lo, (off, n))
        if lo + 1 < len(f) and f[lo][0] + f[lo][1] == f[lo + 1][0]:
            f[lo] = (f[lo][0], f[lo][1] + f[lo + 1][1])
            del f[lo + 1]
        if lo > 0 and f[lo - 1][0] + f[lo - 1][1] == f[lo][0]:
            f[lo - 1] = (f[lo - 1][0], f[lo - 1][1] + f[lo][1])
            del f[lo]


def _opts(shape):
    return rt.trace_opts(**shape) if shape else None


def _live(events):
    return [e for e in events if not e.query()]


# --------------------------------------------------------------- server ----
class Server:
    def __init__(self, scenes, budget_bytes=2 << 30, streams=4, rebuild_streams=0,
                 max_inflight=256, max_batch=128, keep_results=False, trace_shape=None,
                 max_ticks=3, idle_batch=16, max_tick_rays=16 << 20):
        self.scenes = scenes
        self.max_ticks = max_ticks or streams
        self.trace_shape = dict(trace_shape or {})
        self.opts = _opts(self.trace_shape)
        self.keep_results = keep_results
        self.budget = budget_bytes
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        self.idle_batch = min(idle_batch, max_batch) if idle_batch else max_batch
        self.max_tick_rays = max_tick_rays
        dev = rt._dev()
        self.arena = Arena(budget_bytes, dev)
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.rebuild_streams = [torch.cuda.Stream() for _ in range(rebuild_streams)]
        slots = 2 * self.max_ticks + 4  # up to two trace submits per tick
        self.tring = rt.SubmitRing(slots, rt.submit_bytes("trace", max_batch), dev)
        self.bring = rt.SubmitRing(slots, rt.submit_bytes("build", max_batch), dev)
        self.rebuilds = 0
        self.used = 0
        self.peak = 0
        self.launches = 0
        self.build_launches = 0

    # -- admission ------------------------------------------------------------
    def _admit(self, queue, inflight, now):
        """Requests for the next tick, in arrival order.  Memory is admitted
        FIFO: the first ray request whose buffers do not fit ends the scan.
        A grid scene's (GRID_TICK_REBUILDS + 1)-th rebuild of the tick waits
        for the next tick, and so does every later request on that scene
        (requests on one scene stay in arrival order); requests on other
        scenes behind it are still admitted."""
        took = []
        busy = sum(len(reqs) for _, _, reqs in inflight)
        limit = self.max_batch if inflight else self.idle_batch
        rebuilt, held = {}, set()
        rest = deque()
        rays = 0
        while queue and busy + len(took) < self.max_inflight and len(took) < limit:
            if self.max_tick_rays and rays >= self.max_tick_rays:
                break
            r = queue.popleft()
            if r.scene in held:
                rest.append(r)
                continue
            if r.kind == "rebuild":
                if rebuilt.get(r.scene, 0) == GRID_TICK_REBUILDS:
                    held.add(r.scene)
                    rest.append(r)
                    continue
                rebuilt[r.scene] = rebuilt.get(r.scene, 0) + 1
                r.bytes = 0
            else:
                r.bytes = request_bytes(r)
                if r.bytes > self.arena.nbytes:
                    raise RuntimeError(f"request {r.rid} needs {r.bytes} B > budget")
                off = self.arena.alloc(r.bytes)
                if off is None:
                    if not inflight and not took:
                        raise RuntimeError(f"request {r.rid} does not fit the budget")
                    rest.append(r)
                    break
                r.off = off
                rays += r.n
            self.used += r.bytes
            self.peak = max(self.peak, self.used)
            r.admitted = now
            took.append(r)
        rest.extend(queue)
        queue.clear()
        queue.extend(rest)
        return took

    # -- one tick -------------------------------------------------------------
    def _launch(self, took, tick):
        S = self.scenes
        groups = []
        # resolve, in FIFO order, which scene version each request traces or writes
        reb, ray_reqs = [], []  # ray_reqs: (request, rt_bvh_dev record, (grid, set) read)
        for r in took:
            if r.kind == "rebuild":
                g = S.grid(r.scene)
                k = g.next_set
                r.version = (r.scene, k, r.n, r.rid)
                reb.append((r, g, k))
                S.swap_in(r.scene, k, r.n, r.step, r.rid)
            elif S.is_grid(r.scene):
                g = S.grid(r.scene)
                r.version = (r.scene, g.cur, g.n, g.src[g.cur])
                ray_reqs.append((r, S.dev[r.scene].copy(), (r.scene, g.cur)))  # its turn's version
            else:
                r.version = (r.scene, 0, S.tris[r.scene], None)
                ray_reqs.append((r, S.dev[r.scene], None))
        ts = self.streams[tick % len(self.streams)]
        bs = None
        if reb:
            # rebuild_streams = 0: a tick's rebuilds go on the tick's own stream,
            # ahead of its traces (then one trace launch, in stream order after them)
            bs = (self.rebuild_streams[self.rebuilds % len(self.rebuild_streams)]
                  if self.rebuild_streams else ts)
            self.rebuilds += 1
            groups.append(self._launch_rebuilds(reb, bs))
        if ray_reqs and bs is ts:
            waits = {S.grid(key[0]).ready[key[1]] for _, _, key in ray_reqs if key is not None}
            evs = [e for e in waits if e is not None and not e.query()]
            groups.append(self._launch_traces(ray_reqs, ts, evs))
        elif ray_reqs:
            # requests on versions whose rebuild may still run go out after it, in a
            # second launch on the rebuild stream (after this tick's builds); the
            # others go out at once on the tick's stream
            waits = {}
            for _, _, key in ray_reqs:
                if key is not None and key not in waits:
                    e = S.grid(key[0]).ready[key[1]]
                    waits[key] = e if e is not None and not e.query() else None
            free = [x for x in ray_reqs if x[2] is None or waits[x[2]] is None]
            held = [x for x in ray_reqs if x[2] is not None and waits[x[2]] is not None]
            if free:
                groups.append(self._launch_traces(free, ts, []))
            if held:
                evs = list({id(e): e for e in waits.values() if e is not None}.values())
                groups.append(self._launch_traces(held, bs or ts, evs))
        return groups

    def _launch_rebuilds(self, reb, bs):
        S = self.scenes
        items = np.zeros(len(reb), BUILD_DT)
        gen = np.zeros(len(reb), GRID_DT)
        for j, (r, g, k) in enumerate(reb):
            # after the traces still reading the set it overwrites, and after
            # the rebuild that wrote that set last (it may be on another stream)
            for e in _live(g.readers[k]) + _live([g.ready[k]] if g.ready[k] else []):
                bs.wait_event(e)
            g.readers[k] = []
            s = g.sets[k]
            big = r.n > rt.FUSED_BUILD_MAX
            items[j] = (s["pos"].data_ptr(), 0, s["nodes"].data_ptr(), s["tris"].data_ptr(),
                        s["info"].data_ptr(), 0, 0, r.n, g.node_cap,
                        s["scratch"].data_ptr() if big else 0, g.scratch_bytes if big else 0)
            gen[j] = (GRID_STEP * r.step, GRID_EXTENT)
        pin, tab, slot = self.bring.take()
        nb = self.bring.slot_bytes
        # medium grids by 256-thread workgroups: they run beside the traces
        # instead of waiting for an idle compute unit
        rt.call("rt_gpu_build_batch_submit", rt.stream_handle(bs), items.ctypes.data,
                gen.ctypes.data, rt.u32(len(reb)), rt.u32(rt.BUILD_BATCH_SHARED), pin,
                C.c_size_t(nb), tab, C.c_size_t(nb))
        self.build_launches += 1
        self.bring.used(slot, bs)
        if self.keep_results:
            with torch.cuda.stream(bs):
                for r, g, k in reb:
                    s = g.sets[k]
                    nb4 = rt.tree_layout(r.n)[2] * rt.NODE_BYTES
                    r.out = (s["nodes"][:nb4].clone(), s["tris"][: r.n * rt.TRI_BYTES].clone(),
                             s["pos"][: r.n * 9].clone())
        ev = torch.cuda.Event()
        ev.record(bs)
        for r, g, k in reb:
            g.ready[k] = ev
        return ev, [r for r, _, _ in reb]

    def _launch_traces(self, group, ts, waits):
        """One rt_trace_batch_submit of `group` ((request, scene record, grid
        set read) triples) on stream ts, after the events `waits`."""
        S = self.scenes
        ray_reqs = [r for r, _, _ in group]
        m = len(ray_reqs)
        sid = np.fromiter((r.scene for r in ray_reqs), np.int64, m)
        n = np.fromiter((r.n for r in ray_reqs), np.uint64, m)
        closest = np.fromiter((r.kind == "closest" for r in ray_reqs), bool, m)
        items = np.zeros(m, ITEM_DT)
        items["bvh"] = np.array([row for _, row, _ in group], BVH_DT)
        if self.keep_results:
            base = np.zeros(m, np.uint64)
            for j, r in enumerate(ray_reqs):
                r.rays = rt.Rays.empty(r.n)
                base[j] = r.rays.data.data_ptr()
                if r.kind == "closest":
                    r.out = (torch.empty(r.n, dtype=torch.int32, device="cuda"),
                             torch.empty(r.n, dtype=torch.float32, device="cuda"))
                    outp, outt = r.out[0].data_ptr(), r.out[1].data_ptr()
                else:
                    r.out = torch.empty(max(1, (r.n + 31) // 32), dtype=torch.int32,
                                        device="cuda")
                    outp, outt = r.out.data_ptr(), 0
                r.keep = [outp, outt]
            res = np.fromiter((r.keep[0] for r in ray_reqs), np.uint64, m)
            rest = np.fromiter((r.keep[1] for r in ray_reqs), np.uint64, m)
        else:
            base = np.uint64(self.arena.base) + np.fromiter((r.off for r in ray_reqs), np.uint64, m)
            res = base + ((n * np.uint64(32) + np.uint64(_ALIGN - 1)) // np.uint64(_ALIGN)
                          * np.uint64(_ALIGN))
            rest = res + n * np.uint64(4)
        rays = items["rays"]
        for f, name in enumerate(rt.Rays.FIELDS):
            rays[name] = base + np.uint64(4 * f) * n
        items["count"] = n.astype(np.uint32)
        items["flags"] = np.where(closest, rt.RT_CLOSEST, rt.RT_ANY).astype(np.uint32)
        items["hits"]["prim"] = np.where(closest, res, 0)
        items["hits"]["t"] = np.where(closest, rest, 0)
        items["occlusion_bits"] = np.where(closest, 0, res)
        gen = np.zeros(m, GEN_DT)
        gen["seed"] = np.fromiter((r.seed for r in ray_reqs), np.uint32, m)
        gen["lo"], gen["hi"] = S.lo[sid], S.hi[sid]
        gen["tmax"] = np.fromiter((TMAX[r.kind] for r in ray_reqs), np.float32, m)
        for e in waits:  # the rebuilds that wrote the versions traced
            ts.wait_event(e)
        pin, tab, slot = self.tring.take()
        nb = self.tring.slot_bytes
        rt.call("rt_trace_batch_submit", rt.stream_handle(ts), items.ctypes.data,
                gen.ctypes.data, rt.u32(m), C.byref(self.opts) if self.opts else None, pin,
                C.c_size_t(nb), tab, C.c_size_t(nb))
        self.launches += 1
        self.tring.used(slot, ts)
        ev = torch.cuda.Event()
        ev.record(ts)
        for key in {key for _, _, key in group if key is not None}:
            g = S.grid(key[0])
            g.readers[key[1]] = _live(g.readers[key[1]]) + [ev]
        return ev, ray_reqs

    # -- the loop ---------------------------------------------------------------
    def run(self, requests):
        """Serve `requests` in real time from the scenes' first versions;
        returns stats."""
        self.scenes.reset()
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
                        if r.off >= 0:
                            self.arena.release(r.off, r.bytes)
                            r.off = -1
                        if not self.keep_results:
                            r.out = r.rays = None
                        r.keep = []
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
        out.update({"trace_launches": server.launches, "build_launches": server.build_launches,
                    "peak_bytes": server.peak, "budget_bytes": server.budget})
    for k in KINDS:
        lk = [r.latency * 1e3 for r in finished if r.kind == k]
        if lk:
            out[f"p99_ms_{k}"] = round(percentile(lk, 0.99), 3)
    return out


def run_plain(scenes, requests, keep_results=False, trace_shape=None):
    """One request at a time from the scenes' first versions, in arrival
    order, waiting for each (arrival times ignored: the plain path's best
    case).  A ray request: its rays (rt_gen_random_rays) and one trace launch
    (rt_trace_ex; trace_shape: its schedule / launch shape, default the
    library's).  A rebuild: the grid generated into the scene's spare set and
    rt_gpu_build there, then swapped in.  Returns stats; with keep_results
    each request's results stay in r.out (rays in r.rays)."""
    S = scenes
    S.reset()
    shape = dict(trace_shape or {})
    t0 = time.perf_counter()
    for r in requests:
        a = time.perf_counter() - t0
        if r.kind == "rebuild":
            g = S.grid(r.scene)
            k = g.next_set
            r.version = (r.scene, k, r.n, r.rid)
            g.build_into(k, r.n, r.step)
            S.swap_in(r.scene, k, r.n, r.step, r.rid)
            if keep_results:
                s = g.sets[k]
                nb4 = rt.tree_layout(r.n)[2] * rt.NODE_BYTES
                r.out = (s["nodes"][:nb4].clone(), s["tris"][: r.n * rt.TRI_BYTES].clone(),
                         s["pos"][: r.n * 9].clone())
        else:
            if S.is_grid(r.scene):
                g = S.grid(r.scene)
                r.version = (r.scene, g.cur, g.n, g.src[g.cur])
            else:
                r.version = (r.scene, 0, S.tris[r.scene], None)
            bvh = S.bvh(r.scene)
            rays = rt.gen_random_rays(r.n, r.seed, tuple(S.lo[r.scene]), tuple(S.hi[r.scene]),
                                      TMAX[r.kind])
            if r.kind == "closest":
                r.out = bvh.intersect(rays, **shape)
            else:
                r.out = bvh.occluded(rays, **shape)
            if keep_results:
                r.rays = rays
        torch.cuda.synchronize()
        r.arrival, r.done = a, time.perf_counter() - t0
        if not keep_results:
            r.out = r.rays = None
    return summarize(requests, time.perf_counter() - t0)


def same_results(a_reqs, b_reqs):
    """Every request's results identical in the two runs (ray results bit for
    bit; rebuilt grids and trees byte for byte) and the same scene versions
    traced.  Returns the mismatching request ids."""
    bad = []
    for a, b in zip(a_reqs, b_reqs):
        ok = a.version == b.version
        if a.kind == "rebuild":
            ok = ok and all(torch.equal(x, y) for x, y in zip(a.out, b.out))
        elif a.kind == "closest":
            ok = ok and torch.equal(a.out[0][: a.n], b.out[0][: b.n]) and \
                torch.equal(a.out[1][: a.n].view(torch.int32), b.out[1][: b.n].view(torch.int32))
        else:
            words = a.n // 32
            ok = ok and torch.equal(a.out[:words], b.out[:words])
            if a.n % 32:
                mask = (1 << (a.n % 32)) - 1
                ok = ok and (int(a.out[words]) & mask) == (int(b.out[words]) & mask)
        if not ok:
            bad.append(a.rid)
    return bad
