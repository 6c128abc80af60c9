"""Headline benchmark: closest-hit + any-hit shadow rays at 2.5 M triangles.

Workload (one step = one frame of the shadow any-hit configuration):
  mesh   UV sphere 1250 x 1000 quads = 2,500,000 FP32 triangles, radius 1,
         seeded random radial displacement +-2 % (generated on the device),
         Box4 BVH built on the device once (inputs resident before timing;
         radix tree, leaves of at most 1-4 triangles and 0, 4 or 8 rounds of
         treelet restructuring, as tuned on whole frames; the tree is built
         once, outside the timed frames)
  step   1920x1080 pinhole camera rays (60 deg vertical FOV, every pixel hits)
         -> closest-hit trace -> 16 shadow rays per pixel towards 16 lights
         uniform in the box 3x the mesh bounds (tmax = light distance)
         -> any-hit trace -> per-pixel 16-bit occlusion masks
  value  (2,073,600 + 33,177,600) rays / step time, Mrays/s
Also reported: device build time at 4,096 triangles (fused single-workgroup
path) and at 2.5 M, p50 / p99 of the per-frame latency, the roofline of the
dominant kernel (the shadow any-hit traversal), and the CPU traversal of the
same tree and rays on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rdna2-accel-raster_amd"))
import rtbvh as rt  # noqa: E402
import autotune  # noqa: E402
from timing import median, percentile  # noqa: E402

METRIC = "Mrays/s any-hit+closest-hit @2.5M tris; BVH build µs @4K tris; p99 req latency"
HBM_PEAK_GBS = 8000.0
W, H, NLIGHTS = 1920, 1080, 16
NU, NV, SEED = 1250, 1000, 7
EPS = 1e-4


class TimingEvent:
    """A HIP event for timing only, created with hipEventDisableSystemFence:
    recording it skips the system-scope fence (L2 write-back and invalidate)
    that a default event performs, which otherwise lands inside the timed
    interval, adds ~3 us to a 15-us kernel and gives the samples a tail (HIP's
    documentation gives this flag for timing).  The reference's GPU timestamp
    queries flush no caches either."""
    _hip = None
    FLAGS = 0x20000000  # hipEventDisableSystemFence

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes
            import glob
            names = ["libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"] + sorted(glob.glob(
                os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so*")))
            for name in names:
                try:
                    cls._hip = ctypes.CDLL(name)
                    break
                except OSError:
                    continue
            if cls._hip is None:
                raise RuntimeError("libamdhip64.so not found")
        return cls._hip

    def __init__(self):
        import ctypes
        self.C = ctypes
        self.ev = ctypes.c_void_p()
        if self.hip().hipEventCreateWithFlags(ctypes.byref(self.ev), ctypes.c_uint(self.FLAGS)):
            raise RuntimeError("hipEventCreateWithFlags failed")

    def record(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream()
        if self.hip().hipEventRecord(self.ev, self.C.c_void_p(s.cuda_stream)):
            raise RuntimeError("hipEventRecord failed")

    def synchronize(self):
        if self.hip().hipEventSynchronize(self.ev):
            raise RuntimeError("hipEventSynchronize failed")

    def elapsed_us(self, end):
        ms = self.C.c_float()
        if self.hip().hipEventElapsedTime(self.C.byref(ms), self.ev, end.ev):
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value * 1000.0

    def __del__(self):
        if self._hip is not None and self.ev:
            self._hip.hipEventDestroy(self.ev)


def build_time_4k(iters=2000, warm=50):
    """Device build of the 4,096-triangle deforming grid
    (generate_dynamic_grid_mesh layout), fused path, into preallocated
    buffers.  Returns the p50 and p99 of one build timed alone on the device
    (two timing events around one dispatch, queued behind a short delay kernel
    so that the host's launch is not inside the interval: the reference's GPU
    timestamps around its dispatch, src/main.cpp:1177-1186), the p99 with the
    host's launch included (the events around one launch issued on an idle
    device), and the per-build time of 200 builds issued back to back
    (builds/s = 1e6 / that).  2,000 samples each: a p99 of 200 is the second
    largest sample."""
    xyz = rt.gen_dynamic_grid(4096, 0.0, 2.0)
    bvh = rt.GpuBVH4.build(xyz, None, algorithm="implicit")
    cfg = rt.build_cfg("implicit")
    args = (rt.stream_handle(), rt.ptr(xyz), None, rt.u32(4096), rt.C.byref(cfg),
            rt.ptr(bvh.nodes), rt.u32(bvh.node_count), rt.ptr(bvh.tris), rt.ptr(bvh.info),
            None, None, None, rt.C.c_size_t(0))
    s, e = TimingEvent(), TimingEvent()
    device, launch = [], []
    for i in range(warm + iters):
        torch.cuda._sleep(200_000)  # ~0.1 ms: the events and the dispatch queue behind it
        s.record()
        rt.call("rt_gpu_build", *args)
        e.record()
        e.synchronize()
        if i >= warm:
            device.append(s.elapsed_us(e))
    for i in range(warm + iters):
        s.record()
        rt.call("rt_gpu_build", *args)
        e.record()
        e.synchronize()
        if i >= warm:
            launch.append(s.elapsed_us(e))
    torch.cuda.synchronize()
    s.record()
    for i in range(200):
        rt.call("rt_gpu_build", *args)
    e.record()
    e.synchronize()
    back_to_back = s.elapsed_us(e) / 200
    return median(device), percentile(device, 0.99), percentile(launch, 0.99), back_to_back


FRAME_TREELET_ROUNDS = (0, 8)  # treelet restructuring rounds the frame's tree is tuned over
FRAME_EXTRA_TREES = ((2, 4), (1, 4))  # (leaf size, rounds) also tried for the frame
AO_TREELET_ROUNDS = (0, 8, 16)  # the same for the AO frame's terrain


def same_frame(a, b):
    """Frame results (primary t, primary triangle, light masks) equal up to
    ties: every primary t bit-identical, the masks identical wherever the
    primary triangle is, and another triangle at under 1e-4 of the pixels
    (two triangles meeting at exactly the hit t: which one a tree reports
    decides the shadow rays' origin triangle)."""
    (t1, p1, m1), (t0, p0, m0) = a, b
    if not torch.equal(t1.view(torch.int32), t0.view(torch.int32)):
        return False
    same_prim = p1 == p0
    if (~same_prim).float().mean().item() >= 1e-4:
        return False
    return torch.equal(m1[same_prim], m0[same_prim])


def same_ao_frame(a, b):
    """AO frame results (primary t, triangle, instance, unfiltered AO) equal
    up to ties: every t bit-identical, another triangle or instance at under
    1e-4 of the pixels, and the AO identical wherever both are the same (the
    AO rays start on the reported triangle; the filtered image spreads each
    such pixel over its 7x7 neighbourhood)."""
    (t1, p1, i1, ao1), (t0, p0, i0, ao0) = a, b
    if not torch.equal(t1.view(torch.int32), t0.view(torch.int32)):
        return False
    same_hit = (p1 == p0) & (i1 == i0)
    if (~same_hit).float().mean().item() >= 1e-4:
        return False
    return torch.equal(ao1[same_hit], ao0[same_hit])


class Frame:
    builder = "radix"

    def __init__(self, tuner=None):
        self.pos, self.idx = rt.gen_sphere(NU, NV, 1.0, 0.02, SEED)
        torch.cuda.synchronize()
        # radix digit width of the sort: measured when tuning (same tree for each)
        self.sort_bits = (tuner.sort_bits(self.pos, self.idx, algorithm=self.builder,
                                          tag="sphere") if tuner else rt.DEFAULT_SORT_BITS)
        self.build_tree(4, 0)
        lo = self.pos.min(0).values.cpu().numpy().astype(np.float64)
        hi = self.pos.max(0).values.cpu().numpy().astype(np.float64)
        c, ext = (lo + hi) / 2, (hi - lo)
        lights = rt.lights_in_box(c - 1.5 * ext, c + 1.5 * ext, NLIGHTS, 1337)
        self.lights = torch.tensor(lights, dtype=torch.float32, device="cuda")
        self.npix = W * H
        self.primary = rt.Rays.empty(self.npix)
        self.shadow = rt.Rays.empty(self.npix * NLIGHTS)
        self.prim = torch.empty(self.npix, dtype=torch.int32, device="cuda")
        self.t = torch.empty(self.npix, dtype=torch.float32, device="cuda")
        self.bits = torch.empty((self.npix * NLIGHTS + 31) // 32, dtype=torch.int32, device="cuda")
        self.masks = torch.empty(self.npix, dtype=torch.int16, device="cuda")
        self.ev = None
        self.primary_opts, self.shadow_opts = {}, {}

    def build_tree(self, leaf, rounds=0):
        """The frame's radix tree with leaves of at most `leaf` triangles and
        `rounds` rounds of treelet restructuring (device build, timed after a
        warm-up build)."""
        rt.GpuBVH4.build(self.pos, self.idx, algorithm=self.builder, sort_bits=self.sort_bits,
                         max_leaf_prims=leaf, treelet_rounds=rounds)  # warm the build path
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        self.bvh = rt.GpuBVH4.build(self.pos, self.idx, algorithm=self.builder, sync=False,
                                    sort_bits=self.sort_bits, max_leaf_prims=leaf,
                                    treelet_rounds=rounds)
        e.record()
        e.synchronize()
        self.build_ms = s.elapsed_time(e)
        self.bvh.node_count, self.bvh.depth = [int(v) for v in self.bvh.info[:2].cpu().tolist()]
        self.leaf, self.rounds = leaf, rounds

    def tune(self, tuner):
        """The tree (leaf size x treelet rounds), timed on whole frames, each
        tree traced with its own tuned schedule / launch shapes (leaf sizes 1
        and 2 are within a few per cent of each other under the default
        shapes, and their order flips once the shapes are tuned).  Every
        shape gives the same hits; a tree that changes the frame's result
        beyond ties (same_frame) is left out."""
        trees = {}

        def use(c):
            leaf, rounds = int(c[0]), int(c[1])
            if (leaf, rounds) not in trees:
                if trees or (leaf, rounds) != (self.leaf, self.rounds):  # __init__'s tree
                    self.build_tree(leaf, rounds)
                self.step()
                torch.cuda.synchronize()
                tag = f"leaf{leaf}-r{rounds}"
                self.primary_opts = tuner.trace(self.bvh, self.primary, "closest",
                                                tag=f"primary-{tag}", grid_width=W)
                self.shadow_opts = tuner.trace(self.bvh, self.shadow, "any",
                                               tag=f"shadow-{tag}", grid_width=W,
                                               hint=(self.prim, self.npix))
                trees[(leaf, rounds)] = (self.bvh, self.build_ms, self.primary_opts,
                                         self.shadow_opts)
            self.bvh, self.build_ms, self.primary_opts, self.shadow_opts = trees[(leaf, rounds)]
            self.leaf, self.rounds = leaf, rounds

        cands = [(leaf, r) for r in FRAME_TREELET_ROUNDS for leaf in autotune.LEAF_SIZES]
        cands += list(FRAME_EXTRA_TREES)
        use(tuner.frame_leaf_size(use, self.step, lambda: (self.t, self.prim, self.masks),
                                  tag="sphere-2p5m", sizes=cands, same=same_frame))
        trees.clear()

    def step(self, events=None):
        rt.gen_camera_rays(W, H, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0,
                           0.0, 1e30, out=self.primary)
        self.bvh.intersect(self.primary, out=(self.prim, self.t), **self.primary_opts)
        rt.gen_shadow_rays(self.pos, self.idx, self.primary, self.prim, self.t, self.lights, EPS,
                           out=self.shadow)
        if events:
            events[0].record()
        self.bvh.occluded(self.shadow, out=self.bits,
                          **autotune.hinted_kwargs(self.shadow_opts, (self.prim, self.npix)))
        if events:
            events[1].record()
        rt.pack_light_masks(self.bits, self.npix, NLIGHTS, out=self.masks)

    @property
    def rays_per_step(self):
        return self.npix + self.npix * NLIGHTS


def plain_build_ms(frame, reps=5):
    """Device build of the 2.5 M sphere's radix tree without treelet
    restructuring (the rebuild path of deforming meshes), the frame's leaf
    size, per build (warm, back to back)."""
    kw = dict(algorithm=frame.builder, sort_bits=frame.sort_bits, max_leaf_prims=frame.leaf)
    b = rt.GpuBVH4.build(frame.pos, frame.idx, **kw)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        rt.GpuBVH4.build(frame.pos, frame.idx, sync=False, scratch=b._scratch, **kw)
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def refit_time_ms(frame, reps=5):
    """Generic refit (rt_gpu_refit_bvh4) of the 2.5M radix tree at the same
    positions (which leaves every byte of the tree as built), per refit."""
    b = frame.bvh
    scratch = torch.empty(max(256, b.node_count * 4), dtype=torch.uint8, device="cuda")
    args = (rt.stream_handle(), rt.ptr(frame.pos), rt.ptr(frame.idx), rt.ptr(b.nodes),
            rt.u32(b.node_count), rt.ptr(b.tris), rt.u32(b.depth), rt.ptr(scratch),
            rt.C.c_size_t(scratch.numel()))
    rt.call("rt_gpu_refit_bvh4", *args)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        rt.call("rt_gpu_refit_bvh4", *args)
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def algorithmic_bytes_shadow(frame):
    """Bytes one shadow any-hit launch must move at minimum: the whole tree in
    the RADV-equivalent size (192 + 128 boxes + 64 N) plus 32 B per SoA ray
    and 1/8 B of occlusion bit per ray."""
    _, _, tb, as_size = rt.tree_layout(frame.bvh.tri_count)
    r = frame.npix * NLIGHTS
    return as_size + r * 32 + r / 8.0


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


# A GPU's share of the host: the CPU baseline runs on every CPU this process
# may use, up to the 16 host CPUs allotted per GPU on a shared node.
CPU_SHARE = 16


def cpu_threads():
    return max(1, min(CPU_SHARE, len(os.sched_getaffinity(0))))


def cpu_baseline(frame, threads):
    """The CPU oracle's BVH4 traversal on the host cores over the same tree and
    the same rays: one whole frame (all primary rays closest-hit, all shadow
    rays any-hit), parallel over rays in 4,096-ray chunks."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nodes = frame.bvh.nodes.cpu().numpy()[: frame.bvh.node_count * 128].view(O.NODE4)
    tris = frame.bvh.tris.cpu().numpy()[: frame.bvh.tri_count * 64].view(O.TRI64)
    prim_rays = O.Rays.stack(frame.primary.data.cpu().numpy())
    sh_rays = O.Rays.stack(frame.shadow.data.cpu().numpy())
    t0 = time.perf_counter()
    O.trace_bvh4(nodes, tris, prim_rays, mode=O.CLOSEST, threads=threads)
    O.trace_bvh4(nodes, tris, sh_rays, mode=O.ANY, threads=threads)
    dt = time.perf_counter() - t0
    n = len(prim_rays) + len(sh_rays)
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "host_cpus_visible": len(os.sched_getaffinity(0)),
            "sample": f"one whole frame of the same workload and tree: {len(prim_rays)} primary "
                      f"closest-hit + {len(sh_rays)} shadow any-hit rays; {dt:.1f} s wall on "
                      f"{threads} threads",
            "build": cpu_build_baseline(frame, O)}


def cpu_build_baseline(frame, O):
    """SoftwareBVH::build + SoftwareBVH4::fromBVH2 on one host core: the
    reference's own header code compiled with its CMake flags
    (oracle/_ref/libref_shipped.so, kind "reference") where that library is
    present, else the oracle's restatement (kind "port").  The 4,096-triangle
    deforming grid (median of 5) and the 2.5 M-triangle sphere of the frame
    (one build)."""
    xyz = O.dynamic_grid(4096, 0.0, 2.0)[0].reshape(-1, 3)
    gidx = np.arange(xyz.shape[0], dtype=np.uint32).reshape(-1, 3)
    spos = frame.pos.cpu().numpy()
    sidx = frame.idx.cpu().numpy().view(np.uint32)
    kind = "reference" if O.ref(shipped=True) is not None else "port"

    def build(p, i):
        if kind == "reference":
            O.ref_build(p, i, 4, shipped=True)
        else:
            O.software_bvh4(p, i)

    times = []
    for _ in range(6):
        t0 = time.perf_counter()
        build(xyz, gidx)
        times.append(time.perf_counter() - t0)
    t4k = sorted(times[1:])[2]
    t0 = time.perf_counter()
    build(spos, sidx)
    t25 = time.perf_counter() - t0
    return {"kind": kind, "cores": 1, "build_us_4k": round(t4k * 1e6, 1),
            "build_ms_2p5m": round(t25 * 1e3, 1),
            "sample": "SoftwareBVH::build + fromBVH2, max 4 triangles per leaf: the 4,096-"
                      "triangle grid (median of 5 after 1 warm-up) and the 2,500,000-triangle "
                      "sphere (one build), host arrays in, Box4 nodes out"}


def batched_leg(meshes=256, seed=4, tuner=None):
    """Batched build and trace: `meshes` seeded random-displaced grids with
    triangle counts log-uniform in [64, 65536], each with its own ray set of
    4k-512k rays (origins uniform in the mesh bounds, directions uniform on
    the sphere), closest-hit or any-hit 50/50.  Batched: every mesh built in
    one persistent launch (up to 4096 triangles in LDS, larger ones by one
    workgroup each), all ray sets traced in one launch replayed from a graph.
    Plain: one build (rt_gpu_build, the same implicit trees) and one trace
    launch per mesh, waiting for each.  Both legs give identical results
    (checked).  With a tuner, the persistent grid and the batched trace's
    launch shape are picked by measurement (the default-shape times are
    reported beside the tuned ones)."""
    rng = np.random.default_rng(seed)
    geo, rays, modes = [], [], []
    for k in range(meshes):
        n = int(np.exp(rng.uniform(np.log(64), np.log(65536))))
        nx = max(1, int(np.sqrt(n / 2)))
        ny = max(1, n // (2 * nx))
        geo.append(rt.gen_heightfield(nx, ny, 1.0, 0.05, 1000 + k))
        nr = int(np.exp(rng.uniform(np.log(4096), np.log(524288))))
        rays.append(rt.gen_random_rays(nr, 5000 + k, (-1.1, -1.1, -0.2), (1.1, 1.1, 0.2),
                                       0.3 if k % 2 else 1e30))
        modes.append("any" if rng.random() < 0.5 else "closest")
    tris = sum(int(i.shape[0]) for _, i in geo)
    nrays = sum(len(r) for r in rays)
    torch.cuda.synchronize()
    small = [k for k, (_, i) in enumerate(geo) if i.shape[0] <= rt.BATCH_BUILD_MAX]
    big = [k for k in range(meshes) if k not in small]
    # batched build: one persistent launch for the small meshes + the rest
    bb = rt.BuildBatch([geo[k] for k in small])
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    trees = [None] * meshes
    for k, t in zip(small, bb.launch()):
        trees[k] = t
    for k in big:
        trees[k] = rt.GpuBVH4.build(*geo[k], algorithm="implicit")
    torch.cuda.synchronize()

    def build_all():
        s.record()
        bb.launch()
        for k in big:
            rt.GpuBVH4.build(*geo[k], algorithm="implicit", sync=False)
        e.record()
        e.synchronize()
        return s.elapsed_time(e)

    build_ms_default = build_all()
    if tuner is not None:
        tuner.build_batch(bb, tag="grids")
        torch.cuda.synchronize()
    build_ms = build_all()
    # batched trace, captured in a graph
    batch = rt.TraceBatch([(trees[k], rays[k], modes[k], None) for k in range(meshes)])
    batch.launch()
    torch.cuda.synchronize()

    def graph_ms(shape, reps=5):
        batch.launch(**shape)  # plans this shape's chunk table outside the capture
        torch.cuda.synchronize()
        gs = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=gs):
            batch.launch(stream=gs, **shape)
        g.replay()
        torch.cuda.synchronize()
        s.record()
        for _ in range(reps):
            g.replay()
        e.record()
        e.synchronize()
        return s.elapsed_time(e) / reps

    trace_ms_default = graph_ms({})
    shape = tuner.trace_batch(batch, tag="grids") if tuner is not None else {}
    trace_ms = graph_ms(shape)
    # the whole batch -- every build and every trace -- as one captured graph
    graph_ms_total = None
    if not big:
        gs = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=gs):
            bb.launch(stream=gs)
            batch.launch(stream=gs, **shape)
        g.replay()
        torch.cuda.synchronize()
        s.record()
        for _ in range(5):
            g.replay()
        e.record()
        e.synchronize()
        graph_ms_total = s.elapsed_time(e) / 5
    # plain: one build + one trace launch per mesh, waiting for each
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_out = []
    for k in range(meshes):
        tk = rt.GpuBVH4.build(*geo[k], algorithm="implicit")
        o = tk.intersect(rays[k]) if modes[k] == "closest" else tk.occluded(rays[k])
        torch.cuda.synchronize()
        plain_out.append(o)
    plain_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    for k in range(meshes):
        _ = trees[k].intersect(rays[k]) if modes[k] == "closest" else trees[k].occluded(rays[k])
        torch.cuda.synchronize()
    plain_trace_ms = (time.perf_counter() - t0) * 1e3
    same = True
    for k, o in enumerate(batch.outputs):
        n = len(rays[k])
        if modes[k] == "closest":
            same &= bool(torch.equal(o[0], plain_out[k][0]) and torch.equal(o[1], plain_out[k][1]))
        else:
            w = (n + 31) // 32
            same &= bool(torch.equal(o[:w], plain_out[k][:w]))
    return {"workload": f"{meshes} random-displaced grids, 64-65536 tris log-uniform "
                        f"({tris} tris), 4k-512k random rays each ({nrays} rays), "
                        "closest/any 50/50",
            "build_ms": round(build_ms, 3), "build_ms_default_shape": round(build_ms_default, 3),
            "meshes_one_launch": len(small),
            "trace_ms": round(trace_ms, 3), "trace_ms_default_shape": round(trace_ms_default, 3),
            "trace_mrays_per_s": round(nrays / trace_ms / 1e3, 1),
            "plain_trace_ms": round(plain_trace_ms, 3),
            "plain_trace_mrays_per_s": round(nrays / plain_trace_ms / 1e3, 1),
            "plain_build_and_trace_ms": round(plain_ms, 3),
            "batched_build_and_trace_ms": round(build_ms + trace_ms, 3),
            "graph_build_and_trace_ms": graph_ms_total and round(graph_ms_total, 3),
            "speedup_build_and_trace": round(plain_ms / (build_ms + trace_ms), 2),
            "speedup_trace": round(plain_trace_ms / trace_ms, 2),
            "identical_results": same}


def batched_2p5m_leg(frame, requests=64, seed=11, tuner=None):
    """Batched vs plain at 2.5 M triangles: `requests` ray sets against the
    2.5M-triangle sphere of the headline frame, 4k-2M rays each
    (log-uniform), kind shadow (any hit, tmax 0.5), AO (any hit, tmax 0.05)
    or closest hit in turn; origins uniform in 1.2x the mesh bounds,
    directions uniform on the sphere.  Batched: one rt_trace_batch launch
    replayed from a graph (tuned shape).  Plain: one rt_trace per request
    with the default shape, waiting for each.  Results compared bit for
    bit."""
    rng = np.random.default_rng(seed)
    lo = tuple(-1.2 * 1.02 for _ in range(3))
    hi = tuple(1.2 * 1.02 for _ in range(3))
    kinds = ("shadow", "ao", "closest")
    tmax = {"shadow": 0.5, "ao": 0.05, "closest": 1e30}
    items, rays = [], []
    for k in range(requests):
        n = int(np.exp(rng.uniform(np.log(4096), np.log(2_000_000))))
        kind = kinds[k % 3]
        r = rt.gen_random_rays(n, 9000 + k, lo, hi, tmax[kind])
        rays.append((r, kind))
        items.append((frame.bvh, r, "closest" if kind == "closest" else "any", None))
    nrays = sum(len(r) for r, _ in rays)
    batch = rt.TraceBatch(items)
    shape = tuner.trace_batch(batch, tag="sphere-2p5m") if tuner is not None else {}
    batch.launch(**shape)
    torch.cuda.synchronize()
    gs = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=gs):
        batch.launch(stream=gs, **shape)
    g.replay()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(3):
        g.replay()
    e.record()
    e.synchronize()
    batched_ms = s.elapsed_time(e) / 3
    plain = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r, kind in rays:
        plain.append(frame.bvh.intersect(r) if kind == "closest" else frame.bvh.occluded(r))
        torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    same = True
    for (r, kind), o, p in zip(rays, batch.outputs, plain):
        n = len(r)
        if kind == "closest":
            same &= bool(torch.equal(o[0][:n], p[0][:n]) and torch.equal(o[1][:n], p[1][:n]))
        else:
            w = (n + 31) // 32
            same &= bool(torch.equal(o[:w], p[:w]))
    return {"workload": f"{requests} ray sets ({nrays} rays, 4k-2M each, shadow / AO / closest) "
                        "against the 2.5M-triangle sphere",
            "batched_ms": round(batched_ms, 3),
            "batched_mrays_per_s": round(nrays / batched_ms / 1e3, 1),
            "plain_ms": round(plain_ms, 3), "plain_mrays_per_s": round(nrays / plain_ms / 1e3, 1),
            "speedup": round(plain_ms / batched_ms, 2), "identical_results": same}


def ao_frame_leg(frames=200, w=1920, h=1080, tuner=None):
    """Dynamic AO frame: a static 999,698-triangle seeded random-displaced
    heightfield and a 4,096-triangle deforming grid (t_k = 0.05 k, the 200
    time steps k = 0..199 of the configuration) as a two-level scene.  Per frame: regenerate the grid and rebuild its tree
    (fused path, same buffers), 1920x1080 primary closest hit, 1 SPP cosine
    AO rays (tmax 0.5), any hit, AO resolve, 7x7 bilateral filter.  With a
    tuner, the static terrain tree (leaf size x treelet rounds) is timed on
    whole frames (a tree that changes the frame beyond ties, same_ao_frame,
    is left out) and both scene traces use their measured launch shapes."""
    tp, ti = rt.gen_heightfield(707, 707, 4.0, 0.3, 31)
    terrain = rt.GpuBVH4.build(tp, ti, algorithm="radix")
    gx = rt.gen_dynamic_grid(4096, 0.0, 2.0)
    grid = rt.GpuBVH4.build(gx, None, algorithm="implicit")
    cfg = rt.build_cfg("implicit")
    m_grid = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 4.6]], np.float32)
    state = {"leaf": 4, "rounds": 0, "scene": rt.Scene([(terrain, tp, ti, None, rt.INST_IDENTITY),
                                           (grid, gx, None, m_grid, rt.INST_TRANSLATE)])}
    eye, look, up = (0.0, -6.0, 4.0), (0.0, 0.0, 0.2), (0.0, 0.0, 1.0)
    npix = w * h
    primary, ao_rays = rt.Rays.empty(npix), rt.Rays.empty(npix)
    prim = torch.empty(npix, dtype=torch.int32, device="cuda")
    t = torch.empty(npix, dtype=torch.float32, device="cuda")
    inst = torch.empty(npix, dtype=torch.int32, device="cuda")
    bits = torch.empty((npix + 31) // 32, dtype=torch.int32, device="cuda")
    ao = torch.empty(npix, dtype=torch.float32, device="cuda")
    depth = torch.empty(npix, dtype=torch.float32, device="cuda")
    out = torch.empty(npix, dtype=torch.float32, device="cuda")

    shapes = [{}, {}]

    def frame(k):
        rt.gen_dynamic_grid(4096, 0.05 * k, 2.0, out=gx)
        rt.call("rt_gpu_build", rt.stream_handle(), rt.ptr(gx), None, rt.u32(4096),
                rt.C.byref(cfg), rt.ptr(grid.nodes), rt.u32(grid.node_count), rt.ptr(grid.tris),
                rt.ptr(grid.info), None, None, None, rt.C.c_size_t(0))
        rt.gen_camera_rays(w, h, eye, look, up, 60.0, 0.0, 1e30, out=primary)
        scene = state["scene"]
        scene.intersect(primary, out=(prim, t, inst), **shapes[0])
        scene.gen_ao_rays(primary, prim, inst, t, w, h, frame=k, offset=0.03, tmin=0.01,
                          tmax=0.5, out=ao_rays)
        scene.occluded(ao_rays, out=bits, **shapes[1])
        rt.ao_resolve(bits, prim, out=ao)
        torch.where(prim == -1, torch.full_like(t, 1e30), t, out=depth)
        rt.blur7(ao, depth, w, h, 0.05, out=out)

    for k in range(3):
        frame(k)
    torch.cuda.synchronize()
    if tuner is not None:
        scenes = {(4, 0): state["scene"]}

        def use(c):
            leaf, rounds = int(c[0]), int(c[1])
            if (leaf, rounds) not in scenes:
                tb = rt.GpuBVH4.build(tp, ti, algorithm="radix", max_leaf_prims=leaf,
                                      treelet_rounds=rounds)
                scenes[(leaf, rounds)] = rt.Scene([(tb, tp, ti, None, rt.INST_IDENTITY),
                                                   (grid, gx, None, m_grid, rt.INST_TRANSLATE)])
            state["scene"], state["leaf"], state["rounds"] = scenes[(leaf, rounds)], leaf, rounds

        cands = [(leaf, r) for r in AO_TREELET_ROUNDS for leaf in autotune.LEAF_SIZES]
        use(tuner.frame_leaf_size(use, lambda: frame(0), lambda: (t, prim, inst, ao),
                                  tag="ao-terrain", sizes=cands, same=same_ao_frame))
        scenes.clear()
        frame(0)
        shapes[0] = tuner.trace_scene(state["scene"], primary, "closest", tag="ao-primary",
                                      grid_width=w)
        shapes[1] = tuner.trace_scene(state["scene"], ao_rays, "any", tag="ao", grid_width=w)
        torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for k in range(frames):
        frame(k)
    e.record()
    e.synchronize()
    ms = s.elapsed_time(e) / frames
    # image quality of the filtered frame against its 1-SPP input
    # (ImageMetrics::compute on the device), and that call's own time
    a2, b2 = out.view(h, w), ao.view(h, w)
    scratch = torch.empty(rt.image_metrics_scratch_bytes(w, h) // 8, dtype=torch.float64,
                          device="cuda")
    q = torch.empty(3, dtype=torch.float64, device="cuda")
    for _ in range(5):
        rt.image_metrics(a2, b2, out=q, scratch=scratch)
    # device time per call: 50 calls captured in a graph, replayed (host-side
    # call overhead excluded)
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    per_graph = 50
    with torch.cuda.stream(cs):
        rt.image_metrics(a2, b2, out=q, scratch=scratch, stream=cs)
        cs.synchronize()
        with torch.cuda.graph(g, stream=cs):
            for _ in range(per_graph):
                rt.image_metrics(a2, b2, out=q, scratch=scratch, stream=cs)
    torch.cuda.current_stream().wait_stream(cs)
    g.replay()
    torch.cuda.synchronize()
    replays = 4
    s.record()
    for _ in range(replays):
        g.replay()
    e.record()
    e.synchronize()
    met_us = s.elapsed_time(e) * 1000.0 / (replays * per_graph)
    mse, psnr, ssim = q.cpu().tolist()
    return {"workload": "dynamic AO: 999,698-tri heightfield + 4,096-tri grid rebuilt per "
                        "frame (two-level scene), 1920x1080 primary + 1 SPP AO (tmax 0.5) + 7x7 "
                        "bilateral", "frames": frames, "ms_per_frame": round(ms, 3),
            "fps": round(1000.0 / ms, 1), "mrays_per_s": round(2 * npix / ms / 1e3, 1),
            "terrain_max_leaf_prims": state["leaf"],
            "terrain_treelet_rounds": state["rounds"],
            "ao_mean": round(float(out.mean()), 4),
            "quality_filtered_vs_1spp": {"mse": round(mse, 6), "psnr_db": round(psnr, 3),
                                         "ssim": round(ssim, 5)},
            "image_metrics_us": round(met_us, 2),
            "image_metrics_gbs": round(2 * 4 * npix / (met_us * 1e3), 1)}


def serving_leg(seed=7, tuner=None):
    """Serving: 32 resident scenes (10k-2.5M tris, log-uniform) plus 8
    deformable grid scenes; requests of 4k-2M rays (shadow, AO, closest) on
    any scene or 4k-64k-triangle deform-and-rebuilds of a grid scene (later
    requests on it trace the new version).  Capacity: 300 requests all
    waiting at t = 0, served (batched, one submit call per tick and kind) and
    plain (one launch per request, waiting for each; the library's default
    launch shape, and again with the served path's shape), three alternating
    runs of each, every run's ratio reported.  Latency: 2,400 Poisson
    arrivals at half the served capacity (about 0.8 s: the p99 of a shorter
    stream swings with its largest burst), p50 / p99 over all requests.  With a tuner,
    the served trace schedule / shape is measured on a batch of 64 sample
    requests (not those served)."""
    import serve as S
    scenes = S.make_scenes(32)
    budget = 4 << 30
    shape = {}
    if tuner is not None:
        sample = [r for r in S.make_requests(96, 0, scenes, seed=seed + 100)
                  if r.kind != "rebuild"][:64]
        items = []
        for r in sample:
            rays = rt.gen_random_rays(r.n, r.seed, tuple(scenes.lo[r.scene]),
                                      tuple(scenes.hi[r.scene]), S.TMAX[r.kind])
            items.append((scenes.bvh(r.scene), rays,
                          "any" if r.kind != "closest" else "closest", None))
        shape = tuner.trace_batch(rt.TraceBatch(items), tag="serving")
        torch.cuda.synchronize()
    S.run_plain(scenes, S.make_requests(40, 0, scenes, seed=seed + 200))  # warm every path
    S.Server(scenes, budget, trace_shape=shape).run(S.make_requests(40, 0, scenes,
                                                                    seed=seed + 300))
    reqs = S.make_requests(300, 0, scenes, seed=seed)
    # three alternating repetitions of each path (one pass of either is a
    # fraction of a second, and the host loop's share varies)
    runs = {"sat": [], "plain": [], "shaped": []}
    for _ in range(3):
        runs["sat"].append(S.Server(scenes, budget, trace_shape=shape).run(
            S.clone_requests(reqs)))
        runs["plain"].append(S.run_plain(scenes, S.clone_requests(reqs)))
        runs["shaped"].append(S.run_plain(scenes, S.clone_requests(reqs), trace_shape=shape))
    ratios = [round(a["mrays_per_s"] / b["mrays_per_s"], 2)
              for a, b in zip(runs["sat"], runs["plain"])]
    sat, plain, plain_shaped = (sorted(runs[k], key=lambda r: r["mrays_per_s"])[1]
                                for k in ("sat", "plain", "shaped"))
    qps = round(0.5 * len(reqs) / sat["span_s"])
    lat = S.Server(scenes, budget, trace_shape=shape).run(
        S.make_requests(2400, qps, scenes, seed=seed + 1))
    return {"workload": f"32 resident scenes ({sum(scenes.tris)} tris, SAH treelet trees) "
                        "+ 8 deformable grids, "
                        "4k-2M-ray shadow/AO/closest requests, 4k-64k-tri deform-and-rebuilds",
            "capacity_mrays_per_s": sat["mrays_per_s"],
            "plain_mrays_per_s": plain["mrays_per_s"],
            "plain_same_shape_mrays_per_s": plain_shaped["mrays_per_s"],
            "speedup": round(sat["mrays_per_s"] / plain["mrays_per_s"], 2),
            "speedup_runs": ratios,
            "speedup_vs_plain_same_shape": round(sat["mrays_per_s"] /
                                                 plain_shaped["mrays_per_s"], 2),
            "qps": qps, "p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"],
            "latency_mrays_per_s": lat["mrays_per_s"],
            "p99_ms_by_kind": {k: lat.get(f"p99_ms_{k}") for k in S.KINDS}}


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain device copy of nbytes (read nbytes, write nbytes), the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--schedule", choices=["refill", "simple"], default="simple",
                    help="traversal schedule when not tuning (waves refilling finished lanes, "
                         "or one ray/thread)")
    ap.add_argument("--no-tune", action="store_true",
                    help="fixed launch shapes instead of the autotuner's measured choices")
    ap.add_argument("--tune-cache", default=None,
                    help="JSON file of tuned choices: read if present (no re-measurement), "
                         "written after tuning")
    ap.add_argument("--rays-per-wave", type=int, default=1024)
    ap.add_argument("--refill-lanes", type=int, default=16)
    ap.add_argument("--builder", choices=["radix", "implicit"], default="radix")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched and serving legs")
    a = ap.parse_args()
    rt.DEFAULT_SCHEDULE = a.schedule
    rt.TRACE_TUNING.update(rays_per_wave=a.rays_per_wave, refill_lanes=a.refill_lanes)
    Frame.builder = a.builder
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the per-GPU shard "
              "is the one-GPU workload and there is no multi-GPU path in this build",
              file=sys.stderr)
        sys.exit(2)
    torch.cuda.set_device(0)
    rt.lib()
    if a.copy_probe:
        copy_probe()

    # traces are tuned on their whole ray sets (the frame's 33 M shadow rays take
    # ~12 ms a launch): sampled subsets rank near-equal shapes unreliably
    tuner = None if a.no_tune else autotune.Tuner(
        cache_path=a.tune_cache, sample_rays=1 << 26,
        log=lambda m: print(m, file=sys.stderr, flush=True))
    b4k_p50, b4k_p99, b4k_p99_launch, b4k_b2b = build_time_4k()
    f = Frame(tuner)
    refit_ms = refit_time_ms(f)
    if tuner is not None:
        f.tune(tuner)
    for _ in range(a.warmup):
        f.step()
    torch.cuda.synchronize()

    step_ms, kern_ms = [], []
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(a.steps)]
    fs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        fs[i][0].record()
        f.step(events=evs[i])
        fs[i][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for i in range(a.steps):
        step_ms.append(fs[i][0].elapsed_time(fs[i][1]))
        kern_ms.append(evs[i][0].elapsed_time(evs[i][1]))

    ms_per_step = wall * 1000.0 / a.steps
    value = f.rays_per_step / (ms_per_step * 1e-3) / 1e6
    kern_avg = sum(kern_ms) / len(kern_ms)
    ab = algorithmic_bytes_shadow(f)
    achieved = ab / (kern_avg * 1e-3) / 1e9
    # HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated on a copy) of
    # the shadow kernel measured for this run's tree and wave tile; null when
    # the profile holds no measurement of that configuration
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic_shadow_anyhit_2p5m.json")
    tile = f.shadow_opts.get("tile")
    hinted = bool(f.shadow_opts.get("hinted"))
    if os.path.exists(tp):
        with open(tp) as fh:
            for e in json.load(fh).get("entries", []):
                if (e["max_leaf_prims"], e["treelet_rounds"]) == (f.leaf, f.rounds) and \
                        tile is not None and e["tile_rows"] == tile[1] and \
                        bool(e.get("hinted", False)) == hinted:
                    traffic = e["hbm_bytes_per_launch"]
    if hinted:
        kernel = "k_hint_pass + k_hint_trace (hinted occlusion of the shadow rays)"
    elif f.shadow_opts.get("schedule", a.schedule) == "refill":
        kernel = "k_trace_refill_bits any-hit (shadow)"
    else:
        kernel = "k_trace any-hit (shadow)"

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "shadow any-hit: 2.5M-tri displaced sphere, 1920x1080 primary "
                               "closest-hit + 16 shadow rays/pixel",
                   "tris": f.bvh.tri_count, "width": W, "height": H, "lights": NLIGHTS,
                   "rays_per_step": f.rays_per_step, "builder": f.builder,
                   "sort_bits": f.sort_bits,
                   "schedule": "hinted" if hinted else f.shadow_opts.get("schedule", a.schedule),
                   "primary_trace": f.primary_opts or "default",
                   "shadow_trace": f.shadow_opts or "default",
                   "bvh_nodes": f.bvh.node_count, "bvh_depth": f.bvh.depth,
                   "max_leaf_prims": f.leaf, "treelet_rounds": f.rounds},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": kernel,
                     "kernel_ms": round(kern_avg, 4), "algorithmic_bytes": int(ab)},
        "build_us_4k_p50": round(b4k_p50, 2),
        "build_us_4k_p99": round(b4k_p99, 2),
        "build_us_4k_p99_with_launch": round(b4k_p99_launch, 2),
        "build_us_4k_back_to_back": round(b4k_b2b, 2),
        "builds_per_s_4k": round(1e6 / b4k_b2b),
        "build_ms_2p5m": round(plain_build_ms(f), 3),
        "frame_tree_build_ms": round(f.build_ms, 3),
        "refit_ms_2p5m": round(refit_ms, 3),
        "frame_latency_ms_p50": round(median(step_ms), 4),
        "frame_latency_ms_p99": round(percentile(step_ms, 0.99), 4),
    }
    if not a.no_extra:
        out["ao_frame"] = ao_frame_leg(tuner=tuner)
        out["batched"] = batched_leg(tuner=tuner)
        out["batched_2p5m"] = batched_2p5m_leg(f, tuner=tuner)
        out["serving"] = serving_leg(tuner=tuner)
        out["p99_request_latency_ms"] = out["serving"]["p99_ms"]
    if tuner is not None:
        # every tuned choice with its time and the default's: on stderr, so the
        # result line stays short enough to be read whole from a log's tail
        print("tuning " + json.dumps(tuner.report()), file=sys.stderr, flush=True)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(f, cpu_threads())
    else:
        out["cpu_baseline"] = None
    line = json.dumps(out)
    if len(line) > 4000:
        print(f"note: result line is {len(line)} bytes", file=sys.stderr)
    print(line)


if __name__ == "__main__":
    main()
