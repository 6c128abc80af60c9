"""Autotuner: picks launch shapes and the radix digit width by measuring the
candidates on the caller's own workload (or a slice of it) on the device.

Every knob below changes only how the work is scheduled, never the result,
so the tuner also checks that each candidate's output equals the default
configuration's bit for bit (a mismatch raises: it would be a kernel bug).

  trace (rt_trace_ex)          schedule simple | refill, block (threads per
                               workgroup, the ray tile), lds_stack (traversal
                               stack entries per lane held in LDS: more
                               entries = fewer private-memory spills, fewer
                               resident waves), and for refill rays_per_wave
                               (the wave's ray tile) and refill_lanes
  batched trace (rt_trace_batch_ex)   the same knobs (the refill chunk is a
                               wave's share of one item)
  hinted occlusion (rt_trace_occluded_hinted)  block, lds_stack, wave
                               tiles, segment slots, refill_lanes (when the
                               caller has occluder hints for the rays)
  two-level scene (rt_trace_scene_ex)  block, lds_stack
  device build (rt_gpu_build)  sort_bits: digit width of the Morton radix
                               sort, 4 / 6 / 8 / 10 bits = 8 / 5 / 4 / 3 passes
  batched build (rt_gpu_build_batch)  workgroups of the persistent launch

Choices are cached per workload signature (kind, mode, tree depth, size
classes); `Tuner(cache_path=...)` also keeps them in a JSON file across
processes.  Timing uses HIP events on the stream the work runs on.
"""
import json
import math
import os
import time

import torch

import rtbvh as rt

TRACE_BLOCKS = (64, 128, 256)
TRACE_LDS = (8, 12, 16, 20, 24, 32)
REFILL_RAYS_PER_WAVE = (256, 1024, 4096)
REFILL_LANES = (8, 16, 32)
# wave tiles of a ray grid: 16x4, 8x8, 32x2 rays.  16x4 first: on the shadow
# frame 8x8 tiles take the same time but miss L2 about twice as often
# (7.1 vs 3.8 GB of L2 misses per launch), so a tie keeps 16x4
TILE_WIDTHS = (16, 8, 32)
PACKET_TILE_WIDTHS = (16, 8)  # packet traversal (occlusion bits) of 16x4 / 8x8 tiles
HINT_SEGMENTS = (256, 512, 1024)  # slots per segment of the hinted trace
HINT_REFILL_LANES = (32, 48)
# a later candidate must be faster than the best so far by more than this
# fraction (the spread of repeated timings of one candidate) to replace it
TIE_MARGIN = 0.01
LEAF_SIZES = (4, 3, 2, 1)  # radix-tree leaf sizes of a frame's tree, the default first
SAMPLE_ROW_BLOCK = 8       # rows per block of a grid sample (a multiple of every tile height)
FRAME_REPS, FRAME_TARGET_MS = 7, 40.0  # timing of a frame-tree candidate


def hinted_kwargs(choice, hint):
    """The keyword arguments of bvh.occluded for a tuned choice: a hinted
    choice ("hinted": True) gets the hints hint = (triangle ids, period)."""
    kw = dict(choice)
    if kw.pop("hinted", False):
        kw.update(hint=hint[0], hint_period=hint[1])
    return kw


def _grid_sample(rays, width, m):
    """About m rays of a grid of `width`-ray rows as blocks of SAMPLE_ROW_BLOCK
    whole rows spread evenly over all rows (every light layer / image region
    is represented, and the sample is a grid of the same width again)."""
    rows = len(rays) // width
    blocks = max(1, min(rows // SAMPLE_ROW_BLOCK, m // (width * SAMPLE_ROW_BLOCK)))
    if blocks * SAMPLE_ROW_BLOCK >= rows:
        return rays
    step = rows // blocks
    starts = torch.arange(blocks, device=rays.data.device) * step
    row = (starts[:, None] + torch.arange(SAMPLE_ROW_BLOCK, device=rays.data.device)).reshape(-1)
    col = (row[:, None] * width + torch.arange(width, device=rays.data.device)).reshape(-1)
    return rt.Rays(rays.data[:, col].contiguous())


SORT_BITS = rt.SORT_BITS
BUILD_WORKGROUPS = (64, 128, 256, 512, 1024)


def _size_class(n):
    return int(math.log2(max(1, n)))


def time_ms(fn, stream=None, reps=3, target_ms=2.0):
    """Median over `reps` trials of fn()'s device time (ms per call), each
    trial a loop of enough calls to last about target_ms, bracketed by HIP
    events recorded on `stream` (the stream fn launches on)."""
    s = stream or torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()  # warm-up (module load, caches)
    ev0.record(s)
    fn()
    ev1.record(s)
    ev1.synchronize()
    one = max(ev0.elapsed_time(ev1), 1e-3)
    inner = max(1, min(200, int(target_ms / one)))
    trials = []
    for _ in range(reps):
        ev0.record(s)
        for _ in range(inner):
            fn()
        ev1.record(s)
        ev1.synchronize()
        trials.append(ev0.elapsed_time(ev1) / inner)
    trials.sort()
    return trials[len(trials) // 2]


def _same(a, b):
    if isinstance(a, (tuple, list)):
        return all(_same(x, y) for x, y in zip(a, b))
    return bool(torch.equal(a, b))


def _clone(o):
    if isinstance(o, (tuple, list)):
        return type(o)(_clone(x) for x in o)
    return o.clone()


class Tuner:
    def __init__(self, cache_path=None, reps=3, sample_rays=1 << 21, log=None):
        self.cache_path = cache_path
        self.reps = reps
        self.sample_rays = sample_rays
        self.log = log
        self.cache = {}
        self.history = []  # (key, candidate, ms) of every measurement
        self.seconds = 0.0
        self._depth = 0
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def _save(self):
        if self.cache_path:
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.cache_path)

    def _pick(self, key, candidates, run, check, skip_changed=False, same=None, reps=None,
              target_ms=2.0):
        """Times run(c) for every candidate c, checks check(c) == the first
        candidate's output (or same(out, first) when given), caches and
        returns the fastest (a later candidate wins only when faster by more
        than TIE_MARGIN: within timing noise the earlier one stays).  A
        candidate whose output differs raises, or with skip_changed (knobs
        that may change results, such as the tree's leaf size) is left out."""
        if key in self.cache:
            return self.cache[key]["choice"]
        t0 = time.perf_counter()
        self._depth += 1  # a candidate's run may tune other knobs: time counted once
        try:
            best, best_ms, table = self._measure(key, candidates, run, check, skip_changed,
                                                 same or _same, reps or self.reps, target_ms)
        finally:
            self._depth -= 1
        if self._depth == 0:
            self.seconds += time.perf_counter() - t0
        default_ms = table[0][1]
        self.cache[key] = {"choice": best, "ms": round(best_ms, 5),
                           "default_ms": round(default_ms, 5), "candidates": len(candidates)}
        if self.log:
            self.log(f"tuned {key}: {best} {best_ms:.4f} ms (default {default_ms:.4f} ms, "
                     f"{len(candidates)} candidates)")
        self._save()
        return best

    def _measure(self, key, candidates, run, check, skip_changed, same, reps, target_ms):
        ref = _clone(check(candidates[0]))
        torch.cuda.synchronize()
        best, best_ms, table = None, float("inf"), []
        for c in candidates:
            out = check(c)
            torch.cuda.synchronize()
            if not same(out, ref):
                if skip_changed:
                    self.history.append((key, c, None))
                    continue
                raise AssertionError(f"tuner: candidate {c} changed the result of {key}")
            ms = time_ms(lambda: run(c), reps=reps, target_ms=target_ms)
            table.append((c, ms))
            self.history.append((key, c, ms))
            if ms < best_ms * (1.0 - TIE_MARGIN) or best is None:
                best, best_ms = c, ms
        return best, best_ms, table

    # -- trace -----------------------------------------------------------------
    def trace(self, bvh, rays, mode, tag="", grid_width=None, hint=None):
        """Schedule and launch shape for bvh.intersect / bvh.occluded on rays
        like these (a leading slice of at most sample_rays is measured).
        grid_width: the rays are rows of that many rays (an image), which
        adds wave tiles of the grid to the candidates (the slice is then
        whole rows).  hint = (triangle ids, period): occluder hints of the
        rays (mode "any", measured on the whole ray set only), which adds the
        hinted trace (rt_trace_occluded_hinted) to the candidates; a hinted
        choice carries "hinted": True, for which the caller passes
        hint= / hint_period= (hinted_kwargs).  Returns the keyword arguments
        for intersect / occluded."""
        n = len(rays)
        m = min(n, self.sample_rays)
        sample = rays if m == n else (_grid_sample(rays, grid_width, m) if grid_width
                                      else rt.Rays(rays.data[:, :m]))
        m = len(sample)
        key = (f"trace|{mode}|{tag}|d{bvh.depth}|t{_size_class(bvh.tri_count)}"
               f"|r{_size_class(m)}" + (f"|g{grid_width}" if grid_width else ""))
        fn = bvh.intersect if mode == "closest" else bvh.occluded
        out = None
        hinted = hint is not None and mode == "any" and m == n

        def run(c):
            nonlocal out
            out = fn(sample, out=out, **hinted_kwargs(c, hint))
            return out

        def check(c):
            return _clone(run(c))

        simple = [{"schedule": "simple", "block": b, "lds_stack": l}
                  for b in TRACE_BLOCKS for l in TRACE_LDS]
        # the default first: its output is the reference
        simple.sort(key=lambda c: (c["block"], c["lds_stack"]) != (256, 20))
        best = self._pick(key + "|simple", simple, run, check)
        refill = [{"schedule": "refill", "block": best["block"],
                   "lds_stack": best["lds_stack"], "rays_per_wave": w, "refill_lanes": r}
                  for w in REFILL_RAYS_PER_WAVE for r in REFILL_LANES]
        # occlusion bits do not depend on the child order: far-first and slot
        # order at the best simple shape compete too
        orders = [] if mode == "closest" else [
            dict(best, any_order=o) for o in rt.ANY_ORDERS if o != "near"]
        tiles = [dict(best, tile=(grid_width, tw)) for tw in TILE_WIDTHS
                 if grid_width and grid_width % tw == 0]
        # occlusion bits of a ray grid: packet traversal of wave tiles too
        packets = [] if mode == "closest" else [
            {"schedule": "packet", "block": b, "tile": (grid_width, tw)}
            for tw in PACKET_TILE_WIDTHS for b in (64, 256)
            if grid_width and grid_width % tw == 0]
        # occluder hints: the hinted trace over segments of HINT_SEGMENTS
        # slots, linear waves and the grid's wave tiles
        tile = [(grid_width, tw) for tw in TILE_WIDTHS[:1] if grid_width and grid_width % tw == 0]
        hints = [] if not hinted else [
            {"hinted": True, "block": b, "lds_stack": best["lds_stack"], "rays_per_wave": w,
             "refill_lanes": r, **({"tile": t} if t else {})}
            for t in tile + [None] for b in dict.fromkeys((best["block"], 128))
            for w in HINT_SEGMENTS for r in HINT_REFILL_LANES
            if t or (b == best["block"] and r == HINT_REFILL_LANES[0])]
        if hinted:
            key += "|hint"
        return self._pick(key, [best] + refill + orders + tiles + packets + hints, run, check)

    def trace_scene(self, scene, rays, mode, tag="", grid_width=None):
        """Launch shape (block, lds_stack, and wave tiles of the ray grid when
        grid_width is given) for Scene.intersect / occluded."""
        n = len(rays)
        m = min(n, self.sample_rays)
        sample = rays if m == n else (_grid_sample(rays, grid_width, m) if grid_width
                                      else rt.Rays(rays.data[:, :m]))
        m = len(sample)
        key = (f"scene|{mode}|{tag}|i{scene.count}|d{scene.depth}|r{_size_class(m)}"
               + (f"|g{grid_width}" if grid_width else ""))
        fn = scene.intersect if mode == "closest" else scene.occluded
        out = None

        def run(c):
            nonlocal out
            out = fn(sample, out=out, **c)
            return out

        def check(c):
            return _clone(run(c))

        cands = [{"block": b, "lds_stack": l} for b in TRACE_BLOCKS for l in TRACE_LDS]
        cands.sort(key=lambda c: (c["block"], c["lds_stack"]) != (256, 20))
        if not grid_width:
            return self._pick(key, cands, run, check)
        best = self._pick(key + "|shape", cands, run, check)
        tiles = [dict(best, tile=(grid_width, tw)) for tw in TILE_WIDTHS
                 if grid_width % tw == 0]
        return self._pick(key, [best] + tiles, run, check)

    def trace_batch(self, batch, tag=""):
        """Schedule and launch shape for TraceBatch.launch: the best simple
        shape (block, lds_stack), then the refill schedule around it."""
        key = (f"batch|{tag}|n{batch.count}|d{batch.depth}"
               f"|c{_size_class(batch.total_chunks)}")
        cands = [{"schedule": "simple", "block": b, "lds_stack": l}
                 for b in TRACE_BLOCKS for l in TRACE_LDS]
        cands.sort(key=lambda c: (c["block"], c["lds_stack"]) != (256, 20))

        def run(c):
            return batch.launch(**c)

        def check(c):
            return _clone(run(c))

        best = self._pick(key + "|simple", cands, run, check)
        refill = [{"schedule": "refill", "block": best["block"],
                   "lds_stack": best["lds_stack"], "rays_per_wave": w, "refill_lanes": r}
                  for w in REFILL_RAYS_PER_WAVE for r in REFILL_LANES]
        return self._pick(key, [best] + refill, run, check)

    # -- build -----------------------------------------------------------------
    def sort_bits(self, pos, idx=None, algorithm="radix", tag=""):
        """Radix digit width for GpuBVH4.build of meshes like this one
        (meshes of more than 4096 triangles; the fused path has one sort)."""
        n = idx.shape[0] if idx is not None else pos.numel() // 9
        key = f"sort|{algorithm}|{tag}|t{_size_class(n)}"
        cands = [8] + [b for b in SORT_BITS if b != 8]
        scratch = torch.empty(rt.build_requirements(n, rt.build_cfg(algorithm))[1] + 256,
                              dtype=torch.uint8, device=pos.device)
        holder = {}

        def run(b):
            holder["bvh"] = rt.GpuBVH4.build(pos, idx, algorithm=algorithm, sort_bits=b,
                                             scratch=scratch, sync=False, keep_order=True)
            return holder["bvh"]

        def check(b):
            # the written part only: a radix tree fills fewer nodes than its capacity
            t = run(b)
            count, depth = [int(v) for v in t.info[:2].cpu().tolist()]
            return (t.nodes[: count * rt.NODE_BYTES].clone(), t.tris.clone(),
                    t.sorted_ids.clone(), t.codes.clone(), torch.tensor([count, depth]))

        return self._pick(key, cands, run, check)

    def frame_leaf_size(self, use, step, output, tag="", sizes=LEAF_SIZES, same=None):
        """Tree of a frame: use(c) makes the frame trace the tree of candidate
        c (a leaf size, or a (leaf size, treelet rounds) pair; built once,
        outside the timing), step() runs one frame, output() is the frame's
        result.  A candidate whose frame result differs from the first one's
        (by same(out, first) when given, else bit for bit) is left out."""

        def run(c):
            use(c)
            return step()

        def check(c):
            use(c)
            step()
            return _clone(output())

        # whole frames differ by a few per cent between trees: each candidate
        # is the median of FRAME_REPS trials of about FRAME_TARGET_MS of frames
        return self._pick(f"leaf|{tag}", list(sizes), run, check, skip_changed=True, same=same,
                          reps=FRAME_REPS, target_ms=FRAME_TARGET_MS)

    def build_batch(self, batch, tag=""):
        """Persistent-grid size for BuildBatch.launch."""
        key = f"buildbatch|{tag}|n{batch.count}"
        cands = [0] + [w for w in BUILD_WORKGROUPS if w < batch.count] + [batch.count]
        cands = list(dict.fromkeys(cands))

        def run(w):
            batch.workgroups = w
            return batch.launch()

        def check(w):
            return [(t.nodes[: t.node_count * rt.NODE_BYTES].clone(), t.tris.clone(),
                     t.info[:2].clone()) for t in run(w)]

        w = self._pick(key, cands, run, check)
        batch.workgroups = w
        return w

    def report(self):
        return {"seconds": round(self.seconds, 2),
                "choices": {k: v for k, v in self.cache.items()}}
