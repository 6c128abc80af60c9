"""Tuning knobs and the autotuner.  Every launch shape (block size, LDS
stack depth, schedule) and every radix digit width must give the oracle's
results exactly: hit IDs, t and counters bit-identical, occlusion bits
identical, Morton codes, sorted order, nodes and leaves byte-identical.
The autotuner's choices are then applied and checked the same way."""
import json

import numpy as np
import pytest
import torch

from conftest import u32, to_np
from test_gpu_build import _cmp_build
from test_gpu_trace import _cmp_any, _cmp_closest, _upload_tree

pytestmark = pytest.mark.gpu

SHAPES = [dict(block=64, lds_stack=4), dict(block=128, lds_stack=12),
          dict(block=256, lds_stack=48), dict(block=64, lds_stack=48),
          dict(block=256, lds_stack=4), dict(block=128, lds_stack=32)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"b{s['block']}-l{s['lds_stack']}")
@pytest.mark.parametrize("schedule", ["simple", "refill"])
def test_launch_shapes_match_oracle(rt, O, shape, schedule):
    pos, idx = O.sphere(120, 90, 1.0, 0.02, 5)
    n4, tris, bvh = _upload_tree(rt, O, pos, idx)
    opts = dict(shape, schedule=schedule)
    if schedule == "refill":
        opts.update(rays_per_wave=512, refill_lanes=8)
    r = O.random_rays(70_001, 21, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
    _cmp_closest(rt, O, n4, tris, bvh, r, opts)
    r = O.random_rays(50_017, 22, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=0.5)
    _cmp_any(rt, O, n4, tris, bvh, r, opts)


@pytest.mark.parametrize("lds", [4, 20])
def test_unknown_depth_large_private_stack(rt, O, lds):
    """depth 0 (unknown) sizes the private overflow for 64 levels (the large
    private variant); results are unchanged."""
    pos, idx = O.sphere(100, 80, 1.0, 0.02, 6)
    n4, tris, bvh = _upload_tree(rt, O, pos, idx)
    bvh.depth = 0
    r = O.random_rays(40_000, 23, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
    _cmp_closest(rt, O, n4, tris, bvh, r, dict(lds_stack=lds, block=128))


def test_batch_launch_shapes(rt, O):
    from test_gpu_batch import _items, _scenes
    trees = _scenes(rt, O)
    items = _items(rt, O, trees)
    batch = rt.TraceBatch(items)
    ref = [tuple(x.clone() for x in o) if isinstance(o, tuple) else o.clone()
           for o in batch.launch()]
    torch.cuda.synchronize()
    refill = [dict(schedule="refill", block=64, lds_stack=8, rays_per_wave=64, refill_lanes=1),
              dict(schedule="refill", block=256, lds_stack=20, rays_per_wave=4096,
                   refill_lanes=64),
              dict(schedule="refill", block=128, lds_stack=16, rays_per_wave=256,
                   refill_lanes=8)]
    for shape in SHAPES + refill:
        outs = batch.launch(**shape)
        torch.cuda.synchronize()
        for (bvh, rays, mode, _), a, b in zip(items, outs, ref):
            n = len(rays)
            if mode == "closest":
                np.testing.assert_array_equal(u32(a[0])[:n], u32(b[0])[:n])
                np.testing.assert_array_equal(u32(a[1])[:n], u32(b[1])[:n])
            else:
                w = (n + 31) // 32
                np.testing.assert_array_equal(u32(a)[:w], u32(b)[:w])


@pytest.mark.parametrize("bits", [4, 6, 10])
@pytest.mark.parametrize("n", [4097, 65_537, 250_000])
def test_sort_bits_match_oracle(rt, O, bits, n):
    """Every digit width sorts to the same order (stable LSD passes over the
    30-bit codes): codes, ids, nodes, leaves equal the oracle's."""
    nu = 50
    nv = max(2, n // (2 * nu) + 1)
    pos, idx = O.sphere(nu, nv, 1.0, 0.02, 9)
    old = rt.DEFAULT_SORT_BITS
    rt.DEFAULT_SORT_BITS = bits
    try:
        _cmp_build(rt, O, pos, idx[:n], "implicit")
        if n == 65_537:
            _cmp_build(rt, O, pos, idx[:n], "radix", 4)
    finally:
        rt.DEFAULT_SORT_BITS = old


@pytest.mark.parametrize("bits", [4, 10])
def test_sort_bits_ties(rt, O, bits):
    """Terrain grid: long runs of equal codes keep triangle-index order for
    every digit width."""
    pos, idx = O.terrain(100_000)
    old = rt.DEFAULT_SORT_BITS
    rt.DEFAULT_SORT_BITS = bits
    try:
        _cmp_build(rt, O, pos, idx, "radix", 4)
    finally:
        rt.DEFAULT_SORT_BITS = old


def test_tuner_choices_keep_results(rt, O, tmp_path):
    import autotune
    cache = tmp_path / "tune.json"
    tuner = autotune.Tuner(cache_path=str(cache), reps=1, sample_rays=1 << 16)
    ps, ids = rt.gen_sphere(200, 150, 1.0, 0.02, 3)
    bvh = rt.GpuBVH4.build(ps, ids, algorithm="radix")
    rays = rt.gen_random_rays(200_000, 9, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2), 1e30)
    c = tuner.trace(bvh, rays, "closest", tag="t")
    assert c["schedule"] in ("simple", "refill") and c["block"] in (64, 128, 256)
    p0, t0 = bvh.intersect(rays)
    p1, t1 = bvh.intersect(rays, **c)
    torch.cuda.synchronize()
    assert torch.equal(p0, p1) and torch.equal(t0, t1)
    a = tuner.trace(bvh, rays, "any", tag="t")
    assert torch.equal(bvh.occluded(rays), bvh.occluded(rays, **a))
    bits = tuner.sort_bits(ps, ids, algorithm="radix", tag="t")
    assert bits in rt.SORT_BITS
    grids = [rt.gen_heightfield(10 + k, 8 + k // 2, 1.0, 0.05, k) for k in range(40)]
    assert max(int(i.shape[0]) for _, i in grids) <= rt.BATCH_BUILD_MAX
    bb = rt.BuildBatch(grids)
    w = tuner.build_batch(bb, tag="t")
    assert w == bb.workgroups
    # cached: same answers without measuring again; the file holds them
    n_hist = len(tuner.history)
    assert tuner.trace(bvh, rays, "closest", tag="t") == c
    assert len(tuner.history) == n_hist
    saved = json.loads(cache.read_text())
    assert any(k.startswith("trace|closest|t|") for k in saved)
    again = autotune.Tuner(cache_path=str(cache), sample_rays=1 << 16)
    assert again.trace(bvh, rays, "closest", tag="t") == c
    assert len(again.history) == 0
    rep = tuner.report()
    assert rep["choices"] and rep["seconds"] > 0


def test_tuner_trace_batch(rt, O):
    import autotune
    from test_gpu_batch import _items, _scenes
    trees = _scenes(rt, O)
    items = _items(rt, O, trees)
    batch = rt.TraceBatch(items)
    tuner = autotune.Tuner(reps=1)
    shape = tuner.trace_batch(batch, tag="t")
    assert shape["block"] in (64, 128, 256) and 4 <= shape["lds_stack"] <= 48
    assert shape["schedule"] in ("simple", "refill")
