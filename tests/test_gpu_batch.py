"""Batched trace (rt_trace_batch): many trees, ray sets and modes in one
launch; every item's results equal rt_trace on that item alone (same
kernel body), and the oracle's on spot checks.  Graph replays of the
captured launch equal the eager results."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((b[i >> 5] >> (i & 31)) & 1).astype(bool)


def _scenes(rt, O):
    out = []
    p, i = O.wave_grid(32, 32, 1.0, 0.0, 0.1, 4.0, 4.0)
    out.append(rt.GpuBVH4.build(torch.from_numpy(p).cuda(),
                                torch.from_numpy(i.view(np.int32)).cuda(), algorithm="implicit"))
    ps, ids = rt.gen_sphere(300, 200, 1.0, 0.02, 11)
    out.append(rt.GpuBVH4.build(ps, ids, algorithm="radix"))
    xyz = rt.gen_dynamic_grid(4096, 0.4, 1.25)
    out.append(rt.GpuBVH4.build(xyz, None, algorithm="implicit"))
    return out


def _items(rt, O, trees):
    counts = [1, 63, 64, 65, 1000, 4097, 100_003, 0, 7]
    items = []
    for k, n in enumerate(counts):
        bvh = trees[k % len(trees)]
        r = O.random_rays(n, 100 + k, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3),
                          tmax=0.4 if k % 2 else 1e30)
        items.append((bvh, rt.Rays.from_numpy(r.packed()), "any" if k % 2 else "closest", None))
    return items


def test_batch_equals_alone(rt, O):
    trees = _scenes(rt, O)
    items = _items(rt, O, trees)
    batch = rt.TraceBatch(items)
    outs = batch.launch()
    torch.cuda.synchronize()
    for (bvh, rays, mode, _), got in zip(items, outs):
        n = len(rays)
        if mode == "closest":
            prim, t = bvh.intersect(rays)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(u32(got[0])[:n], u32(prim)[:n])
            np.testing.assert_array_equal(to_np(got[1])[:n].view(np.uint32),
                                          to_np(t)[:n].view(np.uint32))
        else:
            bits = bvh.occluded(rays)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(_bits(u32(got), n), _bits(u32(bits), n))


def test_batch_matches_oracle(rt, O):
    """One small tree from the oracle's software_bvh build, two items (closest
    and any) in one batch: hit IDs, t and occlusion bits equal the CPU path."""
    pos, idx = O.wave_grid(48, 48, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, tris = O.software_bvh4(pos, idx)
    bvh = rt.GpuBVH4.from_bvh4(n4, tris64=tris, depth=O.bvh4_depth(n4))
    r1 = O.random_rays(50_001, 7, (-1.2, -1.2, -0.3), (1.2, 1.2, 0.3), tmax=1e30)
    r2 = O.random_rays(30_000, 8, (-1.2, -1.2, -0.3), (1.2, 1.2, 0.3), tmax=0.3)
    batch = rt.TraceBatch([(bvh, rt.Rays.from_numpy(r1.packed()), "closest", None),
                           (bvh, rt.Rays.from_numpy(r2.packed()), "any", None)])
    (prim, t), bits = batch.launch()
    torch.cuda.synchronize()
    pc, tc = O.trace_bvh4(n4, tris, r1)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    pa, _ = O.trace_bvh4(n4, tris, r2, mode=O.ANY)
    np.testing.assert_array_equal(_bits(u32(bits), len(r2)), pa != O.MISS)


def test_batch_graph_replays(rt, O):
    """The launch captured in a graph: three replays with eager work and
    synchronisation between them, each equal to the eager launch."""
    trees = _scenes(rt, O)
    items = _items(rt, O, trees)
    batch = rt.TraceBatch(items)
    ref = [tuple(x.clone() for x in o) if isinstance(o, tuple) else o.clone()
           for o in batch.launch()]
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        batch.launch(stream=s)
    for rep in range(3):
        for o in batch.outputs:
            for x in (o if isinstance(o, tuple) else (o,)):
                x.fill_(-7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (bvh, rays, mode, _), got, want in zip(items, batch.outputs, ref):
            n = len(rays)
            if mode == "closest":
                np.testing.assert_array_equal(u32(got[0])[:n], u32(want[0])[:n])
                np.testing.assert_array_equal(u32(got[1])[:n], u32(want[1])[:n])
            else:
                np.testing.assert_array_equal(_bits(u32(got), n), _bits(u32(want), n))
