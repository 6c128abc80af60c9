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


@pytest.mark.parametrize("sched", [dict(), dict(schedule="refill", rays_per_wave=512,
                                                  refill_lanes=8, block=64)],
                         ids=["simple", "refill"])
def test_batch_matches_oracle(rt, O, sched):
    """One small tree from the oracle's software_bvh build, two items (closest
    and any) in one batch: hit IDs, t and occlusion bits equal the CPU path."""
    pos, idx = O.wave_grid(48, 48, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, tris = O.software_bvh4(pos, idx)
    bvh = rt.GpuBVH4.from_bvh4(n4, tris64=tris, depth=O.bvh4_depth(n4))
    r1 = O.random_rays(50_001, 7, (-1.2, -1.2, -0.3), (1.2, 1.2, 0.3), tmax=1e30)
    r2 = O.random_rays(30_000, 8, (-1.2, -1.2, -0.3), (1.2, 1.2, 0.3), tmax=0.3)
    batch = rt.TraceBatch([(bvh, rt.Rays.from_numpy(r1.packed()), "closest", None),
                           (bvh, rt.Rays.from_numpy(r2.packed()), "any", None)])
    (prim, t), bits = batch.launch(**sched)
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


def _meshes(rt, O):
    rng = np.random.default_rng(4)
    out = []
    for k, n in enumerate([4096, 1, 2, 5, 64, 100, 255, 256, 257, 1000, 3001, 4095, 777]):
        if k % 3 == 0:
            nu = max(2, int(np.sqrt(n / 2)))
            pos, idx = O.sphere(nu, max(2, (n + 2 * nu - 1) // (2 * nu)), 1.0, 0.02, k)
            idx = idx[:n]
            out.append((torch.from_numpy(pos).cuda(),
                        torch.from_numpy(np.ascontiguousarray(idx).view(np.int32)).cuda()))
        else:
            xyz = rng.random((n, 9), dtype=np.float32) * 2.0 - 1.0
            out.append((torch.from_numpy(xyz).cuda(), None))
    return out


def test_build_batch_equals_alone(rt, O):
    """One persistent launch over meshes of 1..4096 triangles (indexed and
    not): every tree, leaf record, sorted index and code equals the fused
    build of that mesh alone; one also against the oracle."""
    meshes = _meshes(rt, O)
    for wg in (0, 3, rt.BUILD_BATCH_SHARED):
        bb = rt.BuildBatch(meshes, keep_order=True, workgroups=wg)
        trees = bb.launch()
        torch.cuda.synchronize()
        for (pos, idx), got in zip(meshes, trees):
            want = rt.GpuBVH4.build(pos, idx, algorithm="implicit", keep_order=True)
            torch.cuda.synchronize()
            n = got.tri_count
            assert [int(v) for v in got.info[:2].cpu()] == [want.node_count, want.depth]
            np.testing.assert_array_equal(to_np(got.nodes)[: want.node_count * 128],
                                          to_np(want.nodes)[: want.node_count * 128])
            np.testing.assert_array_equal(to_np(got.tris)[: n * 64], to_np(want.tris)[: n * 64])
            np.testing.assert_array_equal(u32(got.sorted_ids), u32(want.sorted_ids)[:n])
            np.testing.assert_array_equal(u32(got.codes), u32(want.codes)[:n])
    pos, idx = meshes[4]
    nodes_c, tris_c, ids_c, _ = O.lbvh_build(to_np(pos).reshape(-1, 9), None)
    np.testing.assert_array_equal(to_np(trees[4].nodes)[: len(nodes_c) * 128],
                                  nodes_c.view(np.uint8).reshape(-1))


def test_build_batch_graph_replays(rt, O):
    meshes = _meshes(rt, O)
    bb = rt.BuildBatch(meshes)
    trees = bb.launch()
    torch.cuda.synchronize()
    ref = [(t.nodes.clone(), t.tris.clone()) for t in trees]
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        bb.launch(stream=s)
    for rep in range(3):
        for t in trees:
            t.nodes.fill_(0)
            t.tris.fill_(0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for t, (rn, rt_) in zip(trees, ref):
            assert torch.equal(t.nodes, rn) and torch.equal(t.tris, rt_)


def _medium_meshes(rt, O):
    """Meshes above 4096 triangles (one workgroup each in the batch) mixed
    with small ones: indexed spheres, unindexed random triangles, a terrain
    grid with long runs of equal codes, and the 65536 maximum."""
    rng = np.random.default_rng(9)
    out = []
    for k, n in enumerate([65536, 4097, 5000, 16384, 16385, 30001, 100, 4096]):
        if k % 2 == 0:
            nu = max(2, int(np.sqrt(n / 2)))
            pos, idx = O.sphere(nu, max(2, (n + 2 * nu - 1) // (2 * nu)), 1.0, 0.02, k)
            idx = idx[:n]
            out.append((torch.from_numpy(pos).cuda(),
                        torch.from_numpy(np.ascontiguousarray(idx).view(np.int32)).cuda()))
        else:
            xyz = rng.random((n, 9), dtype=np.float32) * 2.0 - 1.0
            out.append((torch.from_numpy(xyz).cuda(), None))
    pos, idx = O.terrain(20_000)
    out.append((torch.from_numpy(pos).cuda(),
                torch.from_numpy(np.ascontiguousarray(idx).view(np.int32)).cuda()))
    return out


def test_build_batch_medium_equals_alone(rt, O):
    """Meshes of 4097..65536 triangles built by one workgroup each inside the
    persistent launch (and, with RT_BUILD_BATCH_SHARED, by 256-thread
    workgroups of a second launch, also mixed with fused-path meshes): nodes,
    leaf records, sorted ids and codes equal the multi-launch implicit build
    of the same mesh alone, and the oracle's."""
    meshes = _medium_meshes(rt, O)
    mixed = meshes + _meshes(rt, O)[:4]
    for wg, ms in ((2 | rt.BUILD_BATCH_SHARED, mixed), (rt.BUILD_BATCH_SHARED, meshes),
                   (2, meshes), (0, meshes)):
        bb = rt.BuildBatch(ms, keep_order=True, workgroups=wg)
        trees = bb.launch()
        torch.cuda.synchronize()
        for (pos, idx), got in zip(ms, trees):
            want = rt.GpuBVH4.build(pos, idx, algorithm="implicit", keep_order=True)
            torch.cuda.synchronize()
            n = got.tri_count
            assert [int(v) for v in got.info[:2].cpu()] == [want.node_count, want.depth]
            np.testing.assert_array_equal(to_np(got.nodes)[: want.node_count * 128],
                                          to_np(want.nodes)[: want.node_count * 128])
            np.testing.assert_array_equal(to_np(got.tris)[: n * 64], to_np(want.tris)[: n * 64])
            np.testing.assert_array_equal(u32(got.sorted_ids), u32(want.sorted_ids)[:n])
            np.testing.assert_array_equal(u32(got.codes), u32(want.codes)[:n])
    for k in (2, len(meshes) - 1):  # unindexed random, terrain ties
        pos, idx = meshes[k]
        p = to_np(pos)
        i = None if idx is None else to_np(idx).view(np.uint32)
        nodes_c, tris_c, ids_c, codes_c = O.lbvh_build(p.reshape(-1, 9) if i is None else p, i)
        np.testing.assert_array_equal(u32(trees[k].codes), codes_c)
        np.testing.assert_array_equal(u32(trees[k].sorted_ids), ids_c)
        np.testing.assert_array_equal(to_np(trees[k].nodes)[: len(nodes_c) * 128],
                                      nodes_c.view(np.uint8).reshape(-1))
        np.testing.assert_array_equal(to_np(trees[k].tris)[: len(tris_c) * 64],
                                      tris_c.view(np.uint8).reshape(-1))


def test_build_batch_medium_graph_replays(rt, O):
    meshes = _medium_meshes(rt, O)
    bb = rt.BuildBatch(meshes)
    trees = bb.launch()
    torch.cuda.synchronize()
    ref = [(t.nodes.clone(), t.tris.clone()) for t in trees]
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        bb.launch(stream=s)
    for rep in range(3):
        for t in trees:
            t.nodes.fill_(0)
            t.tris.fill_(0)
        bb.scratch.fill_(0xFF)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for t, (rn, rt_) in zip(trees, ref):
            assert torch.equal(t.nodes, rn) and torch.equal(t.tris, rt_)
