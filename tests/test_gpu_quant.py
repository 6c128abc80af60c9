"""RT_NODES_Q8 trees on the GPU (rt_gpu_quantize + the quantized traversal)
against the oracle: the 64-B nodes byte-identical to or_bvh4_quantize, and a
trace of the quantized tree bit-identical (ids, t, occlusion bits, per-ray
box / triangle counts) to the oracle's trace of the rt_node4 tree with the
decoded boxes; occlusion bits and closest-hit t equal to the exact tree's."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((b[i >> 5] >> (i & 31)) & 1).astype(bool)


def _trees(rt, O, pos, idx, **kw):
    dpos = torch.from_numpy(np.ascontiguousarray(pos, np.float32)).cuda()
    didx = None if idx is None else torch.from_numpy(
        np.ascontiguousarray(idx, np.uint32).view(np.int32)).cuda()
    bvh = rt.GpuBVH4.build(dpos, didx, **kw)
    q = bvh.quantize()
    torch.cuda.synchronize()
    nodes = to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4)
    tris = to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64)
    qc = O.bvh4_quantize(nodes)
    np.testing.assert_array_equal(to_np(q.nodes)[: q.node_count * 64], qc.view(np.uint8))
    return bvh, q, nodes, O.bvh4_dequantize(qc), tris


CASES = [dict(algorithm="radix", max_leaf_prims=1, treelet_rounds=2),
         dict(algorithm="radix", max_leaf_prims=4), dict(algorithm="implicit")]


@pytest.mark.parametrize("kw", CASES, ids=["radix-leaf1-r2", "radix-leaf4", "implicit"])
@pytest.mark.parametrize("tile", [None, (480, 16)])
def test_quantized_trace_matches_oracle(rt, O, kw, tile):
    pos, idx = O.sphere(120, 90, 1.0, 0.05, 3)
    bvh, q, nodes, dq, tris = _trees(rt, O, pos, idx, **kw)
    r = O.random_rays(480 * 200, 5, (-1.3,) * 3, (1.3,) * 3, tmax=1e30)
    rays = rt.Rays.from_numpy(r.packed())
    shape = dict(tile=tile, block=64) if tile else {}
    prim, t, box, tri = q.intersect(rays, counters=True, **shape)
    torch.cuda.synchronize()
    pc, tc, bc, trc = O.trace_bvh4(dq, tris, r, counters=True)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    np.testing.assert_array_equal(u32(box), bc)
    np.testing.assert_array_equal(u32(tri), trc)
    # the exact tree: the same t everywhere (ids equal but at exact ties)
    pe, te = O.trace_bvh4(nodes, tris, r)
    np.testing.assert_array_equal(te.view(np.uint32), tc.view(np.uint32))
    assert (pe == pc).mean() > 0.9999
    r.tmax[:] = 0.25
    rays = rt.Rays.from_numpy(r.packed())
    bits = q.occluded(rays, **shape)  # bits only: the inner-order kernel
    bits_c, box_a, tri_a = q.occluded(rays, counters=True, **shape)
    torch.cuda.synchronize()
    pa, _, ba, tra = O.trace_bvh4(dq, tris, r, mode=O.ANY, counters=True)
    np.testing.assert_array_equal(_bits(u32(bits), len(r)), pa != O.MISS)
    np.testing.assert_array_equal(_bits(u32(bits_c), len(r)), pa != O.MISS)
    np.testing.assert_array_equal(u32(box_a), ba)
    np.testing.assert_array_equal(u32(tri_a), tra)
    pe, _ = O.trace_bvh4(nodes, tris, r, mode=O.ANY)
    np.testing.assert_array_equal(pa != O.MISS, pe != O.MISS)


def test_quantize_uploaded_softwarebvh4_and_soup(rt, O):
    pos, idx = O.terrain(20_000)
    n4, t64 = O.software_bvh4(pos, idx)
    up = rt.GpuBVH4.from_bvh4(n4, tris64=t64)
    q = up.quantize()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(q.nodes)[: q.node_count * 64],
                                  O.bvh4_quantize(n4).view(np.uint8))
    xyz, _ = O.dynamic_grid(4096, 0.4, 2.0)
    _trees(rt, O, xyz, None, algorithm="radix", max_leaf_prims=1)


def test_quantized_tree_rejected_where_unsupported(rt, O):
    pos, idx = O.sphere(20, 10, 1.0, 0.05, 3)
    _, q, _, _, _ = _trees(rt, O, pos, idx, algorithm="radix")
    r = rt.Rays.from_numpy(O.random_rays(100, 1, (-1,) * 3, (1,) * 3).packed())
    with pytest.raises(rt.RtError):
        q.intersect(r, schedule="refill")
    with pytest.raises(rt.RtError):
        rt.TraceBatch([(q, r, "closest", None)])
    with pytest.raises(ValueError):
        rt.Scene([(q, None, None, None, rt.INST_IDENTITY)])
