"""Any-hit child order (rt_trace_opts.any_order) for occlusion-bits-only
queries: near-first (the reference's order), far-first and slot order give
the same occlusion bits, equal to the oracle's, on shadow rays of a
displaced sphere (camera hits towards 16 lights, the frame's ray kinds) and
on random rays; launch shapes included."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((u32(b)[i >> 5] >> (i & 31)) & 1).astype(bool)


@pytest.fixture(scope="module")
def scene(rt, O):
    pos, idx = rt.gen_sphere(250, 200, 1.0, 0.02, 3)  # 100,000 triangles
    bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix")
    prim_rays = rt.gen_camera_rays(160, 120, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                   60.0, 0.0, 1e30)
    prim, t = bvh.intersect(prim_rays)
    lo, hi = pos.amin(0).cpu().double().numpy(), pos.amax(0).cpu().double().numpy()
    c, ext = (lo + hi) / 2, hi - lo
    lights = torch.tensor(rt.lights_in_box(c - 1.5 * ext, c + 1.5 * ext, 16, 1337),
                          dtype=torch.float32, device="cuda")
    shadow = rt.gen_shadow_rays(pos, idx, prim_rays, prim, t, lights, 1e-4)
    torch.cuda.synchronize()
    nodes = to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4)
    tris = to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64)
    return bvh, shadow, nodes, tris


@pytest.mark.parametrize("order", ["near", "far", "slot"])
@pytest.mark.parametrize("block,lds", [(64, 20), (256, 8)])
def test_orders_give_the_oracle_bits(rt, O, scene, order, block, lds):
    bvh, shadow, nodes, tris = scene
    r = O.Rays.stack(to_np(shadow.data))
    want = O.trace_bvh4(nodes, tris, r, mode=O.ANY)[0] != O.MISS
    assert 0.05 < want.mean() < 0.95
    bits = bvh.occluded(shadow, any_order=order, block=block, lds_stack=lds)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_bits(bits, len(r)), want)
    rnd = O.random_rays(50_000, 8, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2), tmax=0.4)
    b2 = bvh.occluded(rt.Rays.from_numpy(rnd.packed()), any_order=order, block=block,
                      lds_stack=lds)
    torch.cuda.synchronize()
    w2 = O.trace_bvh4(nodes, tris, rnd, mode=O.ANY)[0] != O.MISS
    np.testing.assert_array_equal(_bits(b2, len(rnd)), w2)


def test_other_queries_refuse_another_order(rt, scene):
    bvh, shadow, _, _ = scene
    with pytest.raises(rt.RtError):
        bvh.occluded(shadow, any_order="far", counters=True)
    with pytest.raises(rt.RtError):
        bvh.intersect(shadow, any_order="far")
    with pytest.raises(rt.RtError):
        bvh.occluded(shadow, any_order="slot", schedule="refill")
