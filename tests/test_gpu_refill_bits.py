"""The refill schedule for occlusion bits (RT_SCHED_REFILL, any hit without
counters or hit ids: finished lanes take the wave's next rays, with or
without wave tiles): the bits equal the oracle's any hit on the same tree
and rays, and the one-ray-per-thread kernel's, for a light's shadow rays in
wave tiles of 8, 16 and 32 and in linear chunks, on trees with leaves of 1-4
triangles with and without treelet restructuring, over workgroup sizes,
LDS stack depths down to 4 entries (the stack spilling to private memory),
chunks of 64-4096 rays and refill thresholds of 1-64 lanes; on random rays
of ragged counts; on a grid whose last tile row is partial (the padding of
the last word zero) and on an empty tree."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((u32(b)[i >> 5] >> (i & 31)) & 1).astype(bool)


W, H = 160, 120

SHAPES = [dict(block=64, lds_stack=20, rays_per_wave=256, refill_lanes=8),
          dict(block=256, lds_stack=12, rays_per_wave=1024, refill_lanes=16),
          dict(block=128, lds_stack=4, rays_per_wave=64, refill_lanes=1),
          dict(block=64, lds_stack=16, rays_per_wave=4096, refill_lanes=64)]


def _scene(rt, leaf, rounds):
    pos, idx = rt.gen_sphere(250, 200, 1.0, 0.02, 3)  # 100,000 triangles
    bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix", max_leaf_prims=leaf,
                           treelet_rounds=rounds)
    prim_rays = rt.gen_camera_rays(W, H, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                   60.0, 0.0, 1e30)
    prim, t = bvh.intersect(prim_rays)
    lo, hi = pos.amin(0).cpu().double().numpy(), pos.amax(0).cpu().double().numpy()
    c, ext = (lo + hi) / 2, hi - lo
    lights = torch.tensor(rt.lights_in_box(c - 1.5 * ext, c + 1.5 * ext, 16, 1337),
                          dtype=torch.float32, device="cuda")
    shadow = rt.gen_shadow_rays(pos, idx, prim_rays, prim, t, lights, 1e-4)
    torch.cuda.synchronize()
    return bvh, shadow


def _oracle_tree(O, bvh):
    return (to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4),
            to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64))


@pytest.mark.parametrize("leaf,rounds", [(1, 0), (1, 8), (2, 8), (4, 0)])
def test_refill_bits_equal_oracle_on_shadow_rays(rt, O, leaf, rounds):
    bvh, shadow = _scene(rt, leaf, rounds)
    nodes, tris = _oracle_tree(O, bvh)
    r = O.Rays.stack(to_np(shadow.data))
    want = O.trace_bvh4(nodes, tris, r, mode=O.ANY)[0] != O.MISS
    assert 0.05 < want.mean() < 0.98
    single = bvh.occluded(shadow)
    np.testing.assert_array_equal(_bits(single, len(r)), want)
    for tile in (None, (W, 16), (W, 8), (W, 32)):
        for shape in SHAPES:
            print(f"refill bits leaf={leaf} rounds={rounds} tile={tile} {shape}", flush=True)
            b = torch.full_like(single, -1)
            bvh.occluded(shadow, out=b, schedule="refill", tile=tile, **shape)
            torch.cuda.synchronize()
            # whole words equal: ray bits and the zero padding past n
            assert torch.equal(b, single)


def test_refill_bits_on_random_rays_and_ragged_counts(rt, O):
    bvh, _ = _scene(rt, 2, 4)
    nodes, tris = _oracle_tree(O, bvh)
    for n in (50_000, 50_001, 63, 1):
        print(f"refill bits random rays n={n}", flush=True)
        rnd = O.random_rays(n, 8 + n, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2), tmax=0.4)
        rays = rt.Rays.from_numpy(rnd.packed())
        want = O.trace_bvh4(nodes, tris, rnd, mode=O.ANY)[0] != O.MISS
        for shape in SHAPES[:2]:
            b = torch.full((max(1, (n + 31) // 32),), -1, dtype=torch.int32, device="cuda")
            bvh.occluded(rays, out=b, schedule="refill", **shape)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(_bits(b, n), want)
            assert (u32(b)[n >> 5] >> (n & 31) == 0) if n % 32 else True


def test_refill_bits_partial_tiles_and_empty_tree(rt, O):
    bvh, shadow = _scene(rt, 1, 0)
    nodes, tris = _oracle_tree(O, bvh)
    # 121 rows: tiles of 4 / 8 / 2 rows leave a partial tile row; W * 121 is
    # not a multiple of 64 rays
    r0 = rt.Rays(shadow.data[:, : W * 121].contiguous())
    n = W * 121
    want = O.trace_bvh4(nodes, tris, O.Rays.stack(to_np(r0.data)), mode=O.ANY)[0] != O.MISS
    ref = bvh.occluded(r0)
    for tile in ((W, 16), (W, 8), (W, 32)):
        for shape in SHAPES[:3]:
            print(f"refill bits partial tiles tile={tile} {shape}", flush=True)
            b = torch.full_like(ref, -1)
            bvh.occluded(r0, out=b, schedule="refill", tile=tile, **shape)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(_bits(b, n), want)
            assert torch.equal(b, ref)
    empty = rt.GpuBVH4(bvh.nodes, bvh.tris, 0, 0, 0)  # no nodes: every ray misses
    b = torch.full_like(ref, -1)
    empty.occluded(r0, out=b, schedule="refill", tile=(W, 16), block=64, rays_per_wave=256)
    torch.cuda.synchronize()
    assert not u32(b).any()
