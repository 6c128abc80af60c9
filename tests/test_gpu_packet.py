"""Packet traversal (RT_SCHED_PACKET: a wave walks one node order for its 64
rays, each stack entry carrying the lanes whose own box tests reached it)
for occlusion bits: the bits equal the oracle's any hit on the same tree and
rays, and the single-ray kernel's, for a light's coherent shadow rays in wave
tiles and linear waves, on random (incoherent) rays, on trees with leaves of
1-4 triangles with and without treelet restructuring, with the workgroup
sizes; packets of partly valid rays (the grid's last row of tiles, n not a
multiple of 64) and empty trees; other queries are refused."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((u32(b)[i >> 5] >> (i & 31)) & 1).astype(bool)


W, H = 160, 120


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
def test_packet_bits_equal_oracle_on_shadow_rays(rt, O, leaf, rounds):
    bvh, shadow = _scene(rt, leaf, rounds)
    nodes, tris = _oracle_tree(O, bvh)
    r = O.Rays.stack(to_np(shadow.data))
    want = O.trace_bvh4(nodes, tris, r, mode=O.ANY)[0] != O.MISS
    assert 0.05 < want.mean() < 0.98
    single = _bits(bvh.occluded(shadow), len(r))
    np.testing.assert_array_equal(single, want)
    for tile in (None, (W, 16), (W, 8), (W, 32)):
        for block in (64, 256):
            print(f"packet leaf={leaf} rounds={rounds} tile={tile} block={block}", flush=True)
            b = bvh.occluded(shadow, schedule="packet", block=block, tile=tile)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(_bits(b, len(r)), want)


def test_packet_bits_on_random_rays_and_ragged_counts(rt, O):
    bvh, _ = _scene(rt, 2, 4)
    nodes, tris = _oracle_tree(O, bvh)
    for n in (50_000, 50_001, 63, 1):
        rnd = O.random_rays(n, 8 + n, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2), tmax=0.4)
        b = bvh.occluded(rt.Rays.from_numpy(rnd.packed()), schedule="packet", block=128)
        torch.cuda.synchronize()
        want = O.trace_bvh4(nodes, tris, rnd, mode=O.ANY)[0] != O.MISS
        np.testing.assert_array_equal(_bits(b, n), want)
        # the words' bits past n stay zero, as the single-ray kernel leaves them
        assert (u32(b)[n >> 5] >> (n & 31) == 0) if n % 32 else True


def test_packet_partial_tiles_and_empty_tree(rt, O):
    bvh, shadow = _scene(rt, 1, 0)
    nodes, tris = _oracle_tree(O, bvh)
    # 120 rows of one light: tiles of 4 / 8 / 2 rows; 121 rows leaves a partial tile row
    r0 = rt.Rays(shadow.data[:, : W * 121].contiguous())
    want = O.trace_bvh4(nodes, tris, O.Rays.stack(to_np(r0.data)), mode=O.ANY)[0] != O.MISS
    for tile in ((W, 16), (W, 8), (W, 32)):
        b = bvh.occluded(r0, schedule="packet", tile=tile, block=64)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_bits(b, W * 121), want)
    empty = rt.GpuBVH4(bvh.nodes, bvh.tris, 0, 0, 0)  # no nodes: every ray misses
    b = empty.occluded(r0, schedule="packet")
    torch.cuda.synchronize()
    assert not _bits(b, W * 121).any()


def test_packet_refuses_other_queries(rt):
    bvh, shadow = _scene(rt, 4, 0)
    with pytest.raises(rt.RtError):
        bvh.occluded(shadow, schedule="packet", counters=True)
    with pytest.raises(rt.RtError):
        bvh.intersect(shadow, schedule="packet")
    with pytest.raises(rt.RtError):
        bvh.occluded(shadow, schedule="packet", any_order="far")
    q = bvh.quantize()
    with pytest.raises(rt.RtError):
        q.occluded(shadow, schedule="packet")
