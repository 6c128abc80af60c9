"""BVH2 traversal of a SoftwareBVH on the device (rt_trace_bvh2, the loop of
benchmarks/exp0_probe/shaders/software_bvh_procedural.comp:47-171) against
the oracle's traversal of the same tree (or_trace_bvh2): hit ids, t bits,
any-hit bits and per-ray box / triangle test counts bit-identical.

Trees are the oracle's SoftwareBVH::build (software_bvh.hpp:35-175) of the
exp0 terrain (createTessellatedTerrain) and of a displaced sphere; rays are
the exp0 probe's pcg camera rays (coherent and incoherent) and random rays.
"""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((u32(b)[i >> 5] >> (i & 31)) & 1).astype(bool)


@pytest.fixture(scope="module")
def terrain2(O):
    pos, idx = O.terrain(100000, fused=True)
    n2, t48, order = O.bvh2_build(pos, idx, 4)
    return n2, t48, order


@pytest.mark.parametrize("coherent,tmin", [(True, 0.01), (False, 0.01), (False, 0.001)])
def test_bvh2_exp0_rays_match_oracle(rt, O, terrain2, coherent, tmin):
    """tmin 0.01: software_bvh_procedural.comp; 0.001: its LDS-stack variant
    software_bvh_lds.comp (the same loop, t > 0.001)."""
    n2, t48, order = terrain2
    g = rt.GpuBVH2.from_bvh2(n2, t48, order)
    assert 10 <= g.depth <= 64
    r = O.exp0_rays(1 << 16, coherent, tmin=tmin)
    rays = rt.Rays.from_numpy(r.packed())
    prim, t, box, tri = g.intersect(rays, counters=True)
    bits, abox, atri = g.occluded(rays, counters=True)
    torch.cuda.synchronize()
    pc, tc, bc, trc = O.trace_bvh2(n2, t48, r, order=order, counters=True)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    np.testing.assert_array_equal(u32(box), bc)
    np.testing.assert_array_equal(u32(tri), trc)
    pa, _, ba, tra = O.trace_bvh2(n2, t48, r, order=order, mode=O.ANY, counters=True)
    np.testing.assert_array_equal(_bits(bits, len(r)), pa != O.MISS)
    np.testing.assert_array_equal(u32(abox), ba)
    np.testing.assert_array_equal(u32(atri), tra)
    assert 0.5 < (pc != O.MISS).mean()


@pytest.mark.parametrize("block,lds", [(64, 4), (128, 12), (256, 32)])
def test_bvh2_shapes_and_positions(rt, O, block, lds):
    """Random rays through a displaced sphere under every launch shape; no
    order array reports leaf positions."""
    pos, idx = O.sphere(96, 64, 1.0, 0.05, 9)
    n2, t48, _ = O.bvh2_build(pos, idx, 4)
    g = rt.GpuBVH2.from_bvh2(n2, t48)
    r = O.random_rays(20000, 11, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
    rays = rt.Rays.from_numpy(r.packed())
    prim, t = g.intersect(rays, block=block, lds_stack=lds)
    bits = g.occluded(rays, block=block, lds_stack=lds)
    torch.cuda.synchronize()
    pc, tc = O.trace_bvh2(n2, t48, r)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    pa, _ = O.trace_bvh2(n2, t48, r, mode=O.ANY)
    np.testing.assert_array_equal(_bits(bits, len(r)), pa != O.MISS)


def test_bvh2_edge_cases(rt, O):
    """One triangle (a single-leaf tree), a ragged ray count, tmax cut-off."""
    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    n2, t48, order = O.bvh2_build(pos, idx, 4)
    assert len(n2) == 1
    g = rt.GpuBVH2.from_bvh2(n2, t48, order)
    r = O.Rays(37)
    r.ox[:], r.oy[:], r.oz[:] = np.linspace(-0.5, 0.5, 37), 0.0, -2.0
    r.dx[:], r.dy[:], r.dz[:] = 0.0, 0.0, 1.0
    r.tmin[:] = 0.0
    r.tmax[:] = np.where(np.arange(37) % 3 == 0, 1.5, 10.0)
    rays = rt.Rays.from_numpy(r.packed())
    prim, t = g.intersect(rays)
    torch.cuda.synchronize()
    pc, tc = O.trace_bvh2(n2, t48, r, order=order)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    assert (pc[np.arange(37) % 3 == 0] == O.MISS).all() and (pc[np.arange(37) % 3 != 0] == 0).all()
