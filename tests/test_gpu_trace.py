"""GPU traversal parity: rt_trace (HIP) against the CPU oracle on the same
tree and the same rays, for every schedule (one ray per thread; waves that
walk ray chunks and refill finished lanes, at the extreme chunk sizes and
refill thresholds).  Closest-hit triangle IDs, t and box/triangle counters
must be bit-identical (same traversal order and predicates as
software_bvh_bvh4.comp:96-190); any-hit occlusion bits and counters
identical (slot-order any-hit on both sides)."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[dict(schedule="simple"),
                        dict(schedule="refill"),
                        dict(schedule="refill", rays_per_wave=64, refill_lanes=1),
                        dict(schedule="refill", rays_per_wave=4096, refill_lanes=64)],
                ids=["simple", "refill", "refill-64-1", "refill-4096-64"])
def sched(request):
    return request.param


def _upload_tree(rt, O, pos, idx):
    n4, tris = O.software_bvh4(pos, idx)
    return n4, tris, rt.GpuBVH4.from_bvh4(n4, tris64=tris, depth=O.bvh4_depth(n4))


def _rays_to_dev(rt, r):
    return rt.Rays.from_numpy(r.packed())


def _bits(b, n):
    i = np.arange(n)
    return ((b[i >> 5] >> (i & 31)) & 1).astype(bool)


def _cmp_closest(rt, O, n4, tris, bvh, r, sched):
    prim_c, t_c, box_c, tri_c = O.trace_bvh4(n4, tris, r, counters=True)
    prim_g, t_g, box_g, tri_g = bvh.intersect(_rays_to_dev(rt, r), counters=True,
                                              **sched)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(u32(prim_g), prim_c)
    np.testing.assert_array_equal(to_np(t_g).view(np.uint32), t_c.view(np.uint32))
    np.testing.assert_array_equal(u32(box_g), box_c)
    np.testing.assert_array_equal(u32(tri_g), tri_c)
    return prim_c


def _cmp_any(rt, O, n4, tris, bvh, r, sched):
    prim_c, _, box_c, tri_c = O.trace_bvh4(n4, tris, r, mode=O.ANY, counters=True)
    bits, box_g, tri_g = bvh.occluded(_rays_to_dev(rt, r), counters=True, **sched)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_bits(u32(bits), len(r)), prim_c != O.MISS)
    np.testing.assert_array_equal(u32(box_g), box_c)
    np.testing.assert_array_equal(u32(tri_g), tri_c)
    return prim_c


def test_plain_small_grid_closest(rt, O, sched):
    """Small plain method: 64x64-quad wave grid (8,192 triangles,
    z = 0.1 sin(4x) cos(4y)), software_bvh build, 256x256 pinhole camera
    (60 deg) looking down -z; per-pixel ID and t equal the CPU path."""
    pos, idx = O.wave_grid(64, 64, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, tris, bvh = _upload_tree(rt, O, pos, idx)
    r = O.camera_rays(256, 256, (0.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 0.0, 1e30)
    prim = _cmp_closest(rt, O, n4, tris, bvh, r, sched)
    assert (prim != O.MISS).mean() > 0.5


def test_any_hit_bits_match(rt, O, sched):
    pos, idx = O.wave_grid(64, 64, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, tris, bvh = _upload_tree(rt, O, pos, idx)
    r = O.random_rays(100_003, 5, (-1.2, -1.2, -0.3), (1.2, 1.2, 0.3), tmax=0.4)
    prim_a = _cmp_any(rt, O, n4, tris, bvh, r, sched)
    # any-hit agrees with closest-hit t < tmax on identical rays
    pc, tc = O.trace_bvh4(n4, tris, r)
    np.testing.assert_array_equal(pc != O.MISS, prim_a != O.MISS)


def test_device_built_tree_large(rt, O, sched):
    """Device radix build of a 200k-triangle displaced sphere, random rays
    from inside and outside: both modes equal the oracle on the same tree."""
    pos, idx = rt.gen_sphere(400, 250, 1.0, 0.02, 3)
    bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix")
    torch.cuda.synchronize()
    n4 = to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4)
    tris = to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64)
    r = O.random_rays(120_000, 21, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
    _cmp_closest(rt, O, n4, tris, bvh, r, sched)
    r.tmax[:] = 0.25
    _cmp_any(rt, O, n4, tris, bvh, r, sched)


def test_known_answer_probe_triangle(rt, O, sched):
    """probe_main.cpp:45-50,182-211: triangle (-1,-1,0),(1,-1,0),(0,1,0); rays
    from (0,-0.5,-2) along +z hit at t = 2 (+-0.01) with prim 0; rays from
    (10,10,-2) miss."""
    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    n4, tris, bvh = _upload_tree(rt, O, pos, idx)
    r = O.Rays(1024)
    r.ox[:512], r.oy[:512], r.oz[:512] = 0.0, -0.5, -2.0
    r.ox[512:], r.oy[512:], r.oz[512:] = 10.0, 10.0, -2.0
    r.dz[:] = 1.0
    r.tmin[:] = 0.001
    r.tmax[:] = 100.0
    prim, t = bvh.intersect(_rays_to_dev(rt, r), **sched)
    torch.cuda.synchronize()
    prim, t = u32(prim), to_np(t)
    assert (prim[:512] == 0).all() and (np.abs(t[:512] - 2.0) < 0.01).all()
    assert (prim[512:] == O.MISS).all()


def test_known_answer_forged_triangle(rt, O, sched):
    """src/main.cpp:269-313,410-414,502-507: triangle (-2,-2,5),(2,-2,5),(0,2,5);
    the ray from the origin along +z hits at t = 5 (+-1e-3); (10,10,0) misses."""
    pos = np.array([[-2, -2, 5], [2, -2, 5], [0, 2, 5]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    n4, tris, bvh = _upload_tree(rt, O, pos, idx)
    r = O.Rays(2)
    r.ox[1], r.oy[1] = 10.0, 10.0
    r.dz[:] = 1.0
    r.tmin[:] = 0.001
    r.tmax[:] = 100.0
    prim, t = bvh.intersect(_rays_to_dev(rt, r), **sched)
    torch.cuda.synchronize()
    prim, t = u32(prim), to_np(t)
    assert prim[0] == 0 and abs(t[0] - 5.0) < 1e-3 and prim[1] == O.MISS


@pytest.mark.parametrize("target", [1000, 100000])
def test_exp0_stats_against_published(rt, O, target, sched):
    """The reference's own published traversal statistics
    (results/exp0_ray_stats.csv:2-5) reproduced by the HIP kernel on the
    reference terrain (as built by the reference's -march=native build) with
    the reference's procedural rays: mean box tests, triangle tests and hit
    rate within float noise of the GPU the CSV came from.  The rays are made
    on the device (rt_gen_exp0_rays) and equal the oracle's restatement of
    the shader's generator bit for bit."""
    from test_oracle_known import EXP0
    pos, idx = O.terrain(target, fused=True)
    n4, tris, bvh = _upload_tree(rt, O, pos, idx)
    for coherent in (True, False):
        r = O.exp0_rays(1 << 20, coherent)
        dr = rt.gen_exp0_rays(1 << 20, coherent)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(u32(dr.data), r.packed().view(np.uint32))
        prim, t, box, tri = bvh.intersect(dr, counters=True, **sched)
        torch.cuda.synchronize()
        prim, box, tri = u32(prim), u32(box), u32(tri)
        hit = prim != O.MISS
        ref = EXP0[(len(idx), "Coherent" if coherent else "Incoherent")]
        assert abs(hit.mean() * 100 - ref["hit"]) < 2e-3
        assert abs(box.mean() - ref["box"]) / ref["box"] < 2e-3
        assert abs(tri.mean() - ref["tri"]) / ref["tri"] < 2e-3
        # and exactly the oracle's counts (unbounded stack on both sides)
        pc, tc, bc, trc = O.trace_bvh4(n4, tris, r, counters=True)
        np.testing.assert_array_equal(prim, pc)
        np.testing.assert_array_equal(box, bc)


@pytest.mark.parametrize("n", [1, 31, 33, 64, 65, 1000, 4097])
def test_empty_and_ragged_ray_counts(rt, O, n, sched):
    p2, i2 = O.wave_grid(8, 8)
    n4, tris, bvh = _upload_tree(rt, O, p2, i2)
    r = O.random_rays(n, 3, (-1, -1, -0.2), (1, 1, 0.2), tmax=0.5)
    _cmp_closest(rt, O, n4, tris, bvh, r, sched)
    _cmp_any(rt, O, n4, tris, bvh, r, sched)


def test_zero_rays_and_empty_tree(rt, O, sched):
    p2, i2 = O.wave_grid(2, 2)
    _, _, bvh = _upload_tree(rt, O, p2, i2)
    empty = rt.Rays.empty(0)
    prim, t = bvh.intersect(empty, **sched)
    assert prim.numel() == 0
    # an empty tree: every ray misses, t = tmax
    tree0 = rt.GpuBVH4(bvh.nodes, bvh.tris, 0, 0, 0)
    r = O.random_rays(100, 1, (-1, -1, -1), (1, 1, 1), tmax=3.0)
    prim, t = tree0.intersect(_rays_to_dev(rt, r), **sched)
    bits = tree0.occluded(_rays_to_dev(rt, r), **sched)
    torch.cuda.synchronize()
    assert (u32(prim) == O.MISS).all() and (to_np(t) == 3.0).all()
    assert not _bits(u32(bits), 100).any()


@pytest.mark.parametrize("coherent", [True, False])
@pytest.mark.parametrize("n", [1, 1000, 1024, 1025, 70000])
def test_exp0_rays_device_equal_oracle(rt, O, coherent, n):
    """rt_gen_exp0_rays (software_bvh_bvh4.comp:50-94) bit-identical to the
    oracle's generator, ragged counts (the coherent grid's height is
    ceil(n / 1024)), another camera, tmin / tmax filled."""
    for cam, look in (((0.0, 50.0, -45.0), (0.0, 0.0, 5.0)), ((3.0, 7.0, 2.0), (-1.0, 0.5, 9.0))):
        r = O.exp0_rays(n, coherent, cam=cam, look=look, tmin=0.02, tmax=77.0)
        d = rt.gen_exp0_rays(n, coherent, cam=cam, look=look, tmin=0.02, tmax=77.0)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(u32(d.data), r.packed().view(np.uint32))
