"""The dynamic AO configuration at its own size, against the oracle:

  * scene: the static 999,698-triangle seeded random-displaced heightfield
    (its radix tree as the AO frame's tuner picks it: leaves of 1, 16 rounds
    of treelet restructuring, byte-identical to the oracle's) and the
    4,096-triangle deforming grid (generate_dynamic_grid_mesh layout,
    include/radv_bvh.hpp:224-283) rebuilt by the fused device build at time
    steps k = 0, 57 and 199 of the configuration's 200 (t_k = 0.05 k),
    byte-identical to the oracle's implicit build of the same vertices;
  * per step, the 1920x1080 frame with the bench's launch shapes: primary
    closest hits (ids, t, instance) identical to the oracle's on every pixel,
    1 SPP cosine AO rays (shaders/rq_ao.comp:32-97) within 2e-6 (device sin /
    cos against the C library's), and the AO image after the 7x7 bilateral
    filter (shaders/bilateral_blur.comp:28-49) within 1e-3 RMS of the
    oracle's frame (the stated bar for AO images; AO rays that differ in a
    last bit can flip a pixel's visibility).
"""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
EYE, LOOK, UP = (0.0, -6.0, 4.0), (0.0, 0.0, 0.2), (0.0, 0.0, 1.0)
M_GRID = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 4.6]], np.float32)


def _host_tree(O, b):
    return (to_np(b.nodes)[: b.node_count * 128].view(O.NODE4),
            to_np(b.tris)[: b.tri_count * 64].view(O.TRI64))


@pytest.fixture(scope="module")
def terrain(rt, O):
    tp, ti = rt.gen_heightfield(707, 707, 4.0, 0.3, 31)
    assert ti.shape[0] == 999_698
    print("AO terrain: 999,698 triangles, leaves of 1, 16 treelet rounds", flush=True)
    tb = rt.GpuBVH4.build(tp, ti, algorithm="radix", max_leaf_prims=1, treelet_rounds=16,
                          keep_order=True)
    torch.cuda.synchronize()
    return tp, ti, tb


def test_ao_terrain_tree_matches_oracle(rt, O, terrain):
    tp, ti, tb = terrain
    nodes_c, tris_c, ids_c, _ = O.rlbvh_build(to_np(tp), u32(ti).reshape(-1, 3), max_prims=1,
                                              treelet_rounds=16)
    assert tb.node_count == len(nodes_c) and tb.depth == O.bvh4_depth(nodes_c)
    np.testing.assert_array_equal(u32(tb.sorted_ids)[: len(ids_c)], ids_c)
    np.testing.assert_array_equal(to_np(tb.nodes)[: tb.node_count * 128],
                                  nodes_c.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(tb.tris)[: len(ids_c) * 64],
                                  tris_c.view(np.uint8).reshape(-1))


@pytest.mark.parametrize("k", [0, 57, 199])
def test_ao_frame_full_size(rt, O, terrain, k):
    tp, ti, tb = terrain
    # the grid at time step k, rebuilt in place by the fused build (as every frame)
    gx = rt.gen_dynamic_grid(4096, 0.0, 2.0)
    grid = rt.GpuBVH4.build(gx, None, algorithm="implicit")
    rt.gen_dynamic_grid(4096, 0.05 * k, 2.0, out=gx)
    cfg = rt.build_cfg("implicit")
    rt.call("rt_gpu_build", rt.stream_handle(), rt.ptr(gx), None, rt.u32(4096), rt.C.byref(cfg),
            rt.ptr(grid.nodes), rt.u32(grid.node_count), rt.ptr(grid.tris), rt.ptr(grid.info),
            None, None, None, rt.C.c_size_t(0))
    torch.cuda.synchronize()
    gxh = to_np(gx)
    gn, gt, _, _ = O.lbvh_build(gxh, None)
    np.testing.assert_array_equal(to_np(grid.nodes)[: len(gn) * 128], gn.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(grid.tris)[: 4096 * 64], gt.view(np.uint8).reshape(-1))
    scene = rt.Scene([(tb, tp, ti, None, rt.INST_IDENTITY),
                      (grid, gx, None, M_GRID, rt.INST_TRANSLATE)])
    print(f"AO frame k={k}: 1920x1080 primary + AO", flush=True)
    primary = rt.gen_camera_rays(W, H, EYE, LOOK, UP, 60.0, 0.0, 1e30)
    prim, t, inst = scene.intersect(primary, block=64, lds_stack=12, tile=(W, 8))
    prim_d, t_d, inst_d = scene.intersect(primary)
    ao_r = scene.gen_ao_rays(primary, prim, inst, t, W, H, frame=k, offset=0.03, tmin=0.01,
                             tmax=0.5)
    bits = scene.occluded(ao_r, block=64, lds_stack=16, tile=(W, 32))
    bits_d = scene.occluded(ao_r)
    ao = rt.ao_resolve(bits, prim)
    depth = torch.where(prim == -1, torch.full_like(t, 1e30), t)
    out = rt.blur7(ao, depth, W, H, 0.05)
    torch.cuda.synchronize()
    # launch shapes change nothing
    assert torch.equal(prim, prim_d) and torch.equal(inst, inst_d)
    assert torch.equal(t.view(torch.int32), t_d.view(torch.int32))
    assert torch.equal(bits, bits_d)
    # the oracle's frame
    orc = [(*_host_tree(O, tb), np.eye(3, 4, dtype=np.float32), rt.INST_IDENTITY),
           (*_host_tree(O, grid), M_GRID, rt.INST_TRANSLATE)]
    meshes = [(to_np(tp), u32(ti), np.eye(3, 4, dtype=np.float32), rt.INST_IDENTITY),
              (gxh, None, M_GRID, rt.INST_TRANSLATE)]
    pr = O.camera_rays(W, H, EYE, LOOK, UP, 60.0, 0.0, 1e30)
    np.testing.assert_array_equal(to_np(primary.data).view(np.uint32),
                                  pr.packed().view(np.uint32))
    pc, tc, ic = O.trace_scene(orc, pr)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    np.testing.assert_array_equal(u32(inst), ic)
    assert (ic == 1).sum() > 1000 and (ic == 0).sum() > 1000  # both instances in view
    ar = O.ao_rays_scene(meshes, pr, pc, ic, tc, W, H, frame=k, offset=0.03, tmin=0.01,
                         tmax=0.5)
    got = O.Rays.stack(to_np(ao_r.data))
    for f in O.Rays.FIELDS:
        np.testing.assert_allclose(getattr(got, f), getattr(ar, f), rtol=2e-6, atol=2e-6)
    occ = O.trace_scene(orc, ar, mode=O.ANY)
    on = np.nonzero(occ)[0]
    bits_c = np.zeros((W * H + 31) // 32, np.uint32)
    np.bitwise_or.at(bits_c, on >> 5, (1 << (on & 31)).astype(np.uint32))
    ao_c = O.ao_resolve(bits_c, pc)
    dc = np.where(pc == O.MISS, np.float32(1e30), tc).astype(np.float32)
    out_c = O.blur7(ao_c, dc, W, H, 0.05)
    rms = O.rms(to_np(out), out_c)
    print(f"AO frame k={k}: filtered AO RMS vs oracle {rms:.3g}", flush=True)
    assert rms <= 1e-3
    assert 0.05 < float(ao_c.mean()) < 0.99  # some pixels occluded, not all
