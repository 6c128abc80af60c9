"""Software traversal of RADV images on the device (rt_trace_radv), the
consumer the reference's hardware ray queries are for (rayQuery over the
fast_bvh_builder image):

  * on the Morton-order image of a mesh, hit ids, t bits, any-hit bits and
    per-ray box / triangle counts equal rt_trace on the same tree built as
    Box4 nodes (rt_gpu_build RT_BUILD_IMPLICIT) and the oracle's traversal
    of the oracle's decoding of the image;
  * on the input-order image (the reference builder's leaves), the same
    against the oracle's decoding (hits report the input index);
  * an image with corrupted child ids (outside their section, another node
    type, a box child not after its parent, a box id off a 128-B node)
    traces as the same image with those slots empty: nothing is read outside
    the image.
"""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((u32(b)[i >> 5] >> (i & 31)) & 1).astype(bool)


def _rays(O, n, seed):
    return O.random_rays(n, seed, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)


def _check_vs_oracle(rt, O, img_np, rb, r, rays, **tune):
    prim, t, box, tri = rb.intersect(rays, counters=True, **tune)
    bits, abox, atri = rb.occluded(rays, counters=True, **tune)
    torch.cuda.synchronize()
    nodes, tris = O.radv_decode(img_np)
    pc, tc, bc, trc = O.trace_bvh4(nodes, tris, r, counters=True)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    np.testing.assert_array_equal(u32(box), bc)
    np.testing.assert_array_equal(u32(tri), trc)
    pa, _, ba, tra = O.trace_bvh4(nodes, tris, r, mode=O.ANY, counters=True)
    np.testing.assert_array_equal(_bits(bits, len(r)), pa != O.MISS)
    np.testing.assert_array_equal(u32(abox), ba)
    np.testing.assert_array_equal(u32(atri), tra)
    return prim, t, box, tri, bits


@pytest.mark.parametrize("nu,nv", [(1, 1), (10, 7), (64, 32)])
def test_radv_morton_image_equals_box4_tree(rt, O, nu, nv):
    pos, idx = O.sphere(nu, nv, 1.0, 0.05, 5)
    dpos = torch.from_numpy(pos).cuda()
    didx = torch.from_numpy(idx.view(np.int32)).cuda()
    img, _ = rt.build_radv(dpos, didx, order="morton")
    bvh = rt.GpuBVH4.build(dpos, didx, algorithm="implicit")
    torch.cuda.synchronize()
    rb = rt.RadvBVH(img)
    assert rb.info["box_count"] == bvh.node_count and rb.depth == bvh.depth
    r = _rays(O, 20000, nu + nv)
    rays = rt.Rays.from_numpy(r.packed())
    prim, t, box, tri, bits = _check_vs_oracle(rt, O, to_np(img), rb, r, rays)
    p4, t4, b4, tr4 = bvh.intersect(rays, counters=True)
    bits4 = bvh.occluded(rays)
    torch.cuda.synchronize()
    assert torch.equal(prim, p4) and torch.equal(t.view(torch.int32), t4.view(torch.int32))
    assert torch.equal(box, b4) and torch.equal(tri, tr4) and torch.equal(bits, bits4)


@pytest.mark.parametrize("block,lds", [(64, 4), (256, 20)])
def test_radv_input_order_image(rt, O, block, lds):
    xyz, _ = O.dynamic_grid(4096, 0.3, 1.25)
    img, ids = rt.build_radv(torch.from_numpy(xyz).cuda(), None, order="input")
    torch.cuda.synchronize()
    rb = rt.RadvBVH(img)
    r = O.random_rays(30000, 3, (-1.25, -1.25, 4.0), (1.25, 1.25, 6.0), tmax=1e30)
    # half the rays start above the grid (z <= 5.2) and aim down, half are random
    r.oz[:15000] = 7.0
    r.dx[:15000], r.dy[:15000], r.dz[:15000] = 0.0, 0.0, -1.0
    rays = rt.Rays.from_numpy(r.packed())
    prim = _check_vs_oracle(rt, O, to_np(img), rb, r, rays, block=block, lds_stack=lds)[0]
    p = u32(prim)
    assert (p[:15000] != O.MISS).mean() > 0.9 and (p[p != O.MISS] < 4096).all()


def test_radv_corrupted_ids_are_skipped(rt, O):
    pos, idx = O.sphere(40, 30, 1.0, 0.05, 6)
    xyz9 = pos[idx.astype(np.int64)].reshape(-1, 9)
    img = O.radv_encode(xyz9)
    _, _, tb, _ = O.tree_layout(len(xyz9))
    bvh = 192
    ids = img[bvh: bvh + 128 * tb].view(np.uint32).reshape(tb, 32)[:, :4]
    rng = np.random.default_rng(1)
    bad = img.copy()
    clean = img.copy()
    bad_ids = bad[bvh: bvh + 128 * tb].view(np.uint32).reshape(tb, 32)
    clean_ids = clean[bvh: bvh + 128 * tb].view(np.uint32).reshape(tb, 32)
    tbase = 128 * tb
    patches = [0x7FFFFFF5,               # box32 far past the box section
               (0x80 >> 3) | 4,          # box16: not followed
               (tbase >> 3) | 5,         # box32 id on a triangle node
               ((tbase + 8) >> 3) | 0,   # low bits 1: node type 1, not a RADV type
               (0 >> 3) | 5,             # box32 child = the root: before its parent
               (0x40 >> 3) | 5,          # box32 id off a 128-B node
               ((tbase + 64 * len(xyz9)) >> 3) | 0]  # triangle past the section
    nodes = rng.choice(tb, size=24, replace=False)
    for k, ni in enumerate(nodes):
        slot = int(rng.integers(0, 4))
        if ids[ni, slot] == 0xFFFFFFFF:
            continue
        bad_ids[ni, slot] = patches[k % len(patches)]
        clean_ids[ni, slot] = 0xFFFFFFFF
    dimg = torch.from_numpy(bad).cuda()
    rb = rt.RadvBVH(dimg)
    r = _rays(O, 20000, 9)
    rays = rt.Rays.from_numpy(r.packed())
    _check_vs_oracle(rt, O, clean, rb, r, rays)
