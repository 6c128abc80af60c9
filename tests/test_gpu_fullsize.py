"""Parity at the benchmark's full sizes, through properties that do not
depend on size and through sampled rays against the oracle:

  * the 2.5M-triangle displaced sphere of the shadow configuration, radix
    build: Morton codes non-decreasing, sorted ids a permutation, every leaf
    record holding its triangle's vertices and id, every inner slot's box
    containing the child's occupied slots, the root containing every vertex;
  * 200,000 rays of the frame (camera rays and their shadow rays towards the
    16 lights) traced on that tree: closest-hit ids / t and any-hit bits
    bit-identical to the oracle's traversal of the same tree;
  * the headline frame's own tree (leaves of 1, 8 rounds of treelet
    restructuring) byte-identical to the oracle's, and the frame traced with
    the bench's launch shapes (16 x 4 wave tiles) equal to the oracle on
    sampled rays;
  * the 65,536-triangle maximum of the batched build equal to the
    multi-launch build of the same mesh.
"""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu

NU, NV, W, H, NL = 1250, 1000, 1920, 1080, 16


@pytest.fixture(scope="module", params=[4, 2, 1], ids=["leaf4", "leaf2", "leaf1"])
def sphere(rt, request):
    """The frame's sphere and its radix tree, with leaves of at most 4 (the
    default), 2 or 1 triangles (leaf sizes the frame's tuner may pick)."""
    pos, idx = rt.gen_sphere(NU, NV, 1.0, 0.02, 7)
    bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix", keep_order=True,
                           max_leaf_prims=request.param)
    torch.cuda.synchronize()
    return pos, idx, bvh


def test_full_size_build_properties(rt, sphere):
    pos, idx, bvh = sphere
    n = idx.shape[0]
    assert n == 2 * NU * NV
    codes = bvh.codes[:n].to(torch.int64)
    assert bool((codes[1:] >= codes[:-1]).all()), "codes sorted"
    ids = bvh.sorted_ids[:n].to(torch.int64)
    assert torch.equal(torch.sort(ids).values, torch.arange(n, device=ids.device)), "permutation"
    rec_i = bvh.tris[: n * 64].view(torch.int32).view(n, 16)
    rec_f = bvh.tris[: n * 64].view(torch.float32).view(n, 16)
    assert torch.equal(rec_i[:, 12].to(torch.int64), ids), "record ids in sorted order"
    verts = pos[idx.to(torch.int64)[ids]].reshape(n, 9)
    assert torch.equal(rec_f[:, :9], verts), "record vertices"
    nc = bvh.node_count
    raw = bvh.nodes[: nc * 128]
    node_i = raw.view(torch.int32).view(nc, 32)
    b = raw.view(torch.float32).view(nc, 32)[:, :24].reshape(nc, 6, 4)  # minX maxX ... maxZ
    child, count = node_i[:, 24:28].to(torch.int64), node_i[:, 28:32]
    inner = count < 0
    pi, pk = inner.nonzero(as_tuple=True)
    c = child[pi, pk]
    occ = (count[c] != 0)
    big = torch.tensor(float("inf"), device=b.device)
    cmin = torch.where(occ.unsqueeze(1), b[c][:, [0, 2, 4], :], big).amin(2)
    cmax = torch.where(occ.unsqueeze(1), b[c][:, [1, 3, 5], :], -big).amax(2)
    pmin = b[pi, :, pk][:, [0, 2, 4]]
    pmax = b[pi, :, pk][:, [1, 3, 5]]
    assert bool((pmin <= cmin).all() and (pmax >= cmax).all()), "slot boxes contain children"
    root_occ = count[0] != 0
    rmin = torch.where(root_occ, b[0][[0, 2, 4]], big).amin(1)
    rmax = torch.where(root_occ, b[0][[1, 3, 5]], -big).amax(1)
    assert bool((pos.amin(0) >= rmin).all() and (pos.amax(0) <= rmax).all()), "root bounds"


def test_full_size_frame_rays_match_oracle(rt, O, sphere):
    pos, idx, bvh = sphere
    prim_rays = rt.gen_camera_rays(W, H, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                   60.0, 0.0, 1e30)
    prim, t = bvh.intersect(prim_rays)
    lo, hi = pos.amin(0).cpu().double().numpy(), pos.amax(0).cpu().double().numpy()
    c, ext = (lo + hi) / 2, hi - lo
    lights = torch.tensor(rt.lights_in_box(c - 1.5 * ext, c + 1.5 * ext, NL, 1337),
                          dtype=torch.float32, device="cuda")
    shadow = rt.gen_shadow_rays(pos, idx, prim_rays, prim, t, lights, 1e-4)
    torch.cuda.synchronize()
    nodes = to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4)
    tris = to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64)
    rng = np.random.default_rng(3)
    # closest hit: 100,000 camera rays spread over the frame
    sel = np.sort(rng.choice(W * H, 100_000, replace=False))
    r = O.Rays.stack(to_np(prim_rays.data)[:, sel])
    pc, tc = O.trace_bvh4(nodes, tris, r)
    np.testing.assert_array_equal(u32(prim)[sel], pc)
    np.testing.assert_array_equal(to_np(t)[sel].view(np.uint32), tc.view(np.uint32))
    # any hit: 100,000 shadow rays spread over all 16 lights, gathered into a
    # ray set of their own and traced alone
    sel = np.sort(rng.choice(W * H * NL, 100_000, replace=False))
    sh = rt.Rays(shadow.data[:, torch.from_numpy(sel).cuda()].contiguous())
    bits = bvh.occluded(sh)
    torch.cuda.synchronize()
    ra = O.Rays.stack(to_np(sh.data))
    pa, _ = O.trace_bvh4(nodes, tris, ra, mode=O.ANY)
    b = u32(bits)
    i = np.arange(len(sel))
    np.testing.assert_array_equal(((b[i >> 5] >> (i & 31)) & 1).astype(bool), pa != O.MISS)
    assert 0.05 < (pa != O.MISS).mean() < 0.95  # the sample holds both outcomes


@pytest.fixture(scope="module")
def headline(rt):
    """The headline frame's tree as the bench's tuner picks it: the 2.5M
    sphere's radix tree with leaves of 1 triangle and 8 rounds of treelet
    restructuring."""
    pos, idx = rt.gen_sphere(NU, NV, 1.0, 0.02, 7)
    print("headline tree: 2.5M sphere, leaves of 1, 8 treelet rounds", flush=True)
    bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix", keep_order=True, max_leaf_prims=1,
                           treelet_rounds=8)
    torch.cuda.synchronize()
    return pos, idx, bvh


def test_headline_tree_bytes_match_oracle(rt, O, headline):
    """Every node, leaf record and the sorted order of the headline tree
    byte-identical to the oracle's restructured radix tree of the same
    device-generated mesh."""
    pos, idx, bvh = headline
    hp, hi = to_np(pos), u32(idx).reshape(-1, 3)
    nodes_c, tris_c, ids_c, _ = O.rlbvh_build(hp, hi, max_prims=1, treelet_rounds=8)
    n = hi.shape[0]
    assert bvh.node_count == len(nodes_c)
    assert bvh.depth == O.bvh4_depth(nodes_c)
    np.testing.assert_array_equal(u32(bvh.sorted_ids)[:n], ids_c)
    np.testing.assert_array_equal(to_np(bvh.nodes)[: bvh.node_count * 128],
                                  nodes_c.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(bvh.tris)[: n * 64], tris_c.view(np.uint8).reshape(-1))


def test_headline_frame_with_bench_shapes_matches_oracle(rt, O, headline):
    """The whole frame traced as the bench traces it -- primary rays with
    64-thread blocks (linear waves and 16 x 4 tiles), the 33.2M shadow rays
    with 64-thread blocks, 20 LDS entries and 16 x 4 wave tiles of the
    1920-wide grid, and with the pixels' primary hits as occluder hints --
    gives the same ids, t and occlusion bits as the default linear launch;
    100,000 sampled camera rays and 100,000 sampled shadow rays of that frame
    equal the oracle's traversal of the same tree (ids, t bits, occlusion
    bits)."""
    pos, idx, bvh = headline
    prim_rays = rt.gen_camera_rays(W, H, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                   60.0, 0.0, 1e30)
    prim, t = bvh.intersect(prim_rays, block=64, lds_stack=32)
    prim_t, t_t = bvh.intersect(prim_rays, block=64, lds_stack=24, tile=(W, 16))
    prim_d, t_d = bvh.intersect(prim_rays)
    torch.cuda.synchronize()
    for p2, t2 in ((prim_t, t_t), (prim_d, t_d)):
        assert torch.equal(prim, p2) and torch.equal(t.view(torch.int32), t2.view(torch.int32))
    lo, hi = pos.amin(0).cpu().double().numpy(), pos.amax(0).cpu().double().numpy()
    c, ext = (lo + hi) / 2, hi - lo
    lights = torch.tensor(rt.lights_in_box(c - 1.5 * ext, c + 1.5 * ext, NL, 1337),
                          dtype=torch.float32, device="cuda")
    shadow = rt.gen_shadow_rays(pos, idx, prim_rays, prim, t, lights, 1e-4)
    print("headline frame: 33.2M shadow rays, tiled and linear", flush=True)
    bits = bvh.occluded(shadow, block=64, lds_stack=20, tile=(W, 16))
    bits_lin = bvh.occluded(shadow)
    masks = rt.pack_light_masks(bits, W * H, NL)
    torch.cuda.synchronize()
    assert torch.equal(bits, bits_lin)
    # the hinted trace as the bench's tuner runs it: the pixels' primary hits
    # as occluder hints (with the wave's other hints), 16 x 4 tiles and linear
    # waves, segments of 256-1024 slots
    for tile in ((W, 16), None):
        for seg, refill in ((512, 32), (256, 16), (1024, 16)):
            print(f"headline frame hinted tile={tile} segment={seg} refill={refill}", flush=True)
            hb = torch.full_like(bits, -1)
            bvh.occluded(shadow, out=hb, hint=prim, hint_period=W * H, block=64, lds_stack=16,
                         rays_per_wave=seg, refill_lanes=refill, tile=tile)
            torch.cuda.synchronize()
            assert torch.equal(hb, bits)
    nodes = to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4)
    tris = to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64)
    rng = np.random.default_rng(5)
    sel = np.sort(rng.choice(W * H, 100_000, replace=False))
    pc, tc = O.trace_bvh4(nodes, tris, O.Rays.stack(to_np(prim_rays.data)[:, sel]))
    np.testing.assert_array_equal(u32(prim)[sel], pc)
    np.testing.assert_array_equal(to_np(t)[sel].view(np.uint32), tc.view(np.uint32))
    sel = np.sort(rng.choice(W * H * NL, 100_000, replace=False))
    ra = O.Rays.stack(to_np(shadow.data[:, torch.from_numpy(sel).cuda()]))
    pa, _ = O.trace_bvh4(nodes, tris, ra, mode=O.ANY)
    b = u32(bits)
    got = ((b[sel >> 5] >> (sel & 31)) & 1).astype(bool)
    np.testing.assert_array_equal(got, pa != O.MISS)
    assert 0.05 < (pa != O.MISS).mean() < 0.95
    # the per-pixel masks are the bits regrouped: light l of pixel p = bit l * npix + p
    m = to_np(masks).view(np.uint16).astype(np.int64)
    pix, light = sel % (W * H), sel // (W * H)
    np.testing.assert_array_equal((m[pix] >> light) & 1, got.astype(np.int64))


def test_full_size_batched_build_maximum(rt, O):
    nx, ny = 256, 128  # 65,536 triangles
    pos, idx = rt.gen_heightfield(nx, ny, 1.0, 0.05, 77)
    assert idx.shape[0] == rt.BATCH_BUILD_MAX
    bb = rt.BuildBatch([(pos, idx)], keep_order=True)
    got = bb.launch()[0]
    want = rt.GpuBVH4.build(pos, idx, algorithm="implicit", keep_order=True)
    torch.cuda.synchronize()
    assert torch.equal(got.nodes[: want.node_count * 128], want.nodes[: want.node_count * 128])
    assert torch.equal(got.tris, want.tris[: got.tris.numel()])
    assert torch.equal(got.codes, want.codes[: got.codes.numel()])
