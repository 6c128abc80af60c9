"""GPU build parity: rt_gpu_build (Morton-30 + LDS radix sort, then the
implicit fast_bvh_builder layout or the radix tree collapsed with
SoftwareBVH4's rule) against the oracle's restatements.  Sorted order and
Morton codes bit-exact; node and leaf records byte-identical."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _cmp_build(rt, O, pos, idx, algorithm="implicit", maxp=4):
    n = idx.shape[0] if idx is not None else pos.size // 9
    if algorithm == "implicit":
        nodes_c, tris_c, ids_c, codes_c = O.lbvh_build(pos, idx)
    else:
        nodes_c, tris_c, ids_c, codes_c = O.rlbvh_build(pos, idx, max_prims=maxp)
    dpos = torch.from_numpy(np.ascontiguousarray(pos, np.float32)).cuda()
    didx = None if idx is None else torch.from_numpy(
        np.ascontiguousarray(idx, np.uint32).view(np.int32)).cuda()
    bvh = rt.GpuBVH4.build(dpos, didx, algorithm=algorithm, max_leaf_prims=maxp, keep_order=True)
    if algorithm == "radix":
        assert bvh.depth == O.bvh4_depth(nodes_c)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(u32(bvh.codes)[:n], codes_c)
    np.testing.assert_array_equal(u32(bvh.sorted_ids)[:n], ids_c)
    got_nodes = to_np(bvh.nodes)[: bvh.node_count * 128]
    assert bvh.node_count == len(nodes_c)
    np.testing.assert_array_equal(got_nodes, nodes_c.view(np.uint8).reshape(-1))
    got_tris = to_np(bvh.tris)[: n * 64]
    np.testing.assert_array_equal(got_tris, tris_c.view(np.uint8).reshape(-1))
    return bvh, nodes_c, tris_c


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 63, 64, 1000, 4095, 4096])
def test_small_path_dynamic_grid(rt, O, n):
    if n < 2:
        xyz = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)[:n]
    else:
        xyz, _ = O.dynamic_grid(n - (n % 2), 0.3, 1.25)
        if n % 2:
            xyz = np.concatenate([xyz, xyz[:1] + 0.5])
    _cmp_build(rt, O, xyz, None)


@pytest.mark.parametrize("n", [4097, 10_000, 65_536, 250_000])
def test_large_path_indexed(rt, O, n):
    nu = 50
    nv = max(2, n // (2 * nu))
    pos, idx = O.sphere(nu, nv, 1.0, 0.02, 9)
    idx = idx[:n]
    _cmp_build(rt, O, pos, idx)


def test_large_path_terrain_ties(rt, O):
    """Grid terrain: many equal centroid coordinates -> many equal codes; the
    stable sort must keep triangle-index order inside every tie."""
    pos, idx = O.terrain(100_000)
    _cmp_build(rt, O, pos, idx)


@pytest.mark.parametrize("n", [1, 2, 4, 5, 6, 100, 4096, 4097, 60_000, 250_000])
@pytest.mark.parametrize("maxp", [1, 4])
def test_radix_path_indexed(rt, O, n, maxp):
    nu = 50
    nv = max(2, n // (2 * nu) + 1)
    pos, idx = O.sphere(nu, nv, 1.0, 0.02, 9)
    _cmp_build(rt, O, pos, idx[:n], "radix", maxp)


def test_radix_path_duplicate_codes(rt, O):
    """Terrain grid: long runs of equal Morton codes exercise the position
    tie-break of the radix tree."""
    pos, idx = O.terrain(100_000)
    _cmp_build(rt, O, pos, idx, "radix", 4)
    xyz, _ = O.dynamic_grid(20_000, 0.0, 2.0)
    xyz = np.concatenate([xyz, xyz])  # every triangle twice: equal codes
    _cmp_build(rt, O, xyz, None, "radix", 2)


def test_build_then_trace_matches_brute_force(rt, O):
    pos, idx = O.sphere(64, 48, 1.0, 0.02, 3)
    dpos = torch.from_numpy(pos).cuda()
    didx = torch.from_numpy(idx.view(np.int32)).cuda()
    bvh = rt.GpuBVH4.build(dpos, didx, algorithm="radix")
    r = O.random_rays(20_000, 17, (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5), tmax=1e30)
    prim, t = bvh.intersect(rt.Rays.from_numpy(r.packed()))
    torch.cuda.synchronize()
    nodes_c, tris_c, _, _ = O.rlbvh_build(pos, idx)
    pb, tb = O.trace_brute(tris_c, r)
    prim, t = u32(prim), to_np(t)
    same_t = t.view(np.uint32) == tb.view(np.uint32)
    assert same_t.mean() > 0.9999
    # where t agrees the triangle agrees unless two triangles tie at that t
    assert ((prim == pb) | ~same_t).mean() > 0.999


def test_refit_matches_rebuild_layout(rt, O):
    n = 20_000
    xyz0, _ = O.dynamic_grid(n, 0.0, 2.0)
    xyz1, _ = O.dynamic_grid(n, 0.7, 2.0)
    d0 = torch.from_numpy(xyz0).cuda()
    bvh = rt.GpuBVH4.build(d0, None, algorithm="implicit", keep_order=True)
    nodes_c, tris_c, _, _ = O.lbvh_build(xyz0, None)
    nodes_r, tris_r = O.lbvh_refit(xyz1, None, nodes_c, tris_c)
    d1 = torch.from_numpy(xyz1).cuda()
    bvh.refit(d1, None)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(bvh.nodes)[: bvh.node_count * 128],
                                  nodes_r.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(bvh.tris)[: n * 64], tris_r.view(np.uint8).reshape(-1))


@pytest.mark.parametrize("n", [4096, 3002, 130])
def test_small_refit(rt, O, n):
    """Fused-path refit (one workgroup per 256 sorted positions, partial last
    one included): nodes and records equal the oracle's refit."""
    xyz0, _ = O.dynamic_grid(n, 0.0, 1.25)
    xyz1, _ = O.dynamic_grid(n, 1.1, 1.25)
    bvh = rt.GpuBVH4.build(torch.from_numpy(xyz0).cuda(), None, algorithm="implicit")
    nodes_c, tris_c, _, _ = O.lbvh_build(xyz0, None)
    nodes_r, tris_r = O.lbvh_refit(xyz1, None, nodes_c, tris_c)
    bvh.refit(torch.from_numpy(xyz1).cuda(), None)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(bvh.nodes)[: bvh.node_count * 128],
                                  nodes_r.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(bvh.tris)[: n * 64], tris_r.view(np.uint8).reshape(-1))


@pytest.mark.parametrize("nu,nv", [(64, 32), (40, 20), (7, 3)])
def test_small_path_indexed_sphere(rt, O, nu, nv):
    """Indexed UV sphere: thin pole triangles crowd a few Morton buckets (up
    to 35 of 4096 triangles share the top 13 code bits), so ranks inside a
    bucket decide the order."""
    pos, idx = O.sphere(nu, nv, 1.0, 0.02, 5)
    _cmp_build(rt, O, pos, idx)


@pytest.mark.parametrize("dups", [300, 4096])
def test_small_path_clustered_fallback(rt, O, dups):
    """More than 256 triangles in one bucket sends the fused build to its LSD
    sort: `dups` copies of one triangle (equal codes: index order decides)
    among a grid, and a mesh of 4096 identical triangles."""
    xyz, _ = O.dynamic_grid(4096, 0.2, 1.25)
    xyz = xyz.copy()
    pick = np.random.default_rng(3).choice(4096, size=dups, replace=False)
    xyz[pick] = xyz[17]
    _cmp_build(rt, O, xyz, None)


def test_fast_builder_verification_rays(rt, O):
    """src/main.cpp:1002-1106: for generate_dynamic_grid_mesh(N, 1.25),
    N in {64, 256, 1024, 4096}, one +z ray from (cx, cy, 0) per triangle hits
    with |t - cz| < 1e-2; the ray from (10, 10, 0) misses."""
    for n in (64, 256, 1024, 4096):
        xyz, cen = O.dynamic_grid(n, 0.0, 1.25)
        bvh = rt.GpuBVH4.build(torch.from_numpy(xyz).cuda(), None)
        r = O.Rays(n + 1)
        r.ox[:n], r.oy[:n] = cen[:, 0], cen[:, 1]
        r.ox[n], r.oy[n] = 10.0, 10.0
        r.dz[:] = 1.0
        r.tmin[:] = 0.0
        r.tmax[:] = 1e30
        prim, t = bvh.intersect(rt.Rays.from_numpy(r.packed()))
        torch.cuda.synchronize()
        prim, t = u32(prim), to_np(t)
        assert (prim[:n] != O.MISS).all()
        assert (np.abs(t[:n] - cen[:, 2]) < 1e-2).all()
        assert prim[n] == O.MISS


@pytest.mark.parametrize("n", [1, 3, 4, 5, 64, 255, 1000, 4094, 4096])
@pytest.mark.parametrize("order", ["input", "morton"])
def test_radv_image_bytes(rt, O, n, order):
    """rt_gpu_build_radv: the whole RADV image (header, box32 nodes with
    NaN empty slots, triangle nodes) byte-identical to the oracle's
    restatement of fast_bvh_builder.comp, in input order and in the Morton
    order of the fused build."""
    xyz, _ = O.dynamic_grid(max(2, n + (n % 2)), 0.6, 1.25)
    xyz = np.ascontiguousarray(xyz[:n])
    img, ids = rt.build_radv(torch.from_numpy(xyz).cuda(), None, order=order)
    torch.cuda.synchronize()
    want_ids = np.arange(n, dtype=np.uint32) if order == "input" else O.lbvh_build(xyz, None)[2]
    np.testing.assert_array_equal(u32(ids), want_ids)
    ref = O.radv_encode(xyz, order=None if order == "input" else want_ids)
    np.testing.assert_array_equal(to_np(img), ref)


def test_radv_image_indexed(rt, O):
    pos, idx = O.sphere(40, 30, 1.0, 0.02, 8)
    img, ids = rt.build_radv(torch.from_numpy(pos).cuda(),
                             torch.from_numpy(idx.view(np.int32)).cuda(), order="morton")
    torch.cuda.synchronize()
    xyz = pos[idx.astype(np.int64)].reshape(-1, 9)
    want = O.lbvh_build(pos, idx)[2]
    np.testing.assert_array_equal(u32(ids), want)
    np.testing.assert_array_equal(to_np(img), O.radv_encode(xyz, order=want))


def _deformed(pos, seed):
    rng = np.random.default_rng(seed)
    return (pos * np.float32(1.07) + rng.normal(0, 0.01, pos.shape).astype(np.float32)).astype(
        np.float32)


@pytest.mark.parametrize("n,maxp", [(5, 4), (1000, 1), (60_000, 4), (250_000, 4)])
@pytest.mark.parametrize("sync", [True, False])
def test_generic_refit_radix_matches_oracle(rt, O, n, maxp, sync):
    """rt_gpu_refit_bvh4 on a radix tree: unchanged positions give the built
    bytes back; deformed positions give the oracle refit's bytes."""
    nu = 50
    nv = max(2, n // (2 * nu) + 1)
    pos, idx = O.sphere(nu, nv, 1.0, 0.02, 9)
    idx = np.ascontiguousarray(idx[:n])
    nodes_c, tris_c, _, _ = O.rlbvh_build(pos, idx, max_prims=maxp)
    dpos = torch.from_numpy(pos).cuda()
    didx = torch.from_numpy(idx.view(np.int32)).cuda()
    bvh = rt.GpuBVH4.build(dpos, didx, algorithm="radix", max_leaf_prims=maxp, sync=sync)
    bvh.refit(dpos, didx)
    torch.cuda.synchronize()
    assert bvh.node_count == len(nodes_c)
    np.testing.assert_array_equal(to_np(bvh.nodes)[: len(nodes_c) * 128],
                                  nodes_c.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(bvh.tris)[: n * 64], tris_c.view(np.uint8).reshape(-1))
    new = _deformed(pos, n)
    bvh.refit(torch.from_numpy(new).cuda(), didx)
    torch.cuda.synchronize()
    want_n, want_t = O.bvh4_refit(new, idx, nodes_c, tris_c)
    np.testing.assert_array_equal(to_np(bvh.nodes)[: len(nodes_c) * 128],
                                  want_n.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(bvh.tris)[: n * 64], want_t.view(np.uint8).reshape(-1))
    if n <= 1000:  # the refitted tree traces the deformed mesh exactly
        r = O.random_rays(20_000, 3, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
        prim, t = bvh.intersect(rt.Rays.from_numpy(r.packed()))
        pc, tc = O.trace_bvh4(want_n, want_t, r)
        np.testing.assert_array_equal(u32(prim), pc)
        np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))


def test_generic_refit_implicit_and_uploaded(rt, O):
    """The generic refit equals the implicit-tree refit on an implicit tree,
    and the oracle on an uploaded SoftwareBVH4 tree."""
    xyz, _ = O.dynamic_grid(10_000, 0.2, 1.5)
    d = torch.from_numpy(xyz).cuda()
    a = rt.GpuBVH4.build(d, None, algorithm="implicit")
    b = rt.GpuBVH4.build(d, None, algorithm="implicit")
    moved = torch.from_numpy(_deformed(xyz.reshape(-1, 3), 5).reshape(xyz.shape)).cuda()
    a.refit(moved, None)                 # implicit refit
    b.refit(moved, None, generic=True)   # generic refit
    torch.cuda.synchronize()
    assert torch.equal(a.nodes, b.nodes) and torch.equal(a.tris, b.tris)
    pos, idx = O.wave_grid(40, 40, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, t64 = O.software_bvh4(pos, idx)
    up = rt.GpuBVH4.from_bvh4(n4, tris64=t64, depth=O.bvh4_depth(n4))
    new = _deformed(pos, 6)
    didx = torch.from_numpy(np.ascontiguousarray(idx).view(np.int32)).cuda()
    up.refit(torch.from_numpy(new).cuda(), didx)
    torch.cuda.synchronize()
    want_n, want_t = O.bvh4_refit(new, idx, n4, t64)
    np.testing.assert_array_equal(to_np(up.nodes)[: len(n4) * 128],
                                  want_n.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(up.tris)[: len(t64) * 64],
                                  want_t.view(np.uint8).reshape(-1))


@pytest.mark.parametrize("target", [2, 1000, 100_000])
@pytest.mark.parametrize("fused", [False, True])
def test_terrain_generator(rt, O, target, fused):
    """rt_gen_terrain = createTessellatedTerrain: resolution and indices exact,
    x / z exact (no transcendental), heights within 4e-6 of the oracle: the
    device sinf / cosf against the host library's, each within a few ulp,
    scaled by the summands 5 sin cos and 2 sin (|h| <= 7, ulp(4..8) = 4.8e-7),
    so the bound is absolute, not relative to a height near zero."""
    pos, idx = rt.gen_terrain(target, 100.0, fused=fused)
    torch.cuda.synchronize()
    pc, ic = O.terrain(target, fused=fused)
    p = to_np(pos)
    assert p.shape == pc.shape
    np.testing.assert_array_equal(u32(idx), ic)
    np.testing.assert_array_equal(p[:, 0], pc[:, 0])
    np.testing.assert_array_equal(p[:, 2], pc[:, 2])
    err = np.abs(p[:, 1].astype(np.float64) - pc[:, 1])
    assert err.max() <= 4e-6, err.max()


def test_host_layer_rejects_wrong_arguments(rt, O):
    """The Python host layer checks what the C ABI reads before calling it:
    an int64 index tensor (torch's default integer type), a float64 or
    non-contiguous position tensor and undersized outputs raise ValueError
    instead of building from or writing to misread memory."""
    pos, idx = O.wave_grid(8, 8)
    dpos = torch.from_numpy(pos).cuda()
    didx = torch.from_numpy(idx.view(np.int32)).cuda()
    with pytest.raises(ValueError):
        rt.GpuBVH4.build(dpos, didx.long())
    with pytest.raises(ValueError):
        rt.GpuBVH4.build(dpos.double(), didx)
    with pytest.raises(ValueError):
        rt.GpuBVH4.build(dpos.t().contiguous().t(), didx)
    with pytest.raises(ValueError):
        rt.GpuBVH4.build(dpos, didx.reshape(-1))
    bvh = rt.GpuBVH4.build(dpos, didx)
    rays = rt.Rays.from_numpy(O.random_rays(100, 1, (-1, -1, -1), (1, 1, 1)).packed())
    with pytest.raises(ValueError):
        bvh.intersect(rays, out=(torch.empty(99, dtype=torch.int32, device="cuda"),
                                 torch.empty(100, dtype=torch.float32, device="cuda")))
    with pytest.raises(ValueError):
        bvh.intersect(rays, out=(torch.empty(100, dtype=torch.int32, device="cuda"),
                                 torch.empty(100, dtype=torch.float64, device="cuda")))
    with pytest.raises(ValueError):
        bvh.occluded(rays, out=torch.empty(3, dtype=torch.int32, device="cuda"))
    with pytest.raises(ValueError):
        rt.TraceBatch([(bvh, rays, "any", torch.empty(3, dtype=torch.int32, device="cuda"))])
    with pytest.raises(ValueError):
        bvh.refit(dpos[:-3].contiguous(), didx[:-1].contiguous())
    # the right shapes still work
    bits = bvh.occluded(rays, out=torch.empty(4, dtype=torch.int32, device="cuda"))
    torch.cuda.synchronize()
    assert bits.numel() == 4
