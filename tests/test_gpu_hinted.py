"""Occlusion bits with occluder hints (rt_trace_occluded_hinted): a hint
settles a ray only when its triangle passes the hit predicate in [tmin, tmax)
and its own box passes the slab test, so the bits equal the oracle's any hit
on the same tree and rays (and the plain kernel's, whole words) whatever the
hints: the frame's primary hits for its shadow rays, with and without the
wave's other hints, in wave tiles of 8 / 16 / 32 and linear waves, over
launch shapes (segments of 64-4096 slots, refill thresholds, LDS stacks
that spill); on trees with leaves of 1-4 triangles with and without treelet
restructuring and on an uploaded SoftwareBVH4; with random, out-of-range and
missing hints, hints beyond tmax and the rays' own occluders; on rays that
graze the edges of axis-aligned triangles, where the hit predicate accepts
a triangle whose box the slab test rejects (the hint must not settle those);
on ragged ray counts, a partial last tile row and an empty tree."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF
W, H = 160, 120

SHAPES = [dict(block=64, lds_stack=16, rays_per_wave=512, refill_lanes=32),
          dict(block=256, lds_stack=12, rays_per_wave=1024, refill_lanes=16),
          dict(block=128, lds_stack=4, rays_per_wave=64, refill_lanes=1),
          dict(block=64, lds_stack=20, rays_per_wave=4096, refill_lanes=64)]


def _bits(b, n):
    i = np.arange(n)
    return ((u32(b)[i >> 5] >> (i & 31)) & 1).astype(bool)


def _oracle_tree(O, bvh):
    return (to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4),
            to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64))


def _frame(rt, leaf, rounds):
    pos, idx = rt.gen_sphere(250, 200, 1.0, 0.02, 3)  # 100,000 triangles
    bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix", max_leaf_prims=leaf,
                           treelet_rounds=rounds)
    bvh.node_count, bvh.depth = [int(v) for v in bvh.info[:2].cpu().tolist()]
    prim_rays = rt.gen_camera_rays(W, H, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                   60.0, 0.0, 1e30)
    prim, t = bvh.intersect(prim_rays)
    lo, hi = pos.amin(0).cpu().double().numpy(), pos.amax(0).cpu().double().numpy()
    c, ext = (lo + hi) / 2, hi - lo
    lights = torch.tensor(rt.lights_in_box(c - 1.5 * ext, c + 1.5 * ext, 16, 1337),
                          dtype=torch.float32, device="cuda")
    shadow = rt.gen_shadow_rays(pos, idx, prim_rays, prim, t, lights, 1e-4)
    torch.cuda.synchronize()
    return bvh, shadow, prim


@pytest.mark.parametrize("leaf,rounds", [(1, 0), (1, 8), (2, 8), (4, 0)])
def test_hinted_bits_equal_oracle_on_shadow_frame(rt, O, leaf, rounds):
    bvh, shadow, prim = _frame(rt, leaf, rounds)
    nodes, tris = _oracle_tree(O, bvh)
    r = O.Rays.stack(to_np(shadow.data))
    want = O.trace_bvh4(nodes, tris, r, mode=O.ANY)[0] != O.MISS
    assert 0.05 < want.mean() < 0.98
    plain = bvh.occluded(shadow)
    np.testing.assert_array_equal(_bits(plain, len(r)), want)
    for tile in (None, (W, 16), (W, 8), (W, 32)):
        for wave in (True, False):
            for shape in SHAPES:
                print(f"hinted leaf={leaf} rounds={rounds} tile={tile} wave={wave} {shape}",
                      flush=True)
                b = torch.full_like(plain, -1)
                bvh.occluded(shadow, out=b, hint=prim, hint_period=W * H, hint_wave=wave,
                             tile=tile, **shape)
                torch.cuda.synchronize()
                assert torch.equal(b, plain)


def test_hinted_bits_with_adversarial_hints(rt, O):
    bvh, shadow, prim = _frame(rt, 1, 4)
    nodes, tris = _oracle_tree(O, bvh)
    r = O.Rays.stack(to_np(shadow.data))
    n = len(r)
    plain = bvh.occluded(shadow)
    gen = np.random.default_rng(5)
    ids = tris["triangle_id"]
    # each ray's own first occluder (the closest hit within [tmin, tmax)), the
    # closest hit with no tmax (a triangle the predicate accepts beyond tmax
    # for unoccluded rays), random triangles, out-of-range ids and none
    own = O.trace_bvh4(nodes, tris, r)[0]
    far = r.subset(slice(None))
    far.tmax = np.full(n, 1e30, np.float32)
    beyond = O.trace_bvh4(nodes, tris, far)[0]
    rand = gen.choice(ids, n).astype(np.uint32)
    bad = np.where(gen.random(n) < 0.5, np.uint32(bvh.tri_count + 7), np.uint32(MISS))
    for name, h in (("own", own), ("beyond", beyond), ("random", rand), ("bad", bad)):
        for tile in (None, (W, 16)):
            print(f"hinted adversarial hints={name} tile={tile}", flush=True)
            hint = torch.from_numpy(h.astype(np.uint32).view(np.int32)).cuda()
            b = torch.full_like(plain, -1)
            bvh.occluded(shadow, out=b, hint=hint, tile=tile, **SHAPES[0])
            torch.cuda.synchronize()
            assert torch.equal(b, plain)


def _edge_scene():
    """Axis-aligned right triangles in planes z = c whose legs lie on their
    boxes' x and y faces, and rays aimed at points on those legs (exact and
    one ulp either side), from above the plane at shallow and steep angles."""
    tri = []
    for i in range(8):
        for j in range(8):
            x0, y0, z = np.float32(i * 1.37), np.float32(j * 0.91), np.float32(0.25 * ((i + j) % 3))
            tri.append([[x0, y0, z], [x0 + np.float32(1.0), y0, z], [x0, y0 + np.float32(1.0), z]])
    pos = np.array(tri, np.float32).reshape(-1, 3)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    rays = []
    gen = np.random.default_rng(11)
    for k, t3 in enumerate(tri):
        (x0, y0, z), _, _ = t3
        for s in np.linspace(0.05, 0.95, 12, dtype=np.float32):
            for leg in range(2):
                for eps in (-1, 0, 1):
                    px = np.nextafter(x0, np.float32(x0 + eps)) if (leg == 0 and eps) else x0
                    py = np.nextafter(y0, np.float32(y0 + eps)) if (leg == 1 and eps) else y0
                    if leg == 0:
                        py = np.float32(y0 + s)
                    else:
                        px = np.float32(x0 + s)
                    d = gen.normal(size=3)
                    d[2] = -abs(d[2]) * (0.05 if k % 2 else 3.0)
                    d = (d / np.linalg.norm(d)).astype(np.float32)
                    o = np.array([px, py, z], np.float32) - np.float32(0.7) * d
                    rays.append([o[0], o[1], o[2], 1e-4, d[0], d[1], d[2], 10.0])
    a = np.array(rays, np.float32).T.copy()
    return pos, idx, a


def test_hinted_bits_on_edge_grazing_rays(rt, O):
    pos, idx, a = _edge_scene()
    dpos = torch.from_numpy(pos).cuda()
    didx = torch.from_numpy(idx.view(np.int32)).cuda()
    r = O.Rays.stack(a)
    n = len(r)
    for leaf in (1, 4):
        bvh = rt.GpuBVH4.build(dpos, didx, algorithm="radix", max_leaf_prims=leaf)
        bvh.node_count, bvh.depth = [int(v) for v in bvh.info[:2].cpu().tolist()]
        nodes, tris = _oracle_tree(O, bvh)
        want = O.trace_bvh4(nodes, tris, r, mode=O.ANY)[0] != O.MISS
        brute = O.trace_brute(tris, r, mode=O.ANY)[0] != O.MISS
        # the predicate accepts a triangle the slab test rejects on some rays
        assert (brute & ~want).sum() > 0 and want.sum() > 0
        aimed = (np.arange(n) // (12 * 2 * 3)).astype(np.uint32)  # the triangle aimed at
        rays = rt.Rays.from_numpy(a)
        plain = bvh.occluded(rays)
        np.testing.assert_array_equal(_bits(plain, n), want)
        for tile in (None,):
            for shape in SHAPES[:2]:
                print(f"hinted edge rays leaf={leaf} {shape}", flush=True)
                hint = torch.from_numpy(aimed.view(np.int32)).cuda()
                b = torch.full_like(plain, -1)
                bvh.occluded(rays, out=b, hint=hint, **shape)
                torch.cuda.synchronize()
                np.testing.assert_array_equal(_bits(b, n), want)
                assert torch.equal(b, plain)


def test_hinted_bits_on_uploaded_software_bvh4(rt, O):
    pos, idx = O.sphere(120, 90, 1.0, 0.02, 5)
    n4, t64 = O.software_bvh4(pos, idx, max_prims=4)
    bvh = rt.GpuBVH4.from_bvh4(n4, tris64=t64, depth=O.bvh4_depth(n4))
    nodes, tris = _oracle_tree(O, bvh)
    rnd = O.random_rays(30_001, 3, (-1.1, -1.1, -1.1), (1.1, 1.1, 1.1), tmax=0.5)
    rays = rt.Rays.from_numpy(rnd.packed())
    closest = O.trace_bvh4(nodes, tris, rnd)[0]
    want = O.trace_bvh4(nodes, tris, rnd, mode=O.ANY)[0] != O.MISS
    hint = torch.from_numpy(closest.view(np.int32)).cuda()
    for shape in SHAPES[:3]:
        print(f"hinted uploaded SoftwareBVH4 {shape}", flush=True)
        b = torch.full(((len(rnd) + 31) // 32,), -1, dtype=torch.int32, device="cuda")
        bvh.occluded(rays, out=b, hint=hint, **shape)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_bits(b, len(rnd)), want)


def test_hinted_bits_ragged_counts_partial_tiles_and_empty_tree(rt, O):
    bvh, shadow, prim = _frame(rt, 2, 4)
    nodes, tris = _oracle_tree(O, bvh)
    for n in (W * H * 3 + 1, 63, 1):
        r0 = rt.Rays(shadow.data[:, :n].contiguous())
        want = O.trace_bvh4(nodes, tris, O.Rays.stack(to_np(r0.data)), mode=O.ANY)[0] != O.MISS
        for period in (W * H, 7):
            print(f"hinted ragged n={n} period={period}", flush=True)
            b = torch.full(((n + 31) // 32,), -1, dtype=torch.int32, device="cuda")
            bvh.occluded(r0, out=b, hint=prim, hint_period=min(period, W * H), **SHAPES[0])
            torch.cuda.synchronize()
            np.testing.assert_array_equal(_bits(b, n), want)
            assert (u32(b)[n >> 5] >> (n & 31) == 0) if n % 32 else True
    # 121 rows: tiles of 4 / 8 / 2 rows leave a partial tile row
    n = W * 121
    r0 = rt.Rays(shadow.data[:, :n].contiguous())
    ref = bvh.occluded(r0)
    for tile in ((W, 16), (W, 8), (W, 32)):
        print(f"hinted partial tiles tile={tile}", flush=True)
        b = torch.full_like(ref, -1)
        bvh.occluded(r0, out=b, hint=prim, hint_period=W * H, tile=tile, **SHAPES[1])
        torch.cuda.synchronize()
        assert torch.equal(b, ref)
    empty = rt.GpuBVH4(bvh.nodes, bvh.tris, 0, 0, 0)  # no nodes: every ray misses
    b = torch.full_like(ref, -1)
    empty.occluded(r0, out=b, hint=prim, hint_period=W * H, tile=(W, 16), **SHAPES[0])
    torch.cuda.synchronize()
    assert not u32(b).any()
