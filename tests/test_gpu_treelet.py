"""Treelet restructuring of the radix tree on the device (rt_build_cfg
treelet_rounds, Karras & Aila 2013; no reference counterpart) against the
oracle's restructuring of the same radix tree: node and leaf bytes, node
count and depth equal, for trees that need the one-workgroup pass over the
nodes crossing 4096-position windows (n > 4096), equal Morton codes, leaf
sizes 1 / 2 / 4 and 1, 3, 8 and 16 rounds.  The restructured tree traces to the same
occlusion bits and closest-hit t as the plain radix tree (the hits do not
depend on the tree).
"""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _cmp(rt, O, pos, idx, maxp, rounds):
    n = idx.shape[0] if idx is not None else pos.size // 9
    nodes_c, tris_c, ids_c, _ = O.rlbvh_build(pos, idx, max_prims=maxp, treelet_rounds=rounds)
    dpos = torch.from_numpy(np.ascontiguousarray(pos, np.float32)).cuda()
    didx = None if idx is None else torch.from_numpy(
        np.ascontiguousarray(idx, np.uint32).view(np.int32)).cuda()
    print(f"treelet build n={n} maxp={maxp} rounds={rounds}", flush=True)
    bvh = rt.GpuBVH4.build(dpos, didx, algorithm="radix", max_leaf_prims=maxp,
                           treelet_rounds=rounds, keep_order=True)
    torch.cuda.synchronize()
    assert bvh.node_count == len(nodes_c)
    assert bvh.depth == O.bvh4_depth(nodes_c)
    np.testing.assert_array_equal(u32(bvh.sorted_ids)[:n], ids_c)
    np.testing.assert_array_equal(to_np(bvh.nodes)[: bvh.node_count * 128],
                                  nodes_c.view(np.uint8).reshape(-1))
    np.testing.assert_array_equal(to_np(bvh.tris)[: n * 64], tris_c.view(np.uint8).reshape(-1))
    return bvh, dpos, didx


@pytest.mark.parametrize("n", [7, 100, 4096, 20000, 150000])
@pytest.mark.parametrize("maxp", [1, 2, 4])
@pytest.mark.parametrize("rounds", [1, 3])
def test_treelet_sphere_matches_oracle(rt, O, n, maxp, rounds):
    nu = 50
    nv = max(2, n // (2 * nu) + 1)
    pos, idx = O.sphere(nu, nv, 1.0, 0.02, 9)
    _cmp(rt, O, pos, np.ascontiguousarray(idx[:n]), maxp, rounds)


@pytest.mark.parametrize("n", [20000, 150000])
@pytest.mark.parametrize("maxp", [1, 2, 4])
@pytest.mark.parametrize("rounds", [8, 16])
def test_treelet_many_rounds_match_oracle(rt, O, n, maxp, rounds):
    """The round counts the frames' tuner picks (8 for the shadow frame's
    sphere, 16 for the AO frame's terrain), byte-exact against the oracle."""
    nu = 50
    nv = max(2, n // (2 * nu) + 1)
    pos, idx = O.sphere(nu, nv, 1.0, 0.02, 9)
    _cmp(rt, O, pos, np.ascontiguousarray(idx[:n]), maxp, rounds)


def test_treelet_heightfield_16_rounds(rt, O):
    """A seeded random-displaced heightfield (the AO frame's terrain kind,
    at 80,000 triangles), leaves of 1, 16 rounds."""
    pos, idx = O.heightfield(200, 200, 4.0, 0.3, 31)
    _cmp(rt, O, pos, idx, 1, 16)


def test_treelet_equal_codes_and_soup(rt, O):
    pos, idx = O.terrain(100_000)
    _cmp(rt, O, pos, idx, 2, 2)
    xyz, _ = O.dynamic_grid(20_000, 0.0, 2.0)
    _cmp(rt, O, np.concatenate([xyz, xyz]), None, 2, 2)  # every triangle twice
    rng = np.random.default_rng(5)
    soup = (rng.random((30_000, 9)) * 4 - 2).astype(np.float32)  # long thin overlapping triangles
    _cmp(rt, O, soup, None, 1, 2)


def test_treelet_same_hits_as_radix_tree(rt, O):
    pos, idx = O.sphere(300, 200, 1.0, 0.05, 4)
    plain = _cmp(rt, O, pos, idx, 2, 0)[0]
    tl, dpos, didx = _cmp(rt, O, pos, idx, 2, 2)
    r = O.random_rays(200_000, 8, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
    rays = rt.Rays.from_numpy(r.packed())
    p0, t0 = plain.intersect(rays)
    p1, t1 = tl.intersect(rays)
    b0, b1 = plain.occluded(rays), tl.occluded(rays)
    torch.cuda.synchronize()
    assert torch.equal(t0.view(torch.int32), t1.view(torch.int32))
    assert torch.equal(b0, b1)
    same = (p0 == p1).float().mean().item()
    assert same > 0.9999  # a different id only where two triangles meet at exactly equal t
    assert tl.node_count < plain.node_count + plain.node_count // 20


def test_treelet_rejects_bad_config(rt):
    pos = torch.rand(300, 9, device="cuda")
    with pytest.raises(RuntimeError):
        rt.GpuBVH4.build(pos, None, algorithm="implicit", treelet_rounds=1)
    with pytest.raises(RuntimeError):
        rt.GpuBVH4.build(pos, None, algorithm="radix", treelet_rounds=17)
