"""Pins of the oracle's generic Box4 refit (or_bvh4_refit).

  * on unchanged positions it reproduces each builder's own bytes: the
    implicit (fast_bvh_builder) tree, the radix tree collapsed with
    SoftwareBVH4's rule, and the reference's software_bvh tree (the last up to
    the sign of zero bounds, which the refit writes as +0 like every device
    build);
  * after a deformation it agrees with the implicit-tree refit, a separate
    restatement, byte for byte;
  * every slot box equals the brute-force union of its leaf records or of its
    child's occupied slots, and every record holds its triangle's new vertices;
  * a tree whose child index does not exceed its parent's is refused.
"""
import numpy as np
import pytest


def _deform(pos, seed):
    rng = np.random.default_rng(seed)
    return (pos * np.float32(1.07) + rng.normal(0, 0.01, pos.shape).astype(np.float32)).astype(
        np.float32)


def _check_unions(O, pos, idx, nodes, tris):
    """Brute force: records = new vertices; slot boxes = unions."""
    if idx is None:
        v = pos.reshape(-1, 9)[tris["triangle_id"]]
    else:
        v = pos[idx[tris["triangle_id"]]].reshape(-1, 9)
    np.testing.assert_array_equal(tris["v"], v)
    tmin = np.minimum(np.minimum(v[:, 0:3], v[:, 3:6]), v[:, 6:9]) + np.float32(0)
    tmax = np.maximum(np.maximum(v[:, 0:3], v[:, 3:6]), v[:, 6:9]) + np.float32(0)
    for i in range(len(nodes)):
        nd = nodes[i]
        for k in range(4):
            c = int(nd["count"][k])
            if c == 0:
                continue
            if c > 0:
                f = ~int(nd["child"][k])
                lo = tmin[f:f + c].min(0)
                hi = tmax[f:f + c].max(0)
            else:
                ch = nodes[int(nd["child"][k])]
                occ = ch["count"] != 0
                lo = np.array([ch["minX"][occ].min(), ch["minY"][occ].min(), ch["minZ"][occ].min()])
                hi = np.array([ch["maxX"][occ].max(), ch["maxY"][occ].max(), ch["maxZ"][occ].max()])
            got_lo = np.array([nd["minX"][k], nd["minY"][k], nd["minZ"][k]])
            got_hi = np.array([nd["maxX"][k], nd["maxY"][k], nd["maxZ"][k]])
            np.testing.assert_array_equal(got_lo, lo)
            np.testing.assert_array_equal(got_hi, hi)


@pytest.mark.parametrize("maxp", [1, 4])
def test_refit_reproduces_radix_build(O, maxp):
    pos, idx = O.sphere(60, 40, 1.0, 0.02, 2)
    nodes, tris, _, _ = O.rlbvh_build(pos, idx, max_prims=maxp)
    n2, t2 = O.bvh4_refit(pos, idx, nodes, tris)
    assert n2.tobytes() == nodes.tobytes() and t2.tobytes() == tris.tobytes()


def test_refit_reproduces_implicit_build(O):
    xyz, _ = O.dynamic_grid(6000, 0.3, 1.5)
    nodes, tris, _, _ = O.lbvh_build(xyz, None)
    n2, t2 = O.bvh4_refit(xyz, None, nodes, tris)
    assert n2.tobytes() == nodes.tobytes() and t2.tobytes() == tris.tobytes()


def test_refit_reproduces_software_bvh_up_to_zero_sign(O):
    pos, idx = O.wave_grid(24, 24, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, t64 = O.software_bvh4(pos, idx)
    n2, t2 = O.bvh4_refit(pos, idx, n4, t64)
    assert t2.tobytes() == t64.tobytes()
    a = n2.view(np.float32).reshape(len(n2), 32)[:, :24]
    b = n4.view(np.float32).reshape(len(n4), 32)[:, :24]
    np.testing.assert_array_equal(a, b)  # -0 == +0 compares equal
    assert n2["child"].tobytes() == n4["child"].tobytes()
    assert n2["count"].tobytes() == n4["count"].tobytes()


def test_refit_after_deformation(O):
    pos, idx = O.sphere(50, 30, 1.0, 0.02, 3)
    new = _deform(pos, 1)
    nodes, tris, _, _ = O.rlbvh_build(pos, idx, max_prims=4)
    n2, t2 = O.bvh4_refit(new, idx, nodes, tris)
    _check_unions(O, new, idx, n2, t2)
    assert n2["child"].tobytes() == nodes["child"].tobytes()
    # implicit tree: the generic refit equals the implicit-tree refit
    xyz, _ = O.dynamic_grid(3000, 0.1, 1.5)
    ln, lt, _, _ = O.lbvh_build(xyz, None)
    moved = _deform(xyz.reshape(-1, 3), 2).reshape(xyz.shape)
    a_n, a_t = O.bvh4_refit(moved, None, ln, lt)
    b_n, b_t = O.lbvh_refit(moved, None, ln, lt)
    assert a_n.tobytes() == b_n.tobytes() and a_t.tobytes() == b_t.tobytes()
    _check_unions(O, moved.reshape(-1), None, a_n, a_t)


def test_refit_refuses_child_before_parent(O):
    pos, idx = O.sphere(20, 10, 1.0, 0.02, 4)
    nodes, tris, _, _ = O.rlbvh_build(pos, idx, max_prims=4)
    assert len(nodes) > 2
    bad = nodes.copy()
    k = int(np.nonzero(bad[1]["count"] < 0)[0][0]) if (bad[1]["count"] < 0).any() else None
    i = 1 if k is not None else 0
    k = int(np.nonzero(bad[i]["count"] < 0)[0][0])
    bad[i]["child"][k] = i  # a child that is its own parent
    with pytest.raises(ValueError):
        O.bvh4_refit(pos, idx, bad, tris)


def test_quantize_contains_and_is_tight(O):
    """or_bvh4_quantize on a radix tree with leaves of 1 and 2 treelet rounds:
    every decoded slot box contains the rt_node4 slot box; each code is the
    tightest (one step inward breaks the containment, checked by brute force
    over the 256 codes) and each exponent the least with dec(255) >= the
    node's max bound (brute force over -126..127); child / count unchanged;
    empty slots code 0."""
    pos, idx = O.sphere(60, 40, 1.0, 0.05, 4)
    nodes, tris, _, _ = O.rlbvh_build(pos, idx, max_prims=1, treelet_rounds=2)
    q = O.bvh4_quantize(nodes)
    d = O.bvh4_dequantize(q)
    assert (d["child"] == np.where(nodes["count"] != 0, nodes["child"], 0)).all()
    assert (d["count"] == nodes["count"]).all()
    axes = (("minX", "maxX"), ("minY", "maxY"), ("minZ", "maxZ"))
    rng = np.random.default_rng(1)
    for i in rng.choice(len(nodes), 300, replace=False):
        n, qi = nodes[i], q[i]
        for a, (lo, hi) in enumerate(axes):
            occ = n["count"] != 0
            p = np.float32(n[lo][occ].min())
            m = np.float32(n[hi][occ].max())
            assert qi["p"][a] == p
            dec = lambda c, e: np.float32(p + np.float32(c) * np.float32(2.0 ** e))
            es = [e for e in range(-126, 128) if dec(255, e) >= m]
            assert qi["e"][a] == es[0]
            for k in range(4):
                if not occ[k]:
                    assert qi["qlo"][a][k] == 0 and qi["qhi"][a][k] == 0
                    continue
                codes = [dec(c, es[0]) for c in range(256)]
                assert qi["qlo"][a][k] == max(c for c in range(256) if codes[c] <= n[lo][k])
                assert qi["qhi"][a][k] == min(c for c in range(256) if codes[c] >= n[hi][k])
                assert d[lo][i][k] <= n[lo][k] and d[hi][i][k] >= n[hi][k]


def test_quantized_tree_traces_the_same_triangles(O):
    """Tracing the dequantized tree (the boxes a quantized trace tests)
    gives the exact tree's closest hits and occlusion bits on random rays
    (the boxes only grow: a superset of node visits, the same triangles)."""
    pos, idx = O.sphere(200, 150, 1.0, 0.05, 9)
    nodes, tris, _, _ = O.rlbvh_build(pos, idx, max_prims=2)
    d = O.bvh4_dequantize(O.bvh4_quantize(nodes))
    r = O.random_rays(50_000, 3, (-1.3,) * 3, (1.3,) * 3, tmax=1e30)
    pe, te = O.trace_bvh4(nodes, tris, r)
    pq, tq = O.trace_bvh4(d, tris, r)
    assert (te.view(np.uint32) == tq.view(np.uint32)).all()
    assert (pe == pq).mean() > 0.9999
    r.tmax[:] = 0.3
    ae, _ = O.trace_bvh4(nodes, tris, r, mode=O.ANY)
    aq, _ = O.trace_bvh4(d, tris, r, mode=O.ANY)
    assert ((ae != O.MISS) == (aq != O.MISS)).all()
