"""The oracle's treelet restructuring (or_rlbvh_build_opt, Karras & Aila
2013), pinned without the device:

  * on meshes of 5-7 triangles with one triangle per leaf, one round makes
    the root's treelet the whole tree, so the result's SAH cost must equal
    the least SAH cost over EVERY binary tree on those leaves, found here by
    enumerating all of them (945 / 10,395 trees);
  * on larger meshes each round never raises the tree's SAH cost, every
    node's box is the union of its children's, every leaf range appears
    exactly once, and the Box4 tree traces to the same hits as the plain
    radix tree.
"""
import itertools

import numpy as np
import pytest

INNER = 1.2


def _area(lo, hi):
    d = np.asarray(hi, np.float64) - np.asarray(lo, np.float64)
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def _tree_cost(n2, x=0):
    nd = n2[x]
    a = _area(nd["aabbMin"], nd["aabbMax"])
    if nd["leftChild"] < 0:
        return a * nd["rightChild"]
    return INNER * a + _tree_cost(n2, nd["leftChild"]) + _tree_cost(n2, nd["rightChild"])


def _leaves(n2, x=0, out=None):
    out = [] if out is None else out
    nd = n2[x]
    if nd["leftChild"] < 0:
        out.append((int(~nd["leftChild"]), int(nd["rightChild"])))
    else:
        _leaves(n2, nd["leftChild"], out)
        _leaves(n2, nd["rightChild"], out)
    return out


def _all_trees(items):
    """Every full binary tree over the leaf list (unordered children)."""
    if len(items) == 1:
        yield items[0]
        return
    first, rest = items[0], items[1:]
    for k in range(len(rest)):
        for left_rest in itertools.combinations(rest, k):
            right = [x for x in rest if x not in left_rest]
            for a in _all_trees([first, *left_rest]):
                for b in _all_trees(right):
                    yield (a, b)


def _cost_of(t, boxes):
    """(cost, lo, hi) of a nested-tuple tree over leaf indices."""
    if isinstance(t, int):
        lo, hi = boxes[t]
        return _area(lo, hi), lo, hi
    ca, la, ha = _cost_of(t[0], boxes)
    cb, lb, hb = _cost_of(t[1], boxes)
    lo, hi = np.minimum(la, lb), np.maximum(ha, hb)
    return INNER * _area(lo, hi) + ca + cb, lo, hi


@pytest.mark.parametrize("n,seed", [(5, 1), (6, 2), (7, 3), (7, 4), (7, 5)])
def test_one_round_is_optimal_over_all_trees(O, n, seed):
    rng = np.random.default_rng(seed)
    tris = (rng.random((n, 9)) * np.array([4, 1, 2] * 3) - 1).astype(np.float32)
    _, _, _, _, plain = O.rlbvh_build(tris, None, 1, with_bvh2=True)
    _, _, _, _, n2 = O.rlbvh_build(tris, None, 1, with_bvh2=True, treelet_rounds=1)
    leaves = _leaves(plain)
    assert len(leaves) == n and sorted(_leaves(n2)) == sorted(leaves)
    boxes = {}
    for x in range(len(plain)):
        nd = plain[x]
        if nd["leftChild"] < 0:
            boxes[int(~nd["leftChild"])] = (np.array(nd["aabbMin"], np.float64),
                                            np.array(nd["aabbMax"], np.float64))
    best = min(_cost_of(t, boxes)[0] for t in _all_trees(list(range(n))))
    got = _tree_cost(n2)
    assert got == pytest.approx(best, rel=1e-5)
    assert got <= _tree_cost(plain) * (1 + 1e-6)


def _check_tree(n2, n):
    seen = np.zeros(n, np.int32)
    for first, cnt in _leaves(n2):
        seen[first: first + cnt] += 1
    assert (seen == 1).all()
    for nd in n2:
        if nd["leftChild"] >= 0:
            a, b = n2[nd["leftChild"]], n2[nd["rightChild"]]
            np.testing.assert_array_equal(nd["aabbMin"], np.minimum(a["aabbMin"], b["aabbMin"]))
            np.testing.assert_array_equal(nd["aabbMax"], np.maximum(a["aabbMax"], b["aabbMax"]))


@pytest.mark.parametrize("maxp", [1, 2, 4])
def test_rounds_lower_cost_and_keep_hits(O, maxp):
    pos, idx = O.sphere(80, 60, 1.0, 0.05, 2)
    n = len(idx)
    r = O.random_rays(20_000, 4, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
    ref = None
    last = None
    for rounds in (0, 1, 2, 3):
        n4, tris, _, _, n2 = O.rlbvh_build(pos, idx, maxp, with_bvh2=True,
                                           treelet_rounds=rounds)
        _check_tree(n2, n)
        cost = _tree_cost(n2)
        if last is not None:
            assert cost <= last * (1 + 1e-6)
        last = cost
        prim, t = O.trace_bvh4(n4, tris, r)
        bits = O.trace_bvh4(n4, tris, r, mode=O.ANY)[0] != O.MISS
        if ref is None:
            ref = (prim, t, bits)
            cost0 = cost
        np.testing.assert_array_equal(t.view(np.uint32), ref[1].view(np.uint32))
        np.testing.assert_array_equal(bits, ref[2])
        assert ((prim == ref[0]) | (t == ref[1])).all()
    assert last < 0.97 * cost0  # three rounds lower the sphere tree's SAH cost
