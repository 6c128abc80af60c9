"""Pins of the CPU oracle that need no reference build: the reference's
published traversal statistics, its known-answer rays, brute force, closed
forms and invariants.  Each check is chosen so a plausible slip (a dropped
term, a wrong sign, an off-by-one in an index) fails it."""
import numpy as np
import pytest

# results/exp0_ray_stats.csv:2-7 (SoftwareBVH4 traversal on the reference's
# GPU: SW hit rate %, mean box tests per ray, mean triangle tests per ray).
EXP0 = {
    (968, "Coherent"): dict(hit=77.751350, box=18.966507, tri=7.359565),
    (968, "Incoherent"): dict(hit=99.988083, box=20.437469, tri=8.226969),
    (99458, "Coherent"): dict(hit=77.752686, box=38.740490, tri=6.158594),
    (99458, "Incoherent"): dict(hit=99.987885, box=41.927849, tri=7.011395),
    (999698, "Coherent"): dict(hit=77.751251, box=40.599850, tri=7.351475),
    (999698, "Incoherent"): dict(hit=99.987701, box=43.775719, tri=8.202208),
}


def _exp0(O, target, fused=True):
    pos, idx = O.terrain(target, fused=fused)
    n4, tris = O.software_bvh4(pos, idx)
    out = {}
    for coh in ("Coherent", "Incoherent"):
        r = O.exp0_rays(1 << 20, coh == "Coherent")
        prim, _, box, tri = O.trace_bvh4(n4, tris, r, stack_limit=24, counters=True)
        out[coh] = ((prim != O.MISS).mean() * 100, box.mean(), tri.mean())
    return len(idx), out


@pytest.mark.parametrize("target", [1000, 100000, 1000000])
def test_exp0_published_statistics(O, target):
    """Terrain (createTessellatedTerrain as the reference's -O3 -march=native
    build evaluates it) + SoftwareBVH::build + fromBVH2 + BVH4 traversal with
    the shader's pcg rays and 24-entry stack reproduce the published means to
    ~6 significant digits (remaining noise: the GPU's ray-setup rounding)."""
    ntri, got = _exp0(O, target)
    for coh, (hit, box, tri) in got.items():
        ref = EXP0[(ntri, coh)]
        assert abs(hit - ref["hit"]) < 1e-3, (coh, hit, ref)
        assert abs(box - ref["box"]) / ref["box"] < 2e-5, (coh, box, ref)
        assert abs(tri - ref["tri"]) / ref["tri"] < 2e-5, (coh, tri, ref)


def test_exp0_pin_is_sensitive(O):
    """The same pipeline on the unfused terrain (one rounding per operator)
    moves the 968-triangle coherent box-test mean by ~3e-4 relative, which
    the tolerance above rejects: the pin distinguishes a one-ulp change in
    the mesh."""
    ntri, got = _exp0(O, 1000, fused=False)
    box = got["Coherent"][1]
    ref = EXP0[(ntri, "Coherent")]["box"]
    assert abs(box - ref) / ref > 2e-5


def _rays(O, n, o, d, tmin=0.001, tmax=100.0):
    r = O.Rays(n)
    r.ox[:], r.oy[:], r.oz[:] = o
    r.dx[:], r.dy[:], r.dz[:] = d
    r.tmin[:] = tmin
    r.tmax[:] = tmax
    return r


def test_known_answer_probe(O):
    """probe_main.cpp:45-50,182-211."""
    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    n4, tris = O.software_bvh4(pos, idx)
    p, t = O.trace_bvh4(n4, tris, _rays(O, 4, (0, -0.5, -2), (0, 0, 1)))
    assert (p == 0).all() and (np.abs(t - 2.0) < 0.01).all()
    p, t = O.trace_bvh4(n4, tris, _rays(O, 4, (10, 10, -2), (0, 0, 1)))
    assert (p == O.MISS).all()


def test_known_answer_forged(O):
    """src/main.cpp:269-313,410-414: hit at t = 5 (+-1e-3) from the origin."""
    pos = np.array([[-2, -2, 5], [2, -2, 5], [0, 2, 5]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    n4, tris = O.software_bvh4(pos, idx)
    p, t = O.trace_bvh4(n4, tris, _rays(O, 1, (0, 0, 0), (0, 0, 1)))
    assert p[0] == 0 and abs(t[0] - 5.0) < 1e-3
    p, t = O.trace_bvh4(n4, tris, _rays(O, 1, (10, 10, 0), (0, 0, 1)))
    assert p[0] == O.MISS


@pytest.mark.parametrize("n", [64, 256, 1024, 4096])
def test_fast_builder_verification_rays(O, n):
    """src/main.cpp:1002-1106 on the LBVH restatement: the +z ray from each
    centroid (cx, cy, 0) hits with |t - cz| < 1e-2; (10, 10, 0) misses."""
    xyz, cen = O.dynamic_grid(n, 0.0, 1.25)
    nodes, tris, _, _ = O.lbvh_build(xyz)
    r = O.Rays(n + 1)
    r.ox[:n], r.oy[:n] = cen[:, 0], cen[:, 1]
    r.ox[n], r.oy[n] = 10.0, 10.0
    r.dz[:] = 1.0
    r.tmax[:] = 1e30
    p, t = O.trace_bvh4(nodes, tris, r)
    assert (p[:n] != O.MISS).all() and (np.abs(t[:n] - cen[:, 2]) < 1e-2).all()
    assert p[n] == O.MISS


@pytest.mark.parametrize("builder", ["software", "lbvh", "rlbvh"])
def test_traversal_equals_brute_force(O, builder):
    """Traversal result is independent of the tree: the closest t equals the
    brute-force minimum over every triangle, any-hit equals t < tmax."""
    pos, idx = O.sphere(40, 30, 1.0, 0.02, 5)
    if builder == "software":
        n4, tris = O.software_bvh4(pos, idx)
    elif builder == "lbvh":
        n4, tris, _, _ = O.lbvh_build(pos, idx)
    else:
        n4, tris, _, _ = O.rlbvh_build(pos, idx)
    r = O.random_rays(20000, 3, (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5), tmax=1e30)
    r.tmax[::2] = 0.3
    p, t = O.trace_bvh4(n4, tris, r)
    pb, tb = O.trace_brute(tris, r)
    same_t = t.view(np.uint32) == tb.view(np.uint32)
    assert same_t.mean() > 0.9999
    assert ((p == pb) | ~same_t).mean() > 0.9995
    pa, _ = O.trace_bvh4(n4, tris, r, mode=O.ANY)
    pba, _ = O.trace_brute(tris, r, mode=O.ANY)
    assert ((pa != O.MISS) == (pba != O.MISS)).mean() > 0.9999
    assert ((pa != O.MISS) == (p != O.MISS)).all()


def test_stack_limit_drop_only_loses_hits(O):
    """The shader's 24-entry drop (software_bvh_bvh4.comp:185-187) can only
    lose work: with a tiny limit no ray finds a closer hit than unbounded."""
    pos, idx = O.terrain(20000)
    n4, tris = O.software_bvh4(pos, idx)
    r = O.exp0_rays(50000, False)
    p0, t0 = O.trace_bvh4(n4, tris, r)
    p1, t1 = O.trace_bvh4(n4, tris, r, stack_limit=2)
    assert (t1 >= t0).all()
    assert (p1 != p0).any()


def test_morton_closed_form(O):
    for v in [0, 1, 2, 3, 511, 512, 1023]:
        e = O.expand_bits10(v)
        want = 0
        for b in range(10):
            want |= ((v >> b) & 1) << (3 * b)
        assert e == want
    # three triangles on distinct axes: x gets the top bit of each triple
    xyz = np.zeros((3, 9), np.float32)
    cs = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    for i, c in enumerate(cs):
        xyz[i] = np.tile(np.array(c, np.float32), 3)
    lo = np.zeros(3, np.float32)
    b = np.concatenate([lo, np.ones(3, np.float32)])
    codes = O.morton30(xyz, None, b)
    assert codes[0] == 0b100 << 27 or codes[0] == O.expand_bits10(1023) << 2
    assert codes[1] == O.expand_bits10(1023) << 1
    assert codes[2] == O.expand_bits10(1023)


def test_morton_per_axis_brute(O):
    """Quantisation q = clamp(floor((c - lo) * 1024 / (hi - lo)), 0, 1023)
    against a per-axis numpy evaluation in float32 with the same rounding."""
    rng = np.random.default_rng(1)
    xyz = rng.uniform(-3, 5, size=(5000, 9)).astype(np.float32)
    b = O.centroid_bounds(xyz)
    codes = O.morton30(xyz, None, b)
    c = ((xyz[:, 0:3] + xyz[:, 3:6]) + xyz[:, 6:9]) / np.float32(3)
    lo, hi = b[:3], b[3:]
    s = np.float32(1024) / (hi - lo)
    q = np.clip((c - lo) * s, 0, 1023).astype(np.uint32)
    want = np.zeros(len(c), np.uint32)
    for bit in range(10):
        for a, sh in ((0, 2), (1, 1), (2, 0)):
            want |= ((q[:, a] >> bit) & 1) << (3 * bit + sh)
    np.testing.assert_array_equal(codes, want)
    np.testing.assert_array_equal(b[:3], c.min(0))
    np.testing.assert_array_equal(b[3:], c.max(0))


def test_sort_is_stable_sort(O):
    rng = np.random.default_rng(2)
    k = rng.integers(0, 1 << 30, 100000).astype(np.uint32)
    k[::7] = k[0]
    v = np.arange(len(k), dtype=np.uint32)
    ks, vs = O.sort_pairs(k, v)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(ks, k[order])
    np.testing.assert_array_equal(vs, order.astype(np.uint32))


@pytest.mark.parametrize("n", [0, 1, 4, 5, 16, 17, 4096, 4097, 2_500_000])
def test_tree_layout_closed_form(O, n):
    counts, offs, tb, sz = O.tree_layout(n)
    want = [(n + 3) // 4]
    while want[-1] > 1:
        want.append((want[-1] + 3) // 4)
    assert list(counts) == want[::-1]
    assert list(offs) == list(np.concatenate([[0], np.cumsum(want[::-1])[:-1]]))
    assert tb == sum(want) and sz == 192 + 128 * tb + 64 * n


@pytest.mark.parametrize("n", [1, 5, 4096, 4097, 70000])
def test_lbvh_invariants(O, n):
    pos, idx = O.sphere(100, 360, 1.0, 0.02, 4)
    idx = idx[:n]
    nodes, tris, ids, codes = O.lbvh_build(pos, idx)
    assert sorted(ids.tolist()) == list(range(n))
    assert (np.diff(codes.astype(np.int64)) >= 0).all()
    # ties keep index order
    tie = np.diff(codes.astype(np.int64)) == 0
    assert (np.diff(ids.astype(np.int64))[tie] > 0).all()
    np.testing.assert_array_equal(O.morton30(pos, idx)[ids], codes)
    assert (tris["triangle_id"] == ids).all()
    counts, offs, tb, _ = O.tree_layout(n)
    assert len(nodes) == tb
    # every box slot equals the union of its child's slots
    for lv in range(len(counts) - 1):
        for i in range(counts[lv]):
            nd = nodes[offs[lv] + i]
            for c in range(4):
                if nd["count"][c] == 0:
                    continue
                ch = nodes[nd["child"][c]]
                m = ch["count"] != 0
                assert nd["minX"][c] == ch["minX"][m].min() and nd["maxZ"][c] == ch["maxZ"][m].max()
    # root slots union = bounds of all vertices used
    root = nodes[0]
    m = root["count"] != 0
    v = pos[idx.reshape(-1)]
    assert root["minX"][m].min() == v[:, 0].min() and root["maxY"][m].max() == v[:, 1].max()


def _radix_split_brute(codes, f, l):
    """Split of sorted range [f, l] by definition: the highest bit (of the
    64-bit key code << 32 | position) that differs between f and l; the left
    part is every position whose bit is 0."""
    key = lambda i: (int(codes[i]) << 32) | i
    b = (key(f) ^ key(l)).bit_length() - 1
    s = f
    while s + 1 <= l and not (key(s + 1) >> b) & 1:
        s += 1
    return s


@pytest.mark.parametrize("maxp", [1, 2, 4])
def test_rlbvh_structure(O, maxp):
    """Radix-tree LBVH: every BVH2 split is the highest-differing-bit split of
    its range (brute force on the 64-bit key, position as tie-break), leaves
    hold <= maxp triangles, inner boxes are unions, and the Box4 nodes are
    SoftwareBVH4::collapse of that BVH2 (node count = inner BVH2 nodes
    reached)."""
    pos, idx = O.terrain(3000)
    idx = np.concatenate([idx, idx[:500]])  # duplicate triangles -> equal codes
    n = len(idx)
    n4, tris, ids, codes, n2 = O.rlbvh_build(pos, idx, max_prims=maxp, with_bvh2=True)
    assert sorted(ids.tolist()) == list(range(n))
    np.testing.assert_array_equal(O.morton30(pos, idx)[ids], codes)
    np.testing.assert_array_equal(n4.tobytes(), O.bvh4_collapse(n2).tobytes())

    def rng(u):
        nd = n2[u]
        if nd["leftChild"] < 0:
            f = ~int(nd["leftChild"])
            return f, f + int(nd["rightChild"]) - 1
        return rng(int(nd["leftChild"]))[0], rng(int(nd["rightChild"]))[1]

    stack = [0]
    while stack:
        u = stack.pop()
        nd = n2[u]
        f, l = rng(u)
        if nd["leftChild"] < 0:
            assert int(nd["rightChild"]) <= maxp
            continue
        assert l - f + 1 > maxp
        a, b = int(nd["leftChild"]), int(nd["rightChild"])
        assert rng(a) == (f, _radix_split_brute(codes, f, l)) and rng(b)[1] == l
        np.testing.assert_array_equal(nd["aabbMin"], np.minimum(n2[a]["aabbMin"], n2[b]["aabbMin"]))
        np.testing.assert_array_equal(nd["aabbMax"], np.maximum(n2[a]["aabbMax"], n2[b]["aabbMax"]))
        stack += [a, b]
    assert len(n4) <= n - 1


def test_lbvh_boxes_have_no_negative_zero(O):
    """Box bounds are canonical: a mesh with -0.0 coordinates yields +0 bounds
    (min/max of +0 and -0 may return either zero on a given device)."""
    pos = np.array([[-0.0, 0.0, 1.0], [0.0, -0.0, 1.0], [1.0, 1.0, -0.0],
                    [0.0, 0.0, 0.0], [-0.0, 2.0, -0.0]], np.float32)
    idx = np.array([[0, 1, 2], [1, 3, 4], [2, 4, 0], [0, 3, 2], [4, 1, 0]], np.uint32)
    nodes, _, _, _ = O.lbvh_build(pos, idx)
    for f in ("minX", "maxX", "minY", "maxY", "minZ", "maxZ"):
        v = nodes[f].reshape(-1)
        assert not ((v == 0) & np.signbit(v)).any(), f


def test_lbvh_refit_equals_rebuild_of_same_order(O):
    xyz0, _ = O.dynamic_grid(4096, 0.0, 2.0)
    xyz1, _ = O.dynamic_grid(4096, 1.0, 2.0)
    nodes, tris, ids, _ = O.lbvh_build(xyz0)
    n1, t1 = O.lbvh_refit(xyz1, None, nodes, tris)
    assert (t1["triangle_id"] == ids).all()
    np.testing.assert_array_equal(t1["v"], xyz1[ids])
    # refit with the original positions is the identity
    n0, t0 = O.lbvh_refit(xyz0, None, nodes, tris)
    assert n0.tobytes() == nodes.tobytes() and t0.tobytes() == tris.tobytes()


def test_radv_encode_layout(O):
    """fast_bvh_builder.comp byte image: header constants (:220-256), child
    IDs (rel >> 3) | type (:141-143, :190-191), NaN empty slots (:150-154),
    triangle nodes with ID / opaque flag / id 9 (:128-139)."""
    n = 4096
    xyz, _ = O.dynamic_grid(n, 0.0, 1.25)
    img = O.radv_encode(xyz)
    counts, offs, tb, sz = O.tree_layout(n)
    assert len(img) == sz == 437056 + 0  # README.md:51 size class: 192 + 128*1365 + 64*4096
    h = img[:192].view(np.uint32)
    assert h[0] == 192 and h[0x20 // 4] == 8 and h[0x7C // 4] == 4 and h[0x80 // 4] == 1
    assert img[0x30:0x38].view(np.uint64)[0] == sz and img[0x38:0x40].view(np.uint64)[0] == sz + 128
    assert img[0x60:0x68].view(np.uint64)[0] == 192 + 128 * tb
    assert img[0x68:0x70].view(np.uint64)[0] == n
    assert h[180 // 4] == 5 and h[188 // 4] == 0xFFFFFFFF
    leaf = img[192 + 128 * tb:].view(np.uint32).reshape(n, 16)
    np.testing.assert_array_equal(leaf[:, :9].view(np.float32), xyz)
    assert (leaf[:, 12] == np.arange(n)).all() and (leaf[:, 13] == 0x80000000).all()
    assert (leaf[:, 15] == 9).all()
    box = img[192:192 + 128 * tb].view(np.uint32).reshape(tb, 32)
    # root children point at the 4 level-1 boxes, type 5
    assert [int(x) for x in box[0, :4]] == [((offs[1] + c) * 128 >> 3) | 5 for c in range(4)]
    # first leaf box points at triangles 0..3, type 0
    lb = offs[-1]
    assert [int(x) for x in box[lb, :4]] == [(128 * tb + 64 * c) >> 3 for c in range(4)]
    # scene bounds in the header = all vertices
    v = xyz.reshape(-1, 3)
    np.testing.assert_array_equal(img[8:20].view(np.float32), v.min(0))
    np.testing.assert_array_equal(img[20:32].view(np.float32), v.max(0))
    # a partial leaf box has NaN slots with child 0xFFFFFFFF
    img5 = O.radv_encode(xyz[:5])
    _, o5, tb5, _ = O.tree_layout(5)
    b5 = img5[192:192 + 128 * tb5].view(np.uint32).reshape(tb5, 32)
    last = b5[o5[-1] + 1]
    assert list(last[1:4]) == [0xFFFFFFFF] * 3
    assert np.isnan(last[4 + 6:4 + 24].view(np.float32)).all()


def test_blur_fixed_point_and_weights(O):
    w, h = 37, 23
    img = np.full(w * h, 0.625, np.float32)
    depth = np.linspace(0, 1, w * h).astype(np.float32)
    out = O.blur7(img, depth, w, h, 10.0)
    np.testing.assert_allclose(out, 0.625, rtol=0, atol=1e-6)
    # an isolated depth island is not blurred across
    depth2 = np.zeros(w * h, np.float32)
    img2 = np.zeros(w * h, np.float32)
    c = 11 * w + 18
    depth2[c], img2[c] = 5.0, 1.0
    out2 = O.blur7(img2, depth2, w, h, 0.05)
    assert out2[c] == 1.0 and out2[c + 1] == 0.0
    # impulse response on a flat depth: 2-D weights w|i| w|j| / sum
    img3 = np.zeros(w * h, np.float32)
    img3[c] = 1.0
    out3 = O.blur7(img3, np.zeros(w * h, np.float32), w, h, 0.05)
    wt = np.array([0.227027, 0.1945946, 0.1216216, 0.054054], np.float32)
    tot = wt[[3, 2, 1, 0, 1, 2, 3]].sum() ** 2
    assert abs(out3[c] - wt[0] * wt[0] / tot) < 1e-6
    assert abs(out3[c + 2] - wt[2] * wt[0] / tot) < 1e-6


def test_percentile_convention(O):
    """common/vk_timer.cpp:106-107: sorted[floor(q * (n - 1))]."""
    v = np.arange(100, dtype=np.float32)[::-1].copy()
    assert O.percentile(v, 0.99) == 98.0
    assert O.percentile(v, 0.5) == 49.0
    assert O.percentile(np.arange(200, dtype=np.float32), 0.99) == 197.0


def test_generators_closed_forms(O):
    pos, idx = O.sphere(64, 32, 1.0, 0.02, 7)
    r = np.linalg.norm(pos.astype(np.float64), axis=1)
    assert (r >= 0.98 - 1e-6).all() and (r <= 1.02 + 1e-6).all()
    assert idx.max() < len(pos) and len(idx) == 2 * 64 * 32
    p, i = O.wave_grid(64, 64, 1.0, 0.0, 0.1, 4.0, 4.0)
    x, y = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64)
    np.testing.assert_allclose(p[:, 2], 0.1 * np.sin(4 * x) * np.cos(4 * y), atol=2e-7)
    assert len(i) == 8192
    cam = O.camera_rays(64, 48, (0, 0, 2), (0, 0, 0), (0, 1, 0), 60.0)
    d = np.stack([cam.dx, cam.dy, cam.dz], 1).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # centre rays point down -z; the top row's vertical angle is ~half the FOV
    top = cam.dy[32]
    assert abs(np.degrees(np.arcsin(top)) - 30.0 * (1 - 1 / 48)) < 0.5
    rr = O.random_rays(10000, 3, (-1, -2, -3), (1, 2, 3))
    d = np.stack([rr.dx, rr.dy, rr.dz], 1).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert abs(d.mean(0)).max() < 0.05 and rr.ox.min() >= -1 and rr.oz.max() < 3


def _one_at_a_time(data):
    """Jenkins' one-at-a-time hash of a byte string, as published (Dr. Dobb's
    1997): per byte h += b; h += h << 10; h ^= h >> 6; then h += h << 3;
    h ^= h >> 11; h += h << 15 (mod 2^32)."""
    h = 0
    for b in data:
        h = (h + b) & 0xFFFFFFFF
        h = (h + (h << 10)) & 0xFFFFFFFF
        h ^= h >> 6
    h = (h + (h << 3)) & 0xFFFFFFFF
    h ^= h >> 11
    return (h + (h << 15)) & 0xFFFFFFFF


def test_ao_hash_is_one_at_a_time(O):
    """The AO sampler's hash (rq_ao.comp:18-25) is the one-at-a-time hash of
    its one key value: equal to the published algorithm, itself checked on
    its published vectors, on every one-byte key (each shift and add shows)."""
    assert _one_at_a_time(b"a") == 0xCA2E9442
    assert _one_at_a_time(b"The quick brown fox jumps over the lazy dog") == 0x519E91F5
    for x in range(256):
        assert O.ao_hash(x) == _one_at_a_time(bytes([x])), x
    # a wrong shift in the oracle's copy would fail the loop above: the
    # planted mistake, checked on the Python model
    assert any(_one_at_a_time(bytes([x])) != _oaat_slip(x) for x in range(256))


def _oaat_slip(x):
    h = x
    h = (h + (h << 10)) & 0xFFFFFFFF
    h ^= h >> 5  # 6 in the algorithm
    h = (h + (h << 3)) & 0xFFFFFFFF
    h ^= h >> 11
    return (h + (h << 15)) & 0xFFFFFFFF


def test_shadow_and_ao_rays(O):
    pos, idx = O.sphere(64, 48, 1.0, 0.02, 7)
    cam = O.camera_rays(40, 30, (0, 0, 1.25), (0, 0, 0), (0, 1, 0), 60.0)
    n4, tris = O.software_bvh4(pos, idx)
    prim, t = O.trace_bvh4(n4, tris, cam)
    assert (prim != O.MISS).all()
    lights = np.array([[0, 0, 3.0], [3.0, 0, 0]], np.float32)
    sr = O.shadow_rays(pos, idx, cam, prim, t, lights, 1e-4)
    assert len(sr) == 2 * 1200
    d = np.stack([sr.dx, sr.dy, sr.dz], 1)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # origin + tmax * dir lands on the light (light-major layout)
    end = np.stack([sr.ox, sr.oy, sr.oz], 1) + sr.tmax[:, None] * d
    np.testing.assert_allclose(end[:1200], np.broadcast_to(lights[0], (1200, 3)), atol=1e-5)
    np.testing.assert_allclose(end[1200:], np.broadcast_to(lights[1], (1200, 3)), atol=1e-5)
    ao = O.ao_rays(pos, idx, cam, prim, t, 40, 30, 0, 0.003, 0.001, 0.5)
    # AO directions lie in the hemisphere facing the camera side of the surface
    d = np.stack([ao.dx, ao.dy, ao.dz], 1)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    tri = pos[idx[prim]].astype(np.float64)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cd = np.stack([cam.dx, cam.dy, cam.dz], 1)
    nrm[(nrm * cd).sum(1) > 0] *= -1
    assert ((d * nrm).sum(1) >= -1e-6).all()
    # cosine weighting: E[cos] = 2/3 over the hemisphere
    assert abs((d * nrm).sum(1).mean() - 2 / 3) < 0.03
