"""Serving loop (serve.py) on the GPU.

  * same alone and in the mix: every request's result served in the mix
    (admitted alongside others, traced in shared batched launches, a tick's
    rebuilds in one batched build on a stream of their own) equals its
    result served alone on the plain path (one launch per request, waiting
    for each): ray results bit for bit, rebuilt grids and trees byte for
    byte, and the same scene version traced;
  * rebuilds change resident scenes: a ray request traces the version left
    by the rebuilds of its scene that arrived before it (checked against the
    oracle's tree of that version's vertices);
  * the serving configuration at full size: 32 resident scenes of 10k-2.5M
    triangles plus 8 deformable grids, 4k-2M-ray shadow / AO / closest
    requests and 4k-64k-triangle deform-and-rebuilds, Poisson arrivals with a
    tight budget and all at once; sampled rays of at least 16 requests of
    each kind and at least 16 rebuilt trees compared with the oracle;
  * all requests finish, the memory budget holds, a request larger than the
    budget is refused.
"""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serve(rt):
    import serve as S
    return S


@pytest.fixture(scope="module")
def scenes(serve):
    return serve.make_scenes(count=4, lo=10_000, hi=200_000, seed=3, grids=3)


REFILL = dict(schedule="refill", rays_per_wave=256, refill_lanes=8, block=64, lds_stack=16)


def _initial_grids(scenes):
    """Vertices of every grid scene's first version (set 0 after reset)."""
    scenes.reset()
    return {len(scenes.static) + j: to_np(g.sets[0]["pos"][: g.n0 * 9]).reshape(-1, 9).copy()
            for j, g in enumerate(scenes.grids)}


class OracleTrees:
    """The oracle's view of each scene version a request traced: static
    scenes downloaded once; a grid version rebuilt by the oracle (implicit
    LBVH) from that version's vertices."""

    def __init__(self, O, scenes, reqs, initial):
        self.O, self.scenes, self.initial = O, scenes, initial
        self.by_rid = {r.rid: r for r in reqs}
        self.cache = {}

    def tree(self, version):
        sid, _, n, src = version
        key = (sid, src)
        if key not in self.cache:
            O = self.O
            if not self.scenes.is_grid(sid):
                b = self.scenes.static[sid][0]
                self.cache[key] = (to_np(b.nodes)[: b.node_count * 128].view(O.NODE4),
                                   to_np(b.tris)[: b.tri_count * 64].view(O.TRI64))
            else:
                pos = self.initial[sid] if src is None else \
                    to_np(self.by_rid[src].out[2]).reshape(-1, 9)
                assert pos.shape[0] == n
                nodes, tris, _, _ = O.lbvh_build(pos, None)
                self.cache[key] = (nodes, tris)
        return self.cache[key]


def _check_against_oracle(O, serve, scenes, reqs, initial, per_kind, rays_per_req=2048):
    """Sampled rays of per_kind requests of each ray kind and per_kind
    rebuilt trees against the oracle; returns the counts checked."""
    orc = OracleTrees(O, scenes, reqs, initial)
    rng = np.random.default_rng(11)
    done = {k: 0 for k in serve.KINDS}
    for r in reqs:
        if done[r.kind] >= per_kind:
            continue
        if r.kind == "rebuild":
            nodes_g, tris_g, pos_g = (to_np(x) for x in r.out)
            nodes_c, tris_c, _, _ = O.lbvh_build(pos_g.reshape(-1, 9), None)
            np.testing.assert_array_equal(nodes_g, nodes_c.view(np.uint8).reshape(-1))
            np.testing.assert_array_equal(tris_g, tris_c.view(np.uint8).reshape(-1))
            # the deform step: the grid of generate_dynamic_grid_mesh at the request's time
            want, _ = O.dynamic_grid(r.n, serve.GRID_STEP * r.step, serve.GRID_EXTENT)
            np.testing.assert_allclose(pos_g.reshape(-1, 9), want, rtol=0, atol=2e-6)
        else:
            nodes, tris = orc.tree(r.version)
            sel = np.sort(rng.choice(r.n, min(r.n, rays_per_req), replace=False))
            rays = O.Rays.stack(to_np(r.rays.data)[:, sel])
            if r.kind == "closest":
                pc, tc = O.trace_bvh4(nodes, tris, rays)
                np.testing.assert_array_equal(u32(r.out[0])[sel], pc)
                np.testing.assert_array_equal(to_np(r.out[1])[sel].view(np.uint32),
                                              tc.view(np.uint32))
            else:
                pa, _ = O.trace_bvh4(nodes, tris, rays, mode=O.ANY)
                b = u32(r.out)
                np.testing.assert_array_equal(((b[sel >> 5] >> (sel & 31)) & 1).astype(bool),
                                              pa != O.MISS)
        done[r.kind] += 1
    return done


@pytest.mark.parametrize("qps,budget,shape", [(0, 1 << 30, None), (4000, 64 << 20, None),
                                              (0, 1 << 30, REFILL)],
                         ids=["saturated", "poisson-tight-budget", "saturated-refill"])
def test_same_alone_and_in_mix(serve, scenes, qps, budget, shape):
    reqs = serve.make_requests(64, qps, scenes, seed=21, rays=(1_000, 300_000),
                               rebuild_tris=(1_000, 40_000))
    srv = serve.Server(scenes, budget_bytes=budget, keep_results=True, trace_shape=shape)
    print(f"serving {len(reqs)} requests qps={qps} budget={budget}", flush=True)
    st = srv.run(reqs)
    assert st["requests"] == len(reqs)
    assert srv.peak <= budget and srv.used == 0
    alone = serve.clone_requests(reqs)
    serve.run_plain(scenes, alone, keep_results=True)
    torch.cuda.synchronize()
    assert serve.same_results(reqs, alone) == []
    assert st["trace_launches"] < sum(1 for r in reqs if r.kind != "rebuild")
    assert st["rebuilds"] > 0


def test_rebuilds_change_the_resident_scene(serve, scenes, O):
    """A fixed order on one grid scene, all arriving at once: trace, rebuild
    (size n1, step a), trace, rebuild, rebuild, trace, trace, rebuild,
    trace, and a static-scene trace between.  Each trace sees the version of
    the rebuilds before it (a grid's third rebuild in one tick waits for
    the next tick); results equal the oracle's tree of that version."""
    g = len(scenes.static)
    R = serve.Request
    seq = [R(0, 0.0, "closest", g, 50_000, 1), R(1, 0.0, "rebuild", g, 12_000, 57),
           R(2, 0.0, "closest", g, 50_000, 2), R(3, 0.0, "shadow", 0, 30_000, 3),
           R(4, 0.0, "rebuild", g, 40_000, 199), R(5, 0.0, "rebuild", g, 4_096, 3),
           R(6, 0.0, "ao", g, 70_000, 4), R(7, 0.0, "closest", g, 9_000, 5),
           R(8, 0.0, "rebuild", g, 6_000, 77), R(9, 0.0, "closest", g, 5_000, 6)]
    assert serve.GRID_TICK_REBUILDS == 2
    initial = _initial_grids(scenes)
    srv = serve.Server(scenes, budget_bytes=1 << 30, keep_results=True)
    srv.run(seq)
    torch.cuda.synchronize()
    assert [seq[k].version[3] for k in (0, 2, 6, 7, 9)] == [None, 1, 5, 5, 8]
    assert [seq[k].version[2] for k in (0, 2, 6, 7, 9)] == [scenes.grids[0].n0, 12_000, 4_096,
                                                            4_096, 6_000]
    assert srv.build_launches == 2  # rebuilds 5 and 8 went out in the next tick
    alone = serve.clone_requests(seq)
    serve.run_plain(scenes, alone, keep_results=True)
    torch.cuda.synchronize()
    assert serve.same_results(seq, alone) == []
    done = _check_against_oracle(O, serve, scenes, seq, initial, per_kind=8)
    assert done["rebuild"] == 4 and done["closest"] == 4


def test_budget_too_small_is_refused(serve, scenes):
    reqs = serve.make_requests(2, 0, scenes, seed=1, rays=(100_000, 100_001), kinds=("closest",))
    with pytest.raises(RuntimeError):
        serve.Server(scenes, budget_bytes=1 << 20).run(reqs)


@pytest.fixture(scope="module")
def full_scenes(serve):
    print("serving configuration: 32 scenes of 10k-2.5M triangles + 8 grids", flush=True)
    s = serve.make_scenes(count=32, lo=10_000, hi=2_500_000, seed=5, grids=8)
    assert max(s.tris) > 1_000_000 and min(s.tris) < 30_000
    return s


@pytest.mark.parametrize("qps,budget", [(0, 4 << 30), (3000, 512 << 20)],
                         ids=["all-at-once", "poisson-tight-budget"])
def test_serving_configuration_full_size(serve, full_scenes, O, qps, budget):
    """The serving configuration's own workload: 240 requests of every kind
    over the 40 resident scenes; every request the same alone and in the
    mix; sampled rays of 16 requests of each ray kind and 16 rebuilt trees
    equal the oracle's."""
    scenes = full_scenes
    reqs = serve.make_requests(240, qps, scenes, seed=17)
    assert max(r.n for r in reqs if r.kind != "rebuild") > 1_000_000
    initial = _initial_grids(scenes)
    srv = serve.Server(scenes, budget_bytes=budget, keep_results=True)
    print(f"full-size serving qps={qps} budget={budget >> 20} MiB", flush=True)
    st = srv.run(reqs)
    print(st, flush=True)
    assert st["requests"] == len(reqs) and srv.used == 0 and srv.peak <= budget
    alone = serve.clone_requests(reqs)
    serve.run_plain(scenes, alone, keep_results=True)
    torch.cuda.synchronize()
    assert serve.same_results(reqs, alone) == []
    done = _check_against_oracle(O, serve, scenes, reqs, initial, per_kind=16)
    assert all(done[k] == 16 for k in serve.KINDS), done
    # rays traced scene versions written by earlier rebuilds
    assert any(r.version[3] is not None for r in reqs if r.kind != "rebuild")
