"""Serving loop: every request's result is the same served in the mix
(admitted alongside others, traced in shared batched launches, a tick's
rebuilds in one batched build on their own stream) as served alone on the
plain path (one launch per request, waiting for each): ray results bit for
bit, rebuilt trees byte for byte.
Also: all requests finish, the memory budget holds, and a budget too small
for a request is refused."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serve(rt):
    import serve as S
    return S


@pytest.fixture(scope="module")
def scenes(serve):
    return serve.make_scenes(count=4, lo=10_000, hi=200_000, seed=3)


REFILL = dict(schedule="refill", rays_per_wave=256, refill_lanes=8, block=64, lds_stack=16)


@pytest.mark.parametrize("qps,budget,shape", [(0, 1 << 30, None), (4000, 64 << 20, None),
                                              (0, 1 << 30, REFILL)],
                         ids=["saturated", "poisson-tight-budget", "saturated-refill"])
def test_same_alone_and_in_mix(serve, scenes, qps, budget, shape):
    reqs = serve.make_requests(48, qps, scenes, seed=21, rays=(1_000, 300_000),
                               rebuild_tris=(1_000, 40_000))
    srv = serve.Server(scenes, budget_bytes=budget, keep_results=True, trace_shape=shape)
    st = srv.run(reqs)
    assert st["requests"] == len(reqs)
    assert srv.peak <= budget and srv.used == 0
    alone = serve.clone_requests(reqs)
    serve.run_plain(scenes, alone, keep_results=True)
    torch.cuda.synchronize()
    assert serve.same_results(reqs, alone) == []
    assert st["trace_launches"] < sum(1 for r in reqs if r.kind != "rebuild")


def test_budget_too_small_is_refused(serve, scenes):
    reqs = serve.make_requests(2, 0, scenes, seed=1, rays=(100_000, 100_001), kinds=("closest",))
    with pytest.raises(RuntimeError):
        serve.Server(scenes, budget_bytes=1 << 20).run(reqs)
