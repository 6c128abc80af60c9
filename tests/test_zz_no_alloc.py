"""rt_gpu_build_noalloc (the build that allocates nothing: its scratch is a
fixed device array in the library) against rt_gpu_build with a caller-owned
scratch buffer: node, leaf, info, sorted-id and code bytes identical, two
calls in a row with the second smaller (the array keeps the first call's
contents), both tree layouts; sizes beyond the array's limit are refused on
the host."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _build(rt, pos, idx, n, cfg, noalloc):
    cap, need = rt.build_requirements(n, cfg)
    nodes = torch.full((max(cap, 1) * 128,), 0xAB, dtype=torch.uint8, device="cuda")
    tris = torch.full((max(n, 1) * 64,), 0xCD, dtype=torch.uint8, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    ids = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    codes = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    common = (rt.stream_handle(), rt.ptr(pos), rt.ptr(idx), rt.u32(n), C.byref(cfg),
              rt.ptr(nodes), rt.u32(cap), rt.ptr(tris), rt.ptr(info), rt.ptr(ids), rt.ptr(codes))
    if noalloc:
        rt.call("rt_gpu_build_noalloc", *common)
    else:
        scratch = torch.empty(max(need, 256), dtype=torch.uint8, device="cuda")
        rt.call("rt_gpu_build", *common, rt.ptr(scratch), C.c_size_t(scratch.numel()))
        torch.cuda.synchronize()
    return nodes, tris, info, ids, codes


@pytest.mark.parametrize("algo,rounds", [("radix", 0), ("radix", 2), ("implicit", 0)])
def test_noalloc_build_equals_build(rt, algo, rounds):
    cfg = rt.build_cfg(algo, max_leaf_prims=2 if algo == "radix" else 4, treelet_rounds=rounds)
    for nx, ny in ((256, 128), (60, 50)):  # 65,536 (the limit), then 6,000 triangles
        pos, idx = rt.gen_heightfield(nx, ny, 1.0, 0.05, 7 + nx)
        n = idx.shape[0]
        print(f"noalloc build {algo} rounds={rounds} n={n}", flush=True)
        got = _build(rt, pos, idx, n, cfg, True)
        want = _build(rt, pos, idx, n, cfg, False)
        torch.cuda.synchronize()
        nc = int(want[2][0])
        assert int(got[2][0]) == nc and int(got[2][1]) == int(want[2][1])
        assert torch.equal(got[0][: nc * 128], want[0][: nc * 128])
        for g, w in zip(got[1:], want[1:]):
            assert torch.equal(g, w)


def test_noalloc_build_small_and_soup(rt):
    cfg = rt.build_cfg("radix", max_leaf_prims=1)
    for m in (4096, 10, 30_000):
        xyz = rt.gen_dynamic_grid(m, 0.3, 2.0)
        n = xyz.shape[0]
        got = _build(rt, xyz, None, n, cfg, True)
        want = _build(rt, xyz, None, n, cfg, False)
        torch.cuda.synchronize()
        nc = int(want[2][0])
        assert torch.equal(got[0][: nc * 128], want[0][: nc * 128])
        assert torch.equal(got[1], want[1]) and torch.equal(got[3], want[3])


def test_noalloc_build_refuses_too_large(rt):
    cfg = rt.build_cfg("radix")
    with pytest.raises(rt.RtError) as e:
        rt.call("rt_gpu_build_noalloc", rt.stream_handle(), None, None, rt.u32(65537),
                C.byref(cfg), None, rt.u32(0), None, None, None, None)
    assert e.value.code == 2  # RT_ERR_CAPACITY, before any device work
