"""Wave tiles of a ray grid (rt_trace_opts.grid_width / tile_width): every
tile width gives per-ray results identical to the linear launch (closest-hit
ids and t bits, any-hit bits including the zeroed padding of the last word,
box / triangle counters) on a frame's camera rays and its 16-light shadow
rays, on a grid whose rows are not a multiple of the tile height, and with
the any-hit child orders; bad grids, and refill tiles for queries other
than occlusion bits, are refused."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu

W, H, NL = 160, 120, 16


@pytest.fixture(scope="module")
def frame(rt, O):
    pos, idx = rt.gen_sphere(250, 200, 1.0, 0.02, 3)
    bvh = rt.GpuBVH4.build(pos, idx, algorithm="radix")
    prim_rays = rt.gen_camera_rays(W, H, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                   60.0, 0.0, 1e30)
    prim, t = bvh.intersect(prim_rays)
    lo, hi = pos.amin(0).cpu().double().numpy(), pos.amax(0).cpu().double().numpy()
    c, ext = (lo + hi) / 2, hi - lo
    lights = torch.tensor(rt.lights_in_box(c - 1.5 * ext, c + 1.5 * ext, NL, 1337),
                          dtype=torch.float32, device="cuda")
    shadow = rt.gen_shadow_rays(pos, idx, prim_rays, prim, t, lights, 1e-4)
    torch.cuda.synchronize()
    return bvh, prim_rays, shadow


@pytest.mark.parametrize("tw", [8, 16, 32])
@pytest.mark.parametrize("block", [64, 256])
def test_tiles_equal_linear(rt, O, frame, tw, block):
    bvh, prim_rays, shadow = frame
    p0, t0, b0, r0 = bvh.intersect(prim_rays, counters=True, block=block)
    p1, t1, b1, r1 = bvh.intersect(prim_rays, counters=True, block=block, tile=(W, tw))
    bits0 = bvh.occluded(shadow, block=block)
    bits1 = bvh.occluded(shadow, block=block, tile=(W, tw))
    bits2 = bvh.occluded(shadow, block=block, tile=(W, tw), any_order="slot")
    a0, ab0, at0 = bvh.occluded(shadow, counters=True, block=block)
    a1, ab1, at1 = bvh.occluded(shadow, counters=True, block=block, tile=(W, tw))
    torch.cuda.synchronize()
    assert torch.equal(p0, p1) and torch.equal(t0.view(torch.int32), t1.view(torch.int32))
    assert torch.equal(b0, b1) and torch.equal(r0, r1)
    assert torch.equal(bits0, bits1) and torch.equal(bits0, bits2)
    assert torch.equal(a0, a1) and torch.equal(ab0, ab1) and torch.equal(at0, at1)
    occ = u32(bits0)
    assert 0.05 < np.unpackbits(occ.view(np.uint8)).mean() < 0.95


def test_partial_tiles_and_padding(rt, O, frame):
    """48 rays per row, 5 rows (not a multiple of the 8- or 4-row tile
    heights): 240 rays; the last word's 16 padding bits must come out zero
    (the output starts all ones)."""
    bvh, _, _ = frame
    r = O.random_rays(240, 4, (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2), tmax=0.6)
    rays = rt.Rays.from_numpy(r.packed())
    want = bvh.occluded(rays)
    wp, wt = bvh.intersect(rays)
    for tw in (8, 16):
        got = torch.full_like(want, -1)
        bvh.occluded(rays, out=got, tile=(48, tw))
        p, t = bvh.intersect(rays, tile=(48, tw))
        torch.cuda.synchronize()
        assert torch.equal(got, want), tw
        assert torch.equal(p, wp) and torch.equal(t.view(torch.int32), wt.view(torch.int32))
    assert int(u32(want)[-1]) >> 16 == 0


def test_bad_grids_are_refused(rt, frame):
    bvh, prim_rays, _ = frame
    with pytest.raises(rt.RtError):
        bvh.intersect(prim_rays, tile=(W + 8, 8))  # rays not a whole number of rows
    with pytest.raises(rt.RtError):
        bvh.intersect(prim_rays, tile=(W, 64))  # tile width not 8, 16 or 32
    with pytest.raises(rt.RtError):
        bvh.intersect(prim_rays, tile=(40, 16))  # grid width not a multiple of the tile
    # refill with wave tiles: occlusion bits only
    with pytest.raises(rt.RtError):
        bvh.intersect(prim_rays, tile=(W, 8), schedule="refill")
    with pytest.raises(rt.RtError):
        bvh.occluded(prim_rays, tile=(W, 8), schedule="refill", counters=True)
