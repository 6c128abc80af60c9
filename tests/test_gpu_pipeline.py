"""End-to-end shadow any-hit frame on the GPU against the CPU oracle: device
generators (mesh, camera rays, shadow rays), device build, closest-hit and
any-hit traversal, per-pixel light masks.  Ray buffers with no
transcendental functions (camera, shadow) are bit-exact against the oracle's
generators; the sphere mesh uses sin/cos and is checked to a few ulps, and
every later stage consumes the device mesh so it stays bit-exact."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits_to_bool(b, n):
    r = np.arange(n)
    return ((b[r >> 5] >> (r & 31)) & 1).astype(bool)


def _rays_np(O, rays):
    return O.Rays.stack(to_np(rays.data))


def test_sphere_generator_close_to_oracle(rt, O):
    pos, idx = rt.gen_sphere(125, 100, 1.0, 0.02, 7)
    pc, ic = O.sphere(125, 100, 1.0, 0.02, 7)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(u32(idx), ic)
    np.testing.assert_allclose(to_np(pos), pc, rtol=0, atol=4e-7)


def test_grid_generators(rt, O):
    pos, idx = rt.gen_wave_grid(64, 64, 1.0, 0.0, 0.1, 4.0, 4.0)
    pc, ic = O.wave_grid(64, 64, 1.0, 0.0, 0.1, 4.0, 4.0)
    np.testing.assert_array_equal(u32(idx), ic)
    np.testing.assert_array_equal(to_np(pos)[:, :2], pc[:, :2])
    np.testing.assert_allclose(to_np(pos)[:, 2], pc[:, 2], rtol=0, atol=2e-7)
    hp, hi = rt.gen_heightfield(50, 40, 1.0, 0.05, 11)
    hc, hic = O.heightfield(50, 40, 1.0, 0.05, 11)
    np.testing.assert_array_equal(to_np(hp), hc)
    np.testing.assert_array_equal(u32(hi), hic)
    xyz = rt.gen_dynamic_grid(4096, 0.25, 1.25)
    xc, _ = O.dynamic_grid(4096, 0.25, 1.25)
    got = to_np(xyz)
    np.testing.assert_array_equal(got[:, [0, 1, 3, 4, 6, 7]], xc[:, [0, 1, 3, 4, 6, 7]])
    np.testing.assert_allclose(got[:, [2, 5, 8]], xc[:, [2, 5, 8]], rtol=0, atol=1e-6)


def test_camera_rays_bit_exact(rt, O):
    g = rt.gen_camera_rays(320, 180, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0,
                           0.0, 1e30)
    c = O.camera_rays(320, 180, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 0.0,
                      1e30)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(g.data).view(np.uint32), c.packed().view(np.uint32))


def test_random_rays_close(rt, O):
    g = rt.gen_random_rays(50_000, 9, (-1, -2, -3), (1, 2, 3), 7.0)
    c = O.random_rays(50_000, 9, (-1, -2, -3), (1, 2, 3), 7.0)
    got = to_np(g.data)
    np.testing.assert_array_equal(got[[0, 1, 2, 3, 6, 7]], c.packed()[[0, 1, 2, 3, 6, 7]])
    np.testing.assert_allclose(got[[4, 5]], c.packed()[[4, 5]], rtol=0, atol=1e-6)


@pytest.mark.parametrize("nu,nv,w,h", [(250, 200, 320, 180), (125, 100, 97, 61)])
def test_shadow_frame_matches_oracle(rt, O, nu, nv, w, h):
    pos, idx = rt.gen_sphere(nu, nv, 1.0, 0.02, 7)
    bvh = rt.GpuBVH4.build(pos, idx)
    cam = rt.gen_camera_rays(w, h, (0.0, 0.0, 1.25), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0,
                             0.0, 1e30)
    prim, t = bvh.intersect(cam)
    lights = rt.lights_in_box((-3.06, -3.06, -3.06), (3.06, 3.06, 3.06), 16, 1337)
    dl = torch.tensor(lights, dtype=torch.float32, device="cuda")
    sh = rt.gen_shadow_rays(pos, idx, cam, prim, t, dl, 1e-4)
    bits = bvh.occluded(sh)
    masks = rt.pack_light_masks(bits, w * h, 16)
    torch.cuda.synchronize()

    pos_h, idx_h = to_np(pos), u32(idx)
    nodes = to_np(bvh.nodes)[: bvh.node_count * 128].view(O.NODE4)
    tris = to_np(bvh.tris)[: bvh.tri_count * 64].view(O.TRI64)
    cam_h = _rays_np(O, cam)
    pc, tc = O.trace_bvh4(nodes, tris, cam_h)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    if nu >= 250:  # fine enough that the facets stay outside the 0.98 radius cone
        assert (pc != O.MISS).all(), "camera inside the silhouette: every pixel hits"

    sc = O.shadow_rays(pos_h, idx_h, cam_h, pc, tc, np.array(lights, np.float32), 1e-4)
    np.testing.assert_array_equal(to_np(sh.data).view(np.uint32), sc.packed().view(np.uint32))
    pa, _ = O.trace_bvh4(nodes, tris, sc, mode=O.ANY)
    occ = _bits_to_bool(u32(bits), len(sc))
    np.testing.assert_array_equal(occ, pa != O.MISS)
    want = np.zeros(w * h, np.uint32)
    for l in range(16):
        want |= occ[l * w * h:(l + 1) * w * h].astype(np.uint32) << l
    np.testing.assert_array_equal(to_np(masks).view(np.uint16), want.astype(np.uint16))
    # half the lights are behind the visible hemisphere on average
    assert 0.2 < occ.mean() < 0.9
