"""Two-level scene, AO rays, AO resolve and the 7x7 bilateral filter on the
GPU against the oracle: scene traversal (instances with identity,
translation and general affine maps) bit-exact in hit ID, t and instance;
occlusion bits exact; AO resolve and the filter bit-exact; AO rays within
float tolerance (device sin / cos vs the C library's); the whole AO frame
within 1e-3 RMS."""
import numpy as np
import pytest
import torch

from conftest import u32, to_np

pytestmark = pytest.mark.gpu


def _bits(b, n):
    i = np.arange(n)
    return ((b[i >> 5] >> (i & 31)) & 1).astype(bool)


def _rot(deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


@pytest.fixture(scope="module")
def scene(rt, O):
    """terrain (indexed, identity) + deforming grid (9 floats per triangle,
    translated down) + sphere (rotated, scaled, translated)."""
    tp, ti = rt.gen_heightfield(120, 100, 3.0, 0.3, 17)
    terrain = rt.GpuBVH4.build(tp, ti, algorithm="radix")
    gx = rt.gen_dynamic_grid(4096, 0.7, 2.0)
    grid = rt.GpuBVH4.build(gx, None, algorithm="implicit")
    sp, si = rt.gen_sphere(80, 60, 1.0, 0.02, 5)
    sph = rt.GpuBVH4.build(sp, si, algorithm="radix")
    torch.cuda.synchronize()
    m_grid = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 4.6]], np.float32)
    # world -> object of the sphere: object = (R^T (x - c)) / s
    R, c, sc = _rot(30.0), np.array([1.5, -0.5, 0.9]), 0.6
    lin = R.T / sc
    m_sph = np.concatenate([lin, (-lin @ c)[:, None]], axis=1).astype(np.float32)
    inst = [(terrain, tp, ti, None, rt.INST_IDENTITY),
            (grid, gx, None, m_grid, rt.INST_TRANSLATE),
            (sph, sp, si, m_sph, rt.INST_GENERAL)]
    dev = rt.Scene(inst)
    def host(b):
        n4 = to_np(b.nodes)[: b.node_count * 128].view(O.NODE4)
        t64 = to_np(b.tris)[: b.tri_count * 64].view(O.TRI64)
        return n4, t64
    oracle = [(*host(b), (np.eye(3, 4, dtype=np.float32) if m is None else m), k)
              for b, _, _, m, k in inst]
    meshes = [(to_np(p), None if i is None else u32(i), (np.eye(3, 4, dtype=np.float32)
               if m is None else m), k) for _, p, i, m, k in inst]
    return dev, oracle, meshes


@pytest.mark.parametrize("shape", [dict(), dict(block=64, lds_stack=4),
                                   dict(block=128, lds_stack=32), dict(block=256, lds_stack=48)],
                         ids=["default", "b64-l4", "b128-l32", "b256-l48"])
def test_scene_closest_and_any(rt, O, scene, shape):
    dev, orc, _ = scene
    r = O.random_rays(150_001, 12, (-3.2, -3.2, -0.5), (3.2, 3.2, 2.0), tmax=1e30)
    prim, t, inst = dev.intersect(rt.Rays.from_numpy(r.packed()), **shape)
    torch.cuda.synchronize()
    pc, tc, ic = O.trace_scene(orc, r)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    np.testing.assert_array_equal(u32(inst), ic)
    assert len(set(ic[pc != O.MISS].tolist())) == 3  # every instance is hit somewhere
    r.tmax[:] = 0.6
    bits = dev.occluded(rt.Rays.from_numpy(r.packed()), **shape)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_bits(u32(bits), len(r)), O.trace_scene(orc, r, mode=O.ANY))


def test_ao_frame_pipeline(rt, O, scene):
    """camera -> scene closest hit -> AO rays -> scene any hit -> resolve ->
    7x7 bilateral, on the GPU and on the CPU path: primary hits exact, AO
    rays within float tolerance, AO image within 1e-3 RMS, the filter
    bit-exact on the same input."""
    dev, orc, meshes = scene
    w, h = 640, 480  # one AO ray flipped by a last-bit difference is ~3e-4 RMS here
    eye, look, up = (0.0, -5.0, 4.5), (0.0, 0.0, 0.3), (0.0, 0.0, 1.0)
    prim_g = rt.gen_camera_rays(w, h, eye, look, up, 60.0, 0.0, 1e30)
    prim, t, inst = dev.intersect(prim_g)
    ao_r = dev.gen_ao_rays(prim_g, prim, inst, t, w, h, frame=3, offset=0.03, tmin=0.01,
                           tmax=0.5)
    bits = dev.occluded(ao_r)
    ao = rt.ao_resolve(bits, prim)
    depth = torch.where(prim == -1, torch.full_like(t, 1e30), t)
    out = rt.blur7(ao, depth, w, h, 0.05)
    torch.cuda.synchronize()
    pr = O.camera_rays(w, h, eye, look, up, 60.0, 0.0, 1e30)
    pc, tc, ic = O.trace_scene(orc, pr)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(u32(inst), ic)
    ar = O.ao_rays_scene(meshes, pr, pc, ic, tc, w, h, frame=3, offset=0.03, tmin=0.01, tmax=0.5)
    got = O.Rays.stack(to_np(ao_r.data))
    for f in O.Rays.FIELDS:
        np.testing.assert_allclose(getattr(got, f), getattr(ar, f), rtol=2e-6, atol=2e-6)
    occ = O.trace_scene(orc, ar, mode=O.ANY)
    bits_c = np.zeros((w * h + 31) // 32, np.uint32)
    np.bitwise_or.at(bits_c, np.nonzero(occ)[0] >> 5, (1 << (np.nonzero(occ)[0] & 31)).astype(np.uint32))
    ao_c = O.ao_resolve(bits_c, pc)
    dc = np.where(pc == O.MISS, np.float32(1e30), tc).astype(np.float32)
    out_c = O.blur7(ao_c, dc, w, h, 0.05)
    assert O.rms(to_np(out), out_c) <= 1e-3
    assert 0.05 < float(ao_c.mean()) < 1.0  # some pixels occluded, not all
    # resolve and filter alone: bit-exact on the device's own inputs
    np.testing.assert_array_equal(to_np(ao), O.ao_resolve(u32(bits), u32(prim)))
    np.testing.assert_array_equal(to_np(out).view(np.uint32),
                                  O.blur7(to_np(ao), to_np(depth), w, h, 0.05).view(np.uint32))


@pytest.mark.parametrize("w,h", [(77, 53), (16, 16), (1, 1), (300, 7)])
def test_blur7_bit_exact(rt, O, w, h):
    rng = np.random.default_rng(w * 1000 + h)
    img = rng.random(w * h, dtype=np.float32)
    depth = (rng.random(w * h, dtype=np.float32) * 0.2).astype(np.float32)
    depth[rng.random(w * h) < 0.1] = np.float32(1e30)
    out = rt.blur7(torch.from_numpy(img).cuda(), torch.from_numpy(depth).cuda(), w, h, 0.05)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_np(out).view(np.uint32),
                                  O.blur7(img, depth, w, h, 0.05).view(np.uint32))


@pytest.mark.parametrize("tw", [8, 16, 32])
def test_scene_wave_tiles(rt, O, scene, tw):
    """Wave tiles of a 480-ray-wide grid (320 rows): per-ray results equal
    the oracle's (closest: ids, t, instance; any: bits)."""
    dev, orc, _ = scene
    r = O.random_rays(480 * 320, 21, (-3.2, -3.2, -0.5), (3.2, 3.2, 2.0), tmax=1e30)
    prim, t, inst = dev.intersect(rt.Rays.from_numpy(r.packed()), tile=(480, tw), block=64)
    torch.cuda.synchronize()
    pc, tc, ic = O.trace_scene(orc, r)
    np.testing.assert_array_equal(u32(prim), pc)
    np.testing.assert_array_equal(to_np(t).view(np.uint32), tc.view(np.uint32))
    np.testing.assert_array_equal(u32(inst), ic)
    r.tmax[:] = 0.6
    bits = dev.occluded(rt.Rays.from_numpy(r.packed()), tile=(480, tw))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_bits(u32(bits), len(r)), O.trace_scene(orc, r, mode=O.ANY))
    with pytest.raises(rt.RtError):
        dev.occluded(rt.Rays.from_numpy(r.packed()), tile=(500, 8))  # not whole rows
