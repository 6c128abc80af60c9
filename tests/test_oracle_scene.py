"""The oracle's two-level scene (trace_scene / map_rays) pinned without the
device: one identity instance is the plain traversal; duplicated instances
keep the lower instance on ties and do not change hits; known answers
under a translation and under a rotation + scale; any hit = OR over
instances."""
import numpy as np


def _tri_scene(O):
    pos = np.array([[-1, -1, 5], [1, -1, 5], [0, 1, 5]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    return O.software_bvh4(pos, idx)


def _ray(O, o, d, tmax=100.0):
    r = O.Rays(1)
    r.ox[0], r.oy[0], r.oz[0] = o
    r.dx[0], r.dy[0], r.dz[0] = d
    r.tmin[0], r.tmax[0] = 0.0, tmax
    return r


def test_single_identity_instance_is_plain_trace(O):
    pos, idx = O.wave_grid(32, 32, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, tris = O.software_bvh4(pos, idx)
    r = O.random_rays(20_000, 3, (-1.2, -1.2, -0.3), (1.2, 1.2, 0.3), tmax=1e30)
    p, t = O.trace_bvh4(n4, tris, r)
    ps, ts, ins = O.trace_scene([(n4, tris, None, O.INST_IDENTITY)], r)
    np.testing.assert_array_equal(ps, p)
    np.testing.assert_array_equal(ts.view(np.uint32), t.view(np.uint32))
    assert ((ins == 0) == (p != O.MISS)).all()
    occ = O.trace_scene([(n4, tris, None, O.INST_IDENTITY)], r, mode=O.ANY)
    pa, _ = O.trace_bvh4(n4, tris, r, mode=O.ANY)
    np.testing.assert_array_equal(occ, pa != O.MISS)


def test_duplicate_instances_keep_the_first(O):
    pos, idx = O.wave_grid(16, 16, 1.0, 0.0, 0.1, 4.0, 4.0)
    n4, tris = O.software_bvh4(pos, idx)
    r = O.random_rays(5_000, 4, (-1.2, -1.2, -0.3), (1.2, 1.2, 0.3), tmax=1e30)
    one = O.trace_scene([(n4, tris, None, O.INST_IDENTITY)], r)
    two = O.trace_scene([(n4, tris, None, O.INST_IDENTITY)] * 2, r)
    np.testing.assert_array_equal(two[0], one[0])
    np.testing.assert_array_equal(two[1].view(np.uint32), one[1].view(np.uint32))
    assert (two[2][two[0] != O.MISS] == 0).all()


def test_translation_known_answer(O):
    n4, tris = _tri_scene(O)
    # object z = world z + 4: the triangle (object z = 5) sits at world z = 1
    m = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 4]], np.float32)
    inst = [(n4, tris, m, O.INST_TRANSLATE)]
    p, t, i = O.trace_scene(inst, _ray(O, (0.0, -0.5, 0.0), (0.0, 0.0, 1.0)))
    assert p[0] == 0 and t[0] == 1.0 and i[0] == 0
    p, t, i = O.trace_scene(inst, _ray(O, (0.0, -0.5, 0.0), (0.0, 0.0, 1.0), tmax=0.9))
    assert p[0] == O.MISS


def test_rotation_scale_known_answer(O):
    n4, tris = _tri_scene(O)
    # world -> object: object = 2 (wx, wz, -wy), so the triangle's plane
    # (object z = 5) is the world plane y = -2.5
    lin = 2.0 * np.array([[1, 0, 0], [0, 0, 1], [0, -1, 0]], np.float32)
    m = np.concatenate([lin, np.zeros((3, 1), np.float32)], axis=1)
    inst = [(n4, tris, m, O.INST_GENERAL)]
    # world ray from (0, 0, 0.25) along -y: object o = (0, 0.5, 0), d = (0, 0, 2)
    p, t, i = O.trace_scene(inst, _ray(O, (0.0, 0.0, 0.25), (0.0, -1.0, 0.0)))
    assert p[0] == 0 and abs(t[0] - 2.5) < 1e-6
    occ = O.trace_scene(inst + [(n4, tris, None, O.INST_IDENTITY)],
                        _ray(O, (0.0, 0.0, 0.25), (0.0, -1.0, 0.0), tmax=2.0), mode=O.ANY)
    assert not occ[0]
