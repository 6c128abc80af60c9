"""Pins of the oracle's BVH2 traversal (or_trace_bvh2, restating
benchmarks/exp0_probe/shaders/software_bvh_procedural.comp:47-171) on the
reference's own SoftwareBVH trees:

  * against a brute force over every triangle with the shader's predicates
    (|a| < 1e-8 rejected, t > tmin and t < closest accepted), written
    independently in numpy float32 with the same operation order: same
    closest triangle and t bits;
  * any hit: occluded iff the closest hit exists;
  * the probe's known-answer triangle (probe_main.cpp:45-50: t = 2, prim 0).
"""
import numpy as np
import pytest


def _brute(O, t48, rays):
    v0, v1, v2 = t48[:, 0:3], t48[:, 4:7], t48[:, 8:11]
    e1, e2 = v1 - v0, v2 - v0
    out_p = np.full(len(rays), 0xFFFFFFFF, np.uint32)
    out_t = rays.tmax.copy()
    for r in range(len(rays)):
        o = np.array([rays.ox[r], rays.oy[r], rays.oz[r]], np.float32)
        d = np.array([rays.dx[r], rays.dy[r], rays.dz[r]], np.float32)
        hx = d[1] * e2[:, 2] - d[2] * e2[:, 1]
        hy = d[2] * e2[:, 0] - d[0] * e2[:, 2]
        hz = d[0] * e2[:, 1] - d[1] * e2[:, 0]
        a = e1[:, 0] * hx + e1[:, 1] * hy + e1[:, 2] * hz
        ok = np.abs(a) >= np.float32(1e-8)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            f = np.float32(1.0) / a
            sx, sy, sz = o[0] - v0[:, 0], o[1] - v0[:, 1], o[2] - v0[:, 2]
            u = f * (sx * hx + sy * hy + sz * hz)
            ok &= (u >= 0) & (u <= 1)
            qx = sy * e1[:, 2] - sz * e1[:, 1]
            qy = sz * e1[:, 0] - sx * e1[:, 2]
            qz = sx * e1[:, 1] - sy * e1[:, 0]
            v = f * (d[0] * qx + d[1] * qy + d[2] * qz)
            ok &= (v >= 0) & (u + v <= 1)
            t = f * (e2[:, 0] * qx + e2[:, 1] * qy + e2[:, 2] * qz)
        ok &= (t > rays.tmin[r]) & (t < rays.tmax[r])
        if ok.any():
            cand = np.nonzero(ok)[0]
            k = cand[np.argmin(t[cand])]
            out_p[r], out_t[r] = k, t[k]
    return out_p, out_t


def test_bvh2_matches_brute_force(O):
    pos, idx = O.sphere(24, 16, 1.0, 0.05, 4)
    n2, t48, order = O.bvh2_build(pos, idx, 4)
    rays = O.random_rays(3000, 5, (-1.3, -1.3, -1.3), (1.3, 1.3, 1.3), tmax=1e30)
    p, t = O.trace_bvh2(n2, t48, rays)
    bp, bt = _brute(O, t48.reshape(-1, 12), rays)
    np.testing.assert_array_equal(p, bp)
    np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
    assert (p != O.MISS).mean() > 0.3
    # original indices through the leaf order
    po, _ = O.trace_bvh2(n2, t48, rays, order=order)
    hit = p != O.MISS
    np.testing.assert_array_equal(po[hit], order[p[hit]])


def test_bvh2_any_hit_and_counters(O):
    pos, idx = O.terrain(2000)
    n2, t48, order = O.bvh2_build(pos, idx, 4)
    rays = O.exp0_rays(20000, False)
    p, t, box, tri = O.trace_bvh2(n2, t48, rays, counters=True)
    pa, _ = O.trace_bvh2(n2, t48, rays, mode=O.ANY)
    np.testing.assert_array_equal(pa != O.MISS, p != O.MISS)
    assert (box >= 1).all() and (tri <= len(t48)).all()
    # closest hits agree with the BVH4 traversal wherever both predicates agree
    n4, t64 = O.software_bvh4(pos, idx)
    p4, t4 = O.trace_bvh4(n4, t64, rays)
    same = (p != O.MISS) & (p4 != O.MISS)
    assert (order[p[same]] == p4[same]).mean() > 0.999


def test_bvh2_known_answer(O):
    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    n2, t48, order = O.bvh2_build(pos, idx, 4)
    rays = O.Rays(2)
    rays.ox[:] = [0.0, 3.0]
    rays.oy[:] = [0.0, 3.0]
    rays.oz[:] = [2.0, 2.0]
    rays.dx[:] = 0.0
    rays.dy[:] = 0.0
    rays.dz[:] = -1.0
    rays.tmin[:] = 0.01
    rays.tmax[:] = 1000.0
    p, t = O.trace_bvh2(n2, t48, rays)
    assert p[0] == 0 and abs(t[0] - 2.0) < 1e-6
    assert p[1] == O.MISS and t[1] == np.float32(1000.0)
