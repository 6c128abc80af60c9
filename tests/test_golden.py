"""The CPU oracle against the golden fixtures written from the reference's
own build (tests/golden/make_golden.py) -- runs wherever the fixtures are,
with or without the reference checkout -- and, where oracle/_ref was built,
directly against the reference at larger sizes."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, f"ref_{name}.npz"))


@pytest.mark.parametrize("name,fused", [("terrain1000", True), ("terrain1000_strict", False)])
def test_terrain_and_software_bvh(O, name, fused):
    g = _load(name)
    pos, idx = O.terrain(1000, fused=fused)
    np.testing.assert_array_equal(pos.view(np.uint32), g["pos"].view(np.uint32))
    np.testing.assert_array_equal(idx, g["idx"])
    n2, t48, order = O.bvh2_build(pos, idx)
    assert n2.view(np.uint8).tobytes() == g["bvh2"].tobytes()
    np.testing.assert_array_equal(t48.view(np.uint32), g["tris48"].view(np.uint32))
    n4 = O.bvh4_collapse(n2)
    assert n4.view(np.uint8).tobytes() == g["bvh4"].tobytes()
    # order restores the IDs the reference drops: leaf i holds triangle order[i]
    np.testing.assert_array_equal(t48.reshape(-1, 3, 4)[:, :, :3], pos[idx[order]])


@pytest.mark.parametrize("mp", [1, 2, 4])
def test_flat_grid_tie_order(O, mp):
    """A flat grid: every split has many equal centroids, so leaf membership
    is fixed by nth_element's tie handling (libstdc++)."""
    g = _load(f"flatgrid_mp{mp}")
    n2, t48, _ = O.bvh2_build(g["pos"], g["idx"], max_prims=mp)
    assert n2.view(np.uint8).tobytes() == g["bvh2"].tobytes()
    assert O.bvh4_collapse(n2).view(np.uint8).tobytes() == g["bvh4"].tobytes()
    np.testing.assert_array_equal(t48, g["tris48"])


@pytest.mark.parametrize("n", [64, 256, 1024, 4096])
def test_dynamic_grid(O, n):
    g = _load(f"dyngrid{n}")
    xyz, cen = O.dynamic_grid(n, 0.0, 1.25)
    np.testing.assert_array_equal(xyz.view(np.uint32), g["xyz"].view(np.uint32))
    np.testing.assert_array_equal(cen.view(np.uint32), g["cen"].view(np.uint32))


def test_tree_layout_and_push_constants(O):
    g = _load("layout")
    sizes = g["sizes"]
    # software_bvh.hpp / radv_bvh.hpp record sizes (static_asserts :135, :152)
    assert list(sizes[:7]) == [32, 48, 128, 136, 128, 64, 96] and sizes[7] == 48
    for n in g["ns"]:
        c, o, tb, sz = O.tree_layout(int(n))
        np.testing.assert_array_equal(c, g[f"counts_{n}"])
        np.testing.assert_array_equal(o, g[f"offsets_{n}"])
        tb_ref, ta_ref = g[f"scalars_{n}"]
        assert tb == tb_ref and (sz & 0xFFFFFFFF) == ta_ref
        # FastBvhPushConstants keeps at most 8 levels (radv_bvh.hpp:185-188)
        pc = g[f"push_{n}"]
        words = pc[16:].view(np.uint32)
        L = min(len(c), 8)
        assert words[2] == len(c) and list(words[4:4 + L]) == list(o[:L])
        assert list(words[12:12 + L]) == list(c[:L])


# ---- direct comparison with the reference (where oracle/_ref is built) -----

def _need_ref(O, shipped=False):
    if O.ref(shipped) is None:
        pytest.skip("oracle/_ref not built here (no reference checkout)")


@pytest.mark.parametrize("target", [20_000, 300_000])
def test_software_bvh_against_reference_build(O, target):
    _need_ref(O, shipped=True)
    pos, idx = O.ref_terrain(target, shipped=True)
    p2, i2 = O.terrain(target, fused=True)
    np.testing.assert_array_equal(pos.view(np.uint32), p2.view(np.uint32))
    n2r, n4r, t48r = O.ref_build(pos, idx, shipped=True)
    n2, t48, _ = O.bvh2_build(pos, idx)
    assert n2.tobytes() == n2r.tobytes() and t48.tobytes() == t48r.tobytes()
    assert O.bvh4_collapse(n2).tobytes() == n4r.tobytes()


def test_software_bvh_against_reference_sphere(O):
    _need_ref(O)
    pos, idx = O.sphere(300, 200, 1.0, 0.02, 5)
    n2r, n4r, t48r = O.ref_build(pos, idx)
    n2, t48, _ = O.bvh2_build(pos, idx)
    assert n2.tobytes() == n2r.tobytes() and t48.tobytes() == t48r.tobytes()
    assert O.bvh4_collapse(n2).tobytes() == n4r.tobytes()
