"""The RADV image decoder side on the host (no GPU):

* the oracle's decoder (or_radv_decode) inverts the oracle's encoder of
  fast_bvh_builder's image: decode(encode(mesh in LBVH order)) is the
  oracle's implicit Box4 LBVH byte for byte (nodes and 64-B leaves), and it
  rejects images whose child ids leave their section or have another type;
* rt_radv_describe (the header part of dump_bvh_structure,
  include/radv_bvh.hpp:311-390) reads every RadvAccelStructHeader field and
  the root's child ids as the reference's own structs do (oracle/_ref);
* rt_radv_describe / rt_trace_radv argument checks that need no device.
"""
import ctypes as C
import os

import numpy as np
import pytest

REF = os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                  "libref.so"))


def _image(O, nu, nv, order=None):
    pos, idx = O.sphere(nu, nv, 1.0, 0.05, 3)
    xyz9 = pos[idx.astype(np.int64)].reshape(-1, 9)
    return pos, idx, xyz9


@pytest.mark.parametrize("nu,nv", [(1, 1), (3, 1), (10, 7), (40, 30), (64, 32)])
def test_decode_inverts_encode(O, nu, nv):
    pos, idx, xyz9 = _image(O, nu, nv)
    nodes, tris, ids, _ = O.lbvh_build(pos, idx)
    img = O.radv_encode(xyz9, order=ids)
    dn, dt = O.radv_decode(img)
    assert dn.tobytes() == nodes.tobytes()
    assert dt.tobytes() == tris.tobytes()


def test_decode_rejects_malformed(O):
    _, _, xyz9 = _image(O, 10, 7)
    img = O.radv_encode(xyz9)
    bvh = int(img[:4].view(np.uint32)[0])
    tbase = 128 * O.tree_layout(len(xyz9))[2]  # triangle section, from the BVH base
    for patch in (0x7FFFFFF5,  # box32 id far past the box section
                  (0x80 >> 3) | 4,  # box16: not written by this builder
                  (tbase >> 3) | 5,  # box32 id on the first triangle node
                  ((tbase + 8) >> 3) | 0,  # low bits 1: node type 1, not a RADV type
                  (0x40 >> 3) | 5):  # box32 id not on a 128-B node
        bad = img.copy()
        bad[bvh:bvh + 4] = np.array([patch], np.uint32).view(np.uint8)
        with pytest.raises(ValueError):
            O.radv_decode(bad)
    short = img[: len(img) - 64]
    with pytest.raises(ValueError):
        O.radv_decode(short)


@pytest.mark.skipif(not REF, reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("nu,nv", [(1, 1), (10, 7), (64, 32)])
def test_describe_matches_reference_structs(O, nu, nv):
    import rtbvh as rt
    _, _, xyz9 = _image(O, nu, nv)
    img = O.radv_encode(xyz9)
    want = O.ref_radv_header(img)
    got = rt.radv_describe(img)
    for k, v in want.items():
        if k in ("aabb", "root_children"):
            np.testing.assert_array_equal(np.asarray(got[k], v.dtype), v)
        elif k[-1] in "012":
            assert got[k[:-1]][int(k[-1])] == v, k
        else:
            assert got[k] == v, k
    counts, _, tb, size = O.tree_layout(len(xyz9))
    assert got["box_count"] == tb and got["size"] == size
    text = rt.format_radv(img)
    assert "bvh_offset 0xc0" in text and f"primitive_count {len(xyz9)}" in text


def test_describe_and_trace_argument_checks(O):
    import rtbvh as rt
    lib = rt.lib()
    _, _, xyz9 = _image(O, 10, 7)
    img = O.radv_encode(xyz9)
    info = rt.rt_radv_info()
    p = img.ctypes.data_as(C.c_void_p)
    assert lib.rt_radv_describe(p, img.size, C.byref(info)) == 0
    assert lib.rt_radv_describe(p, 100, C.byref(info)) == 1  # shorter than a header
    assert lib.rt_radv_describe(p, img.size - 64, C.byref(info)) == 1  # triangles cut off
    assert info.bvh_offset == 192  # header fields are filled either way
    bad = img.copy()
    bad[0x60:0x68] = np.array([200], np.uint64).view(np.uint8)  # leaf section not on a node
    assert lib.rt_radv_describe(bad.ctypes.data_as(C.c_void_p), bad.size, C.byref(info)) == 1
    assert lib.rt_radv_describe(None, 0, C.byref(info)) == 1
    fake = C.c_void_p(256)
    d = rt.rt_radv_dev(fake, C.c_uint64(img.size), 192, 48, 140, 3)
    r = rt.rt_rays_soa(*([fake] * 8))
    h = rt.rt_hits_soa(fake, fake)
    call = lib.rt_trace_radv
    assert call(None, C.byref(d), C.byref(r), 0, 0, None, C.byref(h), None, None, None) == 0
    d.tri_count = 141  # one triangle past the image
    assert call(None, C.byref(d), C.byref(r), 5, 0, None, C.byref(h), None, None, None) == 2
    d.tri_count = 140
    d.image = C.c_void_p(264)  # not 16-byte aligned
    assert call(None, C.byref(d), C.byref(r), 5, 0, None, C.byref(h), None, None, None) == 1
    d.image = fake
    d.bvh_offset = 136  # not 64-byte aligned (sections still inside the image)
    assert call(None, C.byref(d), C.byref(r), 5, 0, None, C.byref(h), None, None, None) == 1
    d.bvh_offset = 192
    assert call(None, C.byref(d), C.byref(r), 5, 1, None, None, None, None, None) == 1
    assert call(None, C.byref(d), C.byref(r), 5, 0x400, None, C.byref(h), None, None,
                None) == 1
    o = rt.rt_trace_opts(1, 0, 0, 0, 0)
    assert call(None, C.byref(d), C.byref(r), 5, 0, C.byref(o), C.byref(h), None, None,
                None) == 1
    d.depth = 65
    assert call(None, C.byref(d), C.byref(r), 5, 0, None, C.byref(h), None, None, None) == 2
