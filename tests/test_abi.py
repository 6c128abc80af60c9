"""The C ABI without a GPU: librtbvh.so loads, exports every entry point
include/rt_bvh.h declares, answers the host-only queries, and rejects bad
arguments before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rt_bvh.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:rt_status|const char\*)\s+(rt_\w+)\(", text, re.M)))


def test_header_declares_expected_surface():
    names = _declared()
    for must in ("rt_gpu_build", "rt_trace", "rt_tree_layout", "rt_build_requirements",
                 "rt_gpu_refit", "rt_convert_tris48"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import rtbvh
    lib = rtbvh.lib()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(rtbvh.exported_symbols()) >= set(_declared()) - {"rt_version", "rt_last_error"}
    assert lib.rt_version().startswith(b"rtbvh")


def test_host_tree_layout_matches_oracle(O):
    import rtbvh
    for n in (1, 5, 4096, 4097, 2_500_000):
        c, o, tb, sz = rtbvh.tree_layout(n)
        c2, o2, tb2, sz2 = O.tree_layout(n)
        assert c == list(c2) and o == list(o2) and tb == tb2 and sz == sz2


def test_build_requirements():
    import rtbvh
    cap, scratch = rtbvh.build_requirements(4096, rtbvh.build_cfg("implicit"))
    assert cap == 1365 and scratch == 0  # fused single-workgroup path
    cap, scratch = rtbvh.build_requirements(2_500_000, rtbvh.build_cfg("radix"))
    assert cap == 2_499_999 and scratch > 2_500_000 * 60
    cap, _ = rtbvh.build_requirements(10, rtbvh.build_cfg("auto"))
    assert cap == rtbvh.tree_layout(10)[2]
    with pytest.raises(rtbvh.RtError):
        rtbvh.build_requirements(10, rtbvh.rt_build_cfg(7, 4, 0, 0))
    # radix digit widths: 4, 6, 8, 10 (0 = 8); the scratch size does not depend on it
    sizes = {b: rtbvh.build_requirements(100_000, rtbvh.build_cfg("radix", 4, b))[1]
             for b in rtbvh.SORT_BITS}
    assert len(set(sizes.values())) == 1
    for bad in (1, 5, 7, 12):
        with pytest.raises(rtbvh.RtError):
            rtbvh.build_requirements(10, rtbvh.rt_build_cfg(1, 4, bad, 0))


def test_argument_errors_before_device_work():
    import rtbvh
    lib = rtbvh.lib()
    # null bvh / rays / info pointers and unknown flags are RT_ERR_ARG
    assert lib.rt_trace(None, None, None, 10, 0, None, None, None, None) == 1
    b = rtbvh.rt_bvh_dev()
    r = rtbvh.rt_rays_soa()
    assert lib.rt_trace(None, C.byref(b), C.byref(r), 10, 0, None, None, None, None) == 1
    assert lib.rt_trace(None, C.byref(b), C.byref(r), 10, 0x8000, None, None, None, None) == 1
    assert lib.rt_trace(None, C.byref(b), C.byref(r), 0, 0, None, None, None, None) == 0
    b.depth = 65  # deeper than any tree the kernels accept
    fake = C.c_void_p(16)
    r2 = rtbvh.rt_rays_soa(*([fake] * 8))
    h = rtbvh.rt_hits_soa(fake, fake)
    assert lib.rt_trace(None, C.byref(b), C.byref(r2), 4, 0, C.byref(h), None, None, None) == 2
    cfg = rtbvh.build_cfg("radix")
    assert lib.rt_gpu_build(None, None, None, 10, C.byref(cfg), None, 0, None, None, None, None,
                            None, 0) == 1
    assert lib.rt_gen_sphere(None, 2, 2, 1.0, 0.0, 0, None, None) == 1
    assert lib.rt_pack_light_masks(None, fake, 10, 17, fake) == 1


def test_trace_launch_shape_checks():
    """rt_trace_ex / rt_trace_batch_ex reject launch shapes outside block in
    {64, 128, 256} and LDS stack depths outside 4..48 before any device work."""
    import rtbvh
    lib = rtbvh.lib()
    assert C.sizeof(rtbvh.rt_trace_opts) == 32
    b = rtbvh.rt_bvh_dev()
    r = rtbvh.rt_rays_soa()
    for block, lds, ok in [(64, 4, True), (128, 48, True), (256, 20, True), (0, 0, True),
                           (32, 20, False), (512, 20, False), (96, 20, False),
                           (256, 3, False), (256, 49, False)]:
        o = rtbvh.rt_trace_opts(0, 0, 0, block, lds)
        want = 0 if ok else 1
        assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, 0, C.byref(o), None, None, None,
                               None) == want, (block, lds)
        assert lib.rt_trace_batch_ex(None, None, None, 0, 0, 1, C.byref(o)) == want
    o = rtbvh.rt_trace_opts(1, 0, 0, 0, 0)  # refill with default chunks
    assert lib.rt_trace_batch_ex(None, None, None, 0, 0, 1, C.byref(o)) == 0
    for bad in (rtbvh.rt_trace_opts(2, 0, 0, 0, 0), rtbvh.rt_trace_opts(1, 100, 0, 0, 0),
                rtbvh.rt_trace_opts(1, 8192, 0, 0, 0), rtbvh.rt_trace_opts(1, 256, 65, 0, 0)):
        assert lib.rt_trace_batch_ex(None, None, None, 0, 0, 1, C.byref(bad)) == 1
    # chunk tables of other chunk sizes: multiples of 64 up to 4096
    chunk = (C.c_uint32 * 1)()
    assert lib.rt_trace_batch_plan_ex(None, 0, 256, chunk, None) == 0 and chunk[0] == 0
    for bad in (0, 100, 8192):
        assert lib.rt_trace_batch_plan_ex(None, 0, bad, chunk, None) == 1
    with pytest.raises(ValueError):
        rtbvh.trace_opts("bogus")


def test_struct_layouts():
    import rtbvh
    assert C.sizeof(rtbvh.rt_bvh_dev) == 32
    assert C.sizeof(rtbvh.rt_rays_soa) == 64
    assert C.sizeof(rtbvh.rt_build_cfg) == 16
    import oracle
    assert oracle.NODE4.itemsize == rtbvh.NODE_BYTES and oracle.TRI64.itemsize == rtbvh.TRI_BYTES
    # rt_tri64 = RadvBvhTriangleNode: triangle_id at byte 48 (radv_bvh.hpp:146)
    assert oracle.TRI64.fields["triangle_id"][1] == 48
    assert oracle.NODE4.fields["child"][1] == 96 and oracle.NODE4.fields["count"][1] == 112


def test_trace_batch_plan_on_host():
    """rt_trace_batch_plan is host-only: chunk offsets (64-ray chunks per
    item), the deepest tree, and argument checks; the item record is 128 B."""
    import rtbvh
    assert C.sizeof(rtbvh.rt_trace_item) == 128
    lib = rtbvh.lib()
    fake = C.c_void_p(16)
    items = (rtbvh.rt_trace_item * 4)()
    counts = [1, 64, 65, 0]
    for i, n in enumerate(counts):
        it = items[i]
        it.bvh = rtbvh.rt_bvh_dev(fake, fake, 5, 9, 3 + i, 0)
        it.rays = rtbvh.rt_rays_soa(*([fake] * 8))
        it.count = n
        it.flags = rtbvh.RT_ANY if i % 2 else rtbvh.RT_CLOSEST
        it.hits = rtbvh.rt_hits_soa(fake, fake)
        it.occlusion_bits = fake
    chunk = (C.c_uint32 * 5)()
    depth = C.c_uint32(0)
    assert lib.rt_trace_batch_plan(items, 4, chunk, C.byref(depth)) == 0
    assert list(chunk) == [0, 1, 2, 4, 4] and depth.value == 6
    items[2].flags = 0x100  # counters are not part of the batch path
    assert lib.rt_trace_batch_plan(items, 4, chunk, C.byref(depth)) == 1
    items[2].flags = rtbvh.RT_CLOSEST
    items[2].hits = rtbvh.rt_hits_soa(None, None)  # closest without outputs
    assert lib.rt_trace_batch_plan(items, 4, chunk, C.byref(depth)) == 1
    items[2].hits = rtbvh.rt_hits_soa(fake, fake)
    items[0].bvh.depth = 65
    assert lib.rt_trace_batch_plan(items, 4, chunk, C.byref(depth)) == 2
    assert lib.rt_trace_batch(None, None, None, 0, 0, 1) == 0  # nothing to trace
    assert lib.rt_trace_batch(None, None, None, 3, 7, 1) == 1
    assert lib.rt_trace_batch(None, fake, fake, 3, 7, 0) == 1


def test_scene_and_post_entry_points_on_host():
    import rtbvh
    assert C.sizeof(rtbvh.rt_instance) == 128
    lib = rtbvh.lib()
    fake = C.c_void_p(16)
    r = rtbvh.rt_rays_soa(*([fake] * 8))
    h = rtbvh.rt_hits_soa(fake, fake)
    assert lib.rt_trace_scene(None, fake, 1, 0, C.byref(r), 0, 0, C.byref(h), None, None) == 0
    assert lib.rt_trace_scene(None, fake, 1, 0, C.byref(r), 8, 0x100, C.byref(h), None,
                              None) == 1  # counters: not a scene flag
    assert lib.rt_trace_scene(None, fake, 1, 0, C.byref(r), 8, 1, None, None, None) == 1
    assert lib.rt_trace_scene(None, fake, 1, 65, C.byref(r), 8, 0, C.byref(h), None, None) == 2
    for o, want in ((rtbvh.rt_trace_opts(0, 0, 0, 64, 8), 0), (rtbvh.rt_trace_opts(1, 0, 0, 0, 0), 1),
                    (rtbvh.rt_trace_opts(0, 0, 0, 96, 0), 1), (rtbvh.rt_trace_opts(0, 0, 0, 0, 49), 1)):
        assert lib.rt_trace_scene_ex(None, fake, 1, 0, C.byref(r), 0, 0, C.byref(o), C.byref(h),
                                     None, None) == want
    assert lib.rt_blur7(None, fake, fake, 4, 4, C.c_float(0.05), fake) == 1  # in place
    assert lib.rt_blur7(None, None, fake, 4, 4, C.c_float(0.05), fake) == 1
    assert lib.rt_ao_resolve(None, None, None, 0, None) == 0
    assert lib.rt_ao_resolve(None, None, fake, 5, fake) == 1


def test_radv_argument_errors():
    import rtbvh
    lib = rtbvh.lib()
    fake = C.c_void_p(64)
    assert lib.rt_gpu_build_radv(None, fake, None, 0, 0, fake, 1 << 20, None) == 2
    assert lib.rt_gpu_build_radv(None, fake, None, 4097, 0, fake, 1 << 20, None) == 2
    assert lib.rt_gpu_build_radv(None, fake, None, 16, 2, fake, 1 << 20, None) == 1
    assert lib.rt_gpu_build_radv(None, fake, None, 16, 0, C.c_void_p(72), 1 << 20, None) == 1
    assert lib.rt_gpu_build_radv(None, fake, None, 16, 0, fake, 100, None) == 2


def test_tuner_search_space_is_valid():
    """Every candidate the autotuner may pick is a shape the library accepts
    (checked on the host: rt_trace_ex with n = 0 validates the shape only)."""
    import autotune
    import rtbvh
    lib = rtbvh.lib()
    b, r = rtbvh.rt_bvh_dev(), rtbvh.rt_rays_soa()
    for blk in autotune.TRACE_BLOCKS:
        for lds in autotune.TRACE_LDS:
            for w in autotune.REFILL_RAYS_PER_WAVE:
                for lanes in autotune.REFILL_LANES:
                    o = rtbvh.trace_opts("refill", w, lanes, blk, lds)
                    assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, 0, C.byref(o), None,
                                           None, None, None) == 0
            for order in rtbvh.ANY_ORDERS:  # any hit, bits only, simple schedule
                o = rtbvh.trace_opts("simple", block=blk, lds_stack=lds, any_order=order)
                assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, rtbvh.RT_ANY,
                                       C.byref(o), None, None, None, None) == 0
    for bits in autotune.SORT_BITS:
        rtbvh.build_requirements(100_000, rtbvh.build_cfg("radix", 4, bits))
    assert autotune.LEAF_SIZES[0] == 4  # the default leaf size is the reference for the others
    for leaf in autotune.LEAF_SIZES:
        rtbvh.build_requirements(100_000, rtbvh.build_cfg("radix", leaf))


def test_build_batch_plan_on_host():
    """rt_gpu_build_batch_plan is host-only: meshes up to 65536 triangles;
    those above 4096 need the implicit build's scratch (RT_ERR_CAPACITY
    without it), an 80-B item record."""
    import rtbvh
    lib = rtbvh.lib()
    assert C.sizeof(rtbvh.rt_build_item) == 80
    fake = C.c_void_p(256)
    nbytes = C.c_size_t(0)
    assert lib.rt_gpu_build_batch_table_bytes(3, C.byref(nbytes)) == 0
    table = (C.c_uint8 * nbytes.value)()

    def item(n, scratch_bytes):
        cap = rtbvh.build_requirements(n, rtbvh.build_cfg("implicit"))[0]
        return rtbvh.rt_build_item(fake, None, fake, fake, fake, None, None, n, cap,
                                   fake if scratch_bytes else None, scratch_bytes)

    need = rtbvh.build_requirements(65536, rtbvh.build_cfg("implicit"))[1]
    items = (rtbvh.rt_build_item * 3)(item(65536, need), item(4096, 0), item(7, 0))
    assert lib.rt_gpu_build_batch_plan(items, 3, table, nbytes) == 0
    items[0] = item(65536, need - 1)
    assert lib.rt_gpu_build_batch_plan(items, 3, table, nbytes) == 2
    items[0] = item(65537, 1 << 30)
    assert lib.rt_gpu_build_batch_plan(items, 3, table, nbytes) == 2
    items[0] = item(5000, rtbvh.build_requirements(5000, rtbvh.build_cfg("implicit"))[1])
    assert lib.rt_gpu_build_batch_plan(items, 3, table, nbytes) == 0


def test_generic_refit_argument_checks():
    import rtbvh
    lib = rtbvh.lib()
    fake = C.c_void_p(256)
    assert lib.rt_gpu_refit_bvh4(None, None, None, None, 0, None, 0, None, 0) == 0  # empty
    assert lib.rt_gpu_refit_bvh4(None, None, None, fake, 10, fake, 5, fake, 40) == 1
    assert lib.rt_gpu_refit_bvh4(None, fake, None, fake, 10, fake, 65, fake, 40) == 1
    assert lib.rt_gpu_refit_bvh4(None, fake, None, fake, 10, fake, 5, fake, 39) == 2
    assert lib.rt_gpu_refit_bvh4(None, fake, None, fake, 10, fake, 5, None, 40) == 2


def test_terrain_resolution_matches_oracle(O):
    import rtbvh
    lib = rtbvh.lib()
    for target in (0, 2, 7, 968, 1000, 100_000, 1_000_000, 10_000_000):
        res, nv, nt = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        assert lib.rt_terrain_resolution(target, C.byref(res), C.byref(nv), C.byref(nt)) == 0
        assert res.value == O.lib().or_terrain_res(target)
        assert nv.value == res.value ** 2 and nt.value == 2 * (res.value - 1) ** 2
    assert lib.rt_gen_terrain(None, 1000, C.c_float(100.0), 0, None, None) == 1


def _bvh2_depth_py(n2):
    """Node levels of a BVH2 by an explicit walk from the root."""
    deep, st = 0, [(0, 1)]
    while st:
        i, d = st.pop()
        deep = max(deep, d)
        if n2["leftChild"][i] >= 0:
            st += [(int(n2["leftChild"][i]), d + 1), (int(n2["rightChild"][i]), d + 1)]
    return deep


def test_bvh2_depth_and_argument_checks(O):
    import rtbvh
    lib = rtbvh.lib()
    pos, idx = O.sphere(40, 30, 1.0, 0.05, 3)
    for mp in (1, 4, 8):
        n2, _, _ = O.bvh2_build(pos, idx, mp)
        d = C.c_uint32(0)
        assert lib.rt_bvh2_depth(n2.ctypes.data_as(C.c_void_p), len(n2), C.byref(d)) == 0
        assert d.value == _bvh2_depth_py(n2) > 1
    bad = n2.copy()
    bad["rightChild"][0] = len(bad) + 5  # child index out of range
    assert lib.rt_bvh2_depth(bad.ctypes.data_as(C.c_void_p), len(bad), C.byref(d)) == 1
    assert lib.rt_bvh2_depth(None, 0, C.byref(d)) == 0 and d.value == 0
    fake = C.c_void_p(256)
    b = rtbvh.rt_bvh2_dev(fake, fake, None, 10, 10, 65, 0)
    r = rtbvh.rt_rays_soa(*([fake] * 8))
    h = rtbvh.rt_hits_soa(fake, fake)
    assert lib.rt_trace_bvh2(None, C.byref(b), C.byref(r), 0, 0, None, C.byref(h), None, None,
                             None) == 0  # no rays: nothing to do
    assert lib.rt_trace_bvh2(None, C.byref(b), C.byref(r), 5, 0, None, C.byref(h), None, None,
                             None) == 2  # deeper than 64 levels
    b.depth = 8
    assert lib.rt_trace_bvh2(None, C.byref(b), C.byref(r), 5, 1, None, None, None, None,
                             None) == 1  # any hit without bits
    assert lib.rt_trace_bvh2(None, C.byref(b), C.byref(r), 5, 8, None, C.byref(h), None, None,
                             None) == 1  # unknown flag
    o = rtbvh.rt_trace_opts(1, 0, 0, 0, 0)  # refill schedule: not offered here
    assert lib.rt_trace_bvh2(None, C.byref(b), C.byref(r), 5, 0, C.byref(o), C.byref(h), None,
                             None, None) == 1


def test_image_metrics_host_checks():
    import rtbvh
    lib = rtbvh.lib()
    b = C.c_size_t(0)
    assert lib.rt_image_metrics_scratch_bytes(1920, 1080, C.byref(b)) == 0
    # one (squared error, SSIM) double pair per wave of 64 threads, one block each
    assert b.value == 16 * -(-(1920 // 8) * (1080 // 8) // 64)
    assert lib.rt_image_metrics_scratch_bytes(0, 0, C.byref(b)) == 0 and b.value == 16
    assert lib.rt_image_metrics_scratch_bytes(4, 4, None) == 1
    fake = C.c_void_p(256)
    assert lib.rt_image_metrics(None, fake, fake, 8, 8, None, None, fake, 1 << 20) == 1
    assert lib.rt_image_metrics(None, None, fake, 8, 8, fake, None, fake, 1 << 20) == 1
    assert lib.rt_image_metrics(None, fake, fake, 8, 8, fake, None, None, 1 << 20) == 1
    assert lib.rt_image_metrics(None, fake, fake, 8, 8, fake, None, fake, 8) == 2
    assert lib.rt_image_metrics(None, fake, fake, 8, 8, fake, C.c_void_p(260), fake,
                                1 << 20) == 1  # block output not 8-byte aligned


def test_any_order_checks():
    """any_order: rt_trace_ex only, any hit without counters or hit ids, the
    simple schedule; values 0..2; other entry points take 0 only."""
    import rtbvh
    lib = rtbvh.lib()
    assert rtbvh.rt_trace_opts.any_order.offset == 20
    b = rtbvh.rt_bvh_dev()
    r = rtbvh.rt_rays_soa()
    fake = C.c_void_p(256)
    hits = rtbvh.rt_hits_soa(fake, fake)
    for order in (0, 1, 2):
        o = rtbvh.rt_trace_opts(0, 0, 0, 0, 0, order)
        assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, 1, C.byref(o), None, None, None,
                               None) == 0
    o = rtbvh.rt_trace_opts(0, 0, 0, 0, 0, 3)
    assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, 1, C.byref(o), None, None, None,
                           None) == 1
    far = rtbvh.rt_trace_opts(0, 0, 0, 0, 0, 1)
    for flags, h in ((0, None), (1 | 0x100, None), (1, C.byref(hits)), (1 | 0x200, None)):
        assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, flags, C.byref(far), h, None,
                               None, None) == 1, flags
    refill = rtbvh.rt_trace_opts(1, 0, 0, 0, 0, 1)
    assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, 1, C.byref(refill), None, None,
                           None, None) == 1
    assert lib.rt_trace_batch_ex(None, None, None, 0, 0, 1, C.byref(far)) == 1
    assert lib.rt_trace_scene_ex(None, fake, 1, 0, C.byref(r), 0, 0, C.byref(far), None, None,
                                 None) == 1


def test_tile_checks():
    """grid_width / tile_width: rt_trace_ex only, tile width 8 / 16 / 32
    dividing the grid width, whole rows, simple schedule."""
    import rtbvh
    lib = rtbvh.lib()
    assert rtbvh.rt_trace_opts.grid_width.offset == 24
    assert rtbvh.rt_trace_opts.tile_width.offset == 28
    b = rtbvh.rt_bvh_dev()
    r = rtbvh.rt_rays_soa()
    fake = C.c_void_p(256)
    r5 = rtbvh.rt_rays_soa(*([fake] * 8))
    h = rtbvh.rt_hits_soa(fake, fake)
    for gw, tw, ok in [(0, 0, True), (64, 8, True), (64, 16, True), (64, 32, True),
                       (64, 64, False), (64, 4, False), (0, 8, False), (24, 16, False),
                       (64, 0, False)]:
        o = rtbvh.rt_trace_opts(0, 0, 0, 0, 0, 0, gw, tw)
        assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, 0, C.byref(o), None, None, None,
                               None) == (0 if ok else 1), (gw, tw)
    o = rtbvh.rt_trace_opts(0, 0, 0, 0, 0, 0, 64, 8)
    # 100 rays are not whole rows of 64
    assert lib.rt_trace_ex(None, C.byref(b), C.byref(r5), 100, 0, C.byref(o), C.byref(h), None,
                           None, None) == 1
    o.schedule = 1
    assert lib.rt_trace_ex(None, C.byref(b), C.byref(r), 0, 0, C.byref(o), None, None, None,
                           None) == 1
    o.schedule = 0
    assert lib.rt_trace_batch_ex(None, None, None, 0, 0, 1, C.byref(o)) == 1
    # the scene trace takes wave tiles (whole rows) but no child order
    assert lib.rt_trace_scene_ex(None, fake, 1, 0, C.byref(r), 0, 0, C.byref(o), None, None,
                                 None) == 0
    assert lib.rt_trace_scene_ex(None, fake, 1, 0, C.byref(r5), 100, 0, C.byref(o),
                                 C.byref(h), None, None) == 1
