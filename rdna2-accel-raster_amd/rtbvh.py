"""Python host side of librtbvh.so (ctypes over the C ABI in include/rt_bvh.h).

Mirrors the reference's software_bvh interface for this path:
  SoftwareBVH::build / SoftwareBVH4::fromBVH2 (common/software_bvh.hpp:35,193)
      -> GpuBVH4.build (device Morton/Box4 build) and GpuBVH4.from_bvh4 (upload
         of a host-built SoftwareBVH4 tree)
  software_bvh_bvh4.comp traversal and the inline ray queries
      -> GpuBVH4.intersect / GpuBVH4.occluded (rt_trace)
PyTorch supplies device memory and the HIP stream; all compute is the HIP
library.  There is no CPU fallback: without librtbvh.so, or without a GPU,
every entry point raises.
"""
import ctypes as C
import math
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtbvh.so")
MISS = 0xFFFFFFFF

RT_CLOSEST, RT_ANY, RT_COUNTERS, RT_TRACE_PERSISTENT = 0x0, 0x1, 0x100, 0x200
_STATUS = {0: "RT_OK", 1: "RT_ERR_ARG", 2: "RT_ERR_CAPACITY", 3: "RT_ERR_HIP",
           4: "RT_ERR_NOMEM", 5: "RT_ERR_UNSUPPORTED"}

NODE_BYTES = 128  # rt_node4 == GpuBvh4Node
MAX_DEPTH = 64    # deepest tree the trace / refit entry points take
TRI_BYTES = 64    # rt_tri64 == RadvBvhTriangleNode


class RtError(RuntimeError):
    def __init__(self, fn, code, detail=""):
        self.code = code
        super().__init__(f"{fn} -> {_STATUS.get(code, code)} {detail}".strip())


vp = C.c_void_p
u32 = C.c_uint32
f32 = C.c_float


class rt_bvh_dev(C.Structure):
    _fields_ = [("nodes", vp), ("tris", vp), ("node_count", u32), ("tri_count", u32),
                ("depth", u32), ("format", u32)]


NODES_F32, NODES_Q8 = 0, 1   # rt_bvh_dev.format
QNODE_BYTES = 64             # rt_qnode4


class rt_rays_soa(C.Structure):
    _fields_ = [(n, vp) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]


class rt_rays_out(C.Structure):
    _fields_ = [(n, vp) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]


class rt_hits_soa(C.Structure):
    _fields_ = [("prim", vp), ("t", vp)]


class rt_bvh2_dev(C.Structure):
    _fields_ = [("nodes", vp), ("tris48", vp), ("order", vp), ("node_count", u32),
                ("tri_count", u32), ("depth", u32), ("reserved", u32)]


class rt_radv_info(C.Structure):
    _fields_ = [("bvh_offset", u32), ("root_flags", u32), ("aabb", C.c_float * 6),
                ("update_dispatch_size", u32 * 3), ("geometry_type", u32),
                ("geometry_count", u32), ("leaf_node_offsets_offset", u32),
                ("build_flags", u32), ("type", u32), ("compacted_size", C.c_uint64),
                ("serialization_size", C.c_uint64), ("size", C.c_uint64),
                ("copy_dispatch_size", u32 * 3), ("box_count", u32),
                ("leaf_nodes_offset", C.c_uint64), ("primitive_count", C.c_uint64),
                ("instance_count", C.c_uint64), ("root_children", u32 * 4)]


class rt_radv_dev(C.Structure):
    _fields_ = [("image", vp), ("bytes", C.c_uint64), ("bvh_offset", u32), ("box_count", u32),
                ("tri_count", u32), ("depth", u32)]


class rt_trace_item(C.Structure):
    _fields_ = [("bvh", rt_bvh_dev), ("rays", rt_rays_soa), ("count", u32), ("flags", u32),
                ("hits", rt_hits_soa), ("occlusion_bits", vp)]


class rt_instance(C.Structure):
    _fields_ = [("blas", rt_bvh_dev), ("pos", vp), ("idx", vp),
                ("world_to_object", f32 * 12), ("kind", u32), ("reserved", u32 * 7)]


INST_GENERAL, INST_TRANSLATE, INST_IDENTITY = 0, 1, 2


class rt_ray_gen(C.Structure):
    _fields_ = [("seed", u32), ("lo", f32 * 3), ("hi", f32 * 3), ("tmax", f32)]


class rt_grid_gen(C.Structure):
    _fields_ = [("t", f32), ("extent", f32)]


class rt_build_item(C.Structure):
    _fields_ = [("pos", vp), ("idx", vp), ("nodes", vp), ("tris", vp), ("info", vp),
                ("sorted_ids", vp), ("codes", vp), ("n", u32), ("node_capacity", u32),
                ("scratch", vp), ("scratch_bytes", C.c_uint64)]


class rt_build_cfg(C.Structure):
    _fields_ = [("algorithm", u32), ("max_leaf_prims", u32), ("sort_bits", u32),
                ("treelet_rounds", u32)]


# Digit widths of the multi-launch Morton radix sort (same output for each);
# DEFAULT_SORT_BITS is used when a build does not choose one.
SORT_BITS = (4, 6, 8, 10)
DEFAULT_SORT_BITS = 8


BUILD_ALGOS = {"auto": 0, "radix": 1, "implicit": 2}


class rt_trace_opts(C.Structure):
    _fields_ = [("schedule", u32), ("rays_per_wave", u32), ("refill_lanes", u32),
                ("block", u32), ("lds_stack", u32), ("any_order", u32), ("grid_width", u32),
                ("tile_width", u32)]


class rt_occluder_hint(C.Structure):
    _fields_ = [("prim", vp), ("record", vp), ("period", u32), ("ids", u32), ("flags", u32)]


RT_HINT_WAVE = 1
# rt_gpu_build_batch workgroups flag: meshes above 4096 triangles built by
# 256-thread workgroups that share compute units with running traces
BUILD_BATCH_SHARED = 0x80000000


# Child order of an any-hit query returning occlusion bits only (same bits
# for every order; the work differs).
ANY_ORDERS = {"near": 0, "far": 1, "slot": 2}


# Traversal schedule used when a call does not choose one: "simple" (one ray
# per thread) or "refill" (waves walk ray chunks, refilling finished lanes),
# with the chunk size, refill threshold, block size and LDS stack depth of
# TRACE_TUNING; "packet" (a wave walks one node order for its rays) for
# occlusion bits only.  Every choice gives the same results (autotune.py
# measures them).
SCHEDULES = {"simple": 0, "refill": 1, "packet": 2}
DEFAULT_SCHEDULE = "simple"
TRACE_TUNING = {"rays_per_wave": 1024, "refill_lanes": 16, "block": 256, "lds_stack": 20}
TRACE_BLOCKS = (64, 128, 256)
LDS_STACK_RANGE = (4, 48)


def trace_opts(schedule=None, rays_per_wave=None, refill_lanes=None, block=None,
               lds_stack=None, any_order=None, tile=None):
    """tile: (grid_width, tile_width) wave tiles of a ray grid, or None."""
    s = schedule or DEFAULT_SCHEDULE
    if s not in SCHEDULES:
        raise ValueError(f"schedule {s!r}")
    gw, tw = tile if tile else (0, 0)
    return rt_trace_opts(SCHEDULES[s],
                         rays_per_wave or TRACE_TUNING["rays_per_wave"],
                         refill_lanes or TRACE_TUNING["refill_lanes"],
                         block or TRACE_TUNING["block"],
                         lds_stack or TRACE_TUNING["lds_stack"],
                         ANY_ORDERS[any_order or "near"], gw, tw)


_lib = None

# name -> argtypes (restype is rt_status = int for all but the string getters)
_SIGS = {
    "rt_tree_layout": [u32, vp, vp, vp, vp, vp],
    "rt_build_requirements": [u32, vp, vp, vp],
    "rt_gpu_build": [vp, vp, vp, u32, vp, vp, u32, vp, vp, vp, vp, vp, C.c_size_t],
    "rt_gpu_refit": [vp, vp, vp, u32, vp, u32, vp, vp, C.c_size_t],
    "rt_gpu_refit_bvh4": [vp, vp, vp, vp, u32, vp, u32, vp, C.c_size_t],
    "rt_convert_tris48": [vp, vp, vp, u32, vp],
    "rt_trace": [vp, vp, vp, u32, u32, vp, vp, vp, vp],
    "rt_trace_ex": [vp, vp, vp, u32, u32, vp, vp, vp, vp, vp],
    "rt_radv_describe": [vp, C.c_uint64, vp],
    "rt_trace_radv": [vp, vp, vp, u32, u32, vp, vp, vp, vp, vp],
    "rt_bvh2_depth": [vp, u32, vp],
    "rt_trace_bvh2": [vp, vp, vp, u32, u32, vp, vp, vp, vp, vp],
    "rt_gen_sphere": [vp, u32, u32, f32, f32, u32, vp, vp],
    "rt_terrain_resolution": [u32, vp, vp, vp],
    "rt_gen_terrain": [vp, u32, f32, C.c_int, vp, vp],
    "rt_gen_wave_grid": [vp, u32, u32, f32, f32, f32, f32, f32, vp, vp],
    "rt_gen_heightfield": [vp, u32, u32, f32, f32, u32, vp, vp],
    "rt_gen_dynamic_grid": [vp, u32, f32, f32, vp],
    "rt_gen_camera_rays": [vp, u32, u32, vp, vp, vp, f32, f32, f32, vp],
    "rt_gen_shadow_rays": [vp, vp, vp, vp, vp, vp, u32, vp, u32, f32, vp],
    "rt_gen_ao_rays": [vp, vp, vp, vp, vp, vp, u32, u32, u32, f32, f32, f32, vp],
    "rt_gen_random_rays": [vp, u32, u32, vp, vp, f32, vp],
    "rt_gen_exp0_rays": [vp, u32, C.c_int, vp, vp, f32, f32, vp],
    "rt_pack_light_masks": [vp, vp, u32, u32, vp],
    "rt_gpu_build_radv": [vp, vp, vp, u32, u32, vp, C.c_uint64, vp],
    "rt_gpu_build_batch_table_bytes": [u32, vp],
    "rt_gpu_build_batch_plan": [vp, u32, vp, C.c_size_t],
    "rt_gpu_build_batch": [vp, vp, u32, u32],
    "rt_trace_batch_plan": [vp, u32, vp, vp],
    "rt_trace_batch": [vp, vp, vp, u32, u32, u32],
    "rt_trace_batch_ex": [vp, vp, vp, u32, u32, u32, vp],
    "rt_trace_batch_plan_ex": [vp, u32, u32, vp, vp],
    "rt_trace_scene": [vp, vp, u32, u32, vp, u32, u32, vp, vp, vp],
    "rt_trace_scene_ex": [vp, vp, u32, u32, vp, u32, u32, vp, vp, vp, vp],
    "rt_gen_ao_rays_scene": [vp, vp, u32, vp, vp, vp, vp, u32, u32, u32, f32, f32, f32, vp],
    "rt_ao_resolve": [vp, vp, vp, u32, vp],
    "rt_blur7": [vp, vp, vp, u32, u32, f32, vp],
    "rt_image_metrics_scratch_bytes": [u32, u32, vp],
    "rt_image_metrics": [vp, vp, vp, u32, u32, vp, vp, vp, C.c_size_t],
    "rt_device_sync": [],
    "rt_gpu_build_noalloc": [vp, vp, vp, u32, vp, vp, u32, vp, vp, vp, vp],
    "rt_gpu_quantize": [vp, vp, u32, vp],
    "rt_trace_batch_submit_bytes": [u32, C.c_int, vp],
    "rt_trace_batch_submit": [vp, vp, vp, u32, vp, vp, C.c_size_t, vp, C.c_size_t],
    "rt_gpu_build_batch_submit_bytes": [u32, C.c_int, vp],
    "rt_gpu_build_batch_submit": [vp, vp, vp, u32, u32, vp, C.c_size_t, vp, C.c_size_t],
    "rt_prim_records": [vp, vp, u32, vp],
    "rt_trace_occluded_hinted_scratch_bytes": [u32, vp, vp, vp],
    "rt_trace_occluded_hinted": [vp, vp, vp, u32, vp, vp, vp, vp, C.c_size_t],
}


def lib():
    """Load librtbvh.so; raises when it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"librtbvh.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.rt_version.restype = C.c_char_p
        L.rt_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS) + ["rt_version", "rt_last_error"]


def _check(fn, code):
    if code != 0:
        detail = lib().rt_last_error().decode() if code == 3 else ""
        raise RtError(fn, code, detail)


def call(name, *args):
    _check(name, getattr(lib(), name)(*args))


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("rtbvh needs a GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return vp(s.cuda_stream)


def ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    return vp(t.data_ptr())


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


_INDEX_DTYPES = (torch.int32, getattr(torch, "uint32", torch.int32))


def check_mesh(pos, idx):
    """Raise ValueError unless pos / idx are what the C ABI reads: pos a
    contiguous float32 device tensor (3 floats per vertex, or 9 per triangle
    when idx is None), idx None or a contiguous int32 / uint32 device tensor
    of shape (T, 3)."""
    if not isinstance(pos, torch.Tensor) or not pos.is_cuda:
        raise ValueError("pos: a device tensor expected")
    if pos.dtype != torch.float32 or not pos.is_contiguous():
        raise ValueError(f"pos: contiguous float32 expected (got {pos.dtype}, "
                         f"contiguous={pos.is_contiguous()})")
    if idx is None:
        if pos.numel() % 9:
            raise ValueError("pos without idx: 9 floats per triangle expected")
        return
    if not isinstance(idx, torch.Tensor) or not idx.is_cuda:
        raise ValueError("idx: a device tensor expected")
    if idx.dtype not in _INDEX_DTYPES or not idx.is_contiguous():
        raise ValueError(f"idx: contiguous int32 / uint32 expected (got {idx.dtype}, "
                         f"contiguous={idx.is_contiguous()})")
    if idx.dim() != 2 or idx.shape[1] != 3:
        raise ValueError(f"idx: shape (T, 3) expected (got {tuple(idx.shape)})")
    if pos.numel() % 3:
        raise ValueError("pos: 3 floats per vertex expected")


def check_out(t, count, dtypes, what):
    """Raise ValueError unless t is a contiguous device tensor of one of
    `dtypes` holding at least `count` elements."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{what}: a device tensor expected")
    if t.dtype not in dtypes or not t.is_contiguous():
        raise ValueError(f"{what}: contiguous {dtypes} expected (got {t.dtype})")
    if t.numel() < count:
        raise ValueError(f"{what}: {t.numel()} elements < {count} needed")


_I32 = _INDEX_DTYPES
_F32 = (torch.float32,)


# ------------------------------------------------------------------ layout --
def tree_layout(n):
    counts = (u32 * 32)()
    offs = (u32 * 32)()
    L, tb = u32(0), u32(0)
    sz = C.c_uint64(0)
    call("rt_tree_layout", u32(n), counts, offs, C.byref(L), C.byref(tb), C.byref(sz))
    return list(counts[:L.value]), list(offs[:L.value]), tb.value, sz.value


def build_cfg(algorithm="auto", max_leaf_prims=4, sort_bits=None, treelet_rounds=0):
    return rt_build_cfg(BUILD_ALGOS[algorithm], max_leaf_prims,
                        sort_bits or DEFAULT_SORT_BITS, treelet_rounds)


def build_requirements(n, cfg=None):
    """(node capacity, scratch bytes) of rt_gpu_build for n triangles."""
    cap = u32(0)
    s = C.c_size_t(0)
    call("rt_build_requirements", u32(n), C.byref(cfg) if cfg else None, C.byref(cap),
         C.byref(s))
    return cap.value, s.value


# --------------------------------------------------------------------- rays --
class Rays:
    """SoA ray batch on the device: one (8, n) float32 tensor, rows
    ox, oy, oz, tmin, dx, dy, dz, tmax (32 B per ray)."""
    FIELDS = ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")

    def __init__(self, data):
        if data.dim() != 2 or data.shape[0] != 8 or data.dtype != torch.float32:
            raise ValueError("Rays: a float32 (8, n) tensor expected")
        self.data = data.contiguous()

    @staticmethod
    def empty(n, device=None):
        return Rays(torch.empty((8, n), dtype=torch.float32, device=device or _dev()))

    @staticmethod
    def from_numpy(a, device=None):
        return Rays(torch.as_tensor(a, dtype=torch.float32).to(device or _dev()))

    def __len__(self):
        return self.data.shape[1]

    def soa(self):
        p = self.data.data_ptr()
        stride = self.data.stride(0) * 4
        return rt_rays_soa(*[vp(p + i * stride) for i in range(8)])

    def out(self):
        p = self.data.data_ptr()
        stride = self.data.stride(0) * 4
        return rt_rays_out(*[vp(p + i * stride) for i in range(8)])


# ---------------------------------------------------------------- the BVH --
class GpuBVH4:
    """Device-resident Box4 BVH: rt_node4 nodes (GpuBvh4Node layout) and
    rt_tri64 leaves (RadvBvhTriangleNode layout, carrying triangle IDs)."""

    def __init__(self, nodes, tris, node_count, tri_count, depth):
        self.nodes = nodes            # uint8 tensor, node_count * 128 B
        self.tris = tris              # uint8 tensor, tri_count * 64 B
        self.node_count = node_count
        self.tri_count = tri_count
        self.depth = depth
        self.sorted_ids = None
        self.codes = None
        self._scratch = None

    # -- build ---------------------------------------------------------------
    @staticmethod
    def build(pos, idx=None, algorithm="auto", max_leaf_prims=4, stream=None, keep_order=False,
              scratch=None, sync=True, sort_bits=None, treelet_rounds=0):
        """Device build (rt_gpu_build) of a Morton-sorted Box4 tree.

        pos: float32 (V, 3) device tensor; idx: uint32/int32 (T, 3) or None
        for 9 floats per triangle (pos then has shape (T, 9) or (3T, 3)).
        algorithm: "auto" | "radix" | "implicit" (see rt_bvh.h).  sort_bits:
        radix digit width of the multi-launch sort (SORT_BITS; default
        DEFAULT_SORT_BITS).  treelet_rounds ("radix" only): rounds of SAH
        treelet restructuring before the Box4 collapse (fewer node visits per
        ray, the same leaves).  With sync=True the node count and depth are read
        back from the device; otherwise node_count is the capacity and depth 0
        (unknown)."""
        check_mesh(pos, idx)
        dev = pos.device
        n = idx.shape[0] if idx is not None else pos.numel() // 9
        cfg = build_cfg(algorithm, max_leaf_prims, sort_bits, treelet_rounds)
        cap, need = build_requirements(n, cfg)
        nodes = torch.empty(max(cap, 1) * NODE_BYTES, dtype=torch.uint8, device=dev)
        tris = torch.empty(max(n, 1) * TRI_BYTES, dtype=torch.uint8, device=dev)
        info = torch.empty(4, dtype=torch.int32, device=dev)  # every build path writes it
        bvh = GpuBVH4(nodes, tris, cap, n, 0)
        bvh.algorithm = algorithm
        ids = codes = None
        if keep_order:
            ids = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
            codes = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        if scratch is None or scratch.numel() < need:
            scratch = torch.empty(max(need, 256), dtype=torch.uint8, device=dev)
        bvh._scratch = scratch
        call("rt_gpu_build", stream_handle(stream), ptr(pos), ptr(idx), u32(n), C.byref(cfg),
             ptr(nodes), u32(cap), ptr(tris), ptr(info), ptr(ids), ptr(codes), ptr(scratch),
             C.c_size_t(scratch.numel()))
        bvh.info = info
        if sync:
            bvh.node_count, bvh.depth = [int(v) for v in info[:2].cpu().tolist()]
            if bvh.depth > MAX_DEPTH:
                raise RtError("rt_gpu_build", 2, f"tree depth {bvh.depth} > {MAX_DEPTH}")
        bvh.sorted_ids, bvh.codes = ids, codes
        bvh.pos, bvh.idx = pos, idx
        return bvh

    def refit(self, pos, idx=None, stream=None, generic=None):
        """Same topology, new vertex positions.  Implicit trees use the
        implicit refit (rt_gpu_refit); any other tree (radix, uploaded
        SoftwareBVH4), or generic=True, the generic one (rt_gpu_refit_bvh4),
        which needs the node count and depth (sync=True builds have them)."""
        check_mesh(pos, idx)
        if _tri_count(pos, idx) != self.tri_count:
            raise ValueError(f"refit: {_tri_count(pos, idx)} triangles, the tree has "
                             f"{self.tri_count}")
        if generic is None:
            generic = getattr(self, "algorithm", None) != "implicit"
        if generic:
            if getattr(self, "info", None) is not None:
                # the real node count and depth (a sync=False build holds the
                # capacity, whose tail nodes are uninitialised)
                self.node_count, self.depth = [int(v) for v in self.info[:2].cpu().tolist()]
            need = max(256, self.node_count * 4)
            if self._scratch is None or self._scratch.numel() < need:
                self._scratch = torch.empty(need, dtype=torch.uint8, device=pos.device)
            call("rt_gpu_refit_bvh4", stream_handle(stream), ptr(pos), ptr(idx), ptr(self.nodes),
                 u32(self.node_count), ptr(self.tris), u32(self.depth), ptr(self._scratch),
                 C.c_size_t(self._scratch.numel()))
            self.pos, self.idx = pos, idx
            return self
        need = build_requirements(self.tri_count, build_cfg("implicit"))[1]
        if self._scratch is None or self._scratch.numel() < need:
            self._scratch = torch.empty(max(need, 256), dtype=torch.uint8, device=pos.device)
        call("rt_gpu_refit", stream_handle(stream), ptr(pos), ptr(idx), u32(self.tri_count),
             ptr(self.nodes), u32(self.node_count), ptr(self.tris), ptr(self._scratch),
             C.c_size_t(self._scratch.numel()))
        self.pos, self.idx = pos, idx
        return self

    @staticmethod
    def from_bvh4(nodes4, tris48=None, order=None, tris64=None, depth=None, device=None):
        """Upload a host-built SoftwareBVH4 (numpy structured/bytes arrays).

        nodes4: GpuBvh4Node records (128 B each).  Either tris64 (64-B leaf
        records) or tris48 (GpuTriangle, 48 B) plus order (leaf-order original
        triangle indices), which rt_convert_tris48 turns into 64-B leaves."""
        dev = device or _dev()
        nb = np.ascontiguousarray(nodes4).view(np.uint8).reshape(-1)
        ncount = nb.size // NODE_BYTES
        nodes = torch.from_numpy(nb.copy()).to(dev) if nb.size else torch.empty(
            NODE_BYTES, dtype=torch.uint8, device=dev)
        if tris64 is not None:
            tb = np.ascontiguousarray(tris64).view(np.uint8).reshape(-1)
            n = tb.size // TRI_BYTES
            tris = torch.from_numpy(tb.copy()).to(dev)
        else:
            t48 = torch.from_numpy(np.ascontiguousarray(tris48, dtype=np.float32).reshape(-1)).to(dev)
            n = t48.numel() // 12
            tris = torch.empty(max(n, 1) * TRI_BYTES, dtype=torch.uint8, device=dev)
            o = None
            if order is not None:
                o = torch.from_numpy(np.ascontiguousarray(order, dtype=np.uint32).view(np.int32)).to(dev)
            call("rt_convert_tris48", stream_handle(), ptr(t48), ptr(o), u32(n), ptr(tris))
            torch.cuda.current_stream().synchronize()
        if depth is None:
            depth = _depth_host(np.ascontiguousarray(nodes4))
        return GpuBVH4(nodes, tris, ncount, n, depth)

    # -- trace -----------------------------------------------------------------
    format = NODES_F32

    def dev(self):
        return rt_bvh_dev(ptr(self.nodes), ptr(self.tris), u32(self.node_count),
                          u32(self.tri_count), u32(self.depth), u32(self.format))

    def quantize(self, stream=None, out=None):
        """The RT_NODES_Q8 copy of this tree (rt_gpu_quantize): 64-B nodes
        with 8-bit child boxes rounded outward, the same leaves (this tree's
        triangle records are shared).  It traces the same triangles as this
        tree with half the node bytes per visit (rt_trace / rt_trace_ex,
        one ray per thread).  Needs the real node count (a sync build)."""
        if self.format != NODES_F32:
            raise ValueError("already quantized")
        if getattr(self, "info", None) is not None and self.depth == 0:
            self.node_count, self.depth = [int(v) for v in self.info[:2].cpu().tolist()]
        q = out if out is not None else torch.empty(max(self.node_count, 1) * QNODE_BYTES,
                                                    dtype=torch.uint8, device=self.nodes.device)
        check_out(q, self.node_count * QNODE_BYTES, (torch.uint8,), "out (quantized nodes)")
        call("rt_gpu_quantize", stream_handle(stream), ptr(self.nodes), u32(self.node_count),
             ptr(q))
        b = GpuBVH4(q, self.tris, self.node_count, self.tri_count, self.depth)
        b.format = NODES_Q8
        b.source = self
        return b

    def intersect(self, rays, counters=False, stream=None, out=None, schedule=None, **tune):
        """Closest hit per ray: (prim uint32-as-int32, t float32[, box, tri]).
        schedule: "simple" | "refill" (default DEFAULT_SCHEDULE); tune:
        block, lds_stack, and rays_per_wave, refill_lanes for "refill"."""
        n = len(rays)
        dev = rays.data.device
        prim = out[0] if out else torch.empty(n, dtype=torch.int32, device=dev)
        t = out[1] if out else torch.empty(n, dtype=torch.float32, device=dev)
        if out:
            check_out(prim, n, _I32, "out[0] (prim)")
            check_out(t, n, _F32, "out[1] (t)")
        box = tri = None
        flags = RT_CLOSEST
        if counters:
            box = torch.empty(n, dtype=torch.int32, device=dev)
            tri = torch.empty(n, dtype=torch.int32, device=dev)
            flags |= RT_COUNTERS
        hits = rt_hits_soa(ptr(prim), ptr(t))
        b = self.dev()
        r = rays.soa()
        o = trace_opts(schedule, **tune)
        call("rt_trace_ex", stream_handle(stream), C.byref(b), C.byref(r), u32(n), u32(flags),
             C.byref(o), C.byref(hits), None, ptr(box), ptr(tri))
        return (prim, t, box, tri) if counters else (prim, t)

    def prim_records(self, stream=None):
        """Device map triangle id -> leaf record index (rt_prim_records) over
        ids 0 .. tri_count - 1, made once per tree object (the hint table of
        occluded(hint=...)); a refit keeps it valid (leaf records keep their
        triangles), a build makes a new tree object."""
        rec = getattr(self, "_prim_records", None)
        if rec is None:
            rec = torch.empty(max(1, self.tri_count), dtype=torch.int32, device=self.tris.device)
            b = self.dev()
            call("rt_prim_records", stream_handle(stream), C.byref(b), u32(self.tri_count),
                 ptr(rec))
            self._prim_records = rec
        return rec

    def occluded(self, rays, stream=None, out=None, counters=False, schedule=None, hint=None,
                 hint_period=0, hint_wave=True, **tune):
        """Any-hit occlusion bits: int32 tensor of ceil(n/32) words.

        hint: int32 device tensor of triangle ids (RT_MISS = none), ray r's
        hint being hint[r % hint_period] (0 = n), e.g. the primary hits of a
        frame for its shadow rays; hint_wave also tries the hints of the
        other rays of each wave.  The bits are the same with or without hints
        (rt_trace_occluded_hinted); tune: block, lds_stack, tile,
        rays_per_wave (slots of a segment), refill_lanes.  The triangle-id ->
        leaf-record map and the scratch of hinted calls are kept with the
        tree, so hinted calls on one tree from several streams must be
        ordered by the caller (one stream, or events between them)."""
        n = len(rays)
        dev = rays.data.device
        bits = out if out is not None else torch.empty(max(1, (n + 31) // 32), dtype=torch.int32,
                                                       device=dev)
        check_out(bits, (n + 31) // 32, _I32, "out (occlusion bits)")
        if hint is not None:
            if counters or schedule not in (None, "refill"):
                raise ValueError("hints: occlusion bits only, refill schedule")
            period = hint_period or n
            check_out(hint, period, _I32, "hint (triangle ids)")
            rec = self.prim_records(stream)
            h = rt_occluder_hint(ptr(hint), ptr(rec), u32(period), u32(self.tri_count),
                                 u32(RT_HINT_WAVE if hint_wave else 0))
            b = self.dev()
            r = rays.soa()
            o = trace_opts("refill", **tune)
            need = C.c_size_t(0)
            call("rt_trace_occluded_hinted_scratch_bytes", u32(n), C.byref(o), C.byref(h),
                 C.byref(need))
            sc = getattr(self, "_hint_scratch", None)
            if sc is None or sc.numel() < need.value:
                sc = torch.empty(max(256, need.value), dtype=torch.uint8, device=dev)
                self._hint_scratch = sc
            call("rt_trace_occluded_hinted", stream_handle(stream), C.byref(b), C.byref(r),
                 u32(n), C.byref(o), C.byref(h), ptr(bits), ptr(sc), C.c_size_t(sc.numel()))
            return bits
        box = tri = None
        flags = RT_ANY
        if counters:
            box = torch.empty(n, dtype=torch.int32, device=dev)
            tri = torch.empty(n, dtype=torch.int32, device=dev)
            flags |= RT_COUNTERS
        b = self.dev()
        r = rays.soa()
        o = trace_opts(schedule, **tune)
        call("rt_trace_ex", stream_handle(stream), C.byref(b), C.byref(r), u32(n), u32(flags),
             C.byref(o), None, ptr(bits), ptr(box), ptr(tri))
        return (bits, box, tri) if counters else bits


class GpuBVH2:
    """Device copy of a SoftwareBVH as its builder lays it out (GpuBvhNode
    32-B nodes, GpuTriangle 48-B triangles in leaf order), traversed by
    rt_trace_bvh2 with the loop of the reference's software BVH shader
    (software_bvh_procedural.comp:47-171).  order: original triangle index
    of each leaf position (reported as the hit id; None reports the
    position)."""

    def __init__(self, nodes, tris48, order, node_count, tri_count, depth):
        self.nodes, self.tris48, self.order = nodes, tris48, order
        self.node_count, self.tri_count, self.depth = node_count, tri_count, depth

    @staticmethod
    def from_bvh2(nodes2, tris48, order=None, device=None):
        dev = device or _dev()
        nb = np.ascontiguousarray(nodes2).view(np.uint8).reshape(-1)
        if nb.size % 32:
            raise ValueError("nodes2: 32-byte GpuBvhNode records expected")
        ncount = nb.size // 32
        depth = u32(0)
        call("rt_bvh2_depth", nb.ctypes.data_as(vp) if ncount else None, u32(ncount),
             C.byref(depth))
        t = np.ascontiguousarray(tris48, dtype=np.float32).reshape(-1)
        if t.size % 12:
            raise ValueError("tris48: 48-byte GpuTriangle records expected")
        nodes = torch.from_numpy(nb.copy()).to(dev) if ncount else torch.empty(
            32, dtype=torch.uint8, device=dev)
        tris = torch.from_numpy(t.copy()).to(dev) if t.size else torch.empty(
            12, dtype=torch.float32, device=dev)
        o = None
        if order is not None:
            o = torch.from_numpy(np.ascontiguousarray(order, dtype=np.uint32).view(np.int32)).to(dev)
        return GpuBVH2(nodes, tris, o, ncount, t.size // 12, depth.value)

    def dev(self):
        return rt_bvh2_dev(ptr(self.nodes), ptr(self.tris48), ptr(self.order),
                           u32(self.node_count), u32(self.tri_count), u32(self.depth), u32(0))

    def _trace(self, rays, flags, outs, counters, stream, tune):
        n = len(rays)
        box = tri = None
        if counters:
            box = torch.empty(n, dtype=torch.int32, device=rays.data.device)
            tri = torch.empty(n, dtype=torch.int32, device=rays.data.device)
            flags |= RT_COUNTERS
        b = self.dev()
        r = rays.soa()
        o = trace_opts("simple", **tune)
        hits, bits = outs
        call("rt_trace_bvh2", stream_handle(stream), C.byref(b), C.byref(r), u32(n), u32(flags),
             C.byref(o), C.byref(hits) if hits is not None else None, ptr(bits), ptr(box),
             ptr(tri))
        return box, tri

    def intersect(self, rays, counters=False, stream=None, **tune):
        """Closest hit per ray: (prim uint32-as-int32, t float32[, box, tri]).
        tune: block, lds_stack."""
        n = len(rays)
        prim = torch.empty(n, dtype=torch.int32, device=rays.data.device)
        t = torch.empty(n, dtype=torch.float32, device=rays.data.device)
        box, tri = self._trace(rays, RT_CLOSEST, (rt_hits_soa(ptr(prim), ptr(t)), None),
                               counters, stream, tune)
        return (prim, t, box, tri) if counters else (prim, t)

    def occluded(self, rays, counters=False, stream=None, **tune):
        """Any-hit occlusion bits: int32 tensor of ceil(n/32) words."""
        n = len(rays)
        bits = torch.empty(max(1, (n + 31) // 32), dtype=torch.int32, device=rays.data.device)
        box, tri = self._trace(rays, RT_ANY, (None, bits), counters, stream, tune)
        return (bits, box, tri) if counters else bits


# Batched builds take meshes of up to BATCH_BUILD_MAX triangles; those of at
# most FUSED_BUILD_MAX are built in LDS (the fused path), larger ones by one
# workgroup each over a scratch buffer.
BATCH_BUILD_MAX = 65536
FUSED_BUILD_MAX = 4096


def build_radv(pos, idx=None, order="input", stream=None, out=None):
    """The RADV acceleration-structure image of fast_bvh_builder.comp
    (rt_gpu_build_radv): returns (uint8 image of total_as_size bytes, int32
    leaf-order triangle indices).  order: "input" (the reference builder's
    leaves of consecutive triangles) or "morton"."""
    n = _tri_count(pos, idx)
    size = tree_layout(n)[3]
    img = out if out is not None else torch.empty(size, dtype=torch.uint8, device=pos.device)
    ids = torch.empty(n, dtype=torch.int32, device=pos.device)
    call("rt_gpu_build_radv", stream_handle(stream), ptr(pos), ptr(idx), u32(n),
         u32(0 if order == "input" else 1), ptr(img), C.c_uint64(img.numel()), ptr(ids))
    return img, ids


class BuildBatch:
    """Many meshes of at most BATCH_BUILD_MAX triangles built in ONE persistent
    launch (rt_gpu_build_batch), each tree identical to GpuBVH4.build(...,
    algorithm="implicit") of that mesh alone.

    meshes: list of (pos, idx or None).  The table is planned and uploaded
    here, largest meshes first; meshes above FUSED_BUILD_MAX triangles get
    their own scratch (one buffer for the batch).  launch() is a single
    kernel with device-resident arguments (capturable in a graph).
    .trees[i] is mesh i's GpuBVH4 (filled by launch)."""

    def __init__(self, meshes, keep_order=False, workgroups=0, device=None, stream=None):
        """stream: allocate on it and upload the table stream-ordered from
        pinned memory (no host wait), for use from a serving loop."""
        dev = device or _dev()
        self.count = len(meshes)
        self.workgroups = workgroups
        self.trees = []
        arr = (rt_build_item * max(1, self.count))()
        order = sorted(range(self.count), key=lambda i: -_tri_count(*meshes[i]))
        need = {}
        for i in order:
            check_mesh(*meshes[i])
            n = _tri_count(*meshes[i])
            if n > BATCH_BUILD_MAX:
                raise ValueError(f"mesh {i}: {n} triangles > BATCH_BUILD_MAX")
            if n > FUSED_BUILD_MAX:
                need[i] = (build_requirements(n, build_cfg("implicit"))[1] + 255) // 256 * 256
        self.scratch = torch.empty(max(256, sum(need.values())), dtype=torch.uint8, device=dev)
        base, off = self.scratch.data_ptr(), 0
        for slot, i in enumerate(order):
            pos, idx = meshes[i]
            n = _tri_count(pos, idx)
            cap = build_requirements(n, build_cfg("implicit"))[0]
            nodes = torch.empty(max(cap, 1) * NODE_BYTES, dtype=torch.uint8, device=dev)
            tris = torch.empty(max(n, 1) * TRI_BYTES, dtype=torch.uint8, device=dev)
            info = torch.empty(4, dtype=torch.int32, device=dev)
            bvh = GpuBVH4(nodes, tris, cap, n, len(tree_layout(n)[0]))
            bvh.info, bvh.pos, bvh.idx, bvh.algorithm = info, pos, idx, "implicit"
            if keep_order:
                bvh.sorted_ids = torch.empty(n, dtype=torch.int32, device=dev)
                bvh.codes = torch.empty(n, dtype=torch.int32, device=dev)
            sp, sb = None, 0
            if i in need:
                sp, sb = vp(base + off), need[i]
                off += sb
            arr[slot] = rt_build_item(ptr(pos), ptr(idx), ptr(nodes), ptr(tris), ptr(info),
                                      ptr(bvh.sorted_ids), ptr(bvh.codes), n, cap, sp, sb)
            self.trees.append((i, bvh))
        self.trees = [b for _, b in sorted(self.trees, key=lambda x: x[0])]
        nbytes = C.c_size_t(0)
        call("rt_gpu_build_batch_table_bytes", u32(self.count), C.byref(nbytes))
        host = (C.c_uint8 * max(1, nbytes.value))()
        call("rt_gpu_build_batch_plan", arr, u32(self.count), host, nbytes)
        h_table = torch.frombuffer(bytearray(bytes(host)), dtype=torch.uint8)
        if stream is None:
            self.d_table = h_table.to(dev)
        else:
            self._pinned = h_table.pin_memory()
            with torch.cuda.stream(stream):
                self.d_table = self._pinned.to(dev, non_blocking=True)

    def launch(self, stream=None):
        call("rt_gpu_build_batch", stream_handle(stream), ptr(self.d_table), u32(self.count),
             u32(self.workgroups))
        return self.trees


def _tri_count(pos, idx):
    return idx.shape[0] if idx is not None else pos.numel() // 9


class TraceBatch:
    """Many (tree, rays, mode) items traced in ONE launch (rt_trace_batch).

    items: list of (bvh, rays, mode, out) with mode "closest" (out = (prim,
    t) int32 / float32 tensors, or None to allocate) or "any" (out = int32
    bits tensor of ceil(n/32) words, or None).  The item table and chunk
    offsets are uploaded once here; launch() is then a single kernel with
    device-resident arguments (no host work: it can be captured in a graph).
    The tensors of every item must stay alive while the batch is used."""

    def __init__(self, items, device=None, stream=None):
        dev = device or _dev()
        self.count = len(items)
        arr = (rt_trace_item * max(1, self.count))()
        self.outputs, self._keep = [], []
        for i, (bvh, rays, mode, out) in enumerate(items):
            n = len(rays)
            it = arr[i]
            it.bvh, it.rays, it.count = bvh.dev(), rays.soa(), n
            if mode == "any":
                bits = out if out is not None else torch.empty(max(1, (n + 31) // 32),
                                                               dtype=torch.int32, device=dev)
                check_out(bits, (n + 31) // 32, _I32, f"item {i} occlusion bits")
                it.flags, it.occlusion_bits = RT_ANY, ptr(bits)
                it.hits = rt_hits_soa(None, None)
                self.outputs.append(bits)
            elif mode == "closest":
                prim, t = out if out is not None else (
                    torch.empty(n, dtype=torch.int32, device=dev),
                    torch.empty(n, dtype=torch.float32, device=dev))
                check_out(prim, n, _I32, f"item {i} prim")
                check_out(t, n, _F32, f"item {i} t")
                it.flags, it.occlusion_bits = RT_CLOSEST, None
                it.hits = rt_hits_soa(ptr(prim), ptr(t))
                self.outputs.append((prim, t))
            else:
                raise ValueError(f"mode {mode!r}")
            self._keep.append((bvh, rays))
        self._arr, self._dev, self._pinned = arr, dev, []
        self._tables = {}  # rays per chunk -> (device chunk table, total chunks)
        h_items = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
        self.d_items = self._upload(h_items, stream)
        self._table(64, stream)

    def _upload(self, h, stream):
        if stream is None:
            return h.to(self._dev)
        # stream-ordered upload from a pinned copy (no host wait); the pinned
        # buffers live as long as this batch
        p = h.pin_memory()
        self._pinned.append(p)
        with torch.cuda.stream(stream):
            return p.to(self._dev, non_blocking=True)

    def _table(self, rays_per_chunk, stream=None):
        """Chunk table for chunks of rays_per_chunk rays (planned once)."""
        if rays_per_chunk not in self._tables:
            chunk = (u32 * (self.count + 1))()
            depth = u32(0)
            call("rt_trace_batch_plan_ex", self._arr, u32(self.count), u32(rays_per_chunk),
                 chunk, C.byref(depth))
            self.depth = depth.value
            h = torch.tensor(list(chunk), dtype=torch.int64).to(torch.int32)
            self._tables[rays_per_chunk] = (self._upload(h, stream), chunk[self.count])
        return self._tables[rays_per_chunk]

    @property
    def total_chunks(self):
        return self._tables[64][1]

    def launch(self, stream=None, schedule="simple", block=None, lds_stack=None,
               rays_per_wave=None, refill_lanes=None):
        """One launch over every item (same results for every schedule and
        shape; defaults TRACE_TUNING's).  schedule "refill": each wave walks
        rays_per_wave rays of one item, refilling finished lanes."""
        o = trace_opts(schedule, rays_per_wave=rays_per_wave, refill_lanes=refill_lanes,
                       block=block, lds_stack=lds_stack)
        d_chunk, total = self._table(o.rays_per_wave if schedule == "refill" else 64, stream)
        call("rt_trace_batch_ex", stream_handle(stream), ptr(self.d_items), ptr(d_chunk),
             u32(self.count), u32(total), u32(self.depth), C.byref(o))
        return self.outputs


class Scene:
    """Two-level scene (rt_trace_scene): instances [(bvh, pos, idx, m, kind)]
    with m a 3x4 row-major world-to-object map (None for the identity) and
    kind INST_GENERAL / INST_TRANSLATE / INST_IDENTITY.  The instance table
    is uploaded once; rebuild a Scene when an instance's tree changes
    buffers (a refit or a rebuild into the same buffers needs nothing)."""

    def __init__(self, instances, device=None):
        dev = device or _dev()
        self.instances = list(instances)
        arr = (rt_instance * max(1, len(self.instances)))()
        depth = 1
        for k, (bvh, pos, idx, m, kind) in enumerate(self.instances):
            if bvh.format != NODES_F32:
                raise ValueError("scene instances take rt_node4 trees (RT_NODES_F32)")
            arr[k].blas = bvh.dev()
            arr[k].pos, arr[k].idx = ptr(pos), ptr(idx)
            mm = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0] if m is None else \
                [float(v) for v in np.asarray(m, np.float32).reshape(-1)]
            arr[k].world_to_object = (f32 * 12)(*mm)
            arr[k].kind = kind
            depth = max(depth, bvh.depth or 64)
        self.count = len(self.instances)
        self.depth = depth
        self.d_table = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)

    def intersect(self, rays, stream=None, out=None, block=None, lds_stack=None,
                  schedule="simple", tile=None):
        """Closest hit: (prim, t, instance index) tensors (int32, f32, int32).
        block / lds_stack / tile (grid_width, tile_width): launch shape (same
        results for every choice)."""
        n = len(rays)
        dev = rays.data.device
        prim, t, inst = out if out else (torch.empty(n, dtype=torch.int32, device=dev),
                                         torch.empty(n, dtype=torch.float32, device=dev),
                                         torch.empty(n, dtype=torch.int32, device=dev))
        hits = rt_hits_soa(ptr(prim), ptr(t))
        r = rays.soa()
        o = trace_opts(schedule, block=block, lds_stack=lds_stack, tile=tile)
        call("rt_trace_scene_ex", stream_handle(stream), ptr(self.d_table), u32(self.count),
             u32(self.depth), C.byref(r), u32(n), u32(RT_CLOSEST), C.byref(o), C.byref(hits),
             ptr(inst), None)
        return prim, t, inst

    def occluded(self, rays, stream=None, out=None, block=None, lds_stack=None,
                 schedule="simple", tile=None):
        n = len(rays)
        bits = out if out is not None else torch.empty(max(1, (n + 31) // 32),
                                                       dtype=torch.int32, device=rays.data.device)
        r = rays.soa()
        o = trace_opts(schedule, block=block, lds_stack=lds_stack, tile=tile)
        call("rt_trace_scene_ex", stream_handle(stream), ptr(self.d_table), u32(self.count),
             u32(self.depth), C.byref(r), u32(n), u32(RT_ANY), C.byref(o), None, None,
             ptr(bits))
        return bits

    def gen_ao_rays(self, primary, prim, inst, t, w, h, frame=0, offset=0.003, tmin=0.001,
                    tmax=0.5, stream=None, out=None):
        rays = out if out is not None else Rays.empty(w * h)
        p, o = primary.soa(), rays.out()
        call("rt_gen_ao_rays_scene", stream_handle(stream), ptr(self.d_table), u32(self.count),
             C.byref(p), ptr(prim), ptr(inst), ptr(t), u32(w), u32(h), u32(frame), f32(offset),
             f32(tmin), f32(tmax), C.byref(o))
        return rays


def ao_resolve(bits, prim, stream=None, out=None):
    n = prim.numel()
    ao = out if out is not None else torch.empty(n, dtype=torch.float32, device=prim.device)
    call("rt_ao_resolve", stream_handle(stream), ptr(bits), ptr(prim), u32(n), ptr(ao))
    return ao


def blur7(img, depth, w, h, thr=0.05, stream=None, out=None):
    o = out if out is not None else torch.empty_like(img)
    call("rt_blur7", stream_handle(stream), ptr(img), ptr(depth), u32(w), u32(h), f32(thr),
         ptr(o))
    return o


RADV_NODE_TYPES = {0: "triangle", 4: "box16", 5: "box32", 6: "instance", 7: "aabb"}


def radv_describe(image):
    """Decoded header of a RADV image (rt_radv_describe; the header part of
    dump_bvh_structure, include/radv_bvh.hpp:311-390): a dict of the
    RadvAccelStructHeader fields, box_count and the root's children as
    (id, type name, offset from the BVH base) or None for an empty slot.
    image: uint8 numpy array or tensor (a device tensor is copied).  Raises
    RtError when the sections do not fit the image."""
    if isinstance(image, torch.Tensor):
        image = image.cpu().numpy()
    b = np.ascontiguousarray(image, dtype=np.uint8).reshape(-1)
    info = rt_radv_info()
    call("rt_radv_describe", b.ctypes.data_as(vp), C.c_uint64(b.size), C.byref(info))
    d = {name: getattr(info, name) for name, _ in rt_radv_info._fields_}
    for k in ("aabb", "update_dispatch_size", "copy_dispatch_size", "root_children"):
        d[k] = list(d[k])
    d["root"] = [None if c == MISS else (c, RADV_NODE_TYPES.get(c & 7, "?"), (c & ~7) << 3)
                 for c in d["root_children"]]
    return d


def format_radv(image):
    """radv_describe as text: header offsets and values, root children."""
    d = radv_describe(image)
    lines = [f"bvh_offset 0x{d['bvh_offset']:x}  root_flags 0x{d['root_flags']:x}",
             f"aabb min {tuple(d['aabb'][:3])} max {tuple(d['aabb'][3:])}",
             f"size {d['size']}  compacted {d['compacted_size']}  serialization "
             f"{d['serialization_size']}",
             f"geometry_type {d['geometry_type']}  geometry_count {d['geometry_count']}  "
             f"type {d['type']}  build_flags 0x{d['build_flags']:x}",
             f"leaf_nodes_offset 0x{d['leaf_nodes_offset']:x}  primitive_count "
             f"{d['primitive_count']}  box nodes {d['box_count']}"]
    for i, c in enumerate(d["root"]):
        lines.append(f"root child {i}: " + ("empty" if c is None else
                                            f"id 0x{c[0]:08x} {c[1]} at +0x{c[2]:x}"))
    return "\n".join(lines)


class RadvBVH:
    """A RADV image in device memory traversed in software (rt_trace_radv):
    closest / any hit with rt_trace's predicates; hits report triangle_id."""

    def __init__(self, image, info=None, depth=None):
        if image.dtype != torch.uint8 or not image.is_cuda:
            raise ValueError("a uint8 device tensor expected")
        self.image = image.contiguous()
        self.info = info if info is not None else radv_describe(self.image)
        n = self.info["primitive_count"]
        if depth is None:
            # images of this library / fast_bvh_builder: compute_tree_layout levels
            counts, _, tb, _ = tree_layout(n)
            depth = len(counts) if tb == self.info["box_count"] else 0
        self.depth = depth

    def dev(self):
        return rt_radv_dev(ptr(self.image), C.c_uint64(self.image.numel()),
                           u32(self.info["bvh_offset"]), u32(self.info["box_count"]),
                           u32(self.info["primitive_count"]), u32(self.depth))

    def _trace(self, rays, flags, hits, bits, counters, stream, tune):
        n = len(rays)
        box = tri = None
        if counters:
            box = torch.empty(n, dtype=torch.int32, device=rays.data.device)
            tri = torch.empty(n, dtype=torch.int32, device=rays.data.device)
            flags |= RT_COUNTERS
        d = self.dev()
        r = rays.soa()
        o = trace_opts("simple", **tune)
        call("rt_trace_radv", stream_handle(stream), C.byref(d), C.byref(r), u32(n), u32(flags),
             C.byref(o), C.byref(hits) if hits is not None else None, ptr(bits), ptr(box),
             ptr(tri))
        return box, tri

    def intersect(self, rays, counters=False, stream=None, **tune):
        n = len(rays)
        prim = torch.empty(n, dtype=torch.int32, device=rays.data.device)
        t = torch.empty(n, dtype=torch.float32, device=rays.data.device)
        box, tri = self._trace(rays, RT_CLOSEST, rt_hits_soa(ptr(prim), ptr(t)), None, counters,
                               stream, tune)
        return (prim, t, box, tri) if counters else (prim, t)

    def occluded(self, rays, counters=False, stream=None, **tune):
        n = len(rays)
        bits = torch.empty(max(1, (n + 31) // 32), dtype=torch.int32, device=rays.data.device)
        box, tri = self._trace(rays, RT_ANY, None, bits, counters, stream, tune)
        return (bits, box, tri) if counters else bits


def image_metrics_scratch_bytes(w, h):
    b = C.c_size_t(0)
    call("rt_image_metrics_scratch_bytes", u32(w), u32(h), C.byref(b))
    return b.value


def image_metrics(a, b, blocks=False, stream=None, out=None, scratch=None):
    """ImageMetrics::compute on the device (rt_image_metrics): a float64
    device tensor {mse, psnr, ssim} of two (h, w) float32 images; with
    blocks=True also each full 8x8 block's clamped SSIM ((h//8, w//8))."""
    if a.shape != b.shape or a.dim() != 2:
        raise ValueError("two (h, w) images of one shape expected")
    if a.dtype != torch.float32 or b.dtype != torch.float32:
        raise ValueError("float32 images expected")
    a, b = a.contiguous(), b.contiguous()
    h, w = a.shape
    dev = a.device
    res = out if out is not None else torch.empty(3, dtype=torch.float64, device=dev)
    blk = torch.empty((h // 8, w // 8), dtype=torch.float64, device=dev) if blocks else None
    need = image_metrics_scratch_bytes(w, h)
    if scratch is None or scratch.numel() * scratch.element_size() < need:
        scratch = torch.empty(max(1, need // 8), dtype=torch.float64, device=dev)
    call("rt_image_metrics", stream_handle(stream), ptr(a), ptr(b), u32(w), u32(h), ptr(res),
         ptr(blk) if blocks and blk.numel() else None, ptr(scratch),
         C.c_size_t(scratch.numel() * scratch.element_size()))
    return (res, blk) if blocks else res


def _depth_host(nodes4):
    """Node levels on the longest path of a host tree (numpy structured)."""
    if len(nodes4) == 0:
        return 0
    a = nodes4.view(np.uint8).reshape(-1, NODE_BYTES)
    child = a[:, 96:112].copy().view(np.int32)
    count = a[:, 112:128].copy().view(np.int32)
    depth = 0
    stack = [(0, 1)]
    while stack:
        nd, d = stack.pop()
        depth = max(depth, d)
        for k in range(4):
            if count[nd, k] < 0:
                stack.append((int(child[nd, k]), d + 1))
    return depth


# --------------------------------------------------------------- generators --
def gen_terrain(target, size=100.0, fused=False, stream=None):
    """createTessellatedTerrain (rt_gen_terrain): (pos (V, 3) float32, idx (T,
    3) int32) on the device; fused=True evaluates it as the reference's
    -O3 -march=native build does."""
    res, nv, nt = u32(0), u32(0), u32(0)
    call("rt_terrain_resolution", u32(target), C.byref(res), C.byref(nv), C.byref(nt))
    dev = _dev()
    pos = torch.empty((nv.value, 3), dtype=torch.float32, device=dev)
    idx = torch.empty((nt.value, 3), dtype=torch.int32, device=dev)
    call("rt_gen_terrain", stream_handle(stream), u32(target), f32(size), C.c_int(1 if fused else 0),
         ptr(pos), ptr(idx))
    return pos, idx


def gen_sphere(nu, nv, radius=1.0, disp=0.02, seed=7, stream=None):
    dev = _dev()
    pos = torch.empty(((nv + 1) * nu, 3), dtype=torch.float32, device=dev)
    idx = torch.empty((2 * nu * nv, 3), dtype=torch.int32, device=dev)
    call("rt_gen_sphere", stream_handle(stream), u32(nu), u32(nv), f32(radius), f32(disp),
         u32(seed), ptr(pos), ptr(idx))
    return pos, idx


def gen_wave_grid(nx, ny, extent=1.0, t=0.0, amp=0.1, fx=4.0, fy=4.0, stream=None):
    dev = _dev()
    pos = torch.empty(((nx + 1) * (ny + 1), 3), dtype=torch.float32, device=dev)
    idx = torch.empty((2 * nx * ny, 3), dtype=torch.int32, device=dev)
    call("rt_gen_wave_grid", stream_handle(stream), u32(nx), u32(ny), f32(extent), f32(t),
         f32(amp), f32(fx), f32(fy), ptr(pos), ptr(idx))
    return pos, idx


def gen_heightfield(nx, ny, extent=1.0, amp=0.05, seed=11, stream=None):
    dev = _dev()
    pos = torch.empty(((nx + 1) * (ny + 1), 3), dtype=torch.float32, device=dev)
    idx = torch.empty((2 * nx * ny, 3), dtype=torch.int32, device=dev)
    call("rt_gen_heightfield", stream_handle(stream), u32(nx), u32(ny), f32(extent), f32(amp),
         u32(seed), ptr(pos), ptr(idx))
    return pos, idx


def gen_dynamic_grid(n, t=0.0, extent=2.0, stream=None, out=None):
    """generate_dynamic_grid_mesh layout, 9 floats per triangle; two
    triangles per quad, so the result has 2 * (n // 2) rows (an odd n loses
    its last triangle, as in the reference generator)."""
    xyz = out if out is not None else torch.empty((2 * (n // 2), 9), dtype=torch.float32,
                                                   device=_dev())
    call("rt_gen_dynamic_grid", stream_handle(stream), u32(n), f32(t), f32(extent), ptr(xyz))
    return xyz


def gen_camera_rays(w, h, eye, look, up=(0.0, 0.0, 1.0), vfov=60.0, tmin=0.0, tmax=1e30,
                    stream=None, out=None):
    rays = out if out is not None else Rays.empty(w * h)
    o = rays.out()
    call("rt_gen_camera_rays", stream_handle(stream), u32(w), u32(h), _f3(eye), _f3(look), _f3(up),
         f32(vfov), f32(tmin), f32(tmax), C.byref(o))
    return rays


def gen_shadow_rays(pos, idx, primary, prim, t, lights, eps, stream=None, out=None):
    npix = len(primary)
    nl = lights.shape[0]
    rays = out if out is not None else Rays.empty(npix * nl)
    p = primary.soa()
    o = rays.out()
    call("rt_gen_shadow_rays", stream_handle(stream), ptr(pos), ptr(idx), C.byref(p), ptr(prim),
         ptr(t), u32(npix), ptr(lights), u32(nl), f32(eps), C.byref(o))
    return rays


def gen_ao_rays(pos, idx, primary, prim, t, w, h, frame=0, offset=0.003, tmin=0.001, tmax=0.5,
                stream=None, out=None):
    rays = out if out is not None else Rays.empty(w * h)
    p = primary.soa()
    o = rays.out()
    call("rt_gen_ao_rays", stream_handle(stream), ptr(pos), ptr(idx), C.byref(p), ptr(prim), ptr(t),
         u32(w), u32(h), u32(frame), f32(offset), f32(tmin), f32(tmax), C.byref(o))
    return rays


def gen_random_rays(count, seed, lo, hi, tmax=1e30, stream=None, out=None):
    rays = out if out is not None else Rays.empty(count)
    o = rays.out()
    call("rt_gen_random_rays", stream_handle(stream), u32(count), u32(seed), _f3(lo), _f3(hi),
         f32(tmax), C.byref(o))
    return rays


def gen_exp0_rays(count, coherent, cam=(0.0, 50.0, -45.0), look=(0.0, 0.0, 5.0), tmin=0.01,
                  tmax=1000.0, stream=None, out=None):
    """The exp0 probe's procedural rays on the device (rt_gen_exp0_rays,
    software_bvh_bvh4.comp:50-94): a 1024-wide camera grid (coherent) or
    pcg_hash origins and directions; cam / look default to the probe's
    (microbench_main.cpp:290-294)."""
    rays = out if out is not None else Rays.empty(count)
    if len(rays) < count:
        raise ValueError(f"out holds {len(rays)} rays < {count}")
    o = rays.out()
    call("rt_gen_exp0_rays", stream_handle(stream), u32(count), C.c_int(1 if coherent else 0),
         _f3(cam), _f3(look), f32(tmin), f32(tmax), C.byref(o))
    return rays


def pack_light_masks(bits, npix, nlights, stream=None, out=None):
    masks = out if out is not None else torch.empty(npix, dtype=torch.int16, device=bits.device)
    call("rt_pack_light_masks", stream_handle(stream), ptr(bits), u32(npix), u32(nlights),
         ptr(masks))
    return masks


def lights_in_box(lo, hi, count=16, seed=1337):
    """count light positions uniform in the box [lo, hi) from a seeded hash
    (host-side; the same values feed the CPU check)."""
    out = []
    for i in range(count):
        p = []
        for a in range(3):
            h = _pcg(((3 * i + a) * 0x9E3779B9 + seed) & 0xFFFFFFFF)
            u = (h >> 8) * (1.0 / 16777216.0)
            p.append(lo[a] + u * (hi[a] - lo[a]))
        out.append(p)
    return out


def _pcg(v):
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return ((word >> 22) ^ word) & 0xFFFFFFFF


# ------------------------------------------------------ submit staging --
def submit_bytes(kind, count, with_gen=True):
    """Staging / device-table bytes of rt_trace_batch_submit ("trace") or
    rt_gpu_build_batch_submit ("build") for count items."""
    b = C.c_size_t(0)
    fn = "rt_trace_batch_submit_bytes" if kind == "trace" else "rt_gpu_build_batch_submit_bytes"
    call(fn, u32(count), C.c_int(1 if with_gen else 0), C.byref(b))
    return b.value


class SubmitRing:
    """A ring of (pinned host staging, device table) slot pairs for the
    submit entry points, allocated once.  A slot is handed out again only
    after the stream work that read it has passed (its event)."""

    def __init__(self, slots, slot_bytes, device=None):
        dev = device or _dev()
        self.slot_bytes = (slot_bytes + 255) // 256 * 256
        self.pinned = torch.empty(slots * self.slot_bytes, dtype=torch.uint8, pin_memory=True)
        self.table = torch.empty(slots * self.slot_bytes, dtype=torch.uint8, device=dev)
        self.events = [None] * slots
        self.next = 0

    def take(self):
        """(pinned pointer, device pointer, slot index) of the next slot."""
        k = self.next
        self.next = (k + 1) % len(self.events)
        ev = self.events[k]
        if ev is not None and not ev.query():
            ev.synchronize()
        off = k * self.slot_bytes
        return (vp(self.pinned.data_ptr() + off), vp(self.table.data_ptr() + off), k)

    def used(self, k, stream):
        """Record that slot k is read by work on `stream` up to now."""
        ev = torch.cuda.Event()
        ev.record(stream)
        self.events[k] = ev
