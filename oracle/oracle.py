"""ctypes binding of the CPU oracle (liboracle.so) and of oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

Every wrapper returns numpy arrays; device data never passes through here.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

NODE2 = np.dtype([("aabbMin", "<f4", 3), ("leftChild", "<i4"),
                  ("aabbMax", "<f4", 3), ("rightChild", "<i4")])
NODE4 = np.dtype([("minX", "<f4", 4), ("maxX", "<f4", 4), ("minY", "<f4", 4),
                  ("maxY", "<f4", 4), ("minZ", "<f4", 4), ("maxZ", "<f4", 4),
                  ("child", "<i4", 4), ("count", "<i4", 4)])
TRI64 = np.dtype([("v", "<f4", 9), ("reserved", "<u4", 3), ("triangle_id", "<u4"),
                  ("geometry_id_and_flags", "<u4"), ("reserved2", "<u4"), ("id", "<u4")])
QNODE4 = np.dtype([("p", "<f4", 3), ("e", "i1", 3), ("pad0", "u1"), ("qlo", "u1", (3, 4)),
                   ("qhi", "u1", (3, 4)), ("count", "<i2", 4), ("child", "<i4", 4)])
assert NODE2.itemsize == 32 and NODE4.itemsize == 128 and TRI64.itemsize == 64
assert QNODE4.itemsize == 64

MISS = 0xFFFFFFFF
CLOSEST, ANY = 0, 1

_lib = None
_ref = None


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        _lib.or_terrain_res.restype = C.c_uint32
        _lib.or_pcg_hash.restype = C.c_uint32
        _lib.or_ao_hash.restype = C.c_uint32
        _lib.or_expand_bits10.restype = C.c_uint32
        _lib.or_bvh4_depth.restype = C.c_uint32
        _lib.or_rms.restype = C.c_double
        _lib.or_percentile.restype = C.c_float
    return _lib


REF_SHIPPED_PATH = os.path.join(_HERE, "_ref", "libref_shipped.so")
_ref_shipped = None


def ref(shipped=False):
    """The reference's own CPU path, or None where oracle/_ref was not built.

    shipped=False: headers compiled with one rounding per operator.
    shipped=True: compiled with the reference's CMake flags (-O3 and an
    FMA-capable -march), the build that wrote results/*.csv."""
    global _ref, _ref_shipped
    if shipped:
        if _ref_shipped is None and os.path.exists(REF_SHIPPED_PATH):
            _ref_shipped = C.CDLL(REF_SHIPPED_PATH)
        return _ref_shipped
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
    return _ref


u32 = C.c_uint32
f32 = C.c_float


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- meshes ----
def terrain(target, size=100.0, fused=False):
    """createTessellatedTerrain; fused=True evaluates it as the reference's
    own -O3 -march=native build does (the build behind results/*.csv)."""
    res = lib().or_terrain_res(u32(target))
    pos = np.zeros((res * res, 3), np.float32)
    idx = np.zeros(((res - 1) * (res - 1) * 2, 3), np.uint32)
    lib().or_terrain_ex(u32(target), f32(size), C.c_int(1 if fused else 0), _p(pos), _p(idx))
    return pos, idx


def dynamic_grid(n, t=0.0, extent=2.0):
    xyz = np.zeros((n, 9), np.float32)
    cen = np.zeros((n, 3), np.float32)
    lib().or_dynamic_grid(u32(n), f32(t), f32(extent), _p(xyz), _p(cen))
    return xyz, cen


def wave_grid(nx, ny, extent=1.0, t=0.0, amp=0.1, fx=4.0, fy=4.0):
    pos = np.zeros(((nx + 1) * (ny + 1), 3), np.float32)
    idx = np.zeros((nx * ny * 2, 3), np.uint32)
    lib().or_wave_grid(u32(nx), u32(ny), f32(extent), f32(t), f32(amp), f32(fx), f32(fy),
                       _p(pos), _p(idx))
    return pos, idx


def sphere(nu, nv, radius=1.0, disp=0.02, seed=7):
    pos = np.zeros(((nv + 1) * nu, 3), np.float32)
    idx = np.zeros((2 * nu * nv, 3), np.uint32)
    lib().or_sphere(u32(nu), u32(nv), f32(radius), f32(disp), u32(seed), _p(pos), _p(idx))
    return pos, idx


def heightfield(nx, ny, extent=1.0, amp=0.05, seed=11):
    pos = np.zeros(((nx + 1) * (ny + 1), 3), np.float32)
    idx = np.zeros((nx * ny * 2, 3), np.uint32)
    lib().or_heightfield(u32(nx), u32(ny), f32(extent), f32(amp), u32(seed), _p(pos), _p(idx))
    return pos, idx


# ---------------------------------------------------------- software_bvh ----
def bvh2_build(pos, idx, max_prims=4):
    pos, idx = _f(pos), _u(idx)
    n = idx.shape[0]
    cap = 2 * n + 1
    nodes = np.zeros(cap, NODE2)
    cnt = u32(0)
    tris48 = np.zeros((n, 12), np.float32)
    order = np.zeros(n, np.uint32)
    rc = lib().or_bvh2_build(_p(pos), u32(pos.shape[0]), _p(idx), u32(n), u32(max_prims),
                             _p(nodes), u32(cap), C.byref(cnt), _p(tris48), _p(order))
    if rc:
        raise RuntimeError("or_bvh2_build failed")
    return nodes[:cnt.value].copy(), tris48, order


def bvh4_collapse(nodes2):
    nodes2 = np.ascontiguousarray(nodes2)
    cap = max(1, len(nodes2))
    out = np.zeros(cap, NODE4)
    cnt = u32(0)
    rc = lib().or_bvh4_collapse(_p(nodes2), u32(len(nodes2)), _p(out), u32(cap), C.byref(cnt))
    if rc:
        raise RuntimeError("or_bvh4_collapse failed")
    return out[:cnt.value].copy()


def tri64_from48(tris48, order):
    n = tris48.shape[0]
    out = np.zeros(n, TRI64)
    lib().or_tris48_to_tri64(_p(_f(tris48)), _p(_u(order)), u32(n), _p(out))
    return out


def software_bvh4(pos, idx, max_prims=4):
    """SoftwareBVH::build + SoftwareBVH4::fromBVH2, leaves as 64-B records."""
    n2, t48, order = bvh2_build(pos, idx, max_prims)
    n4 = bvh4_collapse(n2)
    return n4, tri64_from48(t48, order)


def bvh4_quantize(nodes):
    """rt_qnode4 records (QNODE4) of a Box4 tree: or_bvh4_quantize."""
    nodes = np.ascontiguousarray(nodes)
    out = np.zeros(len(nodes), QNODE4)
    lib().or_bvh4_quantize(_p(nodes), u32(len(nodes)), _p(out))
    return out


def bvh4_dequantize(q):
    """The Box4 tree (NODE4) with the decoded boxes of quantized nodes."""
    q = np.ascontiguousarray(q)
    out = np.zeros(len(q), NODE4)
    lib().or_bvh4_dequantize(_p(q), u32(len(q)), _p(out))
    return out


def bvh4_depth(nodes):
    nodes = np.ascontiguousarray(nodes)
    return int(lib().or_bvh4_depth(_p(nodes), u32(len(nodes))))


# ------------------------------------------------------------- traversal ----
class Rays:
    """Structure-of-arrays ray batch: 8 float32 arrays of equal length."""
    FIELDS = ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")

    def __init__(self, n=0, **kw):
        for f in self.FIELDS:
            setattr(self, f, _f(kw[f]) if f in kw else np.zeros(n, np.float32))

    def __len__(self):
        return len(self.ox)

    def arrays(self):
        return [getattr(self, f) for f in self.FIELDS]

    def subset(self, sel):
        return Rays(**{f: getattr(self, f)[sel] for f in self.FIELDS})

    @staticmethod
    def stack(a):
        """(8, n) float32 array -> Rays."""
        return Rays(**{f: a[i] for i, f in enumerate(Rays.FIELDS)})

    def packed(self):
        return np.stack(self.arrays()).astype(np.float32)


def default_threads():
    """Worker threads of the parallel oracle routines: the CPUs this process
    may run on, at most 16 (one GPU's share of a shared host)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def trace_bvh4(nodes, tris, rays, mode=CLOSEST, stack_limit=0, threads=None, counters=False):
    nodes = np.ascontiguousarray(nodes)
    tris = np.ascontiguousarray(tris)
    n = len(rays)
    prim = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    box = np.zeros(n, np.uint32) if counters else None
    tri = np.zeros(n, np.uint32) if counters else None
    threads = threads or default_threads()
    lib().or_trace_bvh4(_p(nodes), u32(len(nodes)), _p(tris), *[_p(a) for a in rays.arrays()],
                        u32(n), C.c_int(mode), u32(stack_limit), C.c_int(threads),
                        _p(prim), _p(t), _p(box), _p(tri))
    if counters:
        return prim, t, box, tri
    return prim, t


def trace_bvh2(nodes2, tris48, rays, order=None, mode=CLOSEST, threads=None, counters=False):
    """BVH2 traversal of a SoftwareBVH (software_bvh_procedural.comp): returns
    (prim, t[, box, tri]); prim = order[leaf position] (or the position)."""
    nodes2 = np.ascontiguousarray(nodes2)
    tris48 = _f(tris48)
    n = len(rays)
    prim = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    box = np.zeros(n, np.uint32) if counters else None
    tri = np.zeros(n, np.uint32) if counters else None
    o = None if order is None else np.ascontiguousarray(order, np.uint32)
    lib().or_trace_bvh2(_p(nodes2), u32(len(nodes2)), _p(tris48), _p(o), _p(rays.ox),
                        _p(rays.oy), _p(rays.oz), _p(rays.tmin), _p(rays.dx), _p(rays.dy),
                        _p(rays.dz), _p(rays.tmax), u32(n), C.c_int(mode),
                        C.c_int(threads or default_threads()), _p(prim), _p(t), _p(box),
                        _p(tri))
    return (prim, t, box, tri) if counters else (prim, t)


def trace_brute(tris, rays, mode=CLOSEST, threads=None):
    tris = np.ascontiguousarray(tris)
    n = len(rays)
    prim = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    threads = threads or default_threads()
    lib().or_trace_brute(_p(tris), u32(len(tris)), *[_p(a) for a in rays.arrays()], u32(n),
                         C.c_int(mode), C.c_int(threads), _p(prim), _p(t))
    return prim, t


# ---------------------------------------------------------------- rays -------
def pcg_hash(v):
    return int(lib().or_pcg_hash(u32(v)))


def ao_hash(x):
    """rq_ao.comp's hash (the AO sampler's random numbers)."""
    return int(lib().or_ao_hash(u32(x)))


def exp0_rays(count, coherent, cam=(0.0, 50.0, -45.0), look=(0.0, 0.0, 5.0),
              tmin=0.01, tmax=1000.0):
    r = Rays(count)
    c = _f(cam)
    lk = _f(look)
    lib().or_exp0_rays(u32(count), C.c_int(1 if coherent else 0), _p(c), _p(lk),
                       _p(r.ox), _p(r.oy), _p(r.oz), _p(r.dx), _p(r.dy), _p(r.dz))
    r.tmin[:] = tmin
    r.tmax[:] = tmax
    return r


def camera_rays(w, h, eye, look, up=(0.0, 0.0, 1.0), vfov=60.0, tmin=0.0, tmax=1e30):
    r = Rays(w * h)
    lib().or_camera_rays(u32(w), u32(h), _p(_f(eye)), _p(_f(look)), _p(_f(up)), f32(vfov),
                         f32(tmin), f32(tmax), *[_p(a) for a in r.arrays()])
    return r


def shadow_rays(pos, idx, prim_rays, prim, t, lights, eps):
    npix = len(prim_rays)
    lights = _f(lights).reshape(-1, 3)
    r = Rays(npix * len(lights))
    pr = prim_rays
    lib().or_shadow_rays(_p(_f(pos)), _p(_u(idx)), _p(pr.ox), _p(pr.oy), _p(pr.oz),
                         _p(pr.dx), _p(pr.dy), _p(pr.dz), _p(_u(prim)), _p(_f(t)), u32(npix),
                         _p(lights), u32(len(lights)), f32(eps), *[_p(a) for a in r.arrays()])
    return r


def ao_rays(pos, idx, prim_rays, prim, t, w, h, frame=0, offset=0.003, tmin=0.001,
            tmax=0.5):
    r = Rays(w * h)
    pr = prim_rays
    lib().or_ao_rays(_p(_f(pos)), _p(None if idx is None else _u(idx)), _p(pr.ox), _p(pr.oy),
                     _p(pr.oz),
                     _p(pr.dx), _p(pr.dy), _p(pr.dz), _p(_u(prim)), _p(_f(t)), u32(w), u32(h),
                     u32(frame), f32(offset), f32(tmin), f32(tmax),
                     *[_p(a) for a in r.arrays()])
    return r


class _Inst(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("idx", C.c_void_p), ("world_to_object", C.c_float * 12),
                ("kind", C.c_uint32)]


INST_GENERAL, INST_TRANSLATE, INST_IDENTITY = 0, 1, 2


def map_rays(rays, m, kind):
    """World-to-object map of a ray batch as the device applies it
    (float32, ((m0 x + m1 y) + m2 z) + t; direction by the 3x3 part)."""
    if kind == INST_IDENTITY:
        return rays
    m = np.asarray(m, np.float32).reshape(3, 4)
    o = [rays.ox, rays.oy, rays.oz]
    d = [rays.dx, rays.dy, rays.dz]
    if kind == INST_TRANSLATE:
        mo = [o[a] + m[a, 3] for a in range(3)]
        md = d
    else:
        mo = [m[a, 0] * o[0] + m[a, 1] * o[1] + m[a, 2] * o[2] + m[a, 3] for a in range(3)]
        md = [m[a, 0] * d[0] + m[a, 1] * d[1] + m[a, 2] * d[2] for a in range(3)]
    return Rays(ox=mo[0], oy=mo[1], oz=mo[2], tmin=rays.tmin, dx=md[0], dy=md[1], dz=md[2],
                tmax=rays.tmax)


def trace_scene(instances, rays, mode=CLOSEST):
    """Two-level scene on the CPU path: instances = [(nodes4, tris64, m, kind)]
    visited in order; each BLAS traced (or_trace_bvh4) on the mapped ray with
    tmax = the best t so far (closest), any hit stops at the first occluding
    instance.  Returns (prim, t, inst) for closest, the occluded mask for any."""
    n = len(rays)
    best_t = rays.tmax.copy()
    prim = np.full(n, MISS, np.uint32)
    inst = np.full(n, MISS, np.uint32)
    occ = np.zeros(n, bool)
    for i, (nodes, tris, m, kind) in enumerate(instances):
        r = map_rays(rays, m, kind)
        if mode == ANY:
            live = ~occ
            if not live.any():
                break
            sub = r.subset(live)
            p, _ = trace_bvh4(nodes, tris, sub, mode=ANY)
            occ[np.nonzero(live)[0][p != MISS]] = True
        else:
            r = Rays(ox=r.ox, oy=r.oy, oz=r.oz, tmin=r.tmin, dx=r.dx, dy=r.dy, dz=r.dz,
                     tmax=best_t)
            p, t = trace_bvh4(nodes, tris, r)
            hit = p != MISS
            prim[hit], best_t[hit], inst[hit] = p[hit], t[hit], i
    return occ if mode == ANY else (prim, best_t, inst)


def ao_rays_scene(instances, prim_rays, prim, inst, t, w, h, frame=0, offset=0.003,
                  tmin=0.001, tmax=0.5):
    """instances = [(pos, idx or None, m (3x4), kind)] (the device's
    rt_instance meshes); or_ao_rays_scene."""
    keep = [(_f(pos), None if idx is None else _u(idx)) for pos, idx, _, _ in instances]
    arr = (_Inst * len(instances))()
    for k, (pos, idx, m, kind) in enumerate(instances):
        arr[k].pos = _p(keep[k][0])
        arr[k].idx = _p(keep[k][1])
        arr[k].world_to_object = (C.c_float * 12)(*np.asarray(m, np.float32).reshape(-1))
        arr[k].kind = kind
    r = Rays(w * h)
    pr = prim_rays
    lib().or_ao_rays_scene(arr, u32(len(instances)), _p(pr.ox), _p(pr.oy), _p(pr.oz),
                           _p(pr.dx), _p(pr.dy), _p(pr.dz), _p(_u(prim)), _p(_u(inst)),
                           _p(_f(t)), u32(w), u32(h), u32(frame), f32(offset), f32(tmin),
                           f32(tmax), *[_p(a) for a in r.arrays()])
    return r


def random_rays(count, seed, lo, hi, tmax=1e30):
    r = Rays(count)
    lib().or_random_rays(u32(count), u32(seed), _p(_f(lo)), _p(_f(hi)), f32(tmax),
                         *[_p(a) for a in r.arrays()])
    return r


# ------------------------------------------------------------------ LBVH -----
def centroid_bounds(pos, idx=None, ntris=None):
    pos = _f(pos)
    n = idx.shape[0] if idx is not None else pos.size // 9
    b = np.zeros(6, np.float32)
    lib().or_centroid_bounds(_p(pos), _p(None if idx is None else _u(idx)), u32(n), _p(b))
    return b


def morton30(pos, idx=None, bounds=None):
    pos = _f(pos)
    n = idx.shape[0] if idx is not None else pos.size // 9
    if bounds is None:
        bounds = centroid_bounds(pos, idx)
    codes = np.zeros(n, np.uint32)
    lib().or_morton30(_p(pos), _p(None if idx is None else _u(idx)), u32(n), _p(_f(bounds)),
                      _p(codes))
    return codes


def expand_bits10(v):
    return int(lib().or_expand_bits10(u32(v)))


def sort_pairs(keys, vals):
    k, v = _u(keys).copy(), _u(vals).copy()
    lib().or_sort_pairs(_p(k), _p(v), u32(len(k)))
    return k, v


def tree_layout(n):
    counts = np.zeros(32, np.uint32)
    offs = np.zeros(32, np.uint32)
    L, tb = u32(0), u32(0)
    sz = C.c_uint64(0)
    lib().or_tree_layout(u32(n), _p(counts), _p(offs), C.byref(L), C.byref(tb), C.byref(sz))
    return counts[:L.value].copy(), offs[:L.value].copy(), tb.value, sz.value


def lbvh_build(pos, idx=None):
    pos = _f(pos)
    n = idx.shape[0] if idx is not None else pos.size // 9
    _, _, tb, _ = tree_layout(n) if n else (None, None, 0, 0)
    nodes = np.zeros(max(tb, 1), NODE4)
    tris = np.zeros(n, TRI64)
    ids = np.zeros(n, np.uint32)
    codes = np.zeros(n, np.uint32)
    cnt = u32(0)
    rc = lib().or_lbvh_build(_p(pos), _p(None if idx is None else _u(idx)), u32(n), _p(nodes),
                             u32(len(nodes)), C.byref(cnt), _p(tris), _p(ids), _p(codes))
    if rc:
        raise RuntimeError("or_lbvh_build failed")
    return nodes[:cnt.value].copy(), tris, ids, codes


def rlbvh_build(pos, idx=None, max_prims=4, with_bvh2=False, treelet_rounds=0):
    """Radix-tree LBVH collapsed to Box4: (nodes4, tris64, sorted ids, codes);
    treelet_rounds rounds of SAH treelet restructuring before the collapse."""
    pos = _f(pos)
    n = idx.shape[0] if idx is not None else pos.size // 9
    cap = max(1, 2 * n)
    n2 = np.zeros(cap, NODE2)
    n4 = np.zeros(cap, NODE4)
    tris = np.zeros(n, TRI64)
    ids = np.zeros(n, np.uint32)
    codes = np.zeros(n, np.uint32)
    c2, c4 = u32(0), u32(0)
    rc = lib().or_rlbvh_build_opt(_p(pos), _p(None if idx is None else _u(idx)), u32(n),
                                  u32(max_prims), u32(treelet_rounds), _p(n2), u32(cap),
                                  C.byref(c2), _p(n4), u32(cap), C.byref(c4), _p(tris), _p(ids),
                                  _p(codes))
    if rc:
        raise RuntimeError("or_rlbvh_build failed")
    out = (n4[:c4.value].copy(), tris, ids, codes)
    return out + (n2[:c2.value].copy(),) if with_bvh2 else out


def lbvh_refit(pos, idx, nodes, tris):
    pos = _f(pos)
    n = len(tris)
    nodes = nodes.copy()
    tris = tris.copy()
    lib().or_lbvh_refit(_p(pos), _p(None if idx is None else _u(idx)), u32(n), _p(nodes),
                        u32(len(nodes)), _p(tris))
    return nodes, tris


def bvh4_refit(pos, idx, nodes, tris):
    """Generic refit (or_bvh4_refit) of any Box4 tree with children after
    their parents: returns new (nodes, tris)."""
    pos = _f(pos)
    nodes = nodes.copy()
    tris = tris.copy()
    rc = lib().or_bvh4_refit(_p(pos), _p(None if idx is None else _u(idx)), _p(nodes),
                             u32(len(nodes)), _p(tris), u32(len(tris)))
    if rc:
        raise ValueError("tree is not in parent-before-child order")
    return nodes, tris


def radv_encode(xyz9, order=None):
    xyz9 = _f(xyz9)
    n = xyz9.size // 9
    _, _, _, sz = tree_layout(n)
    out = np.zeros(sz, np.uint8)
    rc = lib().or_radv_encode(_p(xyz9), u32(n), _p(None if order is None else _u(order)), _p(out),
                              C.c_uint64(sz))
    if rc:
        raise RuntimeError("or_radv_encode failed")
    return out


def radv_decode(img):
    """RADV image (uint8 array) -> (Box4 nodes NODE4, 64-B leaves TRI64)."""
    img = np.ascontiguousarray(img, dtype=np.uint8).reshape(-1)
    cap = max(1, img.size // 128)
    nodes = np.zeros(cap, NODE4)
    tris = np.zeros(max(1, img.size // 64), TRI64)
    nc, nt = u32(0), u32(0)
    rc = lib().or_radv_decode(_p(img), C.c_uint64(img.size), _p(nodes), u32(cap), C.byref(nc),
                              _p(tris), u32(len(tris)), C.byref(nt))
    if rc:
        raise ValueError("malformed RADV image")
    return nodes[: nc.value].copy(), tris[: nt.value].copy()


# ------------------------------------------------------------- post ----------
def ao_resolve(bits, prim):
    npix = len(prim)
    ao = np.zeros(npix, np.float32)
    lib().or_ao_resolve(_p(_u(bits)), _p(_u(prim)), u32(npix), _p(ao))
    return ao


def blur7(img, depth, w, h, thr):
    out = np.zeros(w * h, np.float32)
    lib().or_blur7(_p(_f(img)), _p(_f(depth)), u32(w), u32(h), f32(thr), _p(out))
    return out


def rms(a, b):
    a, b = _f(a), _f(b)
    return float(lib().or_rms(_p(a), _p(b), C.c_uint64(a.size)))


def percentile(v, q):
    v = _f(v).copy()
    return float(lib().or_percentile(_p(v), u32(len(v)), f32(q)))


TIMER_FIELDS = ("median", "min", "max", "p5", "p95", "avg", "stddev", "p99")


def timer_stats(v):
    """GpuTimer::runBenchmark statistics (common/vk_timer.cpp:96-121) + p99."""
    v = _f(v).reshape(-1).copy()
    out = np.zeros(8, np.float32)
    lib().or_timer_stats(_p(v), u32(len(v)), _p(out))
    return dict(zip(TIMER_FIELDS, (float(x) for x in out)))


def image_metrics(a, b, blocks=False):
    """ImageMetrics::compute (common/image_metrics.hpp:16-104): (mse, psnr,
    ssim[, per-block clamped SSIM]) of two (h, w) float32 images."""
    a, b = _f(a), _f(b)
    h, w = a.shape
    out = np.zeros(3, np.float64)
    blk = np.zeros(max(1, (w // 8) * (h // 8)), np.float64) if blocks else None
    lib().or_image_metrics(_p(a), _p(b), u32(w), u32(h), _p(out), _p(blk))
    res = tuple(float(x) for x in out)
    return res + (blk[: (w // 8) * (h // 8)],) if blocks else res


RADV_HEADER_FIELDS = ("bvh_offset", "root_flags", "update_dispatch_size0",
                      "update_dispatch_size1", "update_dispatch_size2", "compacted_size",
                      "serialization_size", "copy_dispatch_size0", "copy_dispatch_size1",
                      "copy_dispatch_size2", "size", "geometry_type", "geometry_count",
                      "leaf_nodes_offset", "primitive_count", "instance_count",
                      "leaf_node_offsets_offset", "build_flags", "type")


def ref_radv_header(img):
    """The header and root child ids of a RADV image read through the
    reference's own RadvAccelStructHeader / RadvBvhBox32Node (oracle/_ref)."""
    img = np.ascontiguousarray(img, dtype=np.uint8).reshape(-1)
    out = np.zeros(len(RADV_HEADER_FIELDS), np.uint64)
    aabb = np.zeros(6, np.float32)
    root = np.zeros(4, np.uint32)
    ref().ref_radv_header(_p(img), _p(out), _p(aabb), _p(root))
    d = dict(zip(RADV_HEADER_FIELDS, (int(x) for x in out)))
    d["aabb"] = aabb
    d["root_children"] = root
    return d


def ref_image_metrics(a, b):
    """The reference's own ImageMetrics::compute (oracle/_ref)."""
    a, b = _f(a), _f(b)
    h, w = a.shape
    out = np.zeros(3, np.float64)
    ref().ref_image_metrics(_p(a), _p(b), u32(w), u32(h), _p(out))
    return tuple(float(x) for x in out)


# ------------------------------------------------------------ reference -----
def ref_terrain(target, size=100.0, shipped=False):
    r = ref(shipped)
    nv, nt = u32(0), u32(0)
    r.ref_terrain(u32(target), f32(size), None, None, C.byref(nv), C.byref(nt))
    pos = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.uint32)
    r.ref_terrain(u32(target), f32(size), _p(pos), _p(idx), C.byref(nv), C.byref(nt))
    return pos, idx


def ref_build(pos, idx, max_prims=4, shipped=False):
    r = ref(shipped)
    pos, idx = _f(pos), _u(idx)
    n = idx.shape[0]
    cap = 2 * n + 1
    n2 = np.zeros(cap, NODE2)
    n4 = np.zeros(cap, NODE4)
    t48 = np.zeros((n, 12), np.float32)
    c2, c4 = u32(0), u32(0)
    rc = r.ref_build(_p(pos), u32(pos.shape[0]), _p(idx), u32(n), u32(max_prims), _p(n2),
                     u32(cap), C.byref(c2), _p(n4), u32(cap), C.byref(c4), _p(t48))
    if rc:
        raise RuntimeError("ref_build failed")
    return n2[:c2.value].copy(), n4[:c4.value].copy(), t48


def ref_tree_layout(n):
    r = ref()
    counts = np.zeros(64, np.uint32)
    offs = np.zeros(64, np.uint32)
    L, tb, ta = u32(0), u32(0), u32(0)
    pc = np.zeros(136, np.uint8)
    r.ref_tree_layout(u32(n), _p(counts), _p(offs), C.byref(L), C.byref(tb), C.byref(ta), _p(pc))
    return counts[:L.value].copy(), offs[:L.value].copy(), tb.value, ta.value, pc


def ref_dynamic_grid(n, t=0.0, extent=2.0):
    r = ref()
    xyz = np.zeros((n, 9), np.float32)
    cen = np.zeros((n, 3), np.float32)
    r.ref_dynamic_grid(u32(n), f32(t), f32(extent), _p(xyz), _p(cen))
    return xyz, cen


def ref_sizes():
    out = np.zeros(8, np.uint32)
    ref().ref_sizes(_p(out))
    return out
