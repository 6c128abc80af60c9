"""Write the golden fixtures under tests/golden/ from the reference's own CPU
path, compiled from its headers by oracle/Makefile (`make -C oracle ref`):

  ref_<name>.npz   inputs and the reference's outputs (node / triangle arrays
                   as raw bytes, layouts, push constants)

Only data is stored; no reference source.  Run in a container that has the
reference checkout:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def main():
    if O.ref() is None or O.ref(shipped=True) is None:
        raise SystemExit("oracle/_ref not built (make -C oracle ref)")
    out = {}
    # createTessellatedTerrain(1000) as the reference's own build evaluates it,
    # then SoftwareBVH::build + SoftwareBVH4::fromBVH2 (results/exp0 pipeline)
    pos, idx = O.ref_terrain(1000, shipped=True)
    n2, n4, t48 = O.ref_build(pos, idx, shipped=True)
    out["terrain1000"] = dict(pos=pos, idx=idx, bvh2=n2.view(np.uint8), bvh4=n4.view(np.uint8),
                              tris48=t48)
    # strict build (one rounding per operator) of the same target
    pos, idx = O.ref_terrain(1000)
    n2, n4, t48 = O.ref_build(pos, idx)
    out["terrain1000_strict"] = dict(pos=pos, idx=idx, bvh2=n2.view(np.uint8),
                                     bvh4=n4.view(np.uint8), tris48=t48)
    # a mesh with many centroid ties: 2 triangles per quad on a flat grid,
    # max_prims 1..4 (nth_element tie order)
    gp, gi = O.wave_grid(12, 10, 1.0, 0.0, 0.0, 1.0, 1.0)
    for mp in (1, 2, 4):
        n2, n4, t48 = O.ref_build(gp, gi, max_prims=mp)
        out[f"flatgrid_mp{mp}"] = dict(pos=gp, idx=gi, bvh2=n2.view(np.uint8),
                                       bvh4=n4.view(np.uint8), tris48=t48)
    # generate_dynamic_grid_mesh (radv_bvh.hpp:224) at the builder sweep sizes
    for n in (64, 256, 1024, 4096):
        xyz, cen = O.ref_dynamic_grid(n, 0.0, 1.25)
        out[f"dyngrid{n}"] = dict(xyz=xyz, cen=cen)
    # compute_tree_layout + to_push_constants (radv_bvh.hpp:169-215)
    ns = np.array([1, 3, 4, 5, 16, 17, 64, 256, 1024, 4096, 4097, 65536, 65537, 2_500_000],
                  np.uint32)
    lay = {}
    for n in ns:
        c, o, tb, ta, pc = O.ref_tree_layout(int(n))
        lay[f"counts_{n}"] = c
        lay[f"offsets_{n}"] = o
        lay[f"scalars_{n}"] = np.array([tb, ta], np.uint64)
        lay[f"push_{n}"] = pc
    out["layout"] = dict(ns=ns, sizes=O.ref_sizes(), **lay)
    for name, arrs in out.items():
        path = os.path.join(HERE, f"ref_{name}.npz")
        np.savez(path, **arrs)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
