"""The C++ host interface (rdna2-accel-raster_amd/host/gpu_bvh.hpp) through
its probe program tests/cpp/exp0_probe_gpu: the reference's exp0 traversal
experiment (benchmarks/exp0_probe, results/exp0_ray_stats.csv:2-7) driven
from C++ on the GPU, and the wrapper's error behaviour on the host."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "cpp", "exp0_probe_gpu")


def _probe():
    if not os.path.exists(PROBE):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "tests", "cpp")])
    return PROBE


def test_cpp_host_errors_without_device_work():
    r = subprocess.run([_probe(), "--host-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host-only ok" in r.stdout


@pytest.mark.gpu
def test_cpp_probe_reproduces_published_exp0():
    """Hit rate and mean box / triangle tests per ray of the C++ path on the
    GPU equal the reference's published software-path statistics within the
    tolerance of the Python path's test; a device build refitted at the same
    positions keeps its bytes; the BVH2, RADV-image and image-metrics wrappers
    run on the same terrains."""
    from test_oracle_known import EXP0
    r = subprocess.run([_probe(), "1000", "100000"], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [line.split(",") for line in r.stdout.splitlines()
            if line and line[0].isdigit() and "," in line]
    assert len(rows) == 4
    for row in rows:
        ntri, pattern = int(row[0]), row[1]
        hit, box, _, tri, _, mrays = map(float, row[2:])
        ref = EXP0[(ntri, pattern)]
        assert abs(hit - ref["hit"]) < 2e-3
        assert abs(box - ref["box"]) / ref["box"] < 2e-3
        assert abs(tri - ref["tri"]) / ref["tri"] < 2e-3
        assert mrays > 0
    assert r.stdout.count("refit at the same positions unchanged") == 2
    # occluder hints (own closest hits, and a quarter of them for every quarter) leave
    # the any-hit bits unchanged, with and without the wave's hints
    assert r.stdout.count("hinted occlusion equals any hit") == 4
    # BVH2 shaders on the same terrains and rays (GpuBVH2): hit rates within 0.05 points
    # of the BVH4 path (the shaders' predicates differ only at |a| < 1e-7 vs 1e-8 and the
    # t >= / > tmin edge), both stack variants measured
    hit4 = {(int(row[0]), row[1]): float(row[2]) for row in rows}
    b2 = [line.split(",")[1:] for line in r.stdout.splitlines() if line.startswith("bvh2,")]
    assert len(b2) == 8
    for ntri, pattern, shader, hit, box, tri, mrays in b2:
        assert shader in ("Software_ALU_Scratch", "Software_ALU_LDS")
        assert abs(float(hit) - hit4[(int(ntri), pattern)]) < 0.05
        assert float(box) > 1 and float(tri) > 0 and float(mrays) > 0
    # RADV image of the first terrain traced over its bytes = the implicit Box4 build
    assert "traced over its bytes equals the implicit Box4 tree" in r.stdout
    assert "mse 0 psnr 100" in r.stdout
