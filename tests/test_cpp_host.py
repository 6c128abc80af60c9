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
    positions keeps its bytes."""
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
