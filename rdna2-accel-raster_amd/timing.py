"""Timing statistics with the conventions of the reference's GpuTimer
(common/vk_timer.cpp:96-121), shared by bench.py and the serving loop.

Samples are float32 milliseconds as the reference keeps them.  median =
sorted[n // 2]; p5 / p95 (and p99, the same rule) = sorted[floor(q (n - 1))]
with q and n - 1 multiplied in float32; mean = float32(double sum / n) with
the sum taken in sample order after sorting; standard deviation over n of
float32 differences from that mean, squared and summed in double."""
import math

import numpy as np

FIELDS = ("median", "min", "max", "p5", "p95", "avg", "stddev", "p99")


def _index(q, n):
    return int(np.float32(q) * np.float32(n - 1))


def benchmark_stats(samples):
    """dict of FIELDS (floats; all 0 for no samples) plus run_count."""
    s = np.sort(np.asarray(samples, dtype=np.float32).reshape(-1))
    n = len(s)
    if n == 0:
        return dict({k: 0.0 for k in FIELDS}, run_count=0)
    total = 0.0
    for v in s.tolist():
        total += v
    avg = np.float32(total / n)
    var = 0.0
    for v in (s - avg).tolist():
        var += v * v
    out = {"median": s[n // 2], "min": s[0], "max": s[-1], "p5": s[_index(0.05, n)],
           "p95": s[_index(0.95, n)], "avg": avg, "stddev": np.float32(math.sqrt(var / n)),
           "p99": s[_index(0.99, n)]}
    return dict({k: float(v) for k, v in out.items()}, run_count=n)


def median(samples):
    """sorted[n // 2] of the float32 samples (vk_timer.cpp:103); nan if none."""
    s = np.sort(np.asarray(samples, dtype=np.float32).reshape(-1))
    return float(s[len(s) // 2]) if len(s) else float("nan")


def percentile(samples, q):
    """sorted[floor(q (n - 1))] of the float32 samples (vk_timer.cpp:105-108);
    nan if none."""
    s = np.sort(np.asarray(samples, dtype=np.float32).reshape(-1))
    return float(s[_index(q, len(s))]) if len(s) else float("nan")
