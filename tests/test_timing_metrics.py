"""Timing statistics and image-quality metrics on the host side.

* timing.benchmark_stats (the bench's and the serving loop's p50 / p99)
  equals the oracle's restatement of GpuTimer::runBenchmark's statistics
  (common/vk_timer.cpp:96-121) on random samples, and both are pinned by
  closed forms: median = sorted[n // 2] (the upper middle for even n),
  pq = sorted[floor(q (n - 1))], mean and population standard deviation of
  0..n-1.
* The oracle's ImageMetrics::compute (common/image_metrics.hpp:16-104) is
  pinned by the reference's own header compiled in oracle/_ref (bit-equal
  where built), by closed forms (identical images; a constant offset gives
  mse = offset^2 and the SSIM luminance term alone), and by an independent
  numpy evaluation of the block SSIM.
"""
import math
import os

import numpy as np
import pytest

import timing


def test_timer_stats_closed_forms(O):
    s = O.timer_stats(np.arange(100, dtype=np.float32)[::-1])
    assert s["median"] == 50.0  # sorted[n / 2], not sorted[(n - 1) / 2]
    assert (s["min"], s["max"], s["p5"], s["p95"], s["p99"]) == (0.0, 99.0, 4.0, 94.0, 98.0)
    assert s["avg"] == 49.5
    assert abs(s["stddev"] - math.sqrt((100 ** 2 - 1) / 12)) < 1e-5
    assert O.timer_stats([3.0, 1.0])["median"] == 3.0
    assert O.timer_stats([5.0, 1.0, 3.0])["median"] == 3.0
    assert O.timer_stats(np.arange(200, dtype=np.float32))["p99"] == 197.0
    assert O.timer_stats([])["median"] == 0.0
    one = O.timer_stats([2.5])
    assert one["median"] == one["p5"] == one["p99"] == one["avg"] == 2.5 and one["stddev"] == 0.0


def test_timer_stats_host_equals_oracle(O):
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 10, 21, 100, 101, 999, 4096):
        v = (rng.lognormal(0.0, 0.7, n) * 3.0).astype(np.float32)
        got = timing.benchmark_stats(v)
        want = O.timer_stats(v)
        for k in timing.FIELDS:
            assert np.float32(got[k]) == np.float32(want[k]), (n, k, got[k], want[k])
        assert got["run_count"] == n
        assert timing.median(v) == want["median"]
        assert timing.percentile(v, 0.99) == want["p99"]


def _ssim_numpy(a, b):
    h, w = a.shape
    bh, bw = h // 8, w // 8
    A = a[: bh * 8, : bw * 8].astype(np.float64).reshape(bh, 8, bw, 8).transpose(0, 2, 1, 3)
    B = b[: bh * 8, : bw * 8].astype(np.float64).reshape(bh, 8, bw, 8).transpose(0, 2, 1, 3)
    A, B = A.reshape(bh, bw, 64), B.reshape(bh, bw, 64)
    ma, mb = A.mean(-1), B.mean(-1)
    va = ((A - ma[..., None]) ** 2).sum(-1) / 63
    vb = ((B - mb[..., None]) ** 2).sum(-1) / 63
    cab = ((A - ma[..., None]) * (B - mb[..., None])).sum(-1) / 63
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    s = (2 * ma * mb + c1) * (2 * cab + c2) / ((ma ** 2 + mb ** 2 + c1) * (va + vb + c2))
    return np.clip(s, 0.0, 1.0)


def _images(seed, h, w):
    rng = np.random.default_rng(seed)
    a = rng.random((h, w), dtype=np.float32)
    b = np.clip(a + rng.normal(0, 0.05, (h, w)).astype(np.float32), 0, 1).astype(np.float32)
    return a, b


def test_image_metrics_closed_forms(O):
    a, _ = _images(1, 40, 56)
    mse, psnr, ssim = O.image_metrics(a, a)
    assert (mse, psnr) == (0.0, 100.0) and abs(ssim - 1.0) < 1e-15
    # exact offset: values k/256 plus 1/16 are exact floats
    q = (np.floor(a * 128) / 256).astype(np.float32)
    off = (q + np.float32(1 / 16)).astype(np.float32)
    mse, psnr, ssim, blk = O.image_metrics(q, off, blocks=True)
    assert mse == (1 / 16) ** 2
    assert abs(psnr - 20 * math.log10(16)) < 1e-12
    # equal variances and covariance = variance: SSIM = the luminance term
    Q = q.astype(np.float64).reshape(5, 8, 7, 8).transpose(0, 2, 1, 3).reshape(5, 7, 64)
    m = Q.mean(-1)
    lum = (2 * m * (m + 1 / 16) + 1e-4) / (m ** 2 + (m + 1 / 16) ** 2 + 1e-4)
    np.testing.assert_allclose(blk.reshape(5, 7), lum, rtol=1e-12)
    assert abs(ssim - lum.mean()) < 1e-12
    # no full block; no pixels
    mse, psnr, ssim = O.image_metrics(a[:7, :30], a[:7, :30] * 0.5)
    assert ssim == 1.0 and mse > 0 and psnr < 100
    assert O.image_metrics(np.zeros((0, 8), np.float32), np.zeros((0, 8), np.float32)) == (
        0.0, 100.0, 1.0)


def test_image_metrics_blocks_match_numpy(O):
    a, b = _images(2, 67, 93)  # ragged: 3 right columns and 3 bottom rows outside the blocks
    mse, psnr, ssim, blk = O.image_metrics(a, b, blocks=True)
    want = _ssim_numpy(a, b)
    np.testing.assert_allclose(blk.reshape(want.shape), want, rtol=1e-12, atol=1e-15)
    assert abs(ssim - want.mean()) < 1e-12
    d = a.astype(np.float64) - b.astype(np.float64)
    assert abs(mse - (d * d).mean()) < 1e-15
    assert abs(psnr - 10 * math.log10(1 / mse)) < 1e-12
    # a shifted block grid (a transposed or offset index) is caught
    assert np.abs(_ssim_numpy(a[1:, :], b[1:, :]) - blk.reshape(want.shape)[: 8, :]).max() > 1e-3


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle",
                                                    "_ref", "libref.so")),
                    reason="oracle/_ref not built (no reference checkout)")
def test_image_metrics_equal_reference_header(O):
    for seed, (h, w) in enumerate([(64, 64), (67, 93), (7, 30), (120, 8), (1080 // 4, 1920 // 4)]):
        a, b = _images(10 + seed, h, w)
        assert O.image_metrics(a, b) == O.ref_image_metrics(a, b)
