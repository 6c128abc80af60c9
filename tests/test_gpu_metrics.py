"""rt_image_metrics (ImageMetrics::compute, common/image_metrics.hpp:16-104)
on the device against the oracle's restatement:

  * every 8x8 block's clamped SSIM bit-identical (same per-block order of
    operations, no contraction, IEEE double division);
  * mse and the mean SSIM within the rounding bound of a double sum of n
    non-negative terms taken in two different orders: |difference| <=
    2 (n - 1) 2^-53 of the sum (n = pixel count for the mse, block count for
    the SSIM); psnr = 10 log10(1 / mse) within 1e-9 dB;
  * ragged sizes (pixels outside the full blocks), widths that are not a
    multiple of 4 and unaligned image starts (the scalar-load path), no full
    block, no pixels, identical images, and a 1920 x 1080 frame.
"""
import math

import numpy as np
import pytest
import torch

from conftest import to_np

pytestmark = pytest.mark.gpu


def _images(seed, h, w):
    rng = np.random.default_rng(seed)
    a = rng.random((h, w), dtype=np.float32)
    b = np.clip(a + rng.normal(0, 0.05, (h, w)).astype(np.float32), 0, 1).astype(np.float32)
    return a, b


def _check(rt, O, a, b, da=None, db=None):
    h, w = a.shape
    da = torch.from_numpy(a).cuda() if da is None else da
    db = torch.from_numpy(b).cuda() if db is None else db
    res, blk = rt.image_metrics(da, db, blocks=True)
    torch.cuda.synchronize()
    mse, psnr, ssim = to_np(res).tolist()
    wm, wp, ws, wblk = O.image_metrics(a, b, blocks=True)
    np.testing.assert_array_equal(to_np(blk).reshape(-1).view(np.uint64),
                                  wblk.reshape(-1).view(np.uint64))
    n = h * w
    nb = (h // 8) * (w // 8)
    assert abs(mse - wm) <= 2 * max(n - 1, 1) * 2.0 ** -53 * wm, (mse, wm)
    assert abs(psnr - wp) <= 1e-9, (psnr, wp)
    assert abs(ssim - ws) <= 2 * max(nb - 1, 1) * 2.0 ** -53 * max(ws, 1e-300), (ssim, ws)
    return mse, psnr, ssim


@pytest.mark.parametrize("h,w", [(64, 64), (67, 93), (8, 8), (9, 1001), (1000, 9), (257, 516)])
def test_metrics_match_oracle(rt, O, h, w):
    a, b = _images(h * 7 + w, h, w)
    _check(rt, O, a, b)


def test_metrics_full_hd_frame(rt, O):
    a, b = _images(3, 1080, 1920)
    mse, psnr, ssim = _check(rt, O, a, b)
    assert 0.0 < ssim < 1.0 and 20.0 < psnr < 40.0


def test_metrics_unaligned_and_edge_cases(rt, O):
    a, b = _images(4, 41, 65)
    # image starts one float into a buffer: the scalar-load path
    ba = torch.zeros(a.size + 1, dtype=torch.float32, device="cuda")
    bb = torch.zeros(b.size + 1, dtype=torch.float32, device="cuda")
    ba[1:] = torch.from_numpy(a.reshape(-1)).cuda()
    bb[1:] = torch.from_numpy(b.reshape(-1)).cuda()
    _check(rt, O, a, b, ba[1:].view(41, 65), bb[1:].view(41, 65))
    # no full block: ssim 1
    c, d = _images(5, 7, 300)
    assert _check(rt, O, c, d)[2] == 1.0
    # identical images: mse 0, psnr 100
    mse, psnr, ssim = _check(rt, O, a, a)
    assert (mse, psnr) == (0.0, 100.0) and abs(ssim - 1.0) < 1e-15
    # no pixels
    e = torch.empty((0, 16), dtype=torch.float32, device="cuda")
    res = rt.image_metrics(e, e)
    torch.cuda.synchronize()
    assert to_np(res).tolist() == [0.0, 100.0, 1.0]


def test_metrics_scratch_capacity_error(rt):
    a = torch.zeros((64, 64), dtype=torch.float32, device="cuda")
    out = torch.empty(3, dtype=torch.float64, device="cuda")
    need = rt.image_metrics_scratch_bytes(64, 64)
    s = torch.empty(max(1, need // 8), dtype=torch.float64, device="cuda")
    with pytest.raises(rt.RtError):
        rt.call("rt_image_metrics", rt.stream_handle(), rt.ptr(a), rt.ptr(a), rt.u32(64),
                rt.u32(64), rt.ptr(out), None, rt.ptr(s), rt.C.c_size_t(need - 8))
    assert math.isfinite(need)
