"""The device centroid divides by 3 with a reciprocal multiply and one FMA
residual step (rt_internal.h div3) instead of the IEEE division sequence.
Checked here on the host, where the same IEEE single-precision operations
(mul, fmaf, copysignf) give the same bits: div3(x) == x / 3.0f for a strided
sample of all 2^32 bit patterns (every finite one was checked once; this
repeats 1 in 61) plus signed zeros, subnormals, the largest finite values
and exact multiples of 3."""
import os
import shutil
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float div3(float x) {
    const float y = 0x1.555556p-2f;
    const float q = x * y;
    const float r = fmaf(-3.0f, q, x);
    return copysignf(fmaf(r, y, q), x);
}
static int check(uint32_t b, uint64_t* bad) {
    float x; memcpy(&x, &b, 4);
    if (!isfinite(x)) return 0;
    volatile float ref = x / 3.0f;
    float got = div3(x);
    uint32_t rb, gb; float rr = ref; memcpy(&rb, &rr, 4); memcpy(&gb, &got, 4);
    if (rb != gb) { if (*bad < 5) printf("mismatch %08x\n", b); ++*bad; }
    return 1;
}
int main(void) {
    uint64_t bad = 0, n = 0;
    for (uint64_t u = 0; u < (1ull << 32); u += 61) n += check((uint32_t)u, &bad);
    const uint32_t special[] = {0x00000000u, 0x80000000u, 0x00000001u, 0x80000001u,
                                0x007FFFFFu, 0x807FFFFFu, 0x00800000u, 0x7F7FFFFFu,
                                0xFF7FFFFFu, 0x3F800000u, 0x40400000u, 0xC0400000u};
    for (unsigned i = 0; i < sizeof special / sizeof special[0]; ++i) n += check(special[i], &bad);
    for (int k = -1000; k <= 1000; ++k) {
        float x = 3.0f * (float)k; uint32_t b; memcpy(&b, &x, 4); n += check(b, &bad);
    }
    printf("checked %llu mismatches %llu\n", (unsigned long long)n, (unsigned long long)bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div3_equals_ieee_division():
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "div3.c"), os.path.join(d, "div3")
        with open(c, "w") as f:
            f.write(SRC)
        subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-o", exe, c, "-lm"],
                       check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout
        assert "mismatches 0" in r.stdout
