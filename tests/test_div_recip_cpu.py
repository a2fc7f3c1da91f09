"""The reciprocal-multiply divide used for the cubic-spline kernel's constant
divisors (device_math.hpp div_by_recip): RN32(RN64(a * RN64(1 / b))) equals
the IEEE single-precision a / b whenever |a / b| >= 2^-126; smaller results
take the IEEE divide. Checked with numpy's IEEE float32 divide and float64
multiply on random bit patterns and on the benchmark kernel constants."""

import numpy as np

from emerald_amd import scenes


def _recip_div(a, b):
    inv = np.float64(1.0) / b.astype(np.float64)
    q = a.astype(np.float64) * inv
    out = q.astype(np.float32)
    small = ~(np.abs(q) >= 2.0 ** -126)
    out[small] = (a[small] / b[small])
    return out


def _same_bits(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    both_nan = np.isnan(x) & np.isnan(y)
    return both_nan | (x.view(np.uint32) == y.view(np.uint32))


def test_random_bit_patterns():
    rng = np.random.default_rng(2024)
    with np.errstate(all="ignore"):
        for _ in range(8):
            a = rng.integers(0, 2**32, 2_000_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
            b = rng.integers(0, 2**32, 2_000_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
            ok = _same_bits(_recip_div(a, b), a / b)
            assert ok.all(), (a[~ok][:4], b[~ok][:4])


def test_subnormal_midpoint_takes_the_divide():
    # 3 * 2^-149 / 6 = 2^-150 is a midpoint: ties to even gives +0, while the
    # reciprocal product lands just off it.
    a = np.array([3 * 2.0 ** -149], np.float32)
    b = np.array([6.0], np.float32)
    with np.errstate(all="ignore"):
        assert _same_bits(_recip_div(a, b), a / b).all()
        q = (a.astype(np.float64) * (1.0 / b.astype(np.float64))).astype(np.float32)
    assert abs(float(q[0]) - float((a / b)[0])) <= 2.0 ** -149


def test_kernel_constants_dense():
    # The divisors of W / dW for the benchmark supports (h = spacing):
    # h, h^2, h^3, 4h and pi h^3, over dividends spanning the kernel's range.
    rng = np.random.default_rng(7)
    for s in (0.005, 1.0 / 2560.0, 0.8 / 1000.0):
        h = np.float32(s)
        pi = np.float32(3.14159274101257324218750)
        divisors = [h, h * h, h * h * h, np.float32(4.0) * h, h * h * h * pi]
        for d in divisors:
            b = np.full(1_000_000, d, np.float32)
            a = (rng.random(1_000_000) * 8.0 * float(h)).astype(np.float32)
            a[:1000] = np.arange(1000, dtype=np.float32) * np.float32(h) / np.float32(500.0)
            assert _same_bits(_recip_div(a, b), a / b).all(), float(d)
            w = (rng.standard_normal(1_000_000) * 4.0).astype(np.float32)
            assert _same_bits(_recip_div(w, b), w / b).all(), float(d)
