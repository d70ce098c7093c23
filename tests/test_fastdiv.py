"""The multiply-high division the REGIME kernels use for padded cell
indices (mgpu_regime.hip, make_fastdiv / fdiv), restated in integer
arithmetic: for every divisor and dividend below 2^31 it must equal floor
division. Checked on the divisors of every padded extent up to 4,096 and
their plane products, with dividends at the edges and random between."""
import numpy as np


def make_fastdiv(d):
    if d < 2:
        return d, 0, 0
    l = 0
    while (1 << l) < d:
        l += 1
    m = ((1 << (31 + l)) + d - 1) // d
    assert m < (1 << 32)
    return d, m, l - 1


def fdiv(n, f):
    d, m, s = f
    if d < 2:
        return n
    return ((n * m) >> 32) >> s


def test_fastdiv_is_floor_division():
    rng = np.random.default_rng(0)
    divisors = list(range(1, 4097)) + [n * n for n in range(2, 1300, 7)] + [(1 << k) for k in range(1, 31)] + \
               [(1 << k) + 1 for k in range(1, 30)] + [(1 << k) - 1 for k in range(2, 31)]
    top = (1 << 31) - 1
    for d in divisors:
        f = make_fastdiv(d)
        ns = [0, 1, d - 1, d, d + 1, top, top - 1, (top // d) * d, (top // d) * d - 1]
        ns += [int(x) for x in rng.integers(0, top, size=64)]
        for n in ns:
            if 0 <= n <= top:
                assert fdiv(n, f) == n // d, (n, d)
