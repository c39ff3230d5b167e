"""The exact multiply-shift division behind the flat cell mapping (CPU).

The step kernels find a flat cell's coordinates with q = (n m) >> sh, where
for a divisor d, l = ceil(log2 d), m = ceil(2^(31+l) / d), sh = 31 + l
(psc_common.h make_divider; Granlund and Montgomery 1994, theorem 4.2 with
N = 31).  Checked here against integer division for every n below 5000,
random n up to 2^31 - 1, and the quotient boundaries k d - 1, k d near the top
of the range, over small, power-of-two-adjacent and lattice-shaped divisors;
m must fit in 32 bits.
"""

import numpy as np


def make_divider(d):
    l = 0
    while (1 << l) < d:
        l += 1
    return ((1 << (31 + l)) + d - 1) // d, 31 + l


def test_multiply_shift_division_is_exact_below_2_31():
    rng = np.random.default_rng(0)
    divisors = (list(range(1, 260)) + [511, 512, 513, 1000, 1023, 1024, 1025, 137, 137 * 137, 136 * 136,
                                       512 * 512, 3 ** 13, 2 ** 20 - 1, 2 ** 20 + 1, 2 ** 30 - 1, 2 ** 30]
                + [int(v) for v in rng.integers(2, 2 ** 30, 40)])
    top = 2 ** 31 - 1
    for d in divisors:
        m, sh = make_divider(d)
        assert 0 < m < 2 ** 32 and sh <= 62
        k = top // d
        edges = [k * d - 1, k * d, min(k * d + 1, top), top, d - 1, d, 2 * d - 1, 2 * d]
        n = np.concatenate([np.arange(5000, dtype=np.uint64), rng.integers(0, top + 1, 20000).astype(np.uint64),
                            np.array([v for v in edges if 0 <= v <= top], dtype=np.uint64)])
        q = (n * np.uint64(m)) >> np.uint64(sh)   # n m < 2^63: no wrap in 64 bits
        assert np.array_equal(q, n // np.uint64(d)), d


def test_a_plausible_mistake_fails():
    # floor instead of ceil for m under-divides at multiples of d
    d = 137 * 137
    l = (d - 1).bit_length()
    m_floor = (1 << (31 + l)) // d
    n = np.arange(0, 2 ** 31, d, dtype=np.uint64)[1:2000]
    assert not np.array_equal((n * np.uint64(m_floor)) >> np.uint64(31 + l), n // np.uint64(d))
