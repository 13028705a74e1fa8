"""The numpy restatement of the device reductions' summation order against
exact sums: integer-valued data sums exactly, and random data agrees with
math.fsum to rounding."""
import math

import numpy as np

from fixed_order import block_sum, node_sum


def test_integer_data_is_exact():
    rng = np.random.default_rng(0)
    for n in (0, 1, 255, 256, 257, 70000, 300001):
        x = rng.integers(-1000, 1000, n).astype(np.float64)
        assert node_sum(x) == float(x.sum())
        assert block_sum(x[:5000]) == float(x[:5000].sum())


def test_random_data_close_to_fsum():
    rng = np.random.default_rng(1)
    for n in (3, 1000, 262144, 400000):
        x = 1.0 + 0.01 * rng.standard_normal(n)
        assert abs(node_sum(x) - math.fsum(x)) <= 1e-13 * n


def test_order_is_the_documented_one():
    # 0.5 ulp of 2^53 is lost or kept depending on the order of the adds
    x = np.zeros(600)
    x[0], x[256], x[512] = 2.0 ** 53, 1.0, 1.0  # one thread: (2^53 + 1) + 1 -> 2^53
    assert block_sum(x) == 2.0 ** 53
    y = np.zeros(600)
    y[0], y[1], y[3] = 2.0 ** 53, 1.0, 1.0  # threads 1 and 3 meet first in the tree: 2^53 + (1 + 1)
    assert block_sum(y) == 2.0 ** 53 + 2.0
