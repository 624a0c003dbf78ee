"""The water-level rate sum: the host evaluates Σλ(ν) over n lineages of one
rate as the sequential fold x + x + ... + x (monolithic/mod.rs:139-164), and
the library computes that fold binade by binade (nh_math.h repeated_sum).
Pinned bit for bit against the fold itself, numpy's cumulative sum (a left
fold in float64, round to nearest even), at many n for each x: rates of the
workloads, random rates, rates that are half-integer multiples of a binade's
ulp (the ties), rates below half an ulp of the sum (the sum stops growing)
and subnormal rates.  CPU (the host side of the library)."""
import ctypes

import numpy as np
import pytest

from necsim_hip import _ffi


def fast(x, n):
    out = ctypes.c_double(0.0)
    _ffi.check(_ffi.lib().nh_debug_repeated_sum(x, n, ctypes.byref(out)))
    return out.value


def rates():
    rng = np.random.default_rng(7)
    xs = [0.5 * 1e-6, 0.5 * 1e-3, 0.5, 1.0, 0.1, 1 / 3, 7.0, 1e-300, 5e-324, 3e-320]
    xs += list(2.0 ** rng.uniform(-30, 5, 40))
    for _ in range(40):  # (m + 1/2) ulp of the binade [2^k, 2^(k+1))
        k, m = int(rng.integers(-20, 20)), int(rng.integers(0, 1000))
        xs.append(float(np.ldexp(m + 0.5, k - 52)))
    return xs


@pytest.mark.parametrize("x", rates())
def test_rate_sum_equals_the_fold(x):
    n = 2_000_000 if x > 1e-200 else 200_000
    fold = np.cumsum(np.full(n, x, dtype=np.float64))
    rng = np.random.default_rng(int(np.float64(x).view(np.uint64) % 2**32))
    ns = sorted(set([1, 2, 3, 4, 5, 63, 64, 65, n] + list(rng.integers(1, n + 1, 200))))
    for k in ns:
        assert np.float64(fast(x, k)).view(np.uint64) == fold[k - 1].view(np.uint64), (x, k)
    assert fast(x, 0) == 0.0


def test_rate_sum_of_the_config4_round():
    # 12,566,345 lineages at lambda nu = 0.5e-6 (config 4), against the fold
    n = 12_566_345
    fold = np.cumsum(np.full(n, 0.5e-6, dtype=np.float64))[-1]
    assert np.float64(fast(0.5e-6, n)).view(np.uint64) == fold.view(np.uint64)
