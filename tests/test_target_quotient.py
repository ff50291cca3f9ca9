"""The resampling target u + i/N is formed on the GPU without a float
division: the quotient float(i)/float(N) is taken as
float(double(float(i)) * (1 / double(float(N)))).  For 24-bit operands the
exact quotient lies more than 2^-50 (relative) from every midpoint between
adjacent floats and the double product is within 2^-52 of it, so the two
round to the same float.  These tests check that identity bit for bit
against IEEE float32 division (numpy's), over every i for sizes up to 2^26
and over sampled i for the largest sizes, and that a reciprocal rounded to
float32 instead of float64 breaks it (so the check can fail)."""
import numpy as np
import pytest


def _ieee(i, n):
    return (i.astype(np.float32) / np.float32(n)).view(np.uint32)


def _fast(i, n, recip_dtype=np.float64):
    r = np.float64(1.0) / np.float64(np.float32(n))
    if recip_dtype is np.float32:
        r = np.float64(np.float32(r))
    return (i.astype(np.float32).astype(np.float64) * r).astype(np.float32).view(np.uint32)


def _all_i(n, chunk=1 << 23):
    for lo in range(0, n, chunk):
        yield np.arange(lo, min(n, lo + chunk), dtype=np.int64)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 1000, 4099, 65536, 65537, 1 << 20, 3_000_007,
                               (1 << 24) + 1, 51_200_000, (1 << 26) - 5])
def test_quotient_identity_every_index(n):
    for i in _all_i(n):
        assert np.array_equal(_ieee(i, n), _fast(i, n)), n


@pytest.mark.parametrize("n", [(1 << 28) + 3, 1 << 30, (1 << 31) - 1, 1_234_567_891])
def test_quotient_identity_sampled_indices(n):
    rng = np.random.default_rng(n & 0xFFFF)
    i = np.concatenate([rng.integers(0, n, 1 << 22), np.arange(0, 4096), np.arange(n - 4096, n),
                        np.arange((1 << 24) - 64, (1 << 24) + 64)]).astype(np.int64)
    i = i[(i >= 0) & (i < n)]
    assert np.array_equal(_ieee(i, n), _fast(i, n))


def test_float32_reciprocal_breaks_the_identity():
    bad = 0
    for n in (1000, 65537, 3_000_007):
        i = np.arange(n, dtype=np.int64)
        bad += int(np.count_nonzero(_ieee(i, n) != _fast(i, n, np.float32)))
    assert bad > 0
