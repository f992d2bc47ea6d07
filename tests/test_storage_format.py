"""The E3M12 16-bit population code (storage "e3m12s"), NumPy model in
tests/storage_model.py, pinned by the properties that define the format:
exact block values and spacings, monotone decode, a round trip of every
code, rounding to the nearest code with ties to even, saturation."""

import numpy as np

from tests import storage_model as SM


def test_decode_blocks():
    m = SM.magnitudes()
    assert m[0] == 0.0 and m[4096] == 2.0 ** -10 and m[8192] == 2.0 ** -8 - 2.0 ** -10
    assert m[-1] == 2.0 ** -2 - 2.0 ** -10 - 2.0 ** -15
    d = np.diff(m)
    blk = np.arange(d.size) >> 12
    assert np.all(d == 2.0 ** (np.maximum(blk, 1) - 22 - (blk == 0)))   # 2^-22 in block 0, 2^(k-22) in block k
    neg = SM.decode(np.arange(0x8000, 0x10000, dtype=np.uint32).astype(np.uint16)).astype(np.float64)
    assert np.all(neg == -m) and np.signbit(SM.decode(np.array([0x8000], np.uint16))[0])


def test_round_trip_every_code():
    c = np.arange(0x10000, dtype=np.uint32).astype(np.uint16)
    assert np.array_equal(SM.encode(SM.decode(c)), c)


def test_rounds_to_nearest_code():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-0.249, 0.249, 200000), rng.uniform(-2e-3, 2e-3, 100000),
                        rng.standard_normal(50000) * 1e-6]).astype(np.float32)
    m = SM.magnitudes()
    got = np.abs(SM.decode(SM.encode(x)).astype(np.float64))
    a = np.abs(x.astype(np.float64))
    j = np.clip(np.searchsorted(m, a), 1, m.size - 1)
    best = np.minimum(np.abs(m[j] - a), np.abs(m[j - 1] - a))
    slack = 2.0 ** -23 * (a + 2.0 ** -10)   # the float32 sum |x| + 2^-10 is rounded first (<= 1/2 ulp)
    assert np.all(np.abs(got - a) <= best + slack)
    big = a >= 2.0 ** -10
    assert np.max(np.abs(got - a)[big] / a[big]) <= 2.0 ** -12   # 12+ significant bits over the range
    assert np.max(np.abs(got - a)[~big]) <= 2.0 ** -23 + 2.0 ** -33


def test_ties_to_even_and_saturation():
    lo = 2.0 ** -10
    x = np.array([lo + 2.0 ** -22, lo + 3 * 2.0 ** -22, 0.3, -1.0, np.inf, np.nan, -0.0], dtype=np.float32)
    assert SM.encode(x).tolist() == [4096, 4098, 0x7FFF, 0xFFFF, 0x7FFF, 0x7FFF, 0x8000]
