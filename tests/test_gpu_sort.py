"""The device radix sort behind the admission and contraction orders
(gp_debug_radix_sort) against numpy's stable sort of the same key bits:
sizes on both sides of the one-block variants (4, 8 and 16 waves of 512
keys) and of the multi-block path, bit ranges with many equal keys (so
stability decides the order) and a range whose digits are constant.
The passes take whole 8-bit digits from the first bit, so the keys are
zero above the last sorted bit (as the callers' keys are); bits below the
first sorted bit are random and must not affect the order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 100, 2047, 2048, 2049, 4096, 4097, 8191, 8192, 8193, 100000]
RANGES = [(0, 49), (0, 13), (8, 40), (0, 64)]


def _expect(keys, vals, lo, hi):
    mask = (1 << (hi - lo)) - 1
    field = (keys >> np.uint64(lo)) & np.uint64(mask)
    order = np.argsort(field, kind="stable")
    return keys[order], vals[order]


@pytest.mark.parametrize("n", SIZES)
def test_sort_matches_stable_numpy(gp, ctx, n):
    rng = np.random.default_rng(n)
    for lo, hi in RANGES:
        print("sort n=%d bits [%d, %d)" % (n, lo, hi), flush=True)
        keys = rng.integers(0, 1 << 62, n, dtype=np.uint64)
        # About n / 8 distinct values in the sorted bits: long runs of
        # equal keys.
        keys ^= (keys >> np.uint64(lo)) << np.uint64(lo)
        keys |= (rng.integers(0, max(1, n // 8), n).astype(np.uint64) * np.uint64(0x9E3779B1)) << np.uint64(lo)
        if hi < 64:
            keys &= np.uint64((1 << hi) - 1)
        vals = np.arange(n, dtype=np.uint32)
        got_k, got_v = gp.radix_sort(ctx, keys, vals, lo, hi)
        want_k, want_v = _expect(keys, vals, lo, hi)
        assert np.array_equal(got_v, want_v), (n, lo, hi)
        assert np.array_equal(got_k, want_k), (n, lo, hi)
    # Keys only, and a range whose digits are equal for every key (passes
    # skipped): the order stays the input order.
    keys = rng.integers(0, 1 << 20, n, dtype=np.uint64)
    got_k, _ = gp.radix_sort(ctx, keys, None, 0, 32)
    assert np.array_equal(got_k, np.sort(keys, kind="stable"))
    const = (keys & np.uint64(0xFFFFF)) | np.uint64(5 << 40)
    vals = np.arange(n, dtype=np.uint32)
    got_k, got_v = gp.radix_sort(ctx, const, vals, 32, 48)
    assert np.array_equal(got_v, vals) and np.array_equal(got_k, const)
