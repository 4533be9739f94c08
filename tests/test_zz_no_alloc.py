"""The allocation-free sort entry point (static device scratch) gives the same permutation as
the allocating one, on two calls in a row (the second smaller, so stale scratch from the first
call would show), and refuses more keys than its scratch holds."""
import numpy as np
import pytest

import mcpet


def _keys(n, bits, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2 ** 62, n, dtype=np.uint64)
    keys &= np.uint64((1 << bits) - 1)
    keys[rng.integers(0, n, n // 4)] = keys[rng.integers(0, n, n // 4)]     # many equal keys
    return keys


@pytest.mark.gpu
def test_noalloc_sort_equals_the_allocating_sort_on_two_calls():
    for n, bits, seed in ((1 << 20, 44, 11), (70_001, 40, 12)):
        keys = _keys(n, bits, seed)
        passes = (bits + 7) // 8
        ref = mcpet.selftest_sort(keys, passes)
        got = mcpet.selftest_sort(keys, passes, noalloc=True)
        assert np.array_equal(got, ref), (n, bits)
        assert np.array_equal(got, np.argsort(keys, kind="stable").astype(np.uint32)), (n, bits)


@pytest.mark.gpu
def test_noalloc_sort_rejects_more_keys_than_its_scratch():
    keys = np.zeros((1 << 20) + 1, np.uint64)
    with pytest.raises(mcpet.McpetError) as e:
        mcpet.selftest_sort(keys, 8, noalloc=True)
    assert e.value.status == 8 and "static scratch" in str(e.value)
