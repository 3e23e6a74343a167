"""dsm_intersect_sorted_noalloc (no device allocation; one fixed array per
GPU) against dsm_intersect_sorted on the same inputs, two calls in a row, the
second smaller than the first (its results must not depend on what the
first left in the array), plus the size limit."""
import numpy as np
import pytest

from dsm_amd import engine as E

pytestmark = pytest.mark.gpu


def rand_set(rng, n, hi):
    return np.unique(rng.integers(0, hi, n, dtype=np.int64)).astype(np.uint32)


def test_noalloc_equals_allocating_entry_point():
    print("case: dsm_intersect_sorted_noalloc vs dsm_intersect_sorted, large then small", flush=True)
    rng = np.random.default_rng(8)
    calls = [(rand_set(rng, 600000, 2000000), rand_set(rng, 900000, 2000000)),   # ~3.6M scratch words
             (rand_set(rng, 3000, 20000), rand_set(rng, 5000, 20000))]
    for i, (a, b) in enumerate(calls):
        want = E.intersect_sorted(a, b)
        got = E.intersect_sorted(a, b, noalloc=True)
        assert np.array_equal(got, want), "call %d" % i
        assert np.array_equal(got, np.intersect1d(a, b)), "call %d" % i
        print("  call %d: %d x %d -> %d" % (i, len(a), len(b), len(got)), flush=True)


def test_noalloc_rejects_oversize_input():
    a = np.arange(0, 2 * 3000000, 2, dtype=np.uint32)
    b = np.arange(0, 3000000, dtype=np.uint32)
    with pytest.raises(E.DsmError) as ei:
        E.intersect_sorted(a, b, noalloc=True)
    assert ei.value.code == E.DSM_ELIMIT
