"""The DFS's membership test as a set intersection (dsm_intersect_sorted), on
the device, against numpy's intersect1d (a sorted-set intersection, the
reference's intersect, gpu_match.cu:21-94). Cases: windows the wave stages in
LDS (at most 256 entries) and wider ones searched in place, empty and disjoint
sets, identical sets, a hub-sized second list, sizes that are not multiples of
64, and the largest ids."""
import numpy as np
import pytest

from dsm_amd import engine as E

pytestmark = pytest.mark.gpu


def rand_set(rng, n, hi):
    return np.unique(rng.integers(0, hi, n, dtype=np.int64)).astype(np.uint32)


CASES = [
    ("empty_a", 0, 100, 1000),
    ("empty_b", 100, 0, 1000),
    ("tiny", 3, 5, 10),
    ("dense_small", 1000, 1000, 1500),       # windows of a few entries
    ("sparse_a_dense_b", 200, 200000, 250000),  # windows of ~1,250 entries: searched in place
    ("dense_a_sparse_b", 200000, 200, 250000),
    ("hub", 5000, 1000000, 4000000),
    ("odd_sizes", 64 * 37 + 5, 64 * 11 + 63, 30000),
    ("top_ids", 5000, 5000, 0xFFFFFFFE),
]


@pytest.mark.parametrize("name,na,nb,hi", CASES, ids=[c[0] for c in CASES])
def test_intersection_equals_numpy(name, na, nb, hi):
    print("case: intersect %s (%d x %d)" % (name, na, nb), flush=True)
    rng = np.random.default_rng(hash(name) % 10000)
    a = rand_set(rng, na, hi)
    b = rand_set(rng, nb, hi)
    if name == "top_ids":
        a[-1] = b[-1] = 0xFFFFFFFE
        a, b = np.unique(a), np.unique(b)
    got = E.intersect_sorted(a, b)
    want = np.intersect1d(a, b)
    assert np.array_equal(got, want)


def test_identical_and_disjoint_sets():
    a = np.arange(0, 300000, 3, dtype=np.uint32)
    assert np.array_equal(E.intersect_sorted(a, a), a)
    b = a + 1
    assert E.intersect_sorted(a, b).size == 0
