"""The light-tree split, Rust core's slice::select_nth_unstable_by (Rust
1.97; the reference calls it at crates/trace/src/tree.rs:281-283 and it is
not vendored there): the oracle's restatement pinned by what the
mathematics fixes, and the product's selection equal to it byte for byte.

Pins: the result is a permutation with the selection property under
f32::total_cmp; a slice of at most 16 ends insertion-sorted, which is a
stable sort; index 0 and len-1 swap in the first minimum / first maximum;
one cyclic Lomuto partition equals the textbook swap form of Lomuto's
partition over the input rotated by one. The pivot choice and round loop
for longer slices follow the published source with no Rust output to
compare against (parity unpinned there)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from pfx_amd import _native as N


def total_key(keys):
    """f32::total_cmp as a sortable integer."""
    b = np.asarray(keys, np.float32).view(np.int32).astype(np.int64)
    return np.where(b < 0, b ^ 0x7FFFFFFF, b)


def oracle_select(keys, index, mode=0):
    k = np.array(keys, np.float32)
    ids = np.arange(len(k), dtype=np.uint32)
    r = O.lib().or_select_nth(k.ctypes.data, ids.ctypes.data, len(k), index, mode)
    return k, ids, r


def product_select(keys, index, mode=0):
    k = np.array(keys, np.float32)
    ids = np.arange(len(k), dtype=np.uint32)
    out = C.c_size_t()
    st = N.lib().pfxh_layout_select(k.ctypes.data, ids.ctypes.data, len(k), index, mode, C.byref(out))
    assert st == 0
    return k, ids, out.value


def tie_heavy(rng, n):
    """Keys with many repeats, both zeros and a few negatives."""
    pool = np.array([-1.5, -0.0, 0.0, 0.25, 0.5, 1.0, 2.0, 3.0], np.float32)
    if rng.random() < 0.5:
        return rng.choice(pool, n)
    return rng.standard_normal(n).astype(np.float32)


def check_selection(keys, out_keys, ids, index):
    n = len(keys)
    assert sorted(ids.tolist()) == list(range(n))
    assert np.array_equal(np.asarray(keys, np.float32)[ids].view(np.uint32), out_keys.view(np.uint32))
    t = total_key(out_keys)
    assert t[index] == np.sort(total_key(keys))[index]
    assert np.all(t[:index] <= t[index]) and np.all(t[index + 1:] >= t[index])


def test_selection_property_on_random_inputs():
    rng = np.random.default_rng(11)
    for n in list(range(1, 70)) + [100, 128, 255, 256, 1000, 4097]:
        for _ in range(6 if n < 300 else 2):
            keys = tie_heavy(rng, n)
            index = int(rng.integers(0, n))
            k, ids, _ = oracle_select(keys, index)
            check_selection(keys, k, ids, index)


def test_short_slices_end_stably_sorted():
    """A slice of 16 or fewer (index neither end) is insertion-sorted."""
    rng = np.random.default_rng(12)
    for n in range(3, 17):
        for _ in range(20):
            keys = tie_heavy(rng, n)
            index = int(rng.integers(1, n - 1))
            k, ids, _ = oracle_select(keys, index)
            assert ids.tolist() == np.argsort(total_key(keys), kind="stable").tolist()


def test_end_indices_swap_in_the_first_extreme():
    rng = np.random.default_rng(13)
    for n in (1, 2, 5, 17, 64, 300):
        for _ in range(10):
            keys = tie_heavy(rng, n)
            t = total_key(keys)
            for index, pick in ((0, int(np.argmin(t))), (n - 1, int(np.argmax(t)))):
                expect = list(range(n))
                expect[index], expect[pick] = expect[pick], expect[index]
                _, ids, _ = oracle_select(keys, index)
                assert ids.tolist() == expect


def swap_lomuto(keys, pivot_pos):
    """quicksort.rs partition with the textbook swap form of Lomuto's pass
    over the rest rotated left by one: x1 .. x(n-1), x0."""
    ids = list(range(len(keys)))
    t = total_key(keys)
    ids[0], ids[pivot_pos] = ids[pivot_pos], ids[0]
    rest = ids[2:] + ids[1:2]
    lt = 0
    for r in range(len(rest)):
        smaller = t[rest[r]] < t[ids[0]]
        rest[lt], rest[r] = rest[r], rest[lt]
        lt += int(smaller)
    ids = ids[:1] + rest
    ids[0], ids[lt] = ids[lt], ids[0]
    return ids, lt


def test_partition_equals_swap_lomuto_on_rotated_input():
    rng = np.random.default_rng(14)
    for n in list(range(1, 40)) + [128, 500]:
        for _ in range(8):
            keys = tie_heavy(rng, n)
            p = int(rng.integers(0, n))
            _, ids, lt = oracle_select(keys, p, mode=1)
            expect, expect_lt = swap_lomuto(keys, p)
            assert lt == expect_lt and ids.tolist() == expect


def test_median_of_medians_fallback_selects():
    rng = np.random.default_rng(15)
    for n in (17, 40, 100, 1025, 3000):
        for _ in range(4):
            keys = tie_heavy(rng, n)
            index = int(rng.integers(0, n))
            k, ids, _ = oracle_select(keys, index, mode=2)
            check_selection(keys, k, ids, index)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_product_selection_is_byte_identical(mode):
    rng = np.random.default_rng(20 + mode)
    for n in list(range(1, 80)) + [128, 129, 777, 2048, 5000]:
        for _ in range(4):
            keys = tie_heavy(rng, n)
            index = int(rng.integers(0, n))
            a, b = product_select(keys, index, mode), oracle_select(keys, index, mode)
            assert a[2] == b[2] and np.array_equal(a[1], b[1]) and a[0].tobytes() == b[0].tobytes()


def test_product_selection_rejects_bad_arguments():
    k = np.zeros(4, np.float32)
    ids = np.zeros(4, np.uint32)
    out = C.c_size_t()
    lib = N.lib()
    assert lib.pfxh_layout_select(k.ctypes.data, ids.ctypes.data, 4, 4, 0, C.byref(out)) != 0
    assert lib.pfxh_layout_select(k.ctypes.data, ids.ctypes.data, 4, 0, 3, C.byref(out)) != 0
    assert lib.pfxh_layout_select(None, None, 4, 0, 0, C.byref(out)) != 0
