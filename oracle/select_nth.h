// Oracle (test infrastructure only): Rust core's slice::select_nth_unstable_by,
// restated. The light-tree build calls it to split its emitters around the
// bucket split (crates/trace/src/tree.rs:281-283); the order it leaves on
// each side becomes the order of the tree's leaves and so of the light rows
// the pick indices refer to.
//
// Dependency: the Rust standard library, rust-version 1.97
// (Cargo.toml:7), not vendored under the reference. Restated from its
// published source, library/core/src/slice/:
//   select.rs                    partition_at_index, partition_at_index_loop,
//                                min_index, max_index, median_of_medians,
//                                median_of_ninthers, ninther, median_idx
//   sort/shared/pivot.rs         choose_pivot, median3_rec, median3
//   sort/unstable/quicksort.rs   partition, partition_lomuto_branchless_cyclic
//   sort/shared/smallsort.rs     insertion_sort_shift_left, insert_tail
// The element the tree sorts, `Emitter` (tree.rs:134-139), is 72 bytes in
// Rust (usize + 3 f32 + a 52-byte Measure), at most 96, so `partition`
// instantiates the branchless cyclic Lomuto partition.
//
// What pins this file without a Rust toolchain: the selection property and
// the permutation on many random inputs with heavy ties, the cases that
// reduce to textbook routines (a slice of at most 16 is insertion-sorted,
// i.e. stably sorted; index 0 / len-1 swap in the first minimum / first
// maximum), and the cyclic partition against the textbook swap form of
// Lomuto's partition run on the input rotated by one (x1 .. x(n-1), x0),
// whose result it equals. The choice of pivots and the loop over rounds of
// a slice longer than 16 follow the published algorithm but are compared
// with no Rust output: parity unpinned there.
#pragma once
#include <stddef.h>

#include <utility>

namespace oracle {
namespace rsel {

constexpr size_t INSERTION_SORT_THRESHOLD = 16;    // select.rs
constexpr size_t PSEUDO_MEDIAN_REC_THRESHOLD = 64;  // pivot.rs

// smallsort.rs insert_tail: sorts [begin, tail] given [begin, tail) sorted
template <class T, class L>
void insert_tail(T* begin, T* tail, L& is_less) {
    T* sift = tail - 1;
    if (!is_less(*tail, *sift)) return;
    T tmp = *tail;
    T* gap = tail;
    for (;;) {
        *gap = *sift;
        gap = sift;
        if (sift == begin) break;
        sift = sift - 1;
        if (!is_less(tmp, *sift)) break;
    }
    *gap = tmp;
}
// smallsort.rs insertion_sort_shift_left
template <class T, class L>
void insertion_sort_shift_left(T* v, size_t len, size_t offset, L& is_less) {
    for (T* tail = v + offset; tail != v + len; ++tail) insert_tail(v, tail, is_less);
}

// pivot.rs median3: the median of three, a when it lies between b and c
template <class T, class L>
const T* median3(const T* a, const T* b, const T* c, L& is_less) {
    bool x = is_less(*a, *b);
    bool y = is_less(*a, *c);
    if (x == y) {
        bool z = is_less(*b, *c);
        return (z ^ x) ? c : b;
    }
    return a;
}
// pivot.rs median3_rec: Tukey's ninther, applied recursively
template <class T, class L>
const T* median3_rec(const T* a, const T* b, const T* c, size_t n, L& is_less) {
    if (n * 8 >= PSEUDO_MEDIAN_REC_THRESHOLD) {
        size_t n8 = n / 8;
        a = median3_rec(a, a + n8 * 4, a + n8 * 7, n8, is_less);
        b = median3_rec(b, b + n8 * 4, b + n8 * 7, n8, is_less);
        c = median3_rec(c, c + n8 * 4, c + n8 * 7, n8, is_less);
    }
    return median3(a, b, c, is_less);
}
// pivot.rs choose_pivot (len >= 8; the selection calls it for len > 16)
template <class T, class L>
size_t choose_pivot(const T* v, size_t len, L& is_less) {
    size_t len_div_8 = len / 8;
    const T* a = v;
    const T* b = v + len_div_8 * 4;
    const T* c = v + len_div_8 * 7;
    if (len < PSEUDO_MEDIAN_REC_THRESHOLD) return (size_t)(median3(a, b, c, is_less) - v);
    return (size_t)(median3_rec(a, b, c, len_div_8, is_less) - v);
}

// quicksort.rs partition_lomuto_branchless_cyclic: the first element is
// lifted out, leaving a gap; every later element, then the lifted one, is
// visited in turn: the element at num_lt moves into the gap, the visited
// element moves to num_lt, its old place becomes the gap, and num_lt grows
// by one if it is less than the pivot. (The result equals the swap form of
// Lomuto's partition over x1 .. x(n-1), x0.)
template <class T, class L>
size_t lomuto_cyclic(T* v, size_t len, const T& pivot, L& is_less) {
    if (len == 0) return 0;
    T lifted = v[0];
    T* gap = v;
    size_t num_lt = 0;
    for (size_t r = 1; r <= len; ++r) {
        T* right = (r == len) ? &lifted : v + r;
        bool right_is_lt = is_less(*right, pivot);
        T* left = v + num_lt;
        *gap = *left;
        *left = *right;
        gap = right;
        num_lt += right_is_lt ? 1 : 0;
    }
    return num_lt;
}
// quicksort.rs partition: the pivot goes to the front, the rest is
// partitioned, and the pivot is swapped to the boundary
template <class T, class L>
size_t partition(T* v, size_t len, size_t pivot, L& is_less) {
    if (len == 0) return 0;
    std::swap(v[0], v[pivot]);
    size_t num_lt = lomuto_cyclic(v + 1, len - 1, v[0], is_less);
    std::swap(v[0], v[num_lt]);
    return num_lt;
}

// select.rs min_index / max_index: reduce keeps the accumulator unless the
// next element strictly beats it, so the first of equal minima and the
// first of equal maxima
template <class T, class L>
size_t min_index(const T* v, size_t len, L& is_less) {
    size_t best = 0;
    for (size_t i = 1; i < len; ++i)
        if (is_less(v[i], v[best])) best = i;
    return best;
}
template <class T, class L>
size_t max_index(const T* v, size_t len, L& is_less) {
    size_t best = 0;
    for (size_t i = 1; i < len; ++i)
        if (is_less(v[best], v[i])) best = i;
    return best;
}

// select.rs median_idx
template <class T, class L>
size_t median_idx(const T* v, L& is_less, size_t a, size_t b, size_t c) {
    if (is_less(v[c], v[a])) std::swap(a, c);
    if (is_less(v[c], v[b])) return c;
    if (is_less(v[b], v[a])) return a;
    return b;
}
// select.rs ninther
template <class T, class L>
void ninther(T* v, L& is_less, size_t a, size_t b, size_t c, size_t d, size_t e, size_t f, size_t g, size_t h,
             size_t i) {
    b = median_idx(v, is_less, a, b, c);
    h = median_idx(v, is_less, g, h, i);
    if (is_less(v[h], v[b])) std::swap(b, h);
    if (is_less(v[f], v[d])) std::swap(d, f);
    if (is_less(v[e], v[d])) {
        // d stays
    } else if (is_less(v[f], v[e])) {
        d = f;
    } else {
        if (is_less(v[e], v[b]))
            std::swap(v[e], v[b]);
        else if (is_less(v[h], v[e]))
            std::swap(v[e], v[h]);
        return;
    }
    if (is_less(v[d], v[b]))
        d = b;
    else if (is_less(v[h], v[d]))
        d = h;
    std::swap(v[d], v[e]);
}
template <class T, class L>
void median_of_medians(T* v, size_t len, size_t k, L& is_less);
// select.rs median_of_ninthers
template <class T, class L>
size_t median_of_ninthers(T* v, size_t len, L& is_less) {
    size_t frac = len <= 1024 ? len / 12 : (len <= 128 * 1024 ? len / 64 : len / 1024);
    size_t pivot = frac / 2;
    size_t lo = len / 2 - pivot;
    size_t hi = frac + lo;
    size_t gap = (len - 9 * frac) / 4;
    size_t a = lo - 4 * frac - gap;
    size_t b = hi + gap;
    for (size_t i = lo; i < hi; ++i) {
        ninther(v, is_less, a, i - frac, b, a + 1, i, b + 1, a + 2, i + frac, b + 2);
        a += 3;
        b += 3;
    }
    median_of_medians(v + lo, frac, pivot, is_less);
    return partition(v, len, lo + pivot, is_less);
}
// select.rs median_of_medians: the deterministic fallback after 16 rounds
template <class T, class L>
void median_of_medians(T* v, size_t len, size_t k, L& is_less) {
    for (;;) {
        if (len <= INSERTION_SORT_THRESHOLD) {
            if (len >= 2) insertion_sort_shift_left(v, len, 1, is_less);
            return;
        }
        if (k == len - 1) {
            std::swap(v[max_index(v, len, is_less)], v[k]);
            return;
        } else if (k == 0) {
            std::swap(v[min_index(v, len, is_less)], v[k]);
            return;
        }
        size_t p = median_of_ninthers(v, len, is_less);
        if (p == k) {
            return;
        } else if (p > k) {
            len = p;
        } else {
            v += p + 1;
            len -= p + 1;
            k -= p + 1;
        }
    }
}

// select.rs partition_at_index_loop: introselect on pseudo-median pivots;
// a pivot not greater than the previous round's pivot (the smallest value
// left) splits off the run equal to it instead
template <class T, class L>
void partition_at_index_loop(T* v, size_t len, size_t index, L& is_less) {
    int limit = 16;
    bool has_ancestor = false;
    T ancestor{};
    for (;;) {
        if (len <= INSERTION_SORT_THRESHOLD) {
            if (len >= 2) insertion_sort_shift_left(v, len, 1, is_less);
            return;
        }
        if (limit == 0) {
            median_of_medians(v, len, index, is_less);
            return;
        }
        limit -= 1;
        size_t pivot_pos = choose_pivot(v, len, is_less);
        if (has_ancestor && !is_less(ancestor, v[pivot_pos])) {
            auto less_equal = [&is_less](const T& a, const T& b) { return !is_less(b, a); };
            size_t num_lt = partition(v, len, pivot_pos, less_equal);
            size_t mid = num_lt + 1;
            if (mid > index) return;
            v += mid;
            len -= mid;
            index -= mid;
            has_ancestor = false;
            continue;
        }
        size_t mid = partition(v, len, pivot_pos, is_less);
        if (mid < index) {
            ancestor = v[mid];
            has_ancestor = true;
            v += mid + 1;
            len -= mid + 1;
            index -= mid + 1;
        } else if (mid > index) {
            len = mid;
        } else {
            return;
        }
    }
}

// select.rs partition_at_index (slice::select_nth_unstable_by)
template <class T, class L>
void select_nth_unstable_by(T* v, size_t len, size_t index, L is_less) {
    if (index >= len) return;  // Rust panics; the tree never asks for it
    if (index == len - 1) {
        std::swap(v[max_index(v, len, is_less)], v[index]);
    } else if (index == 0) {
        std::swap(v[min_index(v, len, is_less)], v[index]);
    } else {
        partition_at_index_loop(v, len, index, is_less);
    }
}

}  // namespace rsel
}  // namespace oracle
