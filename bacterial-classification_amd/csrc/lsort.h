// Sort of (score, candidate) pairs by score, descending, that leaves equal
// scores in the same relative order as GCC's std::sort does with the
// comparator `a.score > b.score` (targets/identify.cpp:183-184).
//
// The beam in identify keeps the first BEAM_WIDTH entries after std::sort,
// so when two candidates have bit-equal float scores, which of them survives
// (and in which order the survivors are expanded next) is decided by the
// library's sort, not by the comparator.  std::sort is not stable; its
// result for equal keys is a property of its algorithm: introsort (depth
// limit 2*floor(log2 n)) with median-of-three pivots moved to the front,
// an unguarded Hoare partition, heapsort when the depth limit is reached,
// and a final insertion sort (plain below 17 elements, guarded on the first
// 16 and unguarded after).  This is a restatement of that algorithm as
// host/device code over two parallel arrays; for n <= 16 it is an insertion
// sort and therefore stable.  It is checked against the host's std::sort on
// random, tie-heavy and adversarial inputs.
#pragma once

#if defined(__HIPCC__)
#define HVK_HD __host__ __device__ __forceinline__
#else
#define HVK_HD inline
#endif

namespace hvk {
namespace lsort {

struct Pairs {
    float* s;
    int* c;
    int st;  // element a lives at s[a * st], c[a * st]
    HVK_HD float& S(int a) const { return s[a * st]; }
    HVK_HD int& C(int a) const { return c[a * st]; }
    HVK_HD bool less(int a, int b) const { return S(a) > S(b); }   // comp(*a, *b)
    HVK_HD bool less_v(float v, int b) const { return v > S(b); }  // comp(val, *b)
    HVK_HD bool less_to_v(int a, float v) const { return S(a) > v; }  // comp(*a, val)
    HVK_HD void swap(int a, int b) const {
        const float ts = S(a); S(a) = S(b); S(b) = ts;
        const int tc = C(a); C(a) = C(b); C(b) = tc;
    }
    HVK_HD void move(int dst, int src) const { S(dst) = S(src); C(dst) = C(src); }
    HVK_HD void put(int dst, float vs, int vc) const { S(dst) = vs; C(dst) = vc; }
};

constexpr int kThreshold = 16;

HVK_HD int floor_log2(int n) {
    int k = 0;
    while (n > 1) { n >>= 1; ++k; }
    return k;
}

// insertion of element `last` into the sorted run ending before it; no lower
// bound check (an element not greater than it is known to exist before it)
HVK_HD void unguarded_linear_insert(const Pairs& p, int last) {
    const float vs = p.S(last);
    const int vc = p.C(last);
    int next = last - 1;
    while (p.less_v(vs, next)) {
        p.move(last, next);
        last = next;
        --next;
    }
    p.put(last, vs, vc);
}

HVK_HD void insertion_sort(const Pairs& p, int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i < last; ++i) {
        if (p.less(i, first)) {
            const float vs = p.S(i);
            const int vc = p.C(i);
            for (int j = i; j > first; --j) p.move(j, j - 1);
            p.put(first, vs, vc);
        } else {
            unguarded_linear_insert(p, i);
        }
    }
}

HVK_HD void final_insertion_sort(const Pairs& p, int first, int last) {
    if (last - first > kThreshold) {
        insertion_sort(p, first, first + kThreshold);
        for (int i = first + kThreshold; i < last; ++i) unguarded_linear_insert(p, i);
    } else {
        insertion_sort(p, first, last);
    }
}

// moves the median of *a, *b, *c into *result
HVK_HD void median_to_first(const Pairs& p, int result, int a, int b, int c) {
    if (p.less(a, b)) {
        if (p.less(b, c)) p.swap(result, b);
        else if (p.less(a, c)) p.swap(result, c);
        else p.swap(result, a);
    } else if (p.less(a, c)) {
        p.swap(result, a);
    } else if (p.less(b, c)) {
        p.swap(result, c);
    } else {
        p.swap(result, b);
    }
}

HVK_HD int unguarded_partition(const Pairs& p, int first, int last, int pivot) {
    while (true) {
        while (p.less(first, pivot)) ++first;
        --last;
        while (p.less(pivot, last)) --last;
        if (!(first < last)) return first;
        p.swap(first, last);
        ++first;
    }
}

// heap rooted at `first`, element i at first + i; `hole` receives (vs, vc)
HVK_HD void push_heap(const Pairs& p, int first, int hole, int top, float vs, int vc) {
    int parent = (hole - 1) / 2;
    while (hole > top && p.less_to_v(first + parent, vs)) {
        p.move(first + hole, first + parent);
        hole = parent;
        parent = (hole - 1) / 2;
    }
    p.put(first + hole, vs, vc);
}

HVK_HD void adjust_heap(const Pairs& p, int first, int hole, int len, float vs, int vc) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (p.less(first + child, first + child - 1)) --child;
        p.move(first + hole, first + child);
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        p.move(first + hole, first + child - 1);
        hole = child - 1;
    }
    push_heap(p, first, hole, top, vs, vc);
}

#ifndef HVK_LSORT_PROBE_HEAP
#define HVK_LSORT_PROBE_HEAP() ((void)0)
#endif

HVK_HD void heap_sort(const Pairs& p, int first, int last) {
    HVK_LSORT_PROBE_HEAP();  // a test build counts the calls
    const int len = last - first;
    if (len >= 2) {  // make heap
        int parent = (len - 2) / 2;
        while (true) {
            adjust_heap(p, first, parent, len, p.S(first + parent), p.C(first + parent));
            if (parent == 0) break;
            --parent;
        }
    }
    while (last - first > 1) {  // pop the top to the back
        --last;
        const float vs = p.S(last);
        const int vc = p.C(last);
        p.move(last, first);
        adjust_heap(p, first, 0, last - first, vs, vc);
    }
}

// introsort loop.  The library recurses into [cut, last) and loops on
// [first, cut) with the same remaining depth; the two parts are disjoint, so
// an explicit stack visiting them in any order gives the same array.
HVK_HD void introsort_loop(const Pairs& p, int first, int last) {
    struct Frame { int first, last, depth; };
    Frame stack[64];
    int top = 0;
    stack[top++] = {first, last, 2 * floor_log2(last - first)};
    while (top > 0) {
        Frame f = stack[--top];
        while (f.last - f.first > kThreshold) {
            if (f.depth == 0) {
                heap_sort(p, f.first, f.last);
                break;
            }
            --f.depth;
            const int mid = f.first + (f.last - f.first) / 2;
            median_to_first(p, f.first, f.first + 1, mid, f.last - 1);
            const int cut = unguarded_partition(p, f.first + 1, f.last, f.first);
            stack[top++] = {cut, f.last, f.depth};
            f.last = cut;
        }
    }
}

// std::sort(s, s + n) with comp(a, b) = a > b, carrying c along; element i
// at s[i * stride], c[i * stride]
HVK_HD void sort_desc(float* s, int* c, int n, int stride = 1) {
    if (n <= 1) return;
    const Pairs p{s, c, stride};
    introsort_loop(p, 0, n);
    final_insertion_sort(p, 0, n);
}

}  // namespace lsort
}  // namespace hvk
