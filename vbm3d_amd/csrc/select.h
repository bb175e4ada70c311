// select.h -- partial selection with std::partial_sort's exact result.
//
// The reference ranks candidates with std::partial_sort and a comparator
// on the distance only (src/vbm3d.cpp:1076-1077, 1136-1137, 1171-1172;
// comparator src/Utilities/Utilities.cpp:40-45).  With distinct
// distances the result is simply the `mid` smallest in ascending order;
// with equal distances it depends on the heap algorithm libstdc++ uses
// (heap select over the first `mid` elements, then heap sort).  This
// header implements that binary-heap algorithm so the kernels reproduce
// the reference's order on ties too.  It is compiled for host and device;
// vbm3d_selftest_partial_sort exposes the host copy to the tests.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define VB_HD __host__ __device__
#else
#define VB_HD
#endif

struct VbEnt {
    float d;
    uint32_t i;
};

VB_HD inline bool vb_less(const VbEnt& a, const VbEnt& b) { return a.d < b.d; }

// Move `v` up from `hole` while its parent compares less than it.
VB_HD inline void vb_sift_up(VbEnt* h, int hole, int top, VbEnt v) {
    int parent = (hole - 1) / 2;
    while (hole > top && vb_less(h[parent], v)) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = v;
}

// Floyd's bottom-up re-heapify: walk the hole to a leaf along the larger
// child (the left one on equality), then sift `v` back up.
VB_HD inline void vb_adjust(VbEnt* h, int hole, int len, VbEnt v) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (vb_less(h[child], h[child - 1])) child--;
        h[hole] = h[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        h[hole] = h[child - 1];
        hole = child - 1;
    }
    vb_sift_up(h, hole, top, v);
}

VB_HD inline void vb_make_heap(VbEnt* h, int len) {
    if (len < 2) return;
    for (int parent = (len - 2) / 2;; parent--) {
        VbEnt v = h[parent];
        vb_adjust(h, parent, len, v);
        if (parent == 0) return;
    }
}

// Replace the heap top by *slot and move the old top into *slot.
VB_HD inline void vb_pop_into(VbEnt* h, int len, VbEnt* slot) {
    VbEnt v = *slot;
    *slot = h[0];
    vb_adjust(h, 0, len, v);
}

// a[0..mid) <- the mid smallest of a[0..len), ascending.
VB_HD inline void vb_partial_sort(VbEnt* a, int len, int mid) {
    if (mid <= 0) return;
    vb_make_heap(a, mid);
    for (int i = mid; i < len; ++i)
        if (vb_less(a[i], a[0])) vb_pop_into(a, mid, &a[i]);
    for (int last = mid; last > 1;) {
        --last;
        vb_pop_into(a, last, &a[last]);
    }
}
