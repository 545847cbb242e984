// Host-side test of whether two batch views share an element, used by the
// C-ABI's descriptor validation.  Header-only and free of HIP so that a plain
// C++ compiler can check it against brute-force enumeration.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace rpp_hip {

// The elements a batch view touches: data[b * bs + i * cs] for objects b in
// [0, n) and coefficient indices i in [begin, end).
struct ViewSet {
    const void* data;
    int64_t bs, cs, begin, end;
};

inline int64_t floor_div(int64_t a, int64_t b) {  // b > 0
    int64_t q = a / b;
    if (a % b != 0 && a < 0) --q;
    return q;
}
inline int64_t ceil_div(int64_t a, int64_t b) { return -floor_div(-a, b); }

// Whether views a and b of n objects (elements of `elem` bytes) touch a
// common element.  Views whose byte ranges are disjoint never do.  Two views
// on one lattice (equal strides, an element-aligned offset d between the base
// pointers) are tested exactly when their rows are contiguous: rows of
// objects (coeff stride 1, batch stride >= 0) or rows of coefficients (batch
// stride 1, or a single object).  Rows of a and b then meet iff an integer
// row difference lies in an open interval fixed by d and the two index
// ranges.  Any other pair whose byte ranges intersect counts as overlapping.
inline bool views_overlap(const ViewSet& a, const ViewSet& b, int64_t n, size_t elem) {
    if (n <= 0 || a.begin >= a.end || b.begin >= b.end) return false;
    const intptr_t e = static_cast<intptr_t>(elem);
    // byte extent of a view: the index b*bs + i*cs is extreme at the corners
    auto extent = [&](const ViewSet& v, intptr_t* lo, intptr_t* hi) {
        const int64_t bmin = std::min<int64_t>(0, (n - 1) * v.bs), bmax = std::max<int64_t>(0, (n - 1) * v.bs);
        const int64_t imin = std::min(v.begin * v.cs, (v.end - 1) * v.cs);
        const int64_t imax = std::max(v.begin * v.cs, (v.end - 1) * v.cs);
        *lo = reinterpret_cast<intptr_t>(v.data) + (bmin + imin) * e;
        *hi = reinterpret_cast<intptr_t>(v.data) + (bmax + imax + 1) * e;
    };
    intptr_t alo, ahi, blo, bhi;
    extent(a, &alo, &ahi);
    extent(b, &blo, &bhi);
    if (ahi <= blo || bhi <= alo) return false;
    const intptr_t bytes = reinterpret_cast<intptr_t>(b.data) - reinterpret_cast<intptr_t>(a.data);
    if (a.bs != b.bs || a.cs != b.cs || bytes % e != 0) return true;
    const int64_t d = bytes / e;  // b's element 0 is a's element d
    if (a.cs == 1 && (a.bs >= 0 || n == 1)) {
        const int64_t bs = n == 1 ? 0 : a.bs;
        // one row [begin, end) for every object: two intervals
        if (bs == 0) return d + b.begin < a.end && a.begin < d + b.end;
        // object rows K*bs + [begin, end): rows of a and b, K = k_a - k_b
        // apart, meet iff K*bs lies strictly between d + b.begin - a.end and
        // d + b.end - a.begin, with |K| < n
        const int64_t k0 = std::max(floor_div(d + b.begin - a.end, bs) + 1, -(n - 1));
        const int64_t k1 = std::min(ceil_div(d + b.end - a.begin, bs) - 1, n - 1);
        return k0 <= k1;
    }
    if ((a.bs == 1 || n == 1) && a.cs > 0) {
        // coefficient rows K*cs + [0, n) (one element when n == 1): rows
        // K = i_a - i_b apart meet iff |K*cs - d| < n
        const int64_t k0 = std::max(floor_div(d - n, a.cs) + 1, a.begin - (b.end - 1));
        const int64_t k1 = std::min(ceil_div(d + n, a.cs) - 1, (a.end - 1) - b.begin);
        return k0 <= k1;
    }
    return true;
}

}  // namespace rpp_hip
