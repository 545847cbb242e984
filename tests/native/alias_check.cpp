// Brute-force check of rpp_hip::views_overlap (the C-ABI's aliasing test):
// for random pairs of small batch views, every element address each view
// touches is enumerated and the two byte sets intersected.  The predicate must
// never miss a shared byte, and must equal the enumeration wherever it claims
// to be exact (equal strides, element-aligned offset, contiguous rows of
// objects or of coefficients).  Host-only; built with g++ by the test suite.
#include <cstdint>
#include <cstdio>
#include <random>
#include <set>

#include "alias.hpp"

using rpp_hip::ViewSet;

namespace {

std::set<intptr_t> bytes_of(const ViewSet& v, int64_t n, int64_t elem) {
    std::set<intptr_t> out;
    for (int64_t b = 0; b < n; ++b)
        for (int64_t i = v.begin; i < v.end; ++i)
            for (int64_t k = 0; k < elem; ++k)
                out.insert(reinterpret_cast<intptr_t>(v.data) + (b * v.bs + i * v.cs) * elem + k);
    return out;
}

bool brute(const ViewSet& a, const ViewSet& b, int64_t n, int64_t elem) {
    const auto sa = bytes_of(a, n, elem);
    for (intptr_t x : bytes_of(b, n, elem))
        if (sa.count(x)) return true;
    return false;
}

bool claims_exact(const ViewSet& a, const ViewSet& b, int64_t n, int64_t elem) {
    const intptr_t bytes = reinterpret_cast<intptr_t>(b.data) - reinterpret_cast<intptr_t>(a.data);
    if (a.bs != b.bs || a.cs != b.cs || bytes % elem != 0) return false;
    return (a.cs == 1 && (a.bs >= 0 || n == 1)) || ((a.bs == 1 || n == 1) && a.cs > 0);
}

}  // namespace

int main() {
    std::mt19937_64 rng(20261204);
    auto pick = [&](int64_t lo, int64_t hi) { return std::uniform_int_distribution<int64_t>(lo, hi)(rng); };
    const intptr_t base = intptr_t{1} << 20;
    long cases = 0, exact = 0, overlapping = 0, failures = 0;
    for (int it = 0; it < 200000; ++it) {
        const int64_t elem = pick(0, 1) ? 8 : 4;
        const int64_t n = pick(0, 6);
        auto view = [&](int64_t bs, int64_t cs) {
            const int64_t begin = pick(0, 9);
            const int64_t end = begin + pick(-1, 9);
            const int64_t off = pick(-40, 40) * elem + (pick(0, 7) == 0 ? pick(1, elem - 1) : 0);
            return ViewSet{reinterpret_cast<const void*>(base + off), bs, cs, begin, end};
        };
        const int shape = static_cast<int>(pick(0, 3));
        int64_t bs = pick(0, 14), cs = pick(1, 5);
        if (shape == 0) cs = 1;  // object-major
        if (shape == 1) bs = 1;  // batch-interleaved
        ViewSet a = view(bs, cs);
        ViewSet b = shape == 3 ? view(pick(-3, 14), pick(1, 5)) : view(bs, cs);
        const bool want = brute(a, b, n, elem);
        const bool got = rpp_hip::views_overlap(a, b, n, static_cast<size_t>(elem));
        const bool ex = claims_exact(a, b, n, elem);
        ++cases;
        exact += ex;
        overlapping += want;
        if ((want && !got) || (ex && want != got)) {
            if (++failures <= 10)
                std::printf("FAIL n=%lld elem=%lld a{%+lld bs=%lld cs=%lld [%lld,%lld)} b{%+lld bs=%lld cs=%lld "
                            "[%lld,%lld)}: brute %d predicate %d\n",
                            (long long)n, (long long)elem, (long long)(reinterpret_cast<intptr_t>(a.data) - base),
                            (long long)a.bs, (long long)a.cs, (long long)a.begin, (long long)a.end,
                            (long long)(reinterpret_cast<intptr_t>(b.data) - base), (long long)b.bs,
                            (long long)b.cs, (long long)b.begin, (long long)b.end, want, got);
        }
    }
    std::printf("%s: %ld cases, %ld exact, %ld overlapping, %ld failures\n", failures ? "ALIAS FAIL" : "ALIAS OK",
                cases, exact, overlapping, failures);
    return failures ? 1 : 0;
}
