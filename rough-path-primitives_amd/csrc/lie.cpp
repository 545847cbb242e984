// Host-side Hall-basis metadata: the Hall set, the bracket of two Hall
// elements expressed in the Hall basis, and the Lie->tensor (L2T) and
// tensor->Lie (T2L) matrices, built once per (width, depth) and uploaded to
// each device once per dtype as CSR.
//
// The Hall set is enumerated in the same order as the reference
// (basis/hall_basis.hpp:64-88) so that Lie coordinates agree index for index:
// key 0 is the unused (0,0) slot, keys 1..width are the letters (0, l), and
// each degree lists the pairs (left < right) with left_parent(right) <= left,
// by left degree, then left key, then right key.  Everything else is fixed by
// the mathematics: the expansion of a bracket in a basis is unique, L2T
// column h is the commutator expansion of h, and T2L column w is the right
// bracketing [w1,[w2,[...]]] divided by |w| (Dynkin-Specht-Wever).
#include "lie.hpp"

#include <hip/hip_runtime.h>

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "common.hpp"

namespace rpp_hip {

namespace {

inline uint64_t pair_key(int64_t a, int64_t b) {
    return (static_cast<uint64_t>(a) << 32) | static_cast<uint64_t>(b);
}

using Terms = std::vector<std::pair<int64_t, int64_t>>;  // (key, integer coefficient)

class HallSet {
public:
    HallSet(int width, int depth) : width_(width), depth_(depth) {
        pairs_ = {0, 0};
        degree_begin_ = {0, 1};
        deg_ = {0};
        if (depth >= 1) {
            for (int64_t l = 1; l <= width; ++l) add(0, l, 1);
            degree_begin_.push_back(1 + width);
        }
        for (int d = 2; d <= depth; ++d) {
            for (int ld = 1; 2 * ld <= d; ++ld) {
                const int rd = d - ld;
                for (int64_t left = degree_begin_[ld]; left < degree_begin_[ld + 1]; ++left) {
                    for (int64_t right = std::max(left + 1, degree_begin_[rd]);
                         right < degree_begin_[rd + 1]; ++right) {
                        if (pairs_[2 * right] <= left) add(left, right, d);
                    }
                }
            }
            degree_begin_.push_back(static_cast<int64_t>(deg_.size()));
        }
        if (depth == 0) degree_begin_.resize(2);
    }

    int64_t size() const { return static_cast<int64_t>(deg_.size()); }
    int degree(int64_t key) const { return deg_[static_cast<size_t>(key)]; }
    int64_t left(int64_t key) const { return pairs_[2 * key]; }
    int64_t right(int64_t key) const { return pairs_[2 * key + 1]; }
    const std::vector<int64_t>& pairs() const { return pairs_; }
    const std::vector<int64_t>& degree_begin() const { return degree_begin_; }
    int depth() const { return depth_; }
    int width() const { return width_; }

    int64_t find(int64_t a, int64_t b) const {
        auto it = index_.find(pair_key(a, b));
        return it == index_.end() ? 0 : it->second;
    }

    // [a, b] in the Hall basis (integer coefficients, zero terms dropped).
    const Terms& bracket(int64_t a, int64_t b) {
        static const Terms empty;
        if (a <= 0 || b <= 0 || a == b || a >= size() || b >= size()) return empty;
        if (degree(a) + degree(b) > depth_) return empty;
        const uint64_t k = pair_key(a, b);
        auto it = memo_.find(k);
        if (it != memo_.end()) return it->second;
        Terms out;
        if (int64_t h = find(a, b)) {
            out.emplace_back(h, 1);
        } else if (int64_t h2 = find(b, a)) {
            out.emplace_back(h2, -1);
        } else if (b < a) {
            const Terms rev = bracket(b, a);
            for (auto const& [h, c] : rev) out.emplace_back(h, -c);
        } else {
            // a < b and b = [b1, b2] is not a letter:
            // [a, [b1, b2]] = [[a, b1], b2] - [[a, b2], b1]   (Jacobi)
            const int64_t b1 = left(b), b2 = right(b);
            std::map<int64_t, int64_t> acc;
            auto half = [&](int64_t outer_right, int64_t inner_right, int64_t sign) {
                const Terms outer = bracket(a, outer_right);
                for (auto const& [h, c] : outer) {
                    const Terms inner = bracket(h, inner_right);
                    for (auto const& [g, e] : inner) acc[g] += sign * c * e;
                }
            };
            half(b1, b2, 1);
            half(b2, b1, -1);
            for (auto const& [h, c] : acc)
                if (c != 0) out.emplace_back(h, c);
        }
        return memo_.emplace(k, std::move(out)).first->second;
    }

private:
    void add(int64_t l, int64_t r, int d) {
        index_.emplace(pair_key(l, r), static_cast<int64_t>(deg_.size()));
        pairs_.push_back(l);
        pairs_.push_back(r);
        deg_.push_back(d);
    }

    int width_, depth_;
    std::vector<int64_t> pairs_;
    std::vector<int64_t> degree_begin_;
    std::vector<int> deg_;
    std::unordered_map<uint64_t, int64_t> index_;
    std::unordered_map<uint64_t, Terms> memo_;
};

struct Csc {
    std::vector<double> values;
    std::vector<int64_t> indices;
    std::vector<int64_t> offsets{0};
    void close_column() { offsets.push_back(static_cast<int64_t>(values.size())); }
};

Csr to_csr(const Csc& m, int64_t rows) {
    Csr r;
    r.rows = rows;
    r.offsets.assign(static_cast<size_t>(rows) + 1, 0);
    for (int64_t i : m.indices) ++r.offsets[static_cast<size_t>(i) + 1];
    for (int64_t i = 0; i < rows; ++i) r.offsets[i + 1] += r.offsets[i];
    r.values.resize(m.values.size());
    r.indices.resize(m.values.size());
    std::vector<int64_t> fill(r.offsets.begin(), r.offsets.end() - 1);
    const int64_t cols = static_cast<int64_t>(m.offsets.size()) - 1;
    for (int64_t c = 0; c < cols; ++c) {
        for (int64_t e = m.offsets[c]; e < m.offsets[c + 1]; ++e) {
            const int64_t dst = fill[static_cast<size_t>(m.indices[e])]++;
            r.values[dst] = m.values[e];
            r.indices[dst] = c;
        }
    }
    r.cols = cols;
    return r;
}

}  // namespace

std::shared_ptr<const LieMetadata> build_lie_metadata(int width, int depth) {
    auto md = std::make_shared<LieMetadata>();
    HallSet hall(width, depth);
    md->width = width;
    md->depth = depth;
    md->pairs = hall.pairs();
    md->degree_begin = hall.degree_begin();
    const int64_t true_size = hall.size();
    md->true_size = true_size;

    std::vector<int64_t> tstart(static_cast<size_t>(depth) + 2, 0), tpow(static_cast<size_t>(depth) + 2, 1);
    for (int k = 1; k <= depth + 1; ++k) {
        tpow[k] = tpow[k - 1] * width;
        tstart[k] = tstart[k - 1] + tpow[k - 1];
    }
    const int64_t dim = tstart[depth + 1];
    md->tensor_dim = dim;

    // L2T, CSC: column key-1 is the tensor expansion of Hall element `key`
    // (sparse/l2t_matrix.hpp:58-114): letters map to their word, and
    // [a, b] -> a(x)b - b(x)a, a concatenation of the two columns.
    Csc l2t;
    std::vector<std::map<int64_t, double>> expansion(static_cast<size_t>(true_size));
    for (int64_t key = 1; key < true_size; ++key) {
        auto& col = expansion[static_cast<size_t>(key)];
        if (hall.left(key) == 0) {
            col[hall.right(key)] = 1.0;  // tensor index of the letter word
        } else {
            const int64_t a = hall.left(key), b = hall.right(key);
            const int da = hall.degree(a), db = hall.degree(b);
            const int d = da + db;
            for (auto const& [ia, ca] : expansion[static_cast<size_t>(a)]) {
                for (auto const& [ib, cb] : expansion[static_cast<size_t>(b)]) {
                    const int64_t la = ia - tstart[da], lb = ib - tstart[db];
                    col[tstart[d] + la * tpow[db] + lb] += ca * cb;
                    col[tstart[d] + lb * tpow[da] + la] -= cb * ca;
                }
            }
            for (auto it = col.begin(); it != col.end();) it = it->second == 0.0 ? col.erase(it) : std::next(it);
        }
        for (auto const& [row, v] : col) {
            l2t.values.push_back(v);
            l2t.indices.push_back(row);
        }
        l2t.close_column();
    }
    // T2L, CSC: column w is the right bracketing of the word, over |w|
    // (sparse/t2l_matrix.hpp:57-142); rows are key - 1.
    Csc t2l;
    std::vector<std::vector<std::pair<int64_t, double>>> rb(static_cast<size_t>(dim));
    for (int k = 1; k <= depth; ++k) {
        for (int64_t local = 0; local < tpow[k]; ++local) {
            const int64_t g = tstart[k] + local;
            auto& col = rb[static_cast<size_t>(g)];
            if (k == 1) {
                col.emplace_back(local + 1, 1.0);
                continue;
            }
            const int64_t first = local / tpow[k - 1];
            const int64_t rest = tstart[k - 1] + local % tpow[k - 1];
            std::map<int64_t, double> acc;
            for (auto const& [key, c] : rb[static_cast<size_t>(rest)]) {
                for (auto const& [h, e] : hall.bracket(first + 1, key)) acc[h] += c * static_cast<double>(e);
            }
            for (auto const& [h, v] : acc)
                if (v != 0.0) col.emplace_back(h, v);
        }
    }
    t2l.close_column();  // the empty word
    for (int64_t g = 1; g < dim; ++g) {
        int k = 1;
        while (tstart[k + 1] <= g) ++k;
        const double divisor = 1.0 / static_cast<double>(k);
        for (auto const& [h, v] : rb[static_cast<size_t>(g)]) {
            t2l.values.push_back(k >= 2 ? v * divisor : v);
            t2l.indices.push_back(h - 1);
        }
        t2l.close_column();
    }
    md->l2t_csc_values = l2t.values;
    md->l2t_csc_indices = l2t.indices;
    md->l2t_csc_offsets = l2t.offsets;
    md->t2l_csc_values = t2l.values;
    md->t2l_csc_indices = t2l.indices;
    md->t2l_csc_offsets = t2l.offsets;
    md->l2t = to_csr(l2t, dim);
    md->t2l = to_csr(t2l, std::max<int64_t>(true_size - 1, 0));
    return md;
}

// ------------------------------------------------------------ on-disk form
//
// One little-endian binary file per (width, depth):
//   char[8]  "RPPLIE01"
//   int32    width, depth
//   int64    true_size, tensor_dim, l2t_nnz, t2l_nnz
//   int64    pairs[2 * true_size], degree_begin[depth + 2]
//   L2T CSC  double values[l2t_nnz], int64 rows[l2t_nnz], int64 offsets[true_size]
//   T2L CSC  double values[t2l_nnz], int64 rows[t2l_nnz], int64 offsets[tensor_dim + 1]
//   uint64   FNV-1a of every byte before it
// The values are stored in double, which holds them exactly (integers and
// integers / degree); the per-dtype device copies are derived from them.

namespace {

constexpr char kLieMagic[8] = {'R', 'P', 'P', 'L', 'I', 'E', '0', '1'};

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const auto* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
};

struct Writer {
    std::FILE* f;
    Fnv sum;
    bool ok = true;
    void put(const void* p, size_t n) {
        sum.add(p, n);
        ok = ok && std::fwrite(p, 1, n, f) == n;
    }
    template <typename V>
    void vec(const std::vector<V>& v) {
        if (!v.empty()) put(v.data(), v.size() * sizeof(V));
    }
};

struct Reader {
    std::vector<unsigned char> bytes;
    size_t pos = 0;
    bool get(void* p, size_t n) {
        if (n > bytes.size() - pos) return false;
        std::memcpy(p, bytes.data() + pos, n);
        pos += n;
        return true;
    }
    template <typename V>
    bool vec(std::vector<V>& v, int64_t count) {
        if (count < 0 || static_cast<uint64_t>(count) > (bytes.size() - pos) / sizeof(V)) return false;
        v.resize(static_cast<size_t>(count));
        return count == 0 || get(v.data(), static_cast<size_t>(count) * sizeof(V));
    }
};

// A compressed matrix is well formed: offsets start at 0, never decrease and
// end at nnz, and every inner index lies in [0, inner).
bool valid_compressed(const std::vector<int64_t>& offsets, const std::vector<int64_t>& indices, int64_t inner) {
    if (offsets.empty() || offsets.front() != 0 || offsets.back() != static_cast<int64_t>(indices.size()))
        return false;
    for (size_t i = 1; i < offsets.size(); ++i)
        if (offsets[i] < offsets[i - 1]) return false;
    for (int64_t r : indices)
        if (r < 0 || r >= inner) return false;
    return true;
}

}  // namespace

bool save_lie_metadata(const LieMetadata& md, const std::string& path, std::string* err) {
    static std::atomic<unsigned> serial{0};
    const std::string tmp = path + ".tmp." + std::to_string(static_cast<long long>(getpid())) + "." +
                            std::to_string(serial.fetch_add(1));
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        if (err) *err = "cannot open " + tmp + " for writing";
        return false;
    }
    Writer w{f};
    w.put(kLieMagic, sizeof(kLieMagic));
    const int32_t wd[2] = {md.width, md.depth};
    w.put(wd, sizeof(wd));
    const int64_t sizes[4] = {md.true_size, md.tensor_dim, static_cast<int64_t>(md.l2t_csc_values.size()),
                              static_cast<int64_t>(md.t2l_csc_values.size())};
    w.put(sizes, sizeof(sizes));
    w.vec(md.pairs);
    w.vec(md.degree_begin);
    w.vec(md.l2t_csc_values);
    w.vec(md.l2t_csc_indices);
    w.vec(md.l2t_csc_offsets);
    w.vec(md.t2l_csc_values);
    w.vec(md.t2l_csc_indices);
    w.vec(md.t2l_csc_offsets);
    const uint64_t check = w.sum.h;
    w.ok = w.ok && std::fwrite(&check, 1, sizeof(check), f) == sizeof(check);
    const bool closed = std::fclose(f) == 0;
    // rename is atomic, so concurrent writers and readers see a whole file
    if (!w.ok || !closed || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        if (err) *err = "cannot write " + path;
        return false;
    }
    return true;
}

std::shared_ptr<const LieMetadata> load_lie_metadata(const std::string& path, std::string* err) {
    auto bad = [&](const std::string& why) -> std::shared_ptr<const LieMetadata> {
        if (err) *err = path + ": " + why;
        return nullptr;
    };
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return bad("cannot open");
    Reader r;
    unsigned char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) r.bytes.insert(r.bytes.end(), buf, buf + got);
    std::fclose(f);
    if (r.bytes.size() < sizeof(kLieMagic) + sizeof(uint64_t)) return bad("truncated");
    uint64_t stored;
    std::memcpy(&stored, r.bytes.data() + r.bytes.size() - sizeof(stored), sizeof(stored));
    r.bytes.resize(r.bytes.size() - sizeof(stored));
    Fnv sum;
    sum.add(r.bytes.data(), r.bytes.size());
    if (sum.h != stored) return bad("checksum mismatch");
    char magic[8];
    if (!r.get(magic, sizeof(magic)) || std::memcmp(magic, kLieMagic, sizeof(magic)) != 0)
        return bad("not a Lie metadata file of this format version");
    int32_t wd[2];
    int64_t sizes[4];
    if (!r.get(wd, sizeof(wd)) || !r.get(sizes, sizeof(sizes))) return bad("truncated header");
    auto md = std::make_shared<LieMetadata>();
    md->width = wd[0];
    md->depth = wd[1];
    md->true_size = sizes[0];
    md->tensor_dim = sizes[1];
    if (md->width < 1 || md->depth < 0 || md->depth > kMaxDepth || md->true_size < 1) return bad("bad shape");
    int64_t dim = 0, p = 1;  // p = width^k, at most 2^40
    for (int k = 0; k <= md->depth; ++k) {
        dim += p;
        if (k < md->depth) {
            if (p > (int64_t{1} << 40) / md->width) return bad("bad shape");
            p *= md->width;
        }
    }
    if (md->tensor_dim != dim) return bad("tensor dimension does not match (width, depth)");
    const int64_t ts = md->true_size;
    // every section holds at least its count of 8-byte entries, so counts
    // beyond the file are refused before any arithmetic on them
    const int64_t cap = static_cast<int64_t>(r.bytes.size() / 8);
    if (ts > cap / 2 || dim >= cap) return bad("section sizes do not match the header");
    bool ok = r.vec(md->pairs, 2 * ts) && r.vec(md->degree_begin, md->depth + 2) &&
              r.vec(md->l2t_csc_values, sizes[2]) && r.vec(md->l2t_csc_indices, sizes[2]) &&
              r.vec(md->l2t_csc_offsets, ts) && r.vec(md->t2l_csc_values, sizes[3]) &&
              r.vec(md->t2l_csc_indices, sizes[3]) && r.vec(md->t2l_csc_offsets, dim + 1);
    if (!ok || r.pos != r.bytes.size()) return bad("section sizes do not match the header");
    if (md->degree_begin.front() != 0 || md->degree_begin.back() != ts) return bad("degree offsets inconsistent");
    for (size_t k = 1; k < md->degree_begin.size(); ++k)
        if (md->degree_begin[k] < md->degree_begin[k - 1]) return bad("degree offsets decrease");
    for (int64_t v : md->pairs)
        if (v < 0 || v >= ts) return bad("Hall pair out of range");
    if (!valid_compressed(md->l2t_csc_offsets, md->l2t_csc_indices, dim) ||
        !valid_compressed(md->t2l_csc_offsets, md->t2l_csc_indices, std::max<int64_t>(ts - 1, 0)))
        return bad("matrix structure invalid");
    // both matrices are block diagonal by degree, which the conversions rely on
    std::vector<int64_t> tstart(static_cast<size_t>(md->depth) + 2, 0);
    int64_t pk = 1;
    for (int k = 1; k <= md->depth + 1; ++k) {
        tstart[k] = tstart[k - 1] + pk;
        if (k <= md->depth) pk *= md->width;
    }
    auto degree_in = [](const std::vector<int64_t>& begin, int64_t i) {
        return static_cast<int>(std::upper_bound(begin.begin(), begin.end(), i) - begin.begin()) - 1;
    };
    for (int64_t key = 1; key < ts; ++key) {
        const int k = degree_in(md->degree_begin, key);
        for (int64_t e = md->l2t_csc_offsets[key - 1]; e < md->l2t_csc_offsets[key]; ++e)
            if (degree_in(tstart, md->l2t_csc_indices[e]) != k) return bad("L2T entry outside its degree");
    }
    for (int64_t w = 0; w < dim; ++w) {
        const int k = degree_in(tstart, w);
        for (int64_t e = md->t2l_csc_offsets[w]; e < md->t2l_csc_offsets[w + 1]; ++e)
            if (degree_in(md->degree_begin, md->t2l_csc_indices[e] + 1) != k) return bad("T2L entry outside its degree");
    }
    Csc l2t{md->l2t_csc_values, md->l2t_csc_indices, md->l2t_csc_offsets};
    Csc t2l{md->t2l_csc_values, md->t2l_csc_indices, md->t2l_csc_offsets};
    md->l2t = to_csr(l2t, dim);
    md->t2l = to_csr(t2l, std::max<int64_t>(ts - 1, 0));
    return md;
}

namespace {

std::mutex g_lie_mu;
std::map<std::pair<int, int>, std::shared_ptr<const LieMetadata>> g_lie_cache;

bool same_metadata(const LieMetadata& a, const LieMetadata& b) {
    return a.width == b.width && a.depth == b.depth && a.pairs == b.pairs && a.degree_begin == b.degree_begin &&
           a.l2t_csc_values == b.l2t_csc_values && a.l2t_csc_indices == b.l2t_csc_indices &&
           a.l2t_csc_offsets == b.l2t_csc_offsets && a.t2l_csc_values == b.t2l_csc_values &&
           a.t2l_csc_indices == b.t2l_csc_indices && a.t2l_csc_offsets == b.t2l_csc_offsets;
}

// The cache file of (width, depth) under RPP_HIP_LIE_CACHE_DIR, or "" when
// the variable is unset.
std::string cache_file(int width, int depth) {
    const char* dir = std::getenv("RPP_HIP_LIE_CACHE_DIR");
    if (!dir || !*dir) return {};
    return std::string(dir) + "/rpp_lie_w" + std::to_string(width) + "_d" + std::to_string(depth) + ".bin";
}

}  // namespace

std::shared_ptr<const LieMetadata> lie_metadata(int width, int depth) {
    std::lock_guard<std::mutex> lock(g_lie_mu);
    auto key = std::make_pair(width, depth);
    auto it = g_lie_cache.find(key);
    if (it != g_lie_cache.end()) return it->second;
    std::shared_ptr<const LieMetadata> md;
    const std::string file = cache_file(width, depth);
    if (!file.empty()) {
        md = load_lie_metadata(file, nullptr);
        if (md && (md->width != width || md->depth != depth)) md = nullptr;
    }
    if (!md) {
        md = build_lie_metadata(width, depth);
        // a missing, stale or damaged cache file is rewritten; failure to
        // write only loses the cache
        if (!file.empty()) save_lie_metadata(*md, file, nullptr);
    }
    g_lie_cache.emplace(key, md);
    return md;
}

int install_lie_metadata(std::shared_ptr<const LieMetadata> md, std::string* err) {
    std::lock_guard<std::mutex> lock(g_lie_mu);
    auto key = std::make_pair(md->width, md->depth);
    auto it = g_lie_cache.find(key);
    if (it == g_lie_cache.end()) {
        g_lie_cache.emplace(key, std::move(md));
        return 0;
    }
    if (same_metadata(*it->second, *md)) return 0;
    if (err) *err = "different Lie metadata for this (width, depth) is already in use in this process";
    return RPP_HIP_CONTRACT_VIOLATION;
}

namespace {

struct DeviceMatrices {
    DeviceCsr l2t, t2l;
};

// row_blocks: the first row of each degree and, last, the row count.
template <typename T>
int upload(const Csr& h, const std::vector<int64_t>& row_blocks, DeviceCsr& d) {
    d.rows = h.rows;
    d.cols = h.cols;
    d.nnz = static_cast<int64_t>(h.values.size());
    std::vector<T> vals(h.values.begin(), h.values.end());
    const size_t nv = std::max<size_t>(vals.size(), 1);
    if (hipMalloc(&d.values, nv * sizeof(T)) != hipSuccess) return 5;
    if (hipMalloc(reinterpret_cast<void**>(&d.indices), nv * sizeof(int64_t)) != hipSuccess) return 5;
    if (hipMalloc(reinterpret_cast<void**>(&d.offsets), h.offsets.size() * sizeof(int64_t)) != hipSuccess)
        return 5;
    if (!vals.empty()) {
        if (hipMemcpy(d.values, vals.data(), vals.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return 9;
        if (hipMemcpy(d.indices, h.indices.data(), h.indices.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice) != hipSuccess)
            return 9;
    }
    if (hipMemcpy(d.offsets, h.offsets.data(), h.offsets.size() * sizeof(int64_t), hipMemcpyHostToDevice) !=
        hipSuccess)
        return 9;
    // packed form: value table (its last slot the padding value -0.0) and
    // (table index, column) words, each row padded to whole uint4 groups
    std::vector<T> table;
    std::map<T, uint32_t> slot;
    bool packable = h.cols < (int64_t{1} << 24) - 1 && h.rows < (int64_t{1} << 31) && d.nnz < (int64_t{1} << 31) &&
                    !vals.empty();
    for (size_t e = 0; packable && e < vals.size(); ++e) {
        if (slot.count(vals[e])) continue;
        if (table.size() == 255) {
            packable = false;
            break;
        }
        slot.emplace(vals[e], static_cast<uint32_t>(table.size()));
        table.push_back(vals[e]);
    }
    if (packable) {
        const uint32_t pad = (static_cast<uint32_t>(table.size()) << 24) | 0xffffffu;
        table.push_back(-T(0));
        std::vector<int64_t> group(static_cast<size_t>(h.rows) + 1, 0);  // first group of each row
        for (int64_t r = 0; r < h.rows; ++r) {
            const int64_t len = h.offsets[static_cast<size_t>(r) + 1] - h.offsets[static_cast<size_t>(r)];
            group[static_cast<size_t>(r) + 1] = group[static_cast<size_t>(r)] + (len + 3) / 4;
        }
        if (group.back() + kLiePackSlack >= (int64_t{1} << 31)) return 0;  // no packed form
        // kLiePackSlack padding groups after the last row: the packed kernel
        // loads up to that many groups past a row's end without a test
        std::vector<uint32_t> words(static_cast<size_t>(group.back() + kLiePackSlack) * 4, pad);
        for (int64_t r = 0; r < h.rows; ++r) {
            size_t at = static_cast<size_t>(group[static_cast<size_t>(r)]) * 4;
            for (int64_t e = h.offsets[static_cast<size_t>(r)]; e < h.offsets[static_cast<size_t>(r) + 1]; ++e)
                words[at++] = (slot[vals[static_cast<size_t>(e)]] << 24) | static_cast<uint32_t>(h.indices[static_cast<size_t>(e)]);
        }
        std::vector<int32_t> order(static_cast<size_t>(h.rows));
        for (int64_t r = 0; r < h.rows; ++r) order[static_cast<size_t>(r)] = static_cast<int32_t>(r);
        auto length = [&](int32_t r) { return h.offsets[static_cast<size_t>(r) + 1] - h.offsets[static_cast<size_t>(r)]; };
        // Long rows are taken longest first, so that the critical path
        // starts early and a wavefront's rows end together.  Short rows stay
        // in order: then a wavefront's stores cover consecutive words, which
        // matters when the output (a tensor) is much larger than the argument.
        d.long_rows = h.rows > 0 && group.back() >= 4 * h.rows ? 1 : 0;
        if (d.long_rows)
            for (size_t k = 0; k + 1 < row_blocks.size(); ++k)
                std::stable_sort(order.begin() + row_blocks[k], order.begin() + row_blocks[k + 1],
                                 [&](int32_t a, int32_t b) { return length(a) > length(b); });
        std::vector<int4> desc(order.size());
        for (size_t s = 0; s < order.size(); ++s) {
            const size_t r = static_cast<size_t>(order[s]);
            desc[s] = int4{order[s], static_cast<int32_t>(group[r]), static_cast<int32_t>(group[r + 1]), 0};
        }
        if (hipMalloc(reinterpret_cast<void**>(&d.packed4), std::max<size_t>(words.size(), 4) * sizeof(uint32_t)) !=
                hipSuccess ||
            hipMalloc(&d.table, table.size() * sizeof(T)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&d.rowdesc), std::max<size_t>(desc.size(), 1) * sizeof(int4)) !=
                hipSuccess)
            return 5;
        if ((!words.empty() &&
             hipMemcpy(d.packed4, words.data(), words.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(d.table, table.data(), table.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess ||
            (!desc.empty() &&
             hipMemcpy(d.rowdesc, desc.data(), desc.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess))
            return 9;
        d.table_size = static_cast<int32_t>(table.size());
    }
    return 0;
}

}  // namespace

template <typename T>
int device_lie_matrices(int width, int depth, const DeviceCsr** l2t, const DeviceCsr** t2l) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int>, std::unique_ptr<DeviceMatrices>> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(9, "hipGetDevice failed");
    auto md = lie_metadata(width, depth);
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(width, depth, static_cast<int>(sizeof(T)), dev);
    auto it = cache.find(key);
    if (it == cache.end()) {
        auto dm = std::make_unique<DeviceMatrices>();
        // degree blocks: L2T rows are tensor words, T2L rows Hall keys - 1
        std::vector<int64_t> tensor_blocks, lie_blocks;
        int64_t start = 0, level = 1;
        for (int k = 0; k <= depth; ++k) {
            tensor_blocks.push_back(start);
            start += level;
            level *= width;
        }
        tensor_blocks.push_back(md->l2t.rows);
        for (int k = 1; k <= depth; ++k) lie_blocks.push_back(md->degree_begin[static_cast<size_t>(k)] - 1);
        lie_blocks.push_back(md->t2l.rows);
        if (int rc = upload<T>(md->l2t, tensor_blocks, dm->l2t)) return fail(rc, "uploading the L2T matrix failed");
        if (int rc = upload<T>(md->t2l, lie_blocks, dm->t2l)) return fail(rc, "uploading the T2L matrix failed");
        it = cache.emplace(key, std::move(dm)).first;
    }
    *l2t = &it->second->l2t;
    *t2l = &it->second->t2l;
    return 0;
}

template int device_lie_matrices<float>(int, int, const DeviceCsr**, const DeviceCsr**);
template int device_lie_matrices<double>(int, int, const DeviceCsr**, const DeviceCsr**);

}  // namespace rpp_hip
