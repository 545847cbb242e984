// Host-side robustness check of librpp_hip's C-ABI, built with AddressSanitizer
// and UndefinedBehaviorSanitizer on the host code only (the kernels are the
// library's ordinary objects).  It runs only where no HIP device is present:
// every descriptor it passes holds fake pointers, which the host validation
// must turn into an error (or which fail at the first HIP call for want of a
// device) before any kernel could see them.
//
//   * every descriptor entry point with a NULL descriptor;
//   * random descriptors of every entry point: shapes, counts, strides,
//     windows and pointers drawn from edge values;
//   * the Hall basis metadata: sizes and copies into exactly sized buffers,
//     save and load round trips, and loads of damaged files (truncated,
//     bytes flipped, header fields changed with the checksum recomputed so
//     the structural checks are reached).
//
// Any sanitizer report aborts the run (-fno-sanitize-recover); the exit
// status is 0 and the last line "CAPI SAN OK" when every check passed.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "rpp_hip.h"

namespace {

std::mt19937_64 g_rng(0x5a17c0de);
long g_calls = 0;
int g_failures = 0;

int64_t pick(std::initializer_list<int64_t> xs) {
    std::vector<int64_t> v(xs);
    return v[std::uniform_int_distribution<size_t>(0, v.size() - 1)(g_rng)];
}
int64_t range(int64_t lo, int64_t hi) { return std::uniform_int_distribution<int64_t>(lo, hi)(g_rng); }

void* fake_ptr() {
    switch (range(0, 5)) {
        case 0: return nullptr;
        case 1: return reinterpret_cast<void*>(uintptr_t{8});
        case 2: return reinterpret_cast<void*>(~uintptr_t{0} - 7);
        default: return reinterpret_cast<void*>(uintptr_t{0x7f0000000000} + 256 * static_cast<uintptr_t>(range(0, 64)));
    }
}

rpp_hip_basis basis() {
    return rpp_hip_basis{static_cast<int32_t>(pick({-1, 0, 1, 2, 3, 5, 8, 17, 255, 256, 65535, 65536, 1 << 30})),
                         static_cast<int32_t>(pick({-1, 0, 1, 2, 3, 4, 6, 8, 12, 30, 31, 1 << 30}))};
}

int64_t count() { return pick({-1, 0, 1, 2, 3, 1000, int64_t{1} << 31, int64_t{1} << 40, INT64_MAX}); }

int64_t stride() {
    return pick({-1, 0, 1, 2, 13, 1000, int64_t{1} << 31, int64_t{1} << 40, INT64_MAX / 2, INT64_MAX});
}

rpp_hip_tensor_batch tensor() {
    rpp_hip_tensor_batch t;
    t.data = fake_ptr();
    t.batch_stride = stride();
    t.coeff_stride = pick({-1, 0, 1, 1, 1, 2, 7, INT64_MAX});
    t.min_degree = static_cast<int32_t>(range(-2, 12));
    t.max_degree = static_cast<int32_t>(range(-2, 12));
    return t;
}

rpp_hip_vector_batch vec() {
    return rpp_hip_vector_batch{fake_ptr(), stride(), pick({-1, 0, 1, 2, INT64_MAX}), pick({-1, 0, 1, 5, INT64_MAX}),
                                pick({-1, 0, 1, 13, INT64_MAX})};
}

rpp_hip_sparse_matrix matrix() {
    return rpp_hip_sparse_matrix{static_cast<int32_t>(pick({-1, RPP_HIP_CSR, RPP_HIP_CSC, 7})), fake_ptr(),
                                 static_cast<const int64_t*>(fake_ptr()), static_cast<const int64_t*>(fake_ptr()),
                                 pick({-1, 0, 1, 100, INT64_MAX}), pick({-1, 0, 1, 13, INT64_MAX}),
                                 pick({-1, 0, 1, 13, INT64_MAX})};
}

rpp_hip_graded_basis graded() {
    const rpp_hip_basis b = basis();
    return rpp_hip_graded_basis{static_cast<int32_t>(pick({-1, 0, 1, 2})), b.width, b.depth};
}

void expect(bool ok, const std::string& what) {
    if (!ok) {
        if (++g_failures <= 20) std::printf("FAIL %s\n", what.c_str());
    }
}

// One call: any return code is fine (no device: an accepted descriptor fails
// at its first HIP call), but a non-zero code must come with a message.
template <typename D>
void call(int (*fn)(const D*, void*), const D& d, const char* name) {
    ++g_calls;
    const int rc = fn(&d, nullptr);
    expect(rc >= 0 && rc <= 9, std::string(name) + ": return code out of range");
    if (rc != 0) {
        const char* m = rpp_hip_last_error();
        expect(m != nullptr && m[0] != '\0', std::string(name) + ": error without a message");
    }
}

#define BOTH(NAME, DESC)                     \
    call<DESC>(NAME##_f32, d, #NAME "_f32"); \
    call<DESC>(NAME##_f64, d, #NAME "_f64")

void null_descriptors() {
    auto null = [](auto fn, const char* name) {
        ++g_calls;
        expect(fn(nullptr, nullptr) == RPP_HIP_INVALID_ARGUMENT, std::string(name) + ": NULL descriptor accepted");
    };
#define NULL2(NAME) null(NAME##_f32, #NAME "_f32"), null(NAME##_f64, #NAME "_f64")
    NULL2(rpp_hip_ft_mul);
    NULL2(rpp_hip_ft_fma);
    NULL2(rpp_hip_ft_inplace_fma);
    NULL2(rpp_hip_ft_inplace_mul);
    NULL2(rpp_hip_ft_exp);
    NULL2(rpp_hip_ft_log);
    NULL2(rpp_hip_ft_fmexp);
    NULL2(rpp_hip_ft_fmexp_adjoint);
    NULL2(rpp_hip_ft_log_adjoint);
    NULL2(rpp_hip_tensor_pairing);
    NULL2(rpp_hip_ft_adj_lmul);
    NULL2(rpp_hip_ft_adj_rmul);
    NULL2(rpp_hip_tensor_antipode);
    NULL2(rpp_hip_tensor_reflect);
    NULL2(rpp_hip_sparse_matrix_vector_product);
    NULL2(rpp_hip_tensor_to_lie);
    NULL2(rpp_hip_lie_to_tensor);
    NULL2(rpp_hip_tensor_to_lie_matrix);
    NULL2(rpp_hip_lie_to_tensor_matrix);
    NULL2(rpp_hip_st_mul);
    NULL2(rpp_hip_st_fma);
    NULL2(rpp_hip_st_inplace_fma);
    NULL2(rpp_hip_st_adj_mul);
    NULL2(rpp_hip_left_hs_mul);
    NULL2(rpp_hip_left_hs_adj_lmul);
    NULL2(rpp_hip_left_hs_adj_rmul);
    NULL2(rpp_hip_vector_assign);
    NULL2(rpp_hip_vector_add);
    NULL2(rpp_hip_vector_inplace_add);
    NULL2(rpp_hip_vector_scalar_multiply);
    NULL2(rpp_hip_vector_set_constant);
    NULL2(rpp_hip_tensor_set_identity);
    NULL2(rpp_hip_tensor_add_identity);
    NULL2(rpp_hip_signature);
    NULL2(rpp_hip_signature_backward);
    NULL2(rpp_hip_signature_backward_noalloc);
#undef NULL2
}

void random_descriptors(int rounds) {
    for (int it = 0; it < rounds; ++it) {
        {
            rpp_hip_ft_mul_desc d{basis(), count(), tensor(), tensor(), tensor(), 0.5};
            BOTH(rpp_hip_ft_mul, rpp_hip_ft_mul_desc);
            BOTH(rpp_hip_st_mul, rpp_hip_ft_mul_desc);
        }
        {
            rpp_hip_ft_fma_desc d{basis(), count(), tensor(), tensor(), tensor(), tensor(), 1.0, 2.0};
            BOTH(rpp_hip_ft_fma, rpp_hip_ft_fma_desc);
            BOTH(rpp_hip_st_fma, rpp_hip_ft_fma_desc);
        }
        {
            rpp_hip_ft_inplace_fma_desc d{basis(), count(), static_cast<int32_t>(pick({0, 123, 213, 231, -1})),
                                          tensor(), tensor(), tensor(), 1.0, 1.0};
            BOTH(rpp_hip_ft_inplace_fma, rpp_hip_ft_inplace_fma_desc);
        }
        {
            rpp_hip_ft_inplace_mul_desc d{basis(), count(), tensor(), tensor(), 1.0};
            BOTH(rpp_hip_ft_inplace_mul, rpp_hip_ft_inplace_mul_desc);
        }
        {
            rpp_hip_ft_unary_desc d{basis(), count(), tensor(), tensor()};
            BOTH(rpp_hip_ft_exp, rpp_hip_ft_unary_desc);
            BOTH(rpp_hip_ft_log, rpp_hip_ft_unary_desc);
            BOTH(rpp_hip_tensor_antipode, rpp_hip_ft_unary_desc);
            BOTH(rpp_hip_tensor_reflect, rpp_hip_ft_unary_desc);
        }
        {
            rpp_hip_ft_fmexp_desc d{basis(), count(), tensor(), tensor(), tensor()};
            BOTH(rpp_hip_ft_fmexp, rpp_hip_ft_fmexp_desc);
        }
        {
            rpp_hip_ft_fmexp_adjoint_desc d{basis(), count(), tensor(), tensor(), tensor(), tensor(), tensor()};
            BOTH(rpp_hip_ft_fmexp_adjoint, rpp_hip_ft_fmexp_adjoint_desc);
        }
        {
            rpp_hip_ft_log_adjoint_desc d{basis(), count(), tensor(), tensor(), tensor()};
            BOTH(rpp_hip_ft_log_adjoint, rpp_hip_ft_log_adjoint_desc);
        }
        {
            rpp_hip_tensor_pairing_desc d{basis(), count(), rpp_hip_scalar_batch{fake_ptr(), stride()}, tensor(),
                                          tensor()};
            BOTH(rpp_hip_tensor_pairing, rpp_hip_tensor_pairing_desc);
        }
        {
            rpp_hip_ft_adjoint_desc d{basis(), count(), tensor(), tensor(), tensor()};
            BOTH(rpp_hip_ft_adj_lmul, rpp_hip_ft_adjoint_desc);
            BOTH(rpp_hip_ft_adj_rmul, rpp_hip_ft_adjoint_desc);
            BOTH(rpp_hip_st_adj_mul, rpp_hip_ft_adjoint_desc);
        }
        {
            rpp_hip_spmv_desc d{count(), vec(), vec(), matrix(), 1.0};
            BOTH(rpp_hip_sparse_matrix_vector_product, rpp_hip_spmv_desc);
        }
        {
            rpp_hip_lie_convert_desc d{basis(), count(), tensor(), tensor()};
            BOTH(rpp_hip_tensor_to_lie, rpp_hip_lie_convert_desc);
            BOTH(rpp_hip_lie_to_tensor, rpp_hip_lie_convert_desc);
        }
        {
            rpp_hip_lie_matrix_convert_desc d{basis(), count(), tensor(), tensor(), matrix()};
            BOTH(rpp_hip_tensor_to_lie_matrix, rpp_hip_lie_matrix_convert_desc);
            BOTH(rpp_hip_lie_to_tensor_matrix, rpp_hip_lie_matrix_convert_desc);
        }
        {
            rpp_hip_st_inplace_fma_desc d{basis(), count(), tensor(), tensor(), tensor(), 1.0, 1.0};
            BOTH(rpp_hip_st_inplace_fma, rpp_hip_st_inplace_fma_desc);
        }
        {
            rpp_hip_half_shuffle_desc d{basis(), count(), tensor(), tensor(), tensor(), 1.0};
            BOTH(rpp_hip_left_hs_mul, rpp_hip_half_shuffle_desc);
            BOTH(rpp_hip_left_hs_adj_lmul, rpp_hip_half_shuffle_desc);
            BOTH(rpp_hip_left_hs_adj_rmul, rpp_hip_half_shuffle_desc);
        }
        {
            rpp_hip_vector_assign_desc d{graded(), count(), tensor(), tensor()};
            BOTH(rpp_hip_vector_assign, rpp_hip_vector_assign_desc);
        }
        {
            rpp_hip_vector_add_desc d{graded(), count(), tensor(), tensor(), tensor(), 1.0, -1.0};
            BOTH(rpp_hip_vector_add, rpp_hip_vector_add_desc);
        }
        {
            rpp_hip_vector_inplace_add_desc d{graded(), count(), tensor(), tensor(), 2.0};
            BOTH(rpp_hip_vector_inplace_add, rpp_hip_vector_inplace_add_desc);
        }
        {
            rpp_hip_vector_scalar_desc d{graded(), count(), tensor(), 3.0};
            BOTH(rpp_hip_vector_scalar_multiply, rpp_hip_vector_scalar_desc);
            BOTH(rpp_hip_vector_set_constant, rpp_hip_vector_scalar_desc);
            BOTH(rpp_hip_tensor_set_identity, rpp_hip_vector_scalar_desc);
            BOTH(rpp_hip_tensor_add_identity, rpp_hip_vector_scalar_desc);
        }
        {
            rpp_hip_signature_desc d{};
            d.basis = basis();
            d.num_paths = count();
            d.increments = fake_ptr();
            d.offsets = static_cast<const int64_t*>(fake_ptr());
            d.out = tensor();
            d.checkpoint_every = static_cast<int32_t>(pick({-1, 0, 1, 16}));
            d.checkpoints = fake_ptr();
            d.ckpt_offsets = static_cast<const int64_t*>(fake_ptr());
            d.tile_threads = static_cast<int32_t>(pick({-1, 0, 64, 256, 1 << 20}));
            d.tile_split = static_cast<int32_t>(pick({-1, 0, 1, 2, 99}));
            d.tile_kind = static_cast<int32_t>(pick({-1, 0, 1, 2, 3}));
            d.tile_expand = static_cast<int32_t>(pick({-1, 0, 1, 2, 99}));
            d.order = static_cast<const int64_t*>(fake_ptr());
            d.num_increments = count();
            BOTH(rpp_hip_signature, rpp_hip_signature_desc);
        }
        {
            rpp_hip_signature_backward_desc d{};
            d.basis = basis();
            d.num_paths = count();
            d.increments = fake_ptr();
            d.offsets = static_cast<const int64_t*>(fake_ptr());
            d.signatures = tensor();
            d.cotangent = tensor();
            d.grad_increments = fake_ptr();
            d.checkpoint_every = static_cast<int32_t>(pick({-1, 0, 1, 16}));
            d.checkpoints = fake_ptr();
            d.ckpt_offsets = static_cast<const int64_t*>(fake_ptr());
            d.scratch = fake_ptr();
            d.tile_threads = static_cast<int32_t>(pick({-1, 0, 64, 256}));
            d.tile_split = static_cast<int32_t>(pick({-1, 0, 1, 2}));
            d.order = static_cast<const int64_t*>(fake_ptr());
            d.num_increments = count();
            d.tile_expand = static_cast<int32_t>(pick({-1, 0, 1, 2}));
            d.tile_kind = static_cast<int32_t>(pick({-1, 0, 1, 2}));
            BOTH(rpp_hip_signature_backward, rpp_hip_signature_backward_desc);
            BOTH(rpp_hip_signature_backward_noalloc, rpp_hip_signature_backward_desc);
            ++g_calls;
            (void)rpp_hip_signature_backward_scratch_bytes(&d, static_cast<int32_t>(pick({-1, 4, 8, 16})));
        }
        ++g_calls;
        (void)rpp_hip_signature_checkpoint_every(static_cast<int32_t>(pick({-1, 0, 4, 8})), count());
        int32_t th = 0, sp = 0;
        ++g_calls;
        (void)rpp_hip_fmexp_tile(static_cast<int32_t>(pick({-1, 4, 8})), static_cast<int32_t>(range(-2, 300)),
                                 static_cast<int32_t>(range(-2, 40)), range(0, 1) ? &th : nullptr,
                                 range(0, 1) ? &sp : nullptr);
    }
    // tile listing into exactly sized buffers, including capacity 0
    for (int32_t bytes : {4, 8}) {
        const int n = rpp_hip_signature_tiles(bytes, nullptr, 0);
        expect(n > 0, "signature_tiles count");
        for (int cap : {0, 1, n - 1, n}) {
            std::vector<int32_t> out(static_cast<size_t>(std::max(cap, 0)) * 6);
            ++g_calls;
            expect(rpp_hip_signature_tiles(bytes, cap ? out.data() : nullptr, cap) == n, "signature_tiles");
        }
    }
    expect(rpp_hip_version() != nullptr, "version");
}

// ---------------------------------------------------------------- Lie files

uint64_t fnv(const std::vector<unsigned char>& b, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

bool write_file(const std::string& path, const std::vector<unsigned char>& b) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    return std::fclose(f) == 0 && ok;
}

std::vector<unsigned char> read_file(const std::string& path) {
    std::vector<unsigned char> b;
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return b;
    unsigned char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + got);
    std::fclose(f);
    return b;
}

void reseal(std::vector<unsigned char>& b) {
    if (b.size() < 8) return;
    const uint64_t h = fnv(b, b.size() - 8);
    std::memcpy(b.data() + b.size() - 8, &h, 8);
}

void lie_metadata(const std::string& dir) {
    const int shapes[][2] = {{2, 1}, {2, 4}, {2, 8}, {3, 1}, {3, 5}, {4, 4}, {5, 3}, {8, 3}, {16, 2}, {1, 3}, {0, 2},
                             {2, -1}, {2, 31}};
    for (auto& s : shapes) {
        int64_t sizes[4] = {-1, -1, -1, -1};
        ++g_calls;
        const int rc = rpp_hip_lie_basis_sizes(s[0], s[1], sizes);
        if (rc != 0) continue;
        const int64_t ts = sizes[0], nl = sizes[1], nt = sizes[2], dim = sizes[3];
        std::vector<int64_t> pairs(static_cast<size_t>(2 * ts)), db(static_cast<size_t>(s[1] + 2)),
            li(static_cast<size_t>(nl)), lo(static_cast<size_t>(ts)), ti(static_cast<size_t>(nt)),
            to(static_cast<size_t>(dim + 1));
        std::vector<double> lv(static_cast<size_t>(nl)), tv(static_cast<size_t>(nt));
        ++g_calls;
        expect(rpp_hip_lie_basis_copy(s[0], s[1], pairs.data(), db.data(), lv.data(), li.data(), lo.data(), tv.data(),
                                      ti.data(), to.data()) == 0,
               "lie_basis_copy");
        ++g_calls;
        expect(rpp_hip_lie_basis_copy(s[0], s[1], nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                      nullptr) == 0,
               "lie_basis_copy with NULL outputs");
    }
    // save, load back, then damaged copies of a small and a larger file
    for (auto& s : {std::pair{2, 3}, std::pair{3, 4}}) {
        const std::string path = dir + "/lie_" + std::to_string(s.first) + "_" + std::to_string(s.second) + ".bin";
        ++g_calls;
        expect(rpp_hip_lie_basis_save(s.first, s.second, path.c_str()) == 0, "lie_basis_save");
        int32_t w = 0, d = 0;
        ++g_calls;
        expect(rpp_hip_lie_basis_load(path.c_str(), &w, &d) == 0 && w == s.first && d == s.second,
               "lie_basis_load round trip");
        const std::vector<unsigned char> good = read_file(path);
        expect(good.size() > 64, "saved file size");
        const std::string bad = dir + "/damaged.bin";
        auto load_bad = [&](const std::vector<unsigned char>& b, const char* what, bool must_fail) {
            expect(write_file(bad, b), "writing a damaged file");
            ++g_calls;
            const int rc = rpp_hip_lie_basis_load(bad.c_str(), nullptr, nullptr);
            if (must_fail) expect(rc != 0, std::string("damaged file accepted: ") + what);
        };
        // every truncation of the small file, a stride of the larger one
        const size_t step = good.size() < 4096 ? 1 : 37;
        for (size_t n = 0; n < good.size(); n += step) {
            std::vector<unsigned char> b(good.begin(), good.begin() + static_cast<std::ptrdiff_t>(n));
            load_bad(b, "truncated", true);
            reseal(b);
            load_bad(b, "truncated and resealed", n + 8 < good.size());
        }
        // flipped bytes: caught by the checksum; resealed: by the structure
        // checks, or an equal-valued file (a flip in a value is a different
        // basis, refused as differing from the one in use)
        for (int it = 0; it < 3000; ++it) {
            std::vector<unsigned char> b = good;
            const int flips = static_cast<int>(range(1, 4));
            for (int f = 0; f < flips; ++f)
                b[static_cast<size_t>(range(0, static_cast<int64_t>(b.size()) - 9))] ^=
                    static_cast<unsigned char>(range(1, 255));
            load_bad(b, "flipped", true);
            reseal(b);
            load_bad(b, "flipped and resealed", false);
        }
        // header fields (width, depth at 8..15; sizes at 16..47) set to edge
        // values, resealed
        for (size_t off = 8; off + 4 <= 48; off += 4)
            for (int64_t v : {int64_t{-1}, int64_t{0}, int64_t{1}, int64_t{2}, int64_t{31}, int64_t{65536},
                              int64_t{1} << 31, int64_t{1} << 40, INT64_MAX}) {
                std::vector<unsigned char> b = good;
                if (off < 16) {
                    const int32_t x = static_cast<int32_t>(v);
                    std::memcpy(b.data() + off, &x, 4);
                } else {
                    if ((off - 16) % 8) continue;
                    std::memcpy(b.data() + off, &v, 8);
                }
                reseal(b);
                load_bad(b, "header field", false);
            }
        std::remove(bad.c_str());
        std::remove(path.c_str());
    }
    ++g_calls;
    expect(rpp_hip_lie_basis_load((dir + "/missing.bin").c_str(), nullptr, nullptr) == RPP_HIP_INVALID_ARGUMENT,
           "missing file");
}

}  // namespace

int main(int argc, char** argv) {
    int devices = 0;
    if (hipGetDeviceCount(&devices) == hipSuccess && devices > 0) {
        std::printf("CAPI SAN SKIPPED: a HIP device is present (fake pointers must never reach it)\n");
        return 0;
    }
    const std::string dir = argc > 1 ? argv[1] : ".";
    null_descriptors();
    random_descriptors(argc > 2 ? std::atoi(argv[2]) : 3000);
    lie_metadata(dir);
    std::printf("%s: %ld calls, %d failures\n", g_failures ? "CAPI SAN FAIL" : "CAPI SAN OK", g_calls, g_failures);
    return g_failures ? 1 : 0;
}
