// Hall-basis metadata and its per-device CSR cache (lie.cpp).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

namespace rpp_hip {

struct Csr {
    int64_t rows = 0;
    int64_t cols = 0;
    std::vector<double> values;
    std::vector<int64_t> indices;  // column of each entry
    std::vector<int64_t> offsets;  // rows + 1
};

struct LieMetadata {
    int width = 0;
    int depth = 0;
    int64_t true_size = 0;   // Hall keys incl. the unused slot 0
    int64_t tensor_dim = 0;
    std::vector<int64_t> pairs;         // 2 * true_size
    std::vector<int64_t> degree_begin;  // depth + 2
    // As the reference builders return them (CSC).
    std::vector<double> l2t_csc_values, t2l_csc_values;
    std::vector<int64_t> l2t_csc_indices, l2t_csc_offsets, t2l_csc_indices, t2l_csc_offsets;
    // CSR forms used by the kernels: L2T rows = tensor words, cols = key - 1;
    // T2L rows = key - 1, cols = tensor words.
    Csr l2t, t2l;
};

std::shared_ptr<const LieMetadata> build_lie_metadata(int width, int depth);
std::shared_ptr<const LieMetadata> lie_metadata(int width, int depth);

struct DeviceCsr {
    void* values = nullptr;
    int64_t* indices = nullptr;
    int64_t* offsets = nullptr;
    int64_t rows = 0, cols = 0, nnz = 0;
};

// Device copies for the current device and dtype, uploaded on first use and
// kept for the life of the process.
template <typename T>
int device_lie_matrices(int width, int depth, const DeviceCsr** l2t, const DeviceCsr** t2l);

}  // namespace rpp_hip
