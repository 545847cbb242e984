// Hall-basis metadata and its per-device CSR cache (lie.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "rpp_hip.h"

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
// The process-wide metadata of (width, depth): built on first use, or read
// from RPP_HIP_LIE_CACHE_DIR when that names a directory holding a valid file
// for it (a missing or invalid file is rebuilt and written there).
std::shared_ptr<const LieMetadata> lie_metadata(int width, int depth);

// The on-disk form (lie.cpp): save writes atomically (temporary file, then
// rename); load validates the checksum, the sizes and every index range and
// rebuilds the CSR forms, returning nullptr with a message on failure.
bool save_lie_metadata(const LieMetadata& md, const std::string& path, std::string* err);
std::shared_ptr<const LieMetadata> load_lie_metadata(const std::string& path, std::string* err);
// Makes md the process-wide metadata of its (width, depth); an identical one
// already in use is accepted, a different one is refused (the device copies
// may already be built from it).
int install_lie_metadata(std::shared_ptr<const LieMetadata> md, std::string* err);

// Padding groups after the last row of a packed matrix (DeviceCsr::packed4):
// the packed kernel's prefetch reads groups past a row's end, at indices it
// holds below the row's end group + kLiePackSlack, so never past the array.
constexpr int64_t kLiePackSlack = 16;

struct DeviceCsr {
    void* values = nullptr;
    int64_t* indices = nullptr;
    int64_t* offsets = nullptr;
    int64_t rows = 0, cols = 0, nnz = 0;
    // Packed form, present when the matrix has at most 255 distinct values,
    // fewer than 2^24 - 1 columns and fewer than 2^31 rows and entries: each
    // entry is one word (index into `table` << 24) | column, a quarter of the
    // bytes of the CSR arrays.  Every row is padded to a multiple of four
    // entries with the word (table_size - 1) << 24 | 0xffffff, whose table
    // value is -0.0 (a term that leaves any sum unchanged), and read as
    // uint4 groups.
    // kLiePackSlack padding groups follow the last row.
    uint4* packed4 = nullptr;
    void* table = nullptr;
    int32_t table_size = 0;  // 0: no packed form
    // With the packed form: {row, first group, end group, 0} of the rows in
    // the order the packed kernel assigns them to threads: with long rows
    // (an average of at least four groups, 16 entries), by entry count,
    // longest first, within each degree block (so a slice of whole degrees
    // is a permutation of itself); otherwise in row order.
    int4* rowdesc = nullptr;
    int32_t long_rows = 0;
};

// Device copies for the current device and dtype, uploaded on first use and
// kept for the life of the process.
template <typename T>
int device_lie_matrices(int width, int depth, const DeviceCsr** l2t, const DeviceCsr** t2l);

// out = M[row0 : row0 + rows] arg over a batch with the packed form of M
// (lie_spmv.hip): the slice's row r is out.begin + r, and its entries, which
// must name only columns [col0, col0 + arg.size), read arg.begin + c - col0.
// The slice must consist of whole degrees (DeviceCsr::rowdesc).
// lie_spmv_fits tells whether the packed kernel applies (packed form present,
// one argument segment fits in LDS).
template <typename T>
int launch_lie_spmv(int64_t n, const rpp_hip_vector_batch& out, const rpp_hip_vector_batch& arg, const DeviceCsr& m,
                    int64_t row0, int64_t rows, int64_t col0, hipStream_t stream);
template <typename T>
bool lie_spmv_fits(const DeviceCsr& m, const rpp_hip_vector_batch& arg);

}  // namespace rpp_hip
