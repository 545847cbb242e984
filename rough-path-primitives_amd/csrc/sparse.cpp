// Caller-supplied sparse matrices: the CSC -> CSR preparation of the sparse
// products and the recognition of the Hall T2L/L2T matrices (sparse.hpp).
//
// The reference's CPU CSC product (cpu/single_thread/operations/linalg/
// sparse_matrix_vector.hpp:133-163) zeroes the out view, then for every
// column c in ascending order adds value * (alpha * arg[c]) to each row the
// column names, in entry order.  Row r therefore receives its terms in column
// order, and within one column in entry order; a CSR matrix whose row r lists
// those entries in that order, summed from zero one rounded step at a time,
// gives the same bits.  That CSR is built here once per matrix by a stable
// counting transposition, with no atomics anywhere.
#include "sparse.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "lie.hpp"

namespace rpp_hip {

namespace {

using Key = std::tuple<int, int, int, const void*, const int64_t*, const int64_t*, int64_t, int64_t, int64_t>;

struct Entry {
    int device = 0;
    // host copy of the matrix in the reference's per-row order (values in T,
    // as bytes), made on first need
    bool have_host = false;
    std::vector<unsigned char> values;
    std::vector<int64_t> indices, offsets;
    // library-owned device CSR of a CSC matrix
    void* d_values = nullptr;
    int64_t* d_indices = nullptr;
    int64_t* d_offsets = nullptr;
    // (width, depth, to_lie) -> equals the Hall matrix
    std::map<std::tuple<int, int, bool>, bool> hall;
};

std::mutex g_mu;
std::map<Key, std::unique_ptr<Entry>> g_matrices;

Key key_of(const rpp_hip_sparse_matrix& m, int device, int elem) {
    return Key{device, elem, static_cast<int>(m.format), m.values, m.indices, m.offsets, m.nnz, m.rows, m.cols};
}

// Reads the caller's arrays back and stores them in e in per-row order.
template <typename T>
int load_host(const rpp_hip_sparse_matrix& m, Entry& e, hipStream_t stream) {
    const int64_t outer = m.format == RPP_HIP_CSR ? m.rows : m.cols;
    std::vector<T> vals(static_cast<size_t>(m.nnz));
    std::vector<int64_t> idx(static_cast<size_t>(m.nnz)), off(static_cast<size_t>(outer) + 1);
    if (int rc = check_hip(hipMemcpyAsync(off.data(), m.offsets, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                                          stream),
                           "reading matrix offsets"))
        return rc;
    if (m.nnz > 0) {
        if (int rc = check_hip(hipMemcpyAsync(vals.data(), m.values, vals.size() * sizeof(T), hipMemcpyDeviceToHost,
                                              stream),
                               "reading matrix values"))
            return rc;
        if (int rc = check_hip(hipMemcpyAsync(idx.data(), m.indices, idx.size() * sizeof(int64_t),
                                              hipMemcpyDeviceToHost, stream),
                               "reading matrix indices"))
            return rc;
    }
    if (int rc = check_hip(hipStreamSynchronize(stream), "synchronising for the matrix read-back")) return rc;
    if (off[0] != 0 || off[static_cast<size_t>(outer)] != m.nnz)
        return fail(RPP_HIP_INVALID_ARGUMENT, "matrix offsets must run from 0 to nnz");
    for (int64_t i = 0; i < outer; ++i)
        if (off[i + 1] < off[i]) return fail(RPP_HIP_INVALID_ARGUMENT, "matrix offsets must be non-decreasing");
    const int64_t inner = m.format == RPP_HIP_CSR ? m.cols : m.rows;
    for (int64_t v : idx)
        if (v < 0 || v >= inner) return fail(RPP_HIP_INVALID_ARGUMENT, "matrix index out of range");
    std::vector<T> rv;
    if (m.format == RPP_HIP_CSR) {
        rv = std::move(vals);
        e.indices = std::move(idx);
        e.offsets = std::move(off);
    } else {
        // stable counting transposition: row r collects its entries in the
        // order the column scatter reaches it
        e.offsets.assign(static_cast<size_t>(m.rows) + 1, 0);
        for (int64_t r : idx) ++e.offsets[static_cast<size_t>(r) + 1];
        for (int64_t r = 0; r < m.rows; ++r) e.offsets[r + 1] += e.offsets[r];
        std::vector<int64_t> fill(e.offsets.begin(), e.offsets.end() - 1);
        rv.resize(static_cast<size_t>(m.nnz));
        e.indices.resize(static_cast<size_t>(m.nnz));
        for (int64_t c = 0; c < m.cols; ++c)
            for (int64_t k = off[c]; k < off[c + 1]; ++k) {
                const int64_t dst = fill[static_cast<size_t>(idx[k])]++;
                rv[dst] = vals[k];
                e.indices[dst] = c;
            }
    }
    e.values.resize(rv.size() * sizeof(T));
    if (!rv.empty()) std::memcpy(e.values.data(), rv.data(), e.values.size());
    e.have_host = true;
    return 0;
}

Entry* find_or_add(const rpp_hip_sparse_matrix& m, int elem, int* device) {
    if (hipGetDevice(device) != hipSuccess) return nullptr;
    auto& slot = g_matrices[key_of(m, *device, elem)];
    if (!slot) {
        slot = std::make_unique<Entry>();
        slot->device = *device;
    }
    return slot.get();
}

template <typename T>
int upload(Entry& e, int64_t rows, int64_t nnz) {
    const size_t vb = std::max<size_t>(static_cast<size_t>(nnz) * sizeof(T), 1);
    const size_t ib = std::max<size_t>(static_cast<size_t>(nnz) * sizeof(int64_t), 1);
    const size_t ob = (static_cast<size_t>(rows) + 1) * sizeof(int64_t);
    if (hipMalloc(&e.d_values, vb) != hipSuccess || hipMalloc(reinterpret_cast<void**>(&e.d_indices), ib) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&e.d_offsets), ob) != hipSuccess)
        return fail(RPP_HIP_OUT_OF_RESOURCES, "allocating the CSR form of a CSC matrix");
    if (nnz > 0) {
        if (int rc = check_hip(hipMemcpy(e.d_values, e.values.data(), e.values.size(), hipMemcpyHostToDevice),
                               "uploading the CSR form"))
            return rc;
        if (int rc = check_hip(hipMemcpy(e.d_indices, e.indices.data(), e.indices.size() * sizeof(int64_t),
                                         hipMemcpyHostToDevice),
                               "uploading the CSR form"))
            return rc;
    }
    return check_hip(hipMemcpy(e.d_offsets, e.offsets.data(), ob, hipMemcpyHostToDevice), "uploading the CSR form");
}

void release(Entry& e) {
    if (!e.d_values && !e.d_indices && !e.d_offsets) return;
    int cur = 0;
    const bool switched = hipGetDevice(&cur) == hipSuccess && cur != e.device && hipSetDevice(e.device) == hipSuccess;
    (void)hipDeviceSynchronize();  // kernels queued with these arrays finish first
    if (e.d_values) (void)hipFree(e.d_values);
    if (e.d_indices) (void)hipFree(e.d_indices);
    if (e.d_offsets) (void)hipFree(e.d_offsets);
    e.d_values = nullptr;
    e.d_indices = e.d_offsets = nullptr;
    if (switched) (void)hipSetDevice(cur);
}

}  // namespace

template <typename T>
int prepare_matrix(const rpp_hip_sparse_matrix& m, PreparedMatrix* out, hipStream_t stream) {
    if (m.format == RPP_HIP_CSR) {
        out->csr = m;
        out->csc = false;
        return 0;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    int device = 0;
    Entry* e = find_or_add(m, sizeof(T), &device);
    if (!e) return fail(RPP_HIP_INTERNAL, "hipGetDevice failed");
    if (!e->d_offsets) {
        if (!e->have_host) RPP_TRY(load_host<T>(m, *e, stream));
        if (int rc = upload<T>(*e, m.rows, m.nnz)) {
            release(*e);
            return rc;
        }
    }
    out->csr = rpp_hip_sparse_matrix{RPP_HIP_CSR, e->d_values, e->d_indices, e->d_offsets, m.nnz, m.rows, m.cols};
    out->csc = true;
    return 0;
}

template <typename T>
int matrix_is_hall(const rpp_hip_sparse_matrix& m, int width, int depth, bool to_lie, bool* result,
                   hipStream_t stream) {
    *result = false;
    auto md = lie_metadata(width, depth);
    const Csr& h = to_lie ? md->t2l : md->l2t;
    if (m.rows != h.rows || m.cols != h.cols || m.nnz != static_cast<int64_t>(h.values.size())) return 0;
    std::lock_guard<std::mutex> lock(g_mu);
    int device = 0;
    Entry* e = find_or_add(m, sizeof(T), &device);
    if (!e) return fail(RPP_HIP_INTERNAL, "hipGetDevice failed");
    const auto key = std::make_tuple(width, depth, to_lie);
    auto it = e->hall.find(key);
    if (it != e->hall.end()) {
        *result = it->second;
        return 0;
    }
    if (!e->have_host) RPP_TRY(load_host<T>(m, *e, stream));
    const T* v = reinterpret_cast<const T*>(e->values.data());
    bool same = e->offsets == h.offsets && e->indices == h.indices;
    for (size_t i = 0; same && i < h.values.size(); ++i) same = v[i] == static_cast<T>(h.values[i]);
    e->hall[key] = same;
    *result = same;
    return 0;
}

int forget_matrix(const void* values) {
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto it = g_matrices.begin(); it != g_matrices.end();) {
        if (std::get<3>(it->first) == values) {
            release(*it->second);
            it = g_matrices.erase(it);
        } else {
            ++it;
        }
    }
    return 0;
}

template int prepare_matrix<float>(const rpp_hip_sparse_matrix&, PreparedMatrix*, hipStream_t);
template int prepare_matrix<double>(const rpp_hip_sparse_matrix&, PreparedMatrix*, hipStream_t);
template int matrix_is_hall<float>(const rpp_hip_sparse_matrix&, int, int, bool, bool*, hipStream_t);
template int matrix_is_hall<double>(const rpp_hip_sparse_matrix&, int, int, bool, bool*, hipStream_t);

}  // namespace rpp_hip
