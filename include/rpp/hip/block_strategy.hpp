// rpp::hip::strategies::BlockStrategy -- the MI355X (gfx950) execution
// strategy of RoughPathPrimitives' operation API, beside
// rpp::cpu::strategies::SingleThreadStrategy
// (cpu/single_thread/strategy.hpp:64-128) and in place of the CUDA
// rpp::gpu::strategies::BlockStrategy (gpu/block/strategy.hpp:212-365).
//
// Host-only header.  It is included together with the reference's own
// headers (<rpp/operations.hpp>, <rpp/views/views.hpp>, ...): callers keep
// calling rpp::ops::ft_mul(strategy, config, out, lhs, rhs, basis, n, beta)
// and friends unchanged, and every op below turns the batches into the
// descriptors of include/rpp_hip.h and calls librpp_hip.so.  Batch data are
// device pointers (hipMalloc'd by the caller); layouts::StrideLayout and
// layouts::NoStrideLayout batches are accepted, and so is the batch-
// interleaved layout rpp::hip::layouts::InterleavedLayout below (data an
// rpp::hip::InterleavedIterator); Scalar must equal Accum (float or double).  launch() is
// asynchronous on LaunchConfig::stream and returns rpp::Error<const char*>
// like the reference GPU strategy.
#ifndef RPP_HIP_BLOCK_STRATEGY_HPP
#define RPP_HIP_BLOCK_STRATEGY_HPP

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <tuple>
#include <type_traits>
#include <utility>

#include <rpp/architecture.hpp>
#include <rpp/basis/basis_pack.hpp>
#include <rpp/basis/tensor_basis.hpp>
#include <rpp/operations/basic/ft_adj_lmul.hpp>
#include <rpp/operations/basic/ft_adj_rmul.hpp>
#include <rpp/operations/basic/ft_fma.hpp>
#include <rpp/operations/basic/ft_inplace_fma.hpp>
#include <rpp/operations/basic/ft_inplace_mul.hpp>
#include <rpp/operations/basic/ft_mul.hpp>
#include <rpp/basis/lie_basis.hpp>
#include <rpp/operations/basic/left_hs_adj_lmul.hpp>
#include <rpp/operations/basic/left_hs_adj_rmul.hpp>
#include <rpp/operations/basic/left_hs_mul.hpp>
#include <rpp/operations/basic/lie_to_tensor.hpp>
#include <rpp/operations/basic/st_adj_mul.hpp>
#include <rpp/operations/basic/st_fma.hpp>
#include <rpp/operations/basic/st_inplace_fma.hpp>
#include <rpp/operations/basic/st_mul.hpp>
#include <rpp/operations/basic/tensor_add_identity.hpp>
#include <rpp/operations/basic/tensor_set_identity.hpp>
#include <rpp/operations/basic/tensor_antipode.hpp>
#include <rpp/operations/basic/tensor_generalised_antipode.hpp>
#include <rpp/operations/basic/tensor_pairing.hpp>
#include <rpp/operations/basic/tensor_reflect.hpp>
#include <rpp/operations/basic/tensor_to_lie.hpp>
#include <rpp/operations/intermediate/ft_exp.hpp>
#include <rpp/operations/intermediate/ft_fmexp.hpp>
#include <rpp/operations/intermediate/ft_log.hpp>
#include <rpp/operations/linalg/sparse_matrix_vector.hpp>
#include <rpp/operations/linalg/vector_add.hpp>
#include <rpp/operations/linalg/vector_assign.hpp>
#include <rpp/operations/linalg/vector_inplace_add.hpp>
#include <rpp/operations/linalg/vector_scalar_multiply.hpp>
#include <rpp/operations/linalg/vector_set_constant.hpp>
#include <rpp/sparse/matrix.hpp>
#include <rpp/support/error.hpp>
#include <rpp/views/batch.hpp>
#include <rpp/views/dense_tensor_view.hpp>

#include "rpp_hip.h"

namespace rpp::hip {

// A random-access iterator over one object of a batch-interleaved batch:
// element i is base[i * stride].  The reference's StridedIterator
// (support/strided_iterator.hpp:15-41) has this shape but cannot be
// instantiated as shipped (its second pair of operator< overloads, :200-210,
// redefines the first), so the interleaved layout uses this one.
template <typename T>
class InterleavedIterator {
    T* base_ = nullptr;
    std::ptrdiff_t stride_ = 1;

public:
    using value_type = std::remove_cv_t<T>;
    using difference_type = std::ptrdiff_t;
    using pointer = T*;
    using reference = T&;
    using iterator_category = std::random_access_iterator_tag;

    constexpr InterleavedIterator() noexcept = default;
    constexpr InterleavedIterator(T* base, difference_type stride) noexcept : base_(base), stride_(stride) {}

    constexpr T* base() const noexcept { return base_; }
    constexpr difference_type stride() const noexcept { return stride_; }
    constexpr reference operator*() const noexcept { return *base_; }
    constexpr reference operator[](difference_type i) const noexcept { return base_[i * stride_]; }
    constexpr InterleavedIterator& operator++() noexcept {
        base_ += stride_;
        return *this;
    }
    constexpr InterleavedIterator operator++(int) noexcept {
        InterleavedIterator t(*this);
        base_ += stride_;
        return t;
    }
    constexpr InterleavedIterator& operator--() noexcept {
        base_ -= stride_;
        return *this;
    }
    constexpr InterleavedIterator operator--(int) noexcept {
        InterleavedIterator t(*this);
        base_ -= stride_;
        return t;
    }
    constexpr InterleavedIterator& operator+=(difference_type k) noexcept {
        base_ += k * stride_;
        return *this;
    }
    constexpr InterleavedIterator& operator-=(difference_type k) noexcept {
        base_ -= k * stride_;
        return *this;
    }
    friend constexpr InterleavedIterator operator+(InterleavedIterator it, difference_type k) noexcept {
        return it += k;
    }
    friend constexpr InterleavedIterator operator+(difference_type k, InterleavedIterator it) noexcept {
        return it += k;
    }
    friend constexpr InterleavedIterator operator-(InterleavedIterator it, difference_type k) noexcept {
        return it -= k;
    }
    friend constexpr difference_type operator-(InterleavedIterator const& a, InterleavedIterator const& b) noexcept {
        return (a.base_ - b.base_) / a.stride_;
    }
    friend constexpr bool operator==(InterleavedIterator const& a, InterleavedIterator const& b) noexcept {
        return a.base_ == b.base_;
    }
    friend constexpr bool operator!=(InterleavedIterator const& a, InterleavedIterator const& b) noexcept {
        return a.base_ != b.base_;
    }
    friend constexpr bool operator<(InterleavedIterator const& a, InterleavedIterator const& b) noexcept {
        return a.base_ < b.base_;
    }
    friend constexpr bool operator>(InterleavedIterator const& a, InterleavedIterator const& b) noexcept {
        return a.base_ > b.base_;
    }
    friend constexpr bool operator<=(InterleavedIterator const& a, InterleavedIterator const& b) noexcept {
        return a.base_ <= b.base_;
    }
    friend constexpr bool operator>=(InterleavedIterator const& a, InterleavedIterator const& b) noexcept {
        return a.base_ >= b.base_;
    }
};

namespace layouts {

// The batch-interleaved ("[coefficient][object]") layout that every hip::block
// kernel reads with coalesced loads: coefficient c of object b sits at
// base[b + c * n] for n objects.  Batch data is an InterleavedIterator over the
// base pointer with stride n (the coefficient step inside one object); map()
// moves the base to object b and keeps the stride, so each view walks its
// object's coefficients n apart (a Layout of views/batch.hpp:27-68).
struct InterleavedLayout {
    template <typename T, typename Index>
    constexpr InterleavedIterator<T> map(InterleavedIterator<T> data, Index index) const noexcept {
        return InterleavedIterator<T>(data.base() + index, data.stride());
    }
};

}  // namespace layouts

// A batch of num_objects tensors of `basis` stored batch-interleaved at the
// device pointer data (dim * num_objects scalars), on the window
// [min_degree, max_degree]: the reference's make_tensor_batch
// (views/dense_tensor_view.hpp:38-66) over InterleavedLayout.
template <typename T, typename Basis, typename Degree>
auto make_interleaved_tensor_batch(T* data, std::ptrdiff_t num_objects, Basis const& basis, Degree min_degree,
                                   Degree max_degree) noexcept {
    return rpp::make_tensor_batch(InterleavedIterator<T>(data, num_objects), layouts::InterleavedLayout{}, basis,
                                  min_degree, max_degree);
}

namespace strategies {

struct LaunchConfig {
    void* stream = nullptr;  // hipStream_t
};

template <typename Accum_, typename Architecture_ = arch::DefaultArchitecture>
struct BlockStrategy {
    static_assert(std::is_same_v<Accum_, float> || std::is_same_v<Accum_, double>,
                  "hip::block computes in float or double");
    using Accum = Accum_;
    using Architecture = Architecture_;
    using Size = typename Architecture::Size;
    using Degree = typename Architecture::Degree;
    using Index = typename Architecture::Index;
    using Letter = typename Architecture::Letter;
    using Bitmask = typename Architecture::Bitmask;
    using LaunchConfig = hip::strategies::LaunchConfig;
    struct Context {};  // the device work-groups carry their own context

    // Same contract as SingleThreadStrategy::launch (cpu/single_thread/
    // strategy.hpp:64-90): each op specialisation below implements
    // hip_launch from the batches, bases and extras.
    template <typename Op, typename Batches, typename Bases, typename... Extras>
    RPP_NODISCARD Error<const char*> launch(LaunchConfig config, Batches batches, Bases bases,
                                            Index batch_size, Extras&&... extras) const noexcept {
        return Op::hip_launch(config, batches, bases, static_cast<int64_t>(batch_size),
                              std::forward<Extras>(extras)...);
    }
};

}  // namespace strategies

namespace detail {

template <typename S>
struct is_block : std::false_type {};
template <typename A, typename Arch>
struct is_block<strategies::BlockStrategy<A, Arch>> : std::true_type {};

inline Error<const char*> to_error(int rc) {
    if (rc == 0) return Error<const char*>::success();
    return Error<const char*>(static_cast<ErrorCode>(rc), rpp_hip_last_error());
}

// Step between consecutive objects of a batch, in elements.
template <typename Layout>
int64_t batch_stride_of(Layout const& layout) {
    if constexpr (std::is_same_v<Layout, rpp::layouts::NoStrideLayout> ||
                  std::is_same_v<Layout, hip::layouts::InterleavedLayout>) {
        return 1;
    } else {
        return static_cast<int64_t>(layout.stride_);
    }
}

template <typename Data>
struct is_strided : std::false_type {};
template <typename T>
struct is_strided<InterleavedIterator<T>> : std::true_type {};

// The device pointer of a batch's data (a raw pointer, or the base of an
// InterleavedIterator) and the step between consecutive coefficients of one
// object (1, or the iterator's stride).
template <typename Data>
auto base_pointer_of(Data const& data) {
    if constexpr (is_strided<Data>::value) {
        return data.base();
    } else {
        return data;
    }
}
template <typename Data>
int64_t coeff_stride_of(Data const& data) {
    if constexpr (is_strided<Data>::value) {
        return static_cast<int64_t>(data.stride());
    } else {
        return 1;
    }
}

template <typename Accum, typename Batch>
rpp_hip_tensor_batch tensor_of(Batch const& b) {
    using Data = std::decay_t<decltype(b.data())>;
    using Ptr = decltype(base_pointer_of(std::declval<Data>()));
    static_assert(std::is_pointer_v<Ptr>,
                  "hip::block batches hold raw device pointers (or InterleavedIterators over them)");
    static_assert(std::is_same_v<std::remove_cv_t<std::remove_pointer_t<Ptr>>, Accum>,
                  "hip::block needs Scalar == Accum");
    auto const& md = b.metadata();
    return rpp_hip_tensor_batch{const_cast<void*>(static_cast<const void*>(base_pointer_of(b.data()))),
                                batch_stride_of(std::decay_t<decltype(b.layout())>(b.layout())),
                                coeff_stride_of(b.data()), static_cast<int32_t>(std::get<1>(md)),
                                static_cast<int32_t>(std::get<2>(md))};
}

template <typename Bases>
rpp_hip_basis tensor_basis_of(Bases const& bases) {
    auto const& tb = basis::get_basis(basis::TensorBasisTag{}, bases);
    return rpp_hip_basis{static_cast<int32_t>(tb.width), static_cast<int32_t>(tb.depth)};
}

template <typename Accum>
constexpr bool is_f64 = std::is_same_v<Accum, double>;

#define RPP_HIP_CALL(NAME, DESC, STREAM) \
    (is_f64<Accum> ? NAME##_f64(&(DESC), (STREAM)) : NAME##_f32(&(DESC), (STREAM)))

}  // namespace detail
}  // namespace rpp::hip

namespace rpp::ops {

#define RPP_HIP_OP_COMMON                                                  \
    using Strategy = hip::strategies::BlockStrategy<Accum_, Arch_>;       \
    using Accum = Accum_;                                                  \
    using Context = typename Strategy::Context;                            \
    using LaunchConfig = typename Strategy::LaunchConfig;                  \
    using BaseOperation<Strategy>::BaseOperation;                          \
    static constexpr bool is_implemented = true

// ft_mul (operations/basic/ft_mul.hpp:53-79)
template <typename Accum_, typename Arch_>
class FTMul<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_mul_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                              tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b)),
                              static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_mul, d, cfg.stream));
    }
};

// ft_fma (operations/basic/ft_fma.hpp:54-83)
template <typename Accum_, typename Arch_>
class FTFma<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum alpha = Accum{1}, Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_fma_desc d{tensor_basis_of(bases),           n,
                              tensor_of<Accum>(std::get<0>(b)), tensor_of<Accum>(std::get<1>(b)),
                              tensor_of<Accum>(std::get<2>(b)), tensor_of<Accum>(std::get<3>(b)),
                              static_cast<double>(alpha),       static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_fma, d, cfg.stream));
    }
};

// ft_inplace_fma<FMAType> (operations/basic/ft_inplace_fma.hpp:64-100)
template <typename Accum_, typename Arch_, FTInplaceFMAType FMAType>
class FTInplaceFma<hip::strategies::BlockStrategy<Accum_, Arch_>, FMAType>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum alpha = Accum{1}, Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        constexpr int32_t variant = FMAType == FTInplaceFMAType::AEqualsBCPlusA   ? 231
                                    : FMAType == FTInplaceFMAType::AEqualsABPlusC ? 123
                                                                                  : 213;
        rpp_hip_ft_inplace_fma_desc d{tensor_basis_of(bases),           n,
                                      variant,                          tensor_of<Accum>(std::get<0>(b)),
                                      tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b)),
                                      static_cast<double>(alpha),       static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_inplace_fma, d, cfg.stream));
    }
};

// ft_inplace_mul (operations/basic/ft_inplace_mul.hpp:38-69)
template <typename Accum_, typename Arch_>
class FTInplaceMul<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_inplace_mul_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                      tensor_of<Accum>(std::get<1>(b)), static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_inplace_mul, d, cfg.stream));
    }
};

// ft_exp / ft_log (operations/intermediate/ft_exp.hpp:89-112, ft_log.hpp:91-114)
template <typename Accum_, typename Arch_>
class FTExp<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_unary_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                tensor_of<Accum>(std::get<1>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_exp, d, cfg.stream));
    }
};

template <typename Accum_, typename Arch_>
class FTLog<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_unary_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                tensor_of<Accum>(std::get<1>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_log, d, cfg.stream));
    }
};

// ft_fmexp (operations/intermediate/ft_fmexp.hpp:96-121)
template <typename Accum_, typename Arch_>
class FTFMExp<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_fmexp_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_fmexp, d, cfg.stream));
    }
};

// tensor_antipode / tensor_reflect (operations/basic/tensor_antipode.hpp:15-39,
// tensor_reflect.hpp:15-39; TensorGeneralisedAntipode,
// tensor_generalised_antipode.hpp:17-81)
template <typename Accum_, typename Arch_, TensorAntipodeSigningPolicy Policy>
class TensorGeneralisedAntipode<hip::strategies::BlockStrategy<Accum_, Arch_>, Policy>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_unary_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                tensor_of<Accum>(std::get<1>(b))};
        if constexpr (Policy == TensorAntipodeSigningPolicy::SignByDegree)
            return to_error(RPP_HIP_CALL(rpp_hip_tensor_antipode, d, cfg.stream));
        else
            return to_error(RPP_HIP_CALL(rpp_hip_tensor_reflect, d, cfg.stream));
    }
};

// tensor_pairing (operations/basic/tensor_pairing.hpp:47-73); the result
// batch is make_scalar_batch(device pointer).
template <typename Accum_, typename Arch_>
class TensorPairing<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        auto const& out = std::get<0>(b);
        rpp_hip_scalar_batch s{const_cast<void*>(static_cast<const void*>(out.data())),
                               batch_stride_of(std::decay_t<decltype(out.layout())>(out.layout()))};
        rpp_hip_tensor_pairing_desc d{tensor_basis_of(bases), n, s, tensor_of<Accum>(std::get<1>(b)),
                                      tensor_of<Accum>(std::get<2>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_tensor_pairing, d, cfg.stream));
    }
};

// ft_adj_lmul / ft_adj_rmul (operations/basic/ft_adj_lmul.hpp:49-73,
// ft_adj_rmul.hpp:47-71)
template <typename Accum_, typename Arch_>
class FTAdjLMul<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_adjoint_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                  tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_adj_lmul, d, cfg.stream));
    }
};

template <typename Accum_, typename Arch_>
class FTAdjRMul<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_adjoint_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                  tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_ft_adj_rmul, d, cfg.stream));
    }
};

namespace hip_detail {

// Graded vector batch -> SpMV vector batch: view-relative indexing (R9):
// entry 0 is view.begin() = degree_begin[min_degree] of the view's basis.
template <typename Accum, typename Batch, typename Basis>
rpp_hip_vector_batch vector_of(Batch const& b, Basis const& basis) {
    auto const& md = b.metadata();
    const auto mn = std::get<1>(md);
    const auto mx = std::get<2>(md);
    const int64_t begin = static_cast<int64_t>(basis.start_of_degree(mn));
    const int64_t end = static_cast<int64_t>(basis.end_of_degree(mx));
    return rpp_hip_vector_batch{const_cast<void*>(static_cast<const void*>(hip::detail::base_pointer_of(b.data()))),
                                hip::detail::batch_stride_of(std::decay_t<decltype(b.layout())>(b.layout())),
                                hip::detail::coeff_stride_of(b.data()), begin, end - begin};
}

template <typename Accum, typename Matrix>
rpp_hip_sparse_matrix matrix_of(Matrix const& m) {
    static_assert(sizeof(*m.indices()) == 8 && sizeof(*m.offsets()) == 8,
                  "hip::block sparse matrices use 64-bit indices and offsets");
    constexpr bool csc = sparse::matrix_format_v<Matrix> == sparse::MatrixFormat::CSC;
    return rpp_hip_sparse_matrix{csc ? RPP_HIP_CSC : RPP_HIP_CSR,
                                 static_cast<const void*>(&*m.data()),
                                 reinterpret_cast<const int64_t*>(&*m.indices()),
                                 reinterpret_cast<const int64_t*>(&*m.offsets()),
                                 static_cast<int64_t>(m.nnz()),
                                 static_cast<int64_t>(m.rows()),
                                 static_cast<int64_t>(m.cols())};
}

}  // namespace hip_detail

// sparse_matrix_vector_product (operations/linalg/sparse_matrix_vector.hpp
// :57-96), matrix arrays in device memory.
template <typename Accum_, typename Arch_, sparse::MatrixFormat Format>
class SparseMatrixVectorProduct<hip::strategies::BlockStrategy<Accum_, Arch_>, Format>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases, typename Matrix>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Matrix const& matrix, Accum alpha = Accum{1}) noexcept {
        using namespace hip::detail;
        auto const& ob = basis::get_basis(std::get<0>(std::get<0>(b).metadata()), bases);
        auto const& ab = basis::get_basis(std::get<0>(std::get<1>(b).metadata()), bases);
        rpp_hip_spmv_desc d{n, hip_detail::vector_of<Accum>(std::get<0>(b), ob),
                            hip_detail::vector_of<Accum>(std::get<1>(b), ab),
                            hip_detail::matrix_of<Accum>(matrix), static_cast<double>(alpha)};
        return to_error(RPP_HIP_CALL(rpp_hip_sparse_matrix_vector_product, d, cfg.stream));
    }
};

// tensor_to_lie / lie_to_tensor (operations/basic/tensor_to_lie.hpp:65-103,
// lie_to_tensor.hpp:70-108) with the caller's matrix, CSR or CSC: the
// library recognises the Hall T2L / L2T matrices of the basis (the output of
// the reference's make_t2l_matrix / make_l2t_matrix) and serves them with its
// Lie kernel; any other matrix takes the deterministic sparse product.  Both
// sum every row in the reference's order.
template <typename Accum_, typename Arch_, T2LImplementationType Impl>
class TensorToLie<hip::strategies::BlockStrategy<Accum_, Arch_>, Impl>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases, typename Matrix>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Matrix const& matrix) noexcept {
        using namespace hip::detail;
        rpp_hip_lie_matrix_convert_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<1>(b)),
                                          tensor_of<Accum>(std::get<0>(b)), hip_detail::matrix_of<Accum>(matrix)};
        return to_error(RPP_HIP_CALL(rpp_hip_tensor_to_lie_matrix, d, cfg.stream));
    }
};

template <typename Accum_, typename Arch_, L2TImplementationType Impl>
class LieToTensor<hip::strategies::BlockStrategy<Accum_, Arch_>, Impl>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases, typename Matrix>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Matrix const& matrix) noexcept {
        using namespace hip::detail;
        rpp_hip_lie_matrix_convert_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                          tensor_of<Accum>(std::get<1>(b)), hip_detail::matrix_of<Accum>(matrix)};
        return to_error(RPP_HIP_CALL(rpp_hip_lie_to_tensor_matrix, d, cfg.stream));
    }
};

// The shuffle family (operations/basic/st_*.hpp, left_hs_*.hpp).
template <typename Accum_, typename Arch_>
class STMul<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_mul_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                              tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b)),
                              static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_st_mul, d, cfg.stream));
    }
};

template <typename Accum_, typename Arch_>
class STFma<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum alpha = Accum{1}, Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_fma_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                              tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b)),
                              tensor_of<Accum>(std::get<3>(b)), static_cast<double>(alpha),
                              static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_st_fma, d, cfg.stream));
    }
};

template <typename Accum_, typename Arch_>
class STInplaceFma<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum alpha = Accum{1}, Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_st_inplace_fma_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                      tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b)),
                                      static_cast<double>(alpha), static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_st_inplace_fma, d, cfg.stream));
    }
};

template <typename Accum_, typename Arch_>
class STAdjMul<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_ft_adjoint_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),
                                  tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_st_adj_mul, d, cfg.stream));
    }
};

// left_hs_mul, left_hs_adj_lmul, left_hs_adj_rmul: (out, integrator,
// integrand, beta); the adjoints scale by beta as the reference's GPU
// operators do.
#define RPP_HIP_HALF_SHUFFLE_OP(CLASS, ENTRY)                                                               \
    template <typename Accum_, typename Arch_>                                                              \
    class CLASS<hip::strategies::BlockStrategy<Accum_, Arch_>>                                              \
        : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {                             \
    public:                                                                                                 \
        RPP_HIP_OP_COMMON;                                                                                  \
        template <typename Batches, typename Bases>                                                         \
        static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,        \
                                             int64_t n, Accum beta = Accum{1}) noexcept {                   \
            using namespace hip::detail;                                                                    \
            rpp_hip_half_shuffle_desc d{tensor_basis_of(bases), n, tensor_of<Accum>(std::get<0>(b)),        \
                                        tensor_of<Accum>(std::get<1>(b)), tensor_of<Accum>(std::get<2>(b)), \
                                        static_cast<double>(beta)};                                         \
            return to_error(RPP_HIP_CALL(ENTRY, d, cfg.stream));                                            \
        }                                                                                                   \
    };

RPP_HIP_HALF_SHUFFLE_OP(LeftHSMul, rpp_hip_left_hs_mul)
RPP_HIP_HALF_SHUFFLE_OP(LeftHSAdjLMul, rpp_hip_left_hs_adj_lmul)
RPP_HIP_HALF_SHUFFLE_OP(LeftHSAdjRMul, rpp_hip_left_hs_adj_rmul)
#undef RPP_HIP_HALF_SHUFFLE_OP

namespace hip_detail {

// The grading of a batch: the basis its metadata names, looked up in the
// pack (tensor or Lie; the linear operations use only its degree offsets).
template <typename Batch, typename Bases>
rpp_hip_graded_basis graded_of(Batch const& b, Bases const& bases) {
    auto const& gb = basis::get_basis(std::get<0>(b.metadata()), bases);
    using Tag = typename std::decay_t<decltype(gb)>::Tag;
    static_assert(std::is_same_v<Tag, basis::TensorBasisTag> || std::is_same_v<Tag, basis::LieBasisTag>,
                  "hip::block linear operations take tensor or Lie gradings");
    return rpp_hip_graded_basis{std::is_same_v<Tag, basis::LieBasisTag> ? RPP_HIP_LIE_BASIS : RPP_HIP_TENSOR_BASIS,
                                static_cast<int32_t>(gb.width), static_cast<int32_t>(gb.depth)};
}

}  // namespace hip_detail

// vector_assign (operations/linalg/vector_assign.hpp:36-64)
template <typename Accum_, typename Arch_>
class VectorAssign<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases,
                                         int64_t n) noexcept {
        using namespace hip::detail;
        rpp_hip_vector_assign_desc d{hip_detail::graded_of(std::get<0>(b), bases), n,
                                     tensor_of<Accum>(std::get<0>(b)), tensor_of<Accum>(std::get<1>(b))};
        return to_error(RPP_HIP_CALL(rpp_hip_vector_assign, d, cfg.stream));
    }
};

// vector_add (operations/linalg/vector_add.hpp:40-78)
template <typename Accum_, typename Arch_>
class VectorAdd<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum alpha = Accum{1}, Accum beta = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_vector_add_desc d{hip_detail::graded_of(std::get<0>(b), bases), n,
                                  tensor_of<Accum>(std::get<0>(b)), tensor_of<Accum>(std::get<1>(b)),
                                  tensor_of<Accum>(std::get<2>(b)), static_cast<double>(alpha),
                                  static_cast<double>(beta)};
        return to_error(RPP_HIP_CALL(rpp_hip_vector_add, d, cfg.stream));
    }
};

// vector_inplace_add (operations/linalg/vector_inplace_add.hpp:40-77)
template <typename Accum_, typename Arch_>
class VectorInplaceAdd<hip::strategies::BlockStrategy<Accum_, Arch_>>
    : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {
public:
    RPP_HIP_OP_COMMON;
    template <typename Batches, typename Bases>
    static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, int64_t n,
                                         Accum alpha = Accum{1}) noexcept {
        using namespace hip::detail;
        rpp_hip_vector_inplace_add_desc d{hip_detail::graded_of(std::get<0>(b), bases), n,
                                          tensor_of<Accum>(std::get<0>(b)), tensor_of<Accum>(std::get<1>(b)),
                                          static_cast<double>(alpha)};
        return to_error(RPP_HIP_CALL(rpp_hip_vector_inplace_add, d, cfg.stream));
    }
};

// One batch and one scalar: vector_scalar_multiply, vector_set_constant,
// tensor_set_identity, tensor_add_identity (operations/linalg/
// vector_scalar_multiply.hpp:37-71, vector_set_constant.hpp:38-72,
// operations/basic/tensor_set_identity.hpp:38-73, tensor_add_identity.hpp:37-71)
#define RPP_HIP_SCALAR_OP(CLASS, ENTRY)                                                              \
    template <typename Accum_, typename Arch_>                                                       \
    class CLASS<hip::strategies::BlockStrategy<Accum_, Arch_>>                                       \
        : public BaseOperation<hip::strategies::BlockStrategy<Accum_, Arch_>> {                      \
    public:                                                                                          \
        RPP_HIP_OP_COMMON;                                                                           \
        template <typename Batches, typename Bases>                                                  \
        static Error<const char*> hip_launch(LaunchConfig cfg, Batches const& b, Bases const& bases, \
                                             int64_t n, Accum scalar) noexcept {                     \
            using namespace hip::detail;                                                             \
            rpp_hip_vector_scalar_desc d{hip_detail::graded_of(std::get<0>(b), bases), n,            \
                                         tensor_of<Accum>(std::get<0>(b)), static_cast<double>(scalar)}; \
            return to_error(RPP_HIP_CALL(ENTRY, d, cfg.stream));                                     \
        }                                                                                            \
    };

RPP_HIP_SCALAR_OP(VectorScalarMultiply, rpp_hip_vector_scalar_multiply)
RPP_HIP_SCALAR_OP(VectorSetConstant, rpp_hip_vector_set_constant)
RPP_HIP_SCALAR_OP(TensorSetIdentity, rpp_hip_tensor_set_identity)
RPP_HIP_SCALAR_OP(TensorAddIdentity, rpp_hip_tensor_add_identity)
#undef RPP_HIP_SCALAR_OP

#undef RPP_HIP_OP_COMMON

}  // namespace rpp::ops

#undef RPP_HIP_CALL

#endif  // RPP_HIP_BLOCK_STRATEGY_HPP
