// par.hpp -- the structured-array containers and algorithms the reference's drivers and spatial
// API are written against (parallalg/parallalg.h, parallalg/array.h, parallalg/algorithm.h):
// N-dimensional indices, shapes and strides on a primal (node) or dual (cell) grid, the move-only
// row-major Array, and fill / copy / transform / for_each / for_each_idx / transform_reduce with
// a serial or OpenMP execution policy.  Host containers: the GPU path copies them to and from the
// device through the C-ABI (include/hcamd.h).
#pragma once

#include <array>
#include <cassert>
#include <cstddef>
#include <functional>
#include <type_traits>
#include <utility>
#include <vector>

namespace par {

// grid an array lives on (parallalg/parallalg.h:12-26)
enum struct GridType { Primal, Dual };
inline constexpr GridType Primal = GridType::Primal;
inline constexpr GridType Dual = GridType::Dual;

// N-dimensional index (parallalg.h:28-60)
template <int NDIM, GridType GRIDTYPE = Primal> struct Idx {
    std::array<std::size_t, NDIM> idxs;
    const std::size_t& operator[](const unsigned int i) const { return idxs[i]; }
};
template <int NDIM> using PrimalIdx = Idx<NDIM, Primal>;
template <int NDIM> using DualIdx = Idx<NDIM, Dual>;
using PrimalIdx1 = Idx<1, Primal>;
using PrimalIdx2 = Idx<2, Primal>;
using PrimalIdx3 = Idx<3, Primal>;
using DualIdx1 = Idx<1, Dual>;
using DualIdx2 = Idx<2, Dual>;
using DualIdx3 = Idx<3, Dual>;

// stencil offset, extent and place values (parallalg/array.h:21-60)
template <int NDIM, GridType GRIDTYPE = Primal> struct Offset {
    std::array<int, NDIM> offsets;
    const int& operator[](const unsigned int i) const { return offsets[i]; }
};
template <int NDIM, GridType GRIDTYPE = Primal> struct Shape {
    std::array<std::size_t, NDIM> shape;
    const std::size_t& operator[](const unsigned int i) const { return shape[i]; }
};
template <int NDIM, GridType GRIDTYPE = Primal> struct Stride {
    std::array<std::size_t, NDIM> strides;
    Stride(const Shape<NDIM, GRIDTYPE>& s)
    {
        strides[NDIM - 1] = 1;
        for (int i = NDIM - 2; i >= 0; --i) strides[i] = strides[i + 1] * s[i + 1];
    }
    const std::size_t& operator[](const unsigned int i) const { return strides[i]; }
};
template <int NDIM> using PrimalShape = Shape<NDIM, Primal>;
template <int NDIM> using DualShape = Shape<NDIM, Dual>;
using PrimalShape1 = Shape<1, Primal>;
using PrimalShape2 = Shape<2, Primal>;
using PrimalShape3 = Shape<3, Primal>;
using DualShape1 = Shape<1, Dual>;
using DualShape2 = Shape<2, Dual>;
using DualShape3 = Shape<3, Dual>;
template <int NDIM> using PrimalOffset = Offset<NDIM, Primal>;
template <int NDIM> using DualOffset = Offset<NDIM, Dual>;

template <int NDIM, GridType G> Idx<NDIM, G> operator+(const Idx<NDIM, G>& idx, const Offset<NDIM, G>& off)
{
    Idx<NDIM, G> r(idx);
    for (int i = 0; i < NDIM; ++i) r.idxs[i] += off[i];
    return r;
}
template <int NDIM, GridType G> bool operator==(const Shape<NDIM, G>& a, const Shape<NDIM, G>& b)
{
    bool same = true;
    for (int i = 0; i < NDIM; ++i) same = same && a[i] == b[i];
    return same;
}
template <int NDIM, GridType G> std::size_t operator*(const Stride<NDIM, G>& s, const Idx<NDIM, G>& idx)
{
    std::size_t k = 0;
    for (int m = 0; m < NDIM; ++m) k += s[m] * idx[m];
    return k;
}
template <int NDIM, GridType G> std::size_t operator*(const Idx<NDIM, G>& idx, const Stride<NDIM, G>& s) { return s * idx; }

// node extents from cell extents and back (array.h:209-229)
template <int NDIM> PrimalShape<NDIM> primalShape(const DualShape<NDIM>& d)
{
    PrimalShape<NDIM> p;
    for (int i = 0; i < NDIM; ++i) p.shape[i] = d[i] + 1;
    return p;
}
template <int NDIM> DualShape<NDIM> dualShape(const PrimalShape<NDIM>& p)
{
    DualShape<NDIM> d;
    for (int i = 0; i < NDIM; ++i) d.shape[i] = p[i] - 1;
    return d;
}
template <int NDIM, GridType G> std::size_t length(const Shape<NDIM, G>& s)
{
    std::size_t l = 1;
    for (int i = 0; i < NDIM; ++i) l *= s[i];
    return l;
}

enum struct ArraySizing { Fixed, Dynamic };
constexpr ArraySizing FixedSize = ArraySizing::Fixed;
constexpr ArraySizing DynamicSize = ArraySizing::Dynamic;

// N-dimensional row-major array, last index fastest, move-only (array.h:267-349)
template <typename ElemT, int NDIM, GridType GRIDTYPE = Primal, ArraySizing SIZING = FixedSize> struct Array {
    using ElemType = ElemT;
    static constexpr int nDim = NDIM;
    static constexpr GridType gridType = GRIDTYPE;
    static constexpr ArraySizing arraySizing = SIZING;
    using IdxType = Idx<NDIM, GRIDTYPE>;
    using OffsetType = Offset<NDIM, GRIDTYPE>;
    using ShapeType = Shape<NDIM, GRIDTYPE>;
    using StrideType = Stride<NDIM, GRIDTYPE>;

  private:
    ShapeType shape_array;
    StrideType stride_array;
    std::vector<ElemT> elems_array;

  public:
    Array() = delete;
    Array(const Array&) = delete;
    Array(Array&&) = default;
    Array& operator=(const Array&) = delete;
    Array& operator=(Array&&) = default;

    template <typename... Ints>
        requires(sizeof...(Ints) == NDIM && (std::is_integral_v<Ints> && ...))
    Array(const Ints... is)
        : shape_array{{static_cast<std::size_t>(is)...}}, stride_array(shape_array),
          elems_array((static_cast<std::size_t>(is) * ...))
    {
    }
    Array(const ShapeType& s) : shape_array(s), stride_array(s), elems_array(length(s)) {}
    Array(const ShapeType& s, const ElemT& val) : shape_array(s), stride_array(s), elems_array(length(s), val) {}

    std::size_t flattened_length() const { return elems_array.size(); }
    const ElemT& operator()(const IdxType& idx) const { return elems_array[stride_array * idx]; }
    ElemT& operator()(const IdxType& idx) { return elems_array[stride_array * idx]; }
    const ElemT& flatten(const std::size_t i) const { return elems_array[i]; }
    ElemT& flatten(const std::size_t i) { return elems_array[i]; }
    const std::vector<ElemT>& flatten() const { return elems_array; }
    const std::size_t& shape(const unsigned int i) const { return shape_array[i]; }
    const std::size_t& stride(const unsigned int i) const { return stride_array[i]; }
    const ShapeType& shape() const { return shape_array; }
    const StrideType& stride() const { return stride_array; }
    // contiguous storage, for copies to and from the device
    ElemT* data() { return elems_array.data(); }
    const ElemT* data() const { return elems_array.data(); }

    void resize(const ShapeType& s)
        requires(SIZING == DynamicSize)
    {
        shape_array = s;
        stride_array = StrideType(s);
        elems_array.resize(length(s));
    }
};

template <typename ElemT, int NDIM, ArraySizing S = FixedSize> using PrimalArray = Array<ElemT, NDIM, Primal, S>;
template <typename ElemT, int NDIM, ArraySizing S = FixedSize> using DualArray = Array<ElemT, NDIM, Dual, S>;
template <typename ElemT, ArraySizing S = FixedSize> using DualArray1 = Array<ElemT, 1, Dual, S>;
template <typename ElemT, ArraySizing S = FixedSize> using DualArray2 = Array<ElemT, 2, Dual, S>;
template <typename ElemT, ArraySizing S = FixedSize> using PrimalArray1 = Array<ElemT, 1, Primal, S>;
template <typename ElemT, ArraySizing S = FixedSize> using PrimalArray2 = Array<ElemT, 2, Primal, S>;

// n arrays of one shape (array.h:351-370)
template <typename ElemT, int NDIM, GridType G = Dual>
std::vector<Array<ElemT, NDIM, G>> vec_of_Arrays(const std::size_t n, const Shape<NDIM, G>& s)
{
    std::vector<Array<ElemT, NDIM, G>> v;
    v.reserve(n);
    for (std::size_t k = 0; k < n; ++k) v.emplace_back(s);
    return v;
}

// execution policies (parallalg.h:62-88); the algorithms below run on the host either way
namespace execution {
struct serial_policy {};
inline constexpr serial_policy seq{};
struct openmp_policy {};
inline constexpr openmp_policy omp{};
}   // namespace execution
template <typename T> struct is_execution_policy : std::false_type {};
template <> struct is_execution_policy<execution::serial_policy> : std::true_type {};
template <> struct is_execution_policy<execution::openmp_policy> : std::true_type {};
template <typename T> inline constexpr bool is_execution_policy_v = is_execution_policy<T>::value;
template <typename T>
concept execution_policy = is_execution_policy_v<T>;

namespace detail {
template <typename A> using elem_t = typename std::remove_cvref_t<A>::ElemType;

template <typename A, typename... As> void same_shapes(const A& a, const As&... as)
{
    (assert(a.flattened_length() == as.flattened_length()), ...);
    (void)a;
}

// flattened position -> index of an array of shape s
template <int NDIM, GridType G> Idx<NDIM, G> unflatten(std::size_t k, const Shape<NDIM, G>& s)
{
    Idx<NDIM, G> idx;
    for (int m = NDIM - 1; m >= 0; --m) {
        idx.idxs[m] = k % s[m];
        k /= s[m];
    }
    return idx;
}
template <typename Pol> constexpr bool openmp() { return std::is_same_v<Pol, execution::openmp_policy>; }
}   // namespace detail

// algorithm.h:24-140: deep copies
template <execution_policy Policy, typename T, int N, GridType G, ArraySizing A0, ArraySizing A1>
void copy(Policy, Array<T, N, G, A0>& dst, const Array<T, N, G, A1>& src)
{
    assert(dst.shape() == src.shape());
    const std::size_t n = src.flattened_length();
#pragma omp parallel for if (detail::openmp<Policy>())
    for (std::size_t k = 0; k < n; ++k) dst.flatten(k) = src.flatten(k);
}
template <typename T, int N, GridType G, ArraySizing A0, ArraySizing A1>
void copy(Array<T, N, G, A0>& dst, const Array<T, N, G, A1>& src)
{
    copy(execution::seq, dst, src);
}
template <execution_policy Policy, typename T, int N, GridType G, ArraySizing A>
Array<T, N, G, A> copy(Policy policy, const Array<T, N, G, A>& src)
{
    Array<T, N, G, A> dst(src.shape());
    copy(policy, dst, src);
    return dst;
}
template <typename T, int N, GridType G, ArraySizing A> Array<T, N, G, A> copy(const Array<T, N, G, A>& src)
{
    return copy(execution::seq, src);
}

// algorithm.h:141-200: every element set to one value
template <execution_policy Policy, typename T, int N, GridType G, ArraySizing A>
void fill(Policy, Array<T, N, G, A>& dst, const T& val)
{
    const std::size_t n = dst.flattened_length();
#pragma omp parallel for if (detail::openmp<Policy>())
    for (std::size_t k = 0; k < n; ++k) dst.flatten(k) = val;
}
template <typename T, int N, GridType G, ArraySizing A> void fill(Array<T, N, G, A>& dst, const T& val)
{
    fill(execution::seq, dst, val);
}

// algorithm.h:434-510: func(elements...) on corresponding elements of arrays of one shape
template <execution_policy Policy, typename Func, typename A0, typename... As>
void for_each(Policy, Func&& func, A0&& a0, As&&... as)
{
    detail::same_shapes(a0, as...);
    const std::size_t n = a0.flattened_length();
#pragma omp parallel for if (detail::openmp<Policy>())
    for (std::size_t k = 0; k < n; ++k) func(a0.flatten(k), as.flatten(k)...);
}
template <typename Func, typename A0, typename... As>
    requires(!execution_policy<std::remove_cvref_t<Func>>)
void for_each(Func&& func, A0&& a0, As&&... as)
{
    for_each(execution::seq, std::forward<Func>(func), std::forward<A0>(a0), std::forward<As>(as)...);
}

// algorithm.h:516-900: func(index, elements...) in index order
template <execution_policy Policy, typename Func, typename A0, typename... As>
void for_each_idx(Policy, Func&& func, A0&& a0, As&&... as)
{
    detail::same_shapes(a0, as...);
    const std::size_t n = a0.flattened_length();
    const auto s = a0.shape();
#pragma omp parallel for if (detail::openmp<Policy>())
    for (std::size_t k = 0; k < n; ++k) func(detail::unflatten(k, s), a0.flatten(k), as.flatten(k)...);
}
template <typename Func, typename A0, typename... As>
    requires(!execution_policy<std::remove_cvref_t<Func>>)
void for_each_idx(Func&& func, A0&& a0, As&&... as)
{
    for_each_idx(execution::seq, std::forward<Func>(func), std::forward<A0>(a0), std::forward<As>(as)...);
}

// algorithm.h:927-1050: dst = func(sources...) elementwise
template <execution_policy Policy, typename Func, typename D, typename... As>
void transform(Policy, const Func& func, D& dst, const As&... as)
{
    detail::same_shapes(dst, as...);
    const std::size_t n = dst.flattened_length();
#pragma omp parallel for if (detail::openmp<Policy>())
    for (std::size_t k = 0; k < n; ++k) dst.flatten(k) = func(as.flatten(k)...);
}
template <typename Func, typename D, typename... As>
    requires(!execution_policy<std::remove_cvref_t<Func>>)
void transform(const Func& func, D& dst, const As&... as)
{
    transform(execution::seq, func, dst, as...);
}

// algorithm.h:1121-1162: reduce(init, transform(elements...)) over the flattened arrays in order
template <execution_policy Policy, typename Transform, typename Reduce, typename T, typename A0, typename... As>
T transform_reduce(Policy, const Transform& tr, const Reduce& red, T init, const A0& a0, const As&... as)
{
    detail::same_shapes(a0, as...);
    const std::size_t n = a0.flattened_length();
    T r = init;
    for (std::size_t k = 0; k < n; ++k) r = red(r, tr(a0.flatten(k), as.flatten(k)...));
    return r;
}

}   // namespace par
