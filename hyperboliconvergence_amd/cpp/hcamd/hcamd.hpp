// hcamd.hpp -- C++20 drop-in for the reference's structured-mesh finite-volume hot path.
//
// Same names, template parameters and argument order as the reference's header-only API
// (lacrymose/hyperboliConvergence cpp/inc/): the par:: containers and algorithms (par.hpp),
// laws, variable sets and set2Set / set2State, flux functors, limiters, make_muscl_flux,
// boundary-condition factories, Mesh with its node and cell arrays, SolutionField, the LSQ metric
// types, Runge-Kutta tables, time controls, and the spatial and time-stepping entry points:
// xmetrics, qmetrics, boundaryUpdate, residualCalc, spectralRadius, rungeKuttaAccumulation,
// eulerForwardUpdateGlobal / Local and integrate.
//
// Where the work runs: residualCalc, boundaryUpdate and integrate run on the GPU through the
// C-ABI of include/hcamd.h, on a device session kept per mesh and scheme (geometry is built once
// and reused by later calls with the same mesh nodes, boundary types, species and scheme).  The
// point functors and the per-cell maps whose reference signatures take host arrays only
// (xmetrics, qmetrics, spectralRadius, rungeKuttaAccumulation, eulerForwardUpdate*) evaluate on
// the host with the same point functions the kernels run (hc_point.hpp); the C-ABI offers device
// versions of each of them.  Link with -lhcamd (hyperboliconvergence_amd/libhcamd.so).
#pragma once

#include "hcamd.h"
#include "par.hpp"
#include "../../csrc/hc_point.hpp"

#include <array>
#include <cassert>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

// ---------------- laws, bases, variable sets (conservationLaws/base/base.h:16-83) ----------------

enum struct LawType { ScalarAdvection, Euler };

enum struct EulerBases { Conserved, Characteristic, Primitive, Viscous };
enum struct ScalarAdvectionBases { Conserved, Characteristic };
enum struct EulerBCs { Periodic, Riemann, Entropy, Fixed, InviscidWall };
enum struct ScalarAdvectionBCs { Periodic, Riemann, Fixed };

template <LawType Law> struct BasisTypeHelper;
template <> struct BasisTypeHelper<LawType::Euler> { using type = EulerBases; };
template <> struct BasisTypeHelper<LawType::ScalarAdvection> { using type = ScalarAdvectionBases; };
template <LawType Law> using BasisType = typename BasisTypeHelper<Law>::type;

template <LawType Law> struct BoundaryTypeHelper;
template <> struct BoundaryTypeHelper<LawType::Euler> { using type = EulerBCs; };
template <> struct BoundaryTypeHelper<LawType::ScalarAdvection> { using type = ScalarAdvectionBCs; };
template <LawType Law> using BoundaryType = typename BoundaryTypeHelper<Law>::type;

template <LawType Law, int nDim> inline constexpr int nVar = (Law == LawType::Euler) ? nDim + 2 : nDim + 1;

template <typename Real>
concept floating_point = std::is_floating_point_v<Real>;

template <LawType Law, int nDim, auto Basis, floating_point Real> struct VariableDelta;

// VariableSet: a point of the law's phase space in one basis (base/base.h:22-32)
template <LawType Law, int nDim, auto Basis, floating_point Real> struct VariableSet {
    static constexpr int N = nVar<Law, nDim>;
    std::array<Real, N> elems{};
    Real& operator[](int i) { return elems[i]; }
    const Real& operator[](int i) const { return elems[i]; }
};

// VariableDelta: a displacement in that space (base/base.h:39-49)
template <LawType Law, int nDim, auto Basis, floating_point Real> struct VariableDelta {
    static constexpr int N = nVar<Law, nDim>;
    std::array<Real, N> elems{};
    Real& operator[](int i) { return elems[i]; }
    const Real& operator[](int i) const { return elems[i]; }
};

// affine arithmetic (utils/maths/affineSpace.ipp): d/a multiplies by 1/a
template <LawType L, int D, auto B, floating_point R>
VariableDelta<L, D, B, R> operator-(const VariableSet<L, D, B, R>& a, const VariableSet<L, D, B, R>& b)
{
    VariableDelta<L, D, B, R> d;
    for (int i = 0; i < d.N; ++i) d[i] = a[i] - b[i];
    return d;
}
template <LawType L, int D, auto B, floating_point R>
VariableSet<L, D, B, R> operator+(const VariableSet<L, D, B, R>& a, const VariableDelta<L, D, B, R>& d)
{
    VariableSet<L, D, B, R> r = a;
    for (int i = 0; i < r.N; ++i) r[i] += d[i];
    return r;
}
template <LawType L, int D, auto B, floating_point R>
VariableSet<L, D, B, R> operator-(const VariableSet<L, D, B, R>& a, const VariableDelta<L, D, B, R>& d)
{
    VariableSet<L, D, B, R> r = a;
    for (int i = 0; i < r.N; ++i) r[i] -= d[i];
    return r;
}
template <LawType L, int D, auto B, floating_point R>
VariableDelta<L, D, B, R> operator+(const VariableDelta<L, D, B, R>& a, const VariableDelta<L, D, B, R>& b)
{
    VariableDelta<L, D, B, R> r = a;
    for (int i = 0; i < r.N; ++i) r[i] += b[i];
    return r;
}
template <LawType L, int D, auto B, floating_point R>
VariableDelta<L, D, B, R> operator*(R a, const VariableDelta<L, D, B, R>& d)
{
    VariableDelta<L, D, B, R> r = d;
    for (auto& w : r.elems) w *= a;
    return r;
}
template <LawType L, int D, auto B, floating_point R>
VariableDelta<L, D, B, R> operator/(const VariableDelta<L, D, B, R>& d, R a)
{
    const R a1 = R(1.) / a;
    VariableDelta<L, D, B, R> r = d;
    for (auto& w : r.elems) w *= a1;
    return r;
}

// FluxResult: conserved-basis flux and spectral radius (base/base.h:57-83, fluxResult.ipp:8-22)
template <LawType Law, int nDim, floating_point Real> struct FluxResult {
    using FluxType = VariableDelta<Law, nDim, BasisType<Law>::Conserved, Real>;
    FluxType flux{};
    Real lambda{};
    FluxResult& operator+=(const FluxResult& fr)
    {
        for (int i = 0; i < FluxType::N; ++i) flux[i] += fr.flux[i];
        lambda += fr.lambda;
        return *this;
    }
    FluxResult& operator-=(const FluxResult& fr)
    {
        for (int i = 0; i < FluxType::N; ++i) flux[i] -= fr.flux[i];
        lambda += fr.lambda;
        return *this;
    }
};

// Species (conservationLaws/euler/euler.h:48-60, species.ipp:2-18)
template <LawType Law, floating_point Real> struct Species;
template <floating_point Real> struct Species<LawType::Euler, Real> {
    Real gamma, minf, lref, nu, pr, dt, R, gamma1;
};
template <floating_point Real> struct Species<LawType::ScalarAdvection, Real> {};

template <floating_point Real> Species<LawType::Euler, Real> get_air_species()
{
    Species<LawType::Euler, Real> air;
    air.gamma = 1.4;
    air.minf = 0.0001;
    air.lref = 1.;
    air.pr = 0.7;
    air.nu = 1.81e-5;
    air.dt = std::numeric_limits<Real>::max();
    air.R = 287.058;
    air.gamma1 = 1. / (air.gamma - 1.);
    return air;
}


// traits of a variable set (utils/type-traits.h: law_of_v, dim_of_v, fptype_of_t, vardelta_t)
template <typename T> struct varset_traits;
template <LawType L, int D, auto B, floating_point R> struct varset_traits<VariableSet<L, D, B, R>> {
    static constexpr LawType law = L;
    static constexpr int dim = D;
    static constexpr auto basis = B;
    using real = R;
    using delta = VariableDelta<L, D, B, R>;
};
template <typename T> inline constexpr LawType law_of_v = varset_traits<T>::law;
template <typename T> inline constexpr int dim_of_v = varset_traits<T>::dim;
template <typename T> using fptype_of_t = typename varset_traits<T>::real;
template <typename T> using vardelta_t = typename varset_traits<T>::delta;

// ---------------- geometry (geometry/geometry.h, surface.ipp, volume.ipp) ----------------

namespace geom {
template <int nDim, floating_point Real> struct Point {
    std::array<Real, nDim> elems{};
    Real& operator[](int i) { return elems[i]; }
    const Real& operator[](int i) const { return elems[i]; }
};
template <int nDim, floating_point Real> struct Direction {
    std::array<Real, nDim> elems{};
    Real& operator[](int i) { return elems[i]; }
    const Real& operator[](int i) const { return elems[i]; }
};
template <int nDim, floating_point Real> struct Metric {
    std::array<Direction<nDim, Real>, nDim> m;
    Direction<nDim, Real>& operator[](int i) { return m[i]; }
    const Direction<nDim, Real>& operator[](int i) const { return m[i]; }
};
template <int nDim, floating_point Real> struct Surface {
    Real area;
    Point<nDim, Real> centre;
    Metric<nDim, Real> metric;
};
template <int nDim, floating_point Real> struct Volume {
    Real volume;
    Point<nDim, Real> centre;
};

// surface.ipp:12-29
template <floating_point Real> Surface<2, Real> surface(const Point<2, Real>& p0, const Point<2, Real>& p1)
{
    const Real t0x = p1[0] - p0[0], t0y = p1[1] - p0[1];
    const Real t1x = p0[0] - p1[0], t1y = p0[1] - p1[1];
    Real l2 = 0;
    l2 += t0x * t0x;
    l2 += t0y * t0y;
    const Real l = std::sqrt(l2);
    const Real l1 = 1. / l;
    Surface<2, Real> s;
    s.area = l;
    s.centre[0] = p0[0] + t0x * 0.5;
    s.centre[1] = p0[1] + t0y * 0.5;
    s.metric[0][0] = -(t1y * l1);
    s.metric[0][1] = t1x * l1;
    s.metric[1][0] = t0x * l1;
    s.metric[1][1] = t0y * l1;
    return s;
}
}   // namespace geom


namespace geom {
// volume.ipp:25-66: the bilinear element's |J| and centre at the element centre (the volume is a
// quarter of the area, the centre the vertex average)
template <floating_point Real>
Volume<2, Real> volume(const Point<2, Real>& p0, const Point<2, Real>& p1, const Point<2, Real>& p2,
                       const Point<2, Real>& p3)
{
    const Point<2, Real>* p[4] = {&p0, &p1, &p2, &p3};
    const Real u = Real(0.5) * Real(0.5);
    const Real du[4][2] = {{Real(-0.5) * Real(0.5), Real(0.5) * Real(-0.5)},
                           {Real(0.5) * Real(0.5), Real(0.5) * Real(-0.5)},
                           {Real(-0.5) * Real(0.5), Real(0.5) * Real(0.5)},
                           {Real(0.5) * Real(0.5), Real(0.5) * Real(0.5)}};
    Volume<2, Real> v;
    for (int d = 0; d < 2; ++d) {
        Real c = 0;
        for (int k = 0; k < 4; ++k) c = c + u * (*p[k])[d];
        v.centre[d] = c;
    }
    Real t[2][2];
    for (int a = 0; a < 2; ++a)
        for (int d = 0; d < 2; ++d) {
            Real c = 0;
            for (int k = 0; k < 4; ++k) c = c + du[k][a] * (*p[k])[d];
            t[a][d] = c;
        }
    v.volume = std::fabs(t[0][0] * t[1][1] - t[0][1] * t[1][0]);
    return v;
}
}   // namespace geom

// ---------------- structured mesh (mesh/mesh.h:8-57, mesh/dual.ipp:42-83, generate/twoD.h) ----------------

template <int nDim, floating_point Real> using MeshNodeArray = par::PrimalArray<geom::Point<nDim, Real>, nDim>;
template <int nDim, floating_point Real> using MeshCellArray = par::DualArray<geom::Volume<nDim, Real>, nDim>;

// cells of the dual mesh from the nodes (mesh/dual.ipp:42-83): corners {i,j} {i+1,j} {i,j+1} {i+1,j+1}
template <floating_point Real> void dual(const MeshNodeArray<2, Real>& nodes, MeshCellArray<2, Real>& cells)
{
    const std::size_t ni = cells.shape(0), nj = cells.shape(1);
    assert(nodes.shape(0) == ni + 1 && nodes.shape(1) == nj + 1);
    for (std::size_t i = 0; i < ni; ++i)
        for (std::size_t j = 0; j < nj; ++j)
            cells({i, j}) = geom::volume(nodes({i, j}), nodes({i + 1, j}), nodes({i, j + 1}), nodes({i + 1, j + 1}));
}
template <floating_point Real> MeshCellArray<2, Real> dual(const MeshNodeArray<2, Real>& nodes)
{
    MeshCellArray<2, Real> cells(par::dualShape(nodes.shape()));
    dual(nodes, cells);
    return cells;
}

// node and cell arrays of a structured mesh, move-only (mesh/mesh.h:19-57).  period_i / period_j
// are this build's addition: the translation from a cell to its periodic image across boundary 1
// / 3, which periodic LSQ stencils need (zero for meshes that close on themselves, like an O-grid)
template <int nDim, floating_point Real> struct Mesh {
    using Node = geom::Point<nDim, Real>;
    using Cell = geom::Volume<nDim, Real>;
    using Face = geom::Surface<nDim, Real>;
    using NodeArray = par::PrimalArray<Node, nDim>;
    using CellArray = par::DualArray<Cell, nDim>;
    using NodeShape = par::PrimalShape<nDim>;
    using CellShape = par::DualShape<nDim>;

    NodeShape node_shape;
    CellShape cell_shape;
    NodeArray nodes;
    CellArray cells;
    std::array<Real, 2> period_i{0, 0};
    std::array<Real, 2> period_j{0, 0};

    Mesh() = delete;
    Mesh(const Mesh&) = delete;
    Mesh(Mesh&&) = default;
    Mesh& operator=(const Mesh&) = delete;
    Mesh& operator=(Mesh&&) = default;
    Mesh(const NodeShape& s) : node_shape(s), cell_shape(par::dualShape(s)), nodes(node_shape), cells(cell_shape) {}
    Mesh(const CellShape& s) : node_shape(par::primalShape(s)), cell_shape(s), nodes(node_shape), cells(cell_shape) {}
};

// mesh/generate/twoD.h:15-40: [lox, hix] x [loy, hiy] with linear spacing (periodic translations set)
template <floating_point Real>
Mesh<2, Real> make_linspace_mesh(const par::DualShape2& cellDims, const Real lox, const Real hix, const Real loy,
                                 const Real hiy)
{
    Mesh<2, Real> mesh(cellDims);
    const std::size_t ni = cellDims[0], nj = cellDims[1];
    const Real dx = (hix - lox) / ni, dy = (hiy - loy) / nj;
    for (std::size_t i = 0; i <= ni; ++i)
        for (std::size_t j = 0; j <= nj; ++j) mesh.nodes({i, j}) = {{lox + i * dx, loy + j * dy}};
    dual(mesh.nodes, mesh.cells);
    mesh.period_i = {hix - lox, 0};
    mesh.period_j = {0, hiy - loy};
    return mesh;
}

// mesh/generate/twoD.h:139-179: annulus rc..rb, uniform radial spacing, i = radius, j = theta
template <floating_point Real>
Mesh<2, Real> make_cylinder_mesh(const std::array<Real, 2> lengths, const std::array<std::size_t, 2> ncells)
{
    const std::size_t nr = ncells[0], nth = ncells[1];
    Mesh<2, Real> mesh(par::DualShape2{{nr, nth}});
    const Real dr = (lengths[1] - lengths[0]) / nr;
    const Real dth = (2. * M_PI) / nth;
    for (std::size_t i = 0; i <= nr; ++i)
        for (std::size_t j = 0; j <= nth; ++j) {
            const Real r = lengths[0] + i * dr, th = j * dth;
            mesh.nodes({i, j}) = {{r * std::cos(th), r * std::sin(th)}};
        }
    dual(mesh.nodes, mesh.cells);
    return mesh;
}

// the O-grid of the cylinder cases (new, next to make_cylinder_mesh): geometric radial stretching
// with first spacing dr0 reaching rb exactly, node nth identical to node 0
template <floating_point Real>
Mesh<2, Real> make_ogrid_mesh(const std::array<Real, 2> lengths, const std::array<std::size_t, 2> ncells, Real dr0)
{
    const std::size_t nr = ncells[0], nth = ncells[1];
    const Real L = lengths[1] - lengths[0];
    auto g = [&](Real s) { return dr0 * (std::pow(s, Real(nr)) - 1.) / (s - 1.) - L; };
    Real lo = 1. + 1e-12, hi = 2.;
    while (g(hi) < 0) hi *= 2.;
    for (int it = 0; it < 200; ++it) {
        const Real mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        (g(mid) < 0 ? lo : hi) = mid;
    }
    const Real s = 0.5 * (lo + hi);
    Mesh<2, Real> mesh(par::DualShape2{{nr, nth}});
    for (std::size_t i = 0; i <= nr; ++i) {
        const Real r = (i == nr) ? lengths[1] : lengths[0] + dr0 * (std::pow(s, Real(i)) - 1.) / (s - 1.);
        for (std::size_t j = 0; j <= nth; ++j) {
            const Real th = (2. * M_PI * Real(j == nth ? 0 : j)) / nth;
            mesh.nodes({i, j}) = {{r * std::cos(th), r * std::sin(th)}};
        }
    }
    dual(mesh.nodes, mesh.cells);
    return mesh;
}

// ---------------- solution field (solutionField/solutionField.h:10-140) ----------------

// interior values, and per boundary an array {n_along, 2}: (k, 0) the ghost value, (k, 1) the
// reference value; move-only like the reference's
template <typename VarSet, int nDim> struct SolutionField;
template <typename VarSet> struct SolutionField<VarSet, 2> {
    static constexpr unsigned int nBoundaries = 4;
    static constexpr LawType Law = law_of_v<VarSet>;
    using VarDel = vardelta_t<VarSet>;
    using VarField = par::DualArray<VarSet, 2>;
    using BCType = BoundaryType<Law>;

    VarField interior;
    std::array<VarField, 4> boundary;
    std::array<BCType, 4> bcTypes{};

    SolutionField() = delete;
    SolutionField(const SolutionField&) = delete;
    SolutionField(SolutionField&&) = default;
    SolutionField& operator=(const SolutionField&) = delete;
    SolutionField& operator=(SolutionField&&) = default;
    SolutionField(const par::DualShape<2>& s)
        : interior(s), boundary{VarField(par::DualShape2{{s[1], 2}}), VarField(par::DualShape2{{s[1], 2}}),
                                VarField(par::DualShape2{{s[0], 2}}), VarField(par::DualShape2{{s[0], 2}})}
    {
    }
};

// deep copies (solutionField.h:94-140)
template <par::execution_policy Policy, typename VarSet, int nDim>
void copy(const Policy policy, SolutionField<VarSet, nDim>& dst, const SolutionField<VarSet, nDim>& src)
{
    par::copy(policy, dst.interior, src.interior);
    for (unsigned int b = 0; b < src.nBoundaries; ++b) {
        par::copy(policy, dst.boundary[b], src.boundary[b]);
        dst.bcTypes[b] = src.bcTypes[b];
    }
}
template <par::execution_policy Policy, typename VarSet, int nDim>
SolutionField<VarSet, nDim> copy(const Policy policy, const SolutionField<VarSet, nDim>& src)
{
    SolutionField<VarSet, nDim> dst(src.interior.shape());
    copy(policy, dst, src);
    return dst;
}
template <typename VarSet, int nDim> void copy(SolutionField<VarSet, nDim>& dst, const SolutionField<VarSet, nDim>& src)
{
    copy(par::execution::seq, dst, src);
}
template <typename VarSet, int nDim> SolutionField<VarSet, nDim> copy(const SolutionField<VarSet, nDim>& src)
{
    return copy(par::execution::seq, src);
}

// ---------------- limiters (limiters/limiter.h:14-131) ----------------

namespace Limiters {
template <typename Limiter> struct VariableLimiterInterface {
    template <typename VarDelT> VarDelT operator()(const VarDelT& dq0, const VarDelT& dq1) const
    {
        VarDelT dqlim;
        for (int i = 0; i < VarDelT::N; i++) dqlim[i] = static_cast<const Limiter*>(this)->limit(dq0[i], dq1[i]);
        return dqlim;
    }
};
#define HC_LIMITER(Name, Id)                                                                  \
    struct Name : VariableLimiterInterface<Name> {                                            \
        static constexpr int id = Id;                                                         \
        template <floating_point Real> Real limit(const Real a, const Real b) const           \
        {                                                                                     \
            return hc::limit<Real>(Id, a, b);                                                 \
        }                                                                                     \
    };
HC_LIMITER(NoLimit1, HC_LIM_NOLIMIT1)
HC_LIMITER(NoLimit2, HC_LIM_NOLIMIT2)
HC_LIMITER(NoLimit3, HC_LIM_NOLIMIT3)
HC_LIMITER(VanAlbada2, HC_LIM_VANALBADA2)
HC_LIMITER(MinMod2, HC_LIM_MINMOD2)
HC_LIMITER(Superbee2, HC_LIM_SUPERBEE2)
HC_LIMITER(MonotonizedCentral2, HC_LIM_MC2)
HC_LIMITER(VanLeer2, HC_LIM_VANLEER2)
HC_LIMITER(Cada3, HC_LIM_CADA3)
HC_LIMITER(SmoothVanAlbada2, HC_LIM_SMOOTHVANALBADA)   // new: VanAlbada2 without the a*b<0 switch
#undef HC_LIMITER
}   // namespace Limiters

// ---------------- flux functors (base/base.h:370-472, euler/euler.h:334-405) ----------------

enum struct LowMachScaling { Convective, Acoustic, Adaptive };

namespace hcamd_detail {
template <floating_point Real> hc::Species<Real> species(const Species<LawType::Euler, Real>& s)
{
    return hc::Species<Real>{s.gamma, s.minf, s.lref, s.nu, s.pr, s.dt, s.R, s.gamma1};
}
template <floating_point Real> hc::Species<Real> species(const Species<LawType::ScalarAdvection, Real>&)
{
    return hc::Species<Real>{};
}
template <floating_point Real> hc::FaceG<Real> face(const geom::Surface<2, Real>& f)
{
    return hc::FaceG<Real>{f.area, f.metric[0][0], f.metric[0][1]};
}
template <auto Basis> constexpr int basis_id()
{
    if constexpr (std::is_same_v<decltype(Basis), EulerBases>) return Basis == EulerBases::Primitive ? 1 : 0;
    else return 0;
}
template <typename E> constexpr int scale_id(E s) { return s == E::Convective ? 0 : s == E::Acoustic ? 1 : 2; }
}   // namespace hcamd_detail

// State<Law,nDim,Real>: the derived point-wise quantities of a solution set with the reference's
// accessors (conservationLaws/euler/euler.h:62-83: p, rho, T, H, |u|^2, a^2, u_i;
// scalarAdvection/scalarAdvection.h:48-59: q, u_i), so drivers' output code (writeState of
// script/euler/cylinder2D.cpp:35-63, gresho2D.cpp, shocktube.cpp, script/scalar/cartesian2D.cpp)
// compiles unchanged.  set2State evaluates them with the kernels' point functions.
template <LawType Law, int nDim, floating_point Real> struct State;
template <int nDim, floating_point Real> struct State<LawType::Euler, nDim, Real> {
    std::array<Real, nDim + 6> state{};
    Real& pressure() { return state[0]; }
    Real& density() { return state[1]; }
    Real& temperature() { return state[2]; }
    Real& specificTotalEnthalpy() { return state[3]; }
    Real& velocity2() { return state[4]; }
    Real& speedOfSound2() { return state[5]; }
    Real& velocity(const int i) { return state[6 + i]; }
    const Real& pressure() const { return state[0]; }
    const Real& density() const { return state[1]; }
    const Real& temperature() const { return state[2]; }
    const Real& specificTotalEnthalpy() const { return state[3]; }
    const Real& velocity2() const { return state[4]; }
    const Real& speedOfSound2() const { return state[5]; }
    const Real& velocity(const int i) const { return state[6 + i]; }
};
template <int nDim, floating_point Real> struct State<LawType::ScalarAdvection, nDim, Real> {
    std::array<Real, nVar<LawType::ScalarAdvection, nDim>> state{};
    Real& scalar() { return state[0]; }
    const Real& scalar() const { return state[0]; }
    Real& velocity(const int i) { return state[1 + i]; }
    const Real& velocity(const int i) const { return state[1 + i]; }
};

// set2State (euler/transforms/{conserved,primitive}.ipp, scalarAdvection/transforms/conserved.ipp)
template <LawType Law, int nDim, auto Basis, floating_point Real>
State<Law, nDim, Real> set2State(const Species<Law, Real>& species, const VariableSet<Law, nDim, Basis, Real>& q)
{
    State<Law, nDim, Real> st;
    if constexpr (Law == LawType::Euler) {
        const hc::EState<Real> s = hc::euler_state<Real, hcamd_detail::basis_id<Basis>()>(hcamd_detail::species(species),
                                                                                        q.elems.data());
        st.state = {s.p, s.r, s.t, s.h, s.v2, s.a2, s.u, s.v};
    } else {
        const hc::SState<Real> s = hc::scalar_state(q.elems.data());
        st.state = {s.q, s.u, s.v};
    }
    return st;
}


// state2Set / set2Set (euler/transforms/{conserved,primitive}.ipp, base/base.h:250-273): the
// conserved and primitive sets of a State, and a set in one basis from a set in another
template <typename DstT, LawType Law, int nDim, floating_point Real>
DstT state2Set(const Species<Law, Real>&, const State<Law, nDim, Real>& st)
{
    DstT q;
    if constexpr (Law == LawType::Euler) {
        hc::EState<Real> s;
        s.p = st.pressure();
        s.r = st.density();
        s.t = st.temperature();
        s.h = st.specificTotalEnthalpy();
        s.v2 = st.velocity2();
        s.a2 = st.speedOfSound2();
        s.u = st.velocity(0);
        s.v = st.velocity(1);
        if constexpr (hcamd_detail::basis_id<varset_traits<DstT>::basis>() == 1) hc::euler_to_primitive(s, q.elems.data());
        else hc::euler_to_conserved(s, q.elems.data());
    } else {
        q[0] = st.velocity(0);
        q[1] = st.velocity(1);
        q[2] = st.scalar();
    }
    return q;
}

template <typename DstT, LawType Law, int nDim, auto Basis, floating_point Real>
DstT set2Set(const Species<Law, Real>& species, const VariableSet<Law, nDim, Basis, Real>& src)
{
    if constexpr (std::is_same_v<DstT, VariableSet<Law, nDim, Basis, Real>>) return src;
    else return state2Set<DstT>(species, set2State(species, src));
}

// FluxInterface CRTP: any solution basis -> FluxResult in the conserved basis (base/base.h:370-397)
template <typename Flux, LawType Law> struct FluxInterface {
    template <int nDim, auto Basis, floating_point Real>
    FluxResult<Law, nDim, Real> operator()(const Species<Law, Real>& species, const geom::Surface<nDim, Real>& face,
                                           const VariableSet<Law, nDim, Basis, Real>& ql,
                                           const VariableSet<Law, nDim, Basis, Real>& qr) const
    {
        const auto sp = hcamd_detail::species(species);
        const auto f = hcamd_detail::face(face);
        constexpr int L = Law == LawType::Euler ? hc::LAW_EULER : hc::LAW_SCALAR;
        constexpr int B = hcamd_detail::basis_id<Basis>();
        const hc::Flux4<Real> r = hc::face_flux<Real, L, B, Flux::id>(sp, Flux::vscale, Flux::pscale, f,
                                                                      ql.elems.data(), qr.elems.data());
        FluxResult<Law, nDim, Real> out;
        for (int v = 0; v < nVar<Law, nDim>; ++v) out.flux[v] = r.f[v];
        out.lambda = r.lambda;
        return out;
    }
};

template <LawType Law> struct CentralFlux : FluxInterface<CentralFlux<Law>, Law> {
    static constexpr int id = HC_FLUX_CENTRAL, vscale = 0, pscale = 0;
};
template <LawType Law> struct RusanovFlux : FluxInterface<RusanovFlux<Law>, Law> {
    static constexpr int id = HC_FLUX_RUSANOV, vscale = 0, pscale = 0;
};
template <LawType Law> struct RoeFlux : FluxInterface<RoeFlux<Law>, Law> {
    static_assert(Law == LawType::Euler, "RoeFlux needs the characteristic basis of the Euler equations");
    static constexpr int id = HC_FLUX_ROE, vscale = 0, pscale = 0;
};
// Roe flux with the Roe-average velocity rotated into the face frame (new; the reference's RoeFlux
// keeps it in the global frame, base/fluxes/roeFlux.ipp:40-54)
template <LawType Law> struct RoeFluxFaceFrame : FluxInterface<RoeFluxFaceFrame<Law>, Law> {
    static_assert(Law == LawType::Euler, "RoeFluxFaceFrame needs the characteristic basis of the Euler equations");
    static constexpr int id = HC_FLUX_ROE_FACEFRAME, vscale = 0, pscale = 0;
};
// Roe flux in the Weiss & Smith / Li & Gu form (conservationLaws/euler/euler.h:392-405,
// euler/fluxes/roeUnprecWS.ipp); Euler only, like the reference's
struct RoeUnprecWS : FluxInterface<RoeUnprecWS, LawType::Euler> {
    static constexpr int id = HC_FLUX_ROE_UNPREC_WS, vscale = 0, pscale = 0;
};
template <LowMachScaling VFluxScaling = LowMachScaling::Convective, LowMachScaling PFluxScaling = LowMachScaling::Acoustic>
struct Slau : FluxInterface<Slau<VFluxScaling, PFluxScaling>, LawType::Euler> {
    static constexpr int id = HC_FLUX_SLAU;
    static constexpr int vscale = hcamd_detail::scale_id(VFluxScaling), pscale = hcamd_detail::scale_id(PFluxScaling);
};
template <LowMachScaling VFluxScaling = LowMachScaling::Convective, LowMachScaling PFluxScaling = LowMachScaling::Convective>
struct AusmPlusUP : FluxInterface<AusmPlusUP<VFluxScaling, PFluxScaling>, LawType::Euler> {
    static constexpr int id = HC_FLUX_AUSMPUP;
    static constexpr int vscale = hcamd_detail::scale_id(VFluxScaling), pscale = hcamd_detail::scale_id(PFluxScaling);
};

// ---------------- MUSCL high-order flux (spatial/muscl.h:10-87) ----------------

template <LawType Law, typename Limiter, typename Flux> struct MusclFlux {
    Limiter limiter;
    Flux flux;
    static constexpr LawType law = Law;
    static constexpr int limiter_id = Limiter::id;
    static constexpr int flux_id = Flux::id;
    static constexpr int vscale = Flux::vscale, pscale = Flux::pscale;
};

template <LawType Law, typename Limiter, typename Flux> auto make_muscl_flux(const Limiter& limiter, const Flux& flux)
{
    return MusclFlux<Law, Limiter, Flux>{limiter, flux};
}

// ---------------- boundary conditions (spatial/boundary/boundaryCondition.h:13-136) ----------------

template <LawType Law, BoundaryType<Law> BCType> struct BoundaryCondition {
    constexpr BoundaryType<Law> type() const { return BCType; }
};
template <LawType Law> auto make_periodic_BCond() { return BoundaryCondition<Law, BoundaryType<Law>::Periodic>{}; }
template <LawType Law> auto make_fixed_BCond() { return BoundaryCondition<Law, BoundaryType<Law>::Fixed>{}; }
template <LawType Law, BoundaryType<Law> BCType> auto make_ghostCell_BCond() { return BoundaryCondition<Law, BCType>{}; }
template <LawType Law, BoundaryType<Law> BCType> auto make_flux_BCond() { return BoundaryCondition<Law, BCType>{}; }

// ---------------- time integration (ode.h:87-197, controls.h:11-26) ----------------

namespace ODE {
namespace Explicit {
template <floating_point Real> struct RungeKutta {
    unsigned int order;
    unsigned int nstages;
    Real maxCFL;
    std::array<std::array<Real, 6>, 6> alpha;
    std::array<Real, 6> beta;
};
template <floating_point Real> RungeKutta<Real> ssp11()
{
    RungeKutta<Real> rk{};
    rk.order = 1; rk.nstages = 1; rk.maxCFL = 1.;
    rk.beta[0] = 1.0; rk.alpha[0][0] = 1.0;
    return rk;
}
template <floating_point Real> RungeKutta<Real> ssp22()
{
    RungeKutta<Real> rk{};
    rk.order = 2; rk.nstages = 2; rk.maxCFL = 1.;
    rk.beta[0] = 1.0; rk.beta[1] = 1.0;
    rk.alpha[0][0] = 1.0; rk.alpha[1][0] = 0.5; rk.alpha[1][1] = 0.5;
    return rk;
}
template <floating_point Real> RungeKutta<Real> ssp33()
{
    RungeKutta<Real> rk{};
    rk.order = 3; rk.nstages = 3; rk.maxCFL = 1.;
    rk.beta[0] = 1.0; rk.beta[1] = 0.5; rk.beta[2] = 1.0;
    rk.alpha[0][0] = 1.; rk.alpha[1][0] = 0.5; rk.alpha[1][1] = 0.5;
    rk.alpha[2][0] = 1. / 6.; rk.alpha[2][1] = 1. / 6.; rk.alpha[2][2] = 2. / 3.;
    return rk;
}
template <floating_point Real> RungeKutta<Real> ssp34()
{
    RungeKutta<Real> rk{};
    rk.order = 3; rk.nstages = 4; rk.maxCFL = 2.;
    rk.beta[0] = 0.5; rk.beta[1] = 1.0; rk.beta[2] = 0.5; rk.beta[3] = 1.0;
    rk.alpha[0][0] = 1.0; rk.alpha[1][0] = 0.5; rk.alpha[1][1] = 0.5;
    rk.alpha[2][0] = 1. / 3.; rk.alpha[2][1] = 1. / 3.; rk.alpha[2][2] = 1. / 3.;
    rk.alpha[3][0] = 1. / 6.; rk.alpha[3][1] = 1. / 6.; rk.alpha[3][2] = 1. / 6.; rk.alpha[3][3] = 1. / 2.;
    return rk;
}
// classical RK4 in the same (alpha, beta) form (new)
template <floating_point Real> RungeKutta<Real> rk4()
{
    RungeKutta<Real> rk{};
    rk.order = 4; rk.nstages = 4; rk.maxCFL = 2.;
    rk.beta[0] = 0.5; rk.beta[1] = 0.5; rk.beta[2] = 1.0; rk.beta[3] = 1.0;
    rk.alpha[0][0] = 1.0; rk.alpha[1][1] = 1.0; rk.alpha[2][2] = 1.0;
    rk.alpha[3][0] = 1. / 6.; rk.alpha[3][1] = 1. / 3.; rk.alpha[3][2] = 1. / 3.; rk.alpha[3][3] = 1. / 6.;
    return rk;
}
}   // namespace Explicit
}   // namespace ODE

template <floating_point Real> struct UnsteadyTimeControls {
    std::size_t nTimesteps;
    Real cfl;
};

// controls.h:31-48
template <floating_point Real> struct SteadyTimeControls {
    std::size_t nTimesteps;        // at most this many steps
    Real cfl;
    Real residual_drop;            // stop once the mass residual's L2 norm has dropped by this factor
    bool use_residual_smoothing;   // not built: must be false
    Real smoothing_factor;
};


// ---------------- least-squares metrics (lsq/lsq.h:10-90, xmetric.ipp:53-69, qmetric.ipp:48-70) ----------------

namespace lsq {
// X = sum dx dx^T, upper triangle {xx, xy, yy}
template <int nDim, floating_point Real> struct XMetric {
    std::array<Real, nDim * (nDim + 1) / 2> x{};
    const Real& operator()(const int i) const { return x[i]; }
    Real& operator()(const int i) { return x[i]; }
    XMetric& operator+=(const XMetric& o)
    {
        for (std::size_t k = 0; k < x.size(); ++k) x[k] = x[k] + o.x[k];
        return *this;
    }
};
template <int nDim, floating_point Real> XMetric<nDim, Real> operator+(XMetric<nDim, Real> a, const XMetric<nDim, Real>& b)
{
    return a += b;
}
template <int nDim, floating_point Real> XMetric<nDim, Real> xmetric(const geom::Direction<nDim, Real>& dx)
{
    XMetric<nDim, Real> xm;
    int k = 0;
    for (int i = 0; i < nDim; ++i)
        for (int j = i; j < nDim; ++j) xm(k++) = dx[i] * dx[j];
    return xm;
}

// Q = sum dq (x) dx, q(var, dim)
template <typename VarSet> struct QMetric {
    static constexpr LawType Law = law_of_v<VarSet>;
    static constexpr int nDim = dim_of_v<VarSet>;
    static constexpr int nVar = ::nVar<Law, nDim>;
    using Real = fptype_of_t<VarSet>;
    std::array<Real, nVar * nDim> q{};
    const Real& operator()(const int i, const int j) const { return q[i * nDim + j]; }
    Real& operator()(const int i, const int j) { return q[i * nDim + j]; }
    QMetric& operator+=(const QMetric& o)
    {
        for (std::size_t k = 0; k < q.size(); ++k) q[k] = q[k] + o.q[k];
        return *this;
    }
};
template <typename VarSet> QMetric<VarSet> operator+(QMetric<VarSet> a, const QMetric<VarSet>& b) { return a += b; }
template <typename VarSet, int nDim, floating_point Real>
QMetric<VarSet> qmetric(const geom::Direction<nDim, Real>& dx, const vardelta_t<VarSet>& dq)
{
    QMetric<VarSet> qm;
    for (int i = 0; i < QMetric<VarSet>::nVar; ++i)
        for (int j = 0; j < nDim; ++j) qm(i, j) = dq[i] * dx[j];
    return qm;
}
}   // namespace lsq

namespace hcamd_detail {
template <floating_point Real> geom::Direction<2, Real> centre_diff(const geom::Volume<2, Real>& c0, const geom::Volume<2, Real>& c1)
{
    return geom::Direction<2, Real>{{c1.centre[0] - c0.centre[0], c1.centre[1] - c0.centre[1]}};
}
}   // namespace hcamd_detail

/*
 * xmetrics(policy, cells[, dxdx]) (spatial/lsqMetrics.h:20-62): X of every cell from its face
 * neighbours (each interior i-edge, then each interior j-edge, added to both of its cells) and a
 * mirror point across every side on the boundary (boundary_xmetrics, :104-163).  The reference's
 * definition on the cell array; the device's metric, held with the fused residual, takes the
 * periodic image instead of the mirror point across periodic sides (equal without them).
 */
template <par::execution_policy Policy, floating_point Real>
void xmetrics(const Policy policy, const MeshCellArray<2, Real>& cells, par::DualArray<lsq::XMetric<2, Real>, 2>& dxdx)
{
    using XM = lsq::XMetric<2, Real>;
    assert(cells.shape() == dxdx.shape());
    par::fill(policy, dxdx, XM{});
    const std::size_t ni = cells.shape(0), nj = cells.shape(1);
    for (std::size_t j = 0; j < nj; ++j)
        for (std::size_t i = 0; i + 1 < ni; ++i) {
            const XM x = lsq::xmetric(hcamd_detail::centre_diff(cells({i, j}), cells({i + 1, j})));
            dxdx({i, j}) += x;
            dxdx({i + 1, j}) += x;
        }
    for (std::size_t i = 0; i < ni; ++i)
        for (std::size_t j = 0; j + 1 < nj; ++j) {
            const XM x = lsq::xmetric(hcamd_detail::centre_diff(cells({i, j}), cells({i, j + 1})));
            dxdx({i, j}) += x;
            dxdx({i, j + 1}) += x;
        }
    auto mirror = [&](std::size_t i, std::size_t j, std::size_t i1, std::size_t j1) {
        const auto& c0 = cells({i, j});
        const auto& c1 = cells({i1, j1});
        const Real bx = c0.centre[0] + (c0.centre[0] - c1.centre[0]);
        const Real by = c0.centre[1] + (c0.centre[1] - c1.centre[1]);
        dxdx({i, j}) += lsq::xmetric(geom::Direction<2, Real>{{bx - c0.centre[0], by - c0.centre[1]}});
    };
    for (std::size_t j = 0; j < nj; ++j) {
        mirror(0, j, 1, j);
        mirror(ni - 1, j, ni - 2, j);
    }
    for (std::size_t i = 0; i < ni; ++i) {
        mirror(i, 0, i, 1);
        mirror(i, nj - 1, i, nj - 2);
    }
}
template <par::execution_policy Policy, floating_point Real>
par::DualArray<lsq::XMetric<2, Real>, 2> xmetrics(const Policy policy, const MeshCellArray<2, Real>& cells)
{
    par::DualArray<lsq::XMetric<2, Real>, 2> dxdx(cells.shape());
    xmetrics(policy, cells, dxdx);
    return dxdx;
}

/*
 * qmetrics(policy, cells, q, dqdx) (spatial/lsqMetrics.h:181-213): Q of every cell from the
 * differences across its interior edges (i-edges, then j-edges, added to both cells).  The
 * reference's definition on the arrays; the fused residual adds the periodic images across
 * periodic sides.
 */
template <par::execution_policy Policy, typename SolVarSet, floating_point Real>
void qmetrics(const Policy policy, const MeshCellArray<2, Real>& cells, const par::DualArray<SolVarSet, 2>& q0,
              par::DualArray<lsq::QMetric<SolVarSet>, 2>& dqdx)
{
    using QM = lsq::QMetric<SolVarSet>;
    assert(cells.shape() == q0.shape() && cells.shape() == dqdx.shape());
    par::fill(policy, dqdx, QM{});
    const std::size_t ni = cells.shape(0), nj = cells.shape(1);
    auto edge = [&](par::DualIdx2 a, par::DualIdx2 b) {
        const QM m = lsq::qmetric<SolVarSet>(hcamd_detail::centre_diff(cells(a), cells(b)), q0(b) - q0(a));
        dqdx(a) += m;
        dqdx(b) += m;
    };
    for (std::size_t j = 0; j < nj; ++j)
        for (std::size_t i = 0; i + 1 < ni; ++i) edge({i, j}, {i + 1, j});
    for (std::size_t i = 0; i < ni; ++i)
        for (std::size_t j = 0; j + 1 < nj; ++j) edge({i, j}, {i, j + 1});
}
template <par::execution_policy Policy, typename SolVarSet, floating_point Real>
par::DualArray<lsq::QMetric<SolVarSet>, 2> qmetrics(const Policy policy, const MeshCellArray<2, Real>& cells,
                                                   const par::DualArray<SolVarSet, 2>& q0)
{
    par::DualArray<lsq::QMetric<SolVarSet>, 2> dqdx(cells.shape());
    qmetrics(policy, cells, q0, dqdx);
    return dqdx;
}


// execution policies: the reference's serial and OpenMP policies (par.hpp) select the GPU path
// with its defaults; hip_policy sets its options explicitly
namespace par::execution {
struct hip_policy {
    int device = 0;
    int precision = HC_PRECISION_FP64;   // or HC_PRECISION_MIXED
    int dt_mode = HC_DT_GLOBAL;          // HC_DT_LOCAL: local time stepping with the stage-0 spectral radius
    int graph_steps = 50;
};
inline constexpr hip_policy hip{};
}   // namespace par::execution
namespace par {
template <> struct is_execution_policy<execution::hip_policy> : std::true_type {};
}   // namespace par

// ---------------- device sessions ----------------

namespace hcamd_detail {
inline void check(int rc, hc_context* ctx, const char* what)
{
    if (rc != HC_OK) throw std::runtime_error(std::string(what) + ": " + hc_last_error(ctx));
}

template <typename Policy> constexpr par::execution::hip_policy hip_of(const Policy& p)
{
    if constexpr (std::is_same_v<Policy, par::execution::hip_policy>) return p;
    else return par::execution::hip_policy{};
}

template <typename BC> constexpr int bc_id(const BC& bc) { return static_cast<int>(bc.type()); }

template <typename... BCs> bool has_bc(int type, const std::tuple<BCs...>& bcs)
{
    bool found = false;
    std::apply([&](const auto&... bc) { ((found = found || bc_id(bc) == type), ...); }, bcs);
    return found;
}

inline int euler_bc_code(int t)   // EulerBCs -> HC_BC_*
{
    switch (t) {
    case 0: return HC_BC_PERIODIC;
    case 1: return HC_BC_RIEMANN;
    case 2: return HC_BC_ENTROPY;
    case 3: return HC_BC_FIXED;
    default: return HC_BC_WALL;
    }
}
inline int scalar_bc_code(int t) { return t == 0 ? HC_BC_PERIODIC : t == 1 ? HC_BC_RIEMANN : HC_BC_FIXED; }

// a device context set up for one mesh and one scheme; its key says what it was built from
struct Session {
    hc_context* ctx = nullptr;
    std::uint64_t nodes_hash = 0;
    std::size_t ni = 0, nj = 0;
    hc_case_desc desc{};
    int device = 0, precision = -1, graph_steps = -1;
    ~Session()
    {
        if (ctx) hc_destroy(ctx);
    }
};

inline std::uint64_t hash_bytes(const void* p, std::size_t n, std::uint64_t h = 1469598103934665603ULL)
{
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (std::size_t k = 0; k < n; ++k) h = (h ^ b[k]) * 1099511628211ULL;
    return h;
}

// The sessions of this thread, most recent last; a few are kept so that a time loop alternating
// boundaryUpdate and residualCalc (two schemes on one mesh) rebuilds nothing.
inline std::vector<std::unique_ptr<Session>>& sessions()
{
    thread_local std::vector<std::unique_ptr<Session>> s;
    return s;
}
constexpr std::size_t MAX_SESSIONS = 4;

// The session for (mesh, case descriptor, policy): reused when the mesh nodes, the descriptor and
// the device match, else created (geometry built once on hc_set_case).
inline hc_context* session(const par::execution::hip_policy& pol, const double* nodes, std::size_t ni, std::size_t nj,
                           const hc_case_desc& d)
{
    const std::uint64_t h = hash_bytes(nodes, 2 * (ni + 1) * (nj + 1) * sizeof(double));
    auto& all = sessions();
    for (std::size_t k = 0; k < all.size(); ++k) {
        Session& s = *all[k];
        if (s.nodes_hash == h && s.ni == ni && s.nj == nj && s.device == pol.device &&
            std::memcmp(&s.desc, &d, sizeof(d)) == 0) {
            std::unique_ptr<Session> keep = std::move(all[k]);
            all.erase(all.begin() + k);
            all.push_back(std::move(keep));
            Session& t = *all.back();
            if (t.precision != pol.precision) {
                check(hc_set_precision(t.ctx, pol.precision), t.ctx, "hc_set_precision");
                t.precision = pol.precision;
            }
            if (t.graph_steps != pol.graph_steps) {
                check(hc_set_graph_steps(t.ctx, pol.graph_steps), t.ctx, "hc_set_graph_steps");
                t.graph_steps = pol.graph_steps;
            }
            return t.ctx;
        }
    }
    if (all.size() >= MAX_SESSIONS) all.erase(all.begin());
    auto s = std::make_unique<Session>();
    if (hc_create(&s->ctx, pol.device) != HC_OK) throw std::runtime_error("hc_create failed");
    check(hc_set_mesh(s->ctx, (int)ni, (int)nj, nodes), s->ctx, "hc_set_mesh");
    check(hc_set_case(s->ctx, &d), s->ctx, "hc_set_case");
    check(hc_set_precision(s->ctx, pol.precision), s->ctx, "hc_set_precision");
    check(hc_set_graph_steps(s->ctx, pol.graph_steps), s->ctx, "hc_set_graph_steps");
    s->nodes_hash = h;
    s->ni = ni;
    s->nj = nj;
    s->desc = d;
    s->device = pol.device;
    s->precision = pol.precision;
    s->graph_steps = pol.graph_steps;
    all.push_back(std::move(s));
    return all.back()->ctx;
}

template <floating_point Real> std::vector<double> flat_nodes(const Mesh<2, Real>& mesh)
{
    static_assert(std::is_same_v<Real, double>, "the GPU path takes FP64 geometry");
    const auto& n = mesh.nodes.flatten();
    std::vector<double> out(2 * n.size());
    for (std::size_t k = 0; k < n.size(); ++k) {
        out[2 * k] = n[k][0];
        out[2 * k + 1] = n[k][1];
    }
    return out;
}

// the case descriptor of a scheme: flux / limiter / scalings (make_muscl_flux), the boundary types
// of the field and the tuple that must hold them, the Runge-Kutta table and CFL, the species
template <LawType Law, auto Basis, floating_point Real, typename... BCs>
hc_case_desc describe(int flux_id, int vscale, int pscale, int limiter_id, const std::tuple<BCs...>& bcs,
                      const std::array<BoundaryType<Law>, 4>& bcTypes, const Species<Law, Real>& species,
                      const Mesh<2, Real>& mesh, int nstages, const double (*alpha)[6], const double* beta, double cfl,
                      int dt_mode)
{
    hc_case_desc d;
    std::memset(&d, 0, sizeof(d));
    d.law = Law == LawType::Euler ? HC_LAW_EULER : HC_LAW_SCALAR;
    d.basis = basis_id<Basis>();
    d.flux = flux_id;
    d.vscale = vscale;
    d.pscale = pscale;
    d.limiter = limiter_id;
    for (int b = 0; b < 4; ++b) {
        const int t = static_cast<int>(bcTypes[b]);
        if (!has_bc(t, bcs)) throw std::runtime_error("no boundary condition in the tuple matches bcTypes");
        d.bc[b] = Law == LawType::Euler ? euler_bc_code(t) : scalar_bc_code(t);
    }
    d.nstages = nstages;
    for (int s = 0; s < 6; ++s) {
        d.beta[s] = beta[s];
        for (int k = 0; k < 6; ++k) d.alpha[s][k] = alpha[s][k];
    }
    d.cfl = cfl;
    d.dt_mode = dt_mode;
    d.period_i[0] = mesh.period_i[0];
    d.period_i[1] = mesh.period_i[1];
    d.period_j[0] = mesh.period_j[0];
    d.period_j[1] = mesh.period_j[1];
    if constexpr (Law == LawType::Euler)
        d.sp = hc_species{species.gamma, species.minf, species.lref, species.nu, species.pr, species.dt, species.R,
                          species.gamma1};
    return d;
}

// boundary reference values (k, 1) of the field in the C-ABI's row order
template <typename VarSet> std::vector<double> boundary_refs(const SolutionField<VarSet, 2>& q)
{
    constexpr int nv = VarSet::N;
    std::vector<double> bref;
    for (int b = 0; b < 4; ++b)
        for (std::size_t k = 0; k < q.boundary[b].shape(0); ++k)
            for (int v = 0; v < nv; ++v) bref.push_back(q.boundary[b]({k, 1})[v]);
    return bref;
}

template <typename VarSet> const double* raw(const par::DualArray<VarSet, 2>& a)
{
    static_assert(sizeof(VarSet) == sizeof(double) * VarSet::N, "packed variable sets");
    return reinterpret_cast<const double*>(a.data());
}
template <typename VarSet> double* raw(par::DualArray<VarSet, 2>& a)
{
    static_assert(sizeof(VarSet) == sizeof(double) * VarSet::N, "packed variable sets");
    return reinterpret_cast<double*>(a.data());
}

template <floating_point Real> void rk_arrays(const ODE::Explicit::RungeKutta<Real>& rk, double (*alpha)[6], double* beta)
{
    for (int s = 0; s < 6; ++s) {
        beta[s] = rk.beta[s];
        for (int k = 0; k < 6; ++k) alpha[s][k] = rk.alpha[s][k];
    }
}
}   // namespace hcamd_detail

// ---------------- spatial entry points ----------------

/*
 * boundaryUpdate(mesh, bcs, species, q) (spatial/boundary/boundaryUpdate.h:28-48, 137-266): ghost
 * values q.boundary[b](k, 0) of every non-periodic boundary from the interior and the reference
 * values (Riemann / entropy far field, Fixed, inviscid wall), computed on the GPU.
 */
template <LawType Law, int nDim, floating_point Real, typename SolVarT, typename... BoundaryConds>
void boundaryUpdate(const Mesh<nDim, Real>& mesh, const std::tuple<BoundaryConds...> bcs, const Species<Law, Real>& species,
                    SolutionField<SolVarT, nDim>& q)
{
    static_assert(nDim == 2, "two-dimensional meshes");
    assert(mesh.cells.shape() == q.interior.shape());
    const auto rk = ODE::Explicit::ssp11<Real>();
    double alpha[6][6], beta[6];
    hcamd_detail::rk_arrays(rk, alpha, beta);
    const hc_case_desc d = hcamd_detail::describe<Law, varset_traits<SolVarT>::basis>(
        HC_FLUX_CENTRAL, 0, 0, HC_LIM_NOLIMIT1, bcs, q.bcTypes, species, mesh, 1, alpha, beta, 1.0, HC_DT_GLOBAL);
    const auto nodes = hcamd_detail::flat_nodes(mesh);
    hc_context* ctx = hcamd_detail::session(par::execution::hip_policy{}, nodes.data(), mesh.cells.shape(0),
                                            mesh.cells.shape(1), d);
    constexpr int nv = SolVarT::N;
    const std::vector<double> bref = hcamd_detail::boundary_refs(q);
    std::vector<double> ghost(bref.size());
    std::size_t row = 0;
    for (int b = 0; b < 4; ++b)
        for (std::size_t k = 0; k < q.boundary[b].shape(0); ++k, ++row)
            for (int v = 0; v < nv; ++v) ghost[row * nv + v] = q.boundary[b]({k, 0})[v];
    hcamd_detail::check(hc_boundary_update(ctx, hcamd_detail::raw(q.interior), bref.data(), ghost.data()), ctx,
                        "hc_boundary_update");
    row = 0;
    for (int b = 0; b < 4; ++b)
        for (std::size_t k = 0; k < q.boundary[b].shape(0); ++k, ++row)
            for (int v = 0; v < nv; ++v) q.boundary[b]({k, 0})[v] = ghost[row * nv + v];
}

/*
 * residualCalc(policy, hoflux, bcs, species, mesh, q, dxdx, dqdx, res) (spatial/residualCalc.h:27-69):
 * R_i = -sum_faces F A and lambda_i = sum_faces lambda_f A over the interior and boundary faces in
 * the reference's order, on the GPU.  The fused kernel forms its LSQ metrics itself (the X metric
 * once with the geometry, the Q metric per call from q, with the periodic images across periodic
 * sides); dxdx and dqdx must have the mesh's shape.  The ghost values are computed from the
 * interior and reference values as boundaryUpdate does.
 */
template <typename Policy, LawType Law, int nDim, typename SolVarT, typename FluxRes, typename HighOrderFlux,
          typename... BoundaryConds, floating_point Real>
void residualCalc(const Policy policy, const HighOrderFlux& hoflux, const std::tuple<BoundaryConds...> bcs,
                  const Species<Law, Real>& species, const Mesh<nDim, Real>& mesh, const SolutionField<SolVarT, nDim>& q,
                  const par::DualArray<lsq::XMetric<nDim, Real>, nDim>& dxdx,
                  const par::DualArray<lsq::QMetric<SolVarT>, nDim>& dqdx, par::DualArray<FluxRes, nDim>& res)
{
    static_assert(std::is_same_v<FluxRes, FluxResult<Law, nDim, Real>>, "residual array of FluxResult");
    assert(mesh.cells.shape() == q.interior.shape() && mesh.cells.shape() == dxdx.shape() &&
           mesh.cells.shape() == dqdx.shape() && mesh.cells.shape() == res.shape());
    (void)hoflux;
    const auto rk = ODE::Explicit::ssp11<Real>();
    double alpha[6][6], beta[6];
    hcamd_detail::rk_arrays(rk, alpha, beta);
    const hc_case_desc d = hcamd_detail::describe<Law, varset_traits<SolVarT>::basis>(
        HighOrderFlux::flux_id, HighOrderFlux::vscale, HighOrderFlux::pscale, HighOrderFlux::limiter_id, bcs, q.bcTypes,
        species, mesh, 1, alpha, beta, 1.0, HC_DT_GLOBAL);
    const auto nodes = hcamd_detail::flat_nodes(mesh);
    const auto pol = hcamd_detail::hip_of(policy);
    hc_context* ctx = hcamd_detail::session(pol, nodes.data(), mesh.cells.shape(0), mesh.cells.shape(1), d);
    const std::vector<double> bref = hcamd_detail::boundary_refs(q);
    hcamd_detail::check(hc_set_boundary_refs(ctx, bref.data()), ctx, "hc_set_boundary_refs");
    hcamd_detail::check(hc_set_state(ctx, hcamd_detail::raw(q.interior)), ctx, "hc_set_state");
    static_assert(sizeof(FluxRes) == sizeof(double) * (SolVarT::N + 1), "packed flux results");
    hcamd_detail::check(hc_residual(ctx, reinterpret_cast<double*>(res.data())), ctx, "hc_residual");
}

/*
 * spectralRadius(policy, cells, res) (spatial/spectralRadius.h:19-48): max over cells of
 * lambda / volume, initial value numeric_limits<Real>::min() (:42).
 */
template <typename Policy, int nDim, typename FluxRes, floating_point Real>
Real spectralRadius(const Policy policy, const MeshCellArray<nDim, Real>& cells, const par::DualArray<FluxRes, nDim>& res)
{
    (void)policy;
    return par::transform_reduce(
        par::execution::seq, [](const geom::Volume<nDim, Real>& c, const FluxRes& r) { return r.lambda / c.volume; },
        [](const Real l, const Real r) { return std::max(l, r); }, std::numeric_limits<Real>::min(), cells, res);
}

/*
 * rungeKuttaAccumulation(policy, rk, stg, resStage, resTotal) (timestepping/rungeKutta.h:176-197):
 * resTotal.flux = sum_{k <= stg} alpha[stg][k] resStage[k].flux in that order; lambda stays 0.
 */
template <typename Policy, int nDim, typename FluxRes, floating_point Real>
void rungeKuttaAccumulation(const Policy policy, const ODE::Explicit::RungeKutta<Real>& rk, const std::size_t stg,
                            const std::vector<par::DualArray<FluxRes, nDim>>& resStage,
                            par::DualArray<FluxRes, nDim>& resTotal)
{
    (void)policy;
    par::fill(resTotal, FluxRes{});
    const std::size_t n = resTotal.flattened_length();
    for (std::size_t c = 0; c < n; ++c) {
        FluxRes& t = resTotal.flatten(c);
        for (std::size_t k = 0; k <= stg; ++k) t.flux = t.flux + rk.alpha[stg][k] * resStage[k].flatten(c).flux;
    }
}

namespace hcamd_detail {
template <LawType Law, auto Basis, floating_point Real, typename SolVarT>
SolVarT forward(const Species<Law, Real>& species, const SolVarT& v0, const Real* R, const Real fac)
{
    constexpr int L = Law == LawType::Euler ? hc::LAW_EULER : hc::LAW_SCALAR;
    constexpr int B = basis_id<Basis>();
    const auto sp = hcamd_detail::species(species);
    Real qc[SolVarT::N];
    hc::to_conserved<Real, L, B>(sp, v0.elems.data(), qc);
    for (int v = 0; v < SolVarT::N; ++v) qc[v] = qc[v] + R[v] * fac;
    SolVarT v1;
    hc::from_conserved<Real, L, B>(sp, qc, v1.elems.data());
    return v1;
}
}   // namespace hcamd_detail

/*
 * eulerForwardUpdateGlobal(policy, cells, species, dt, r, q0, q1) (spatial/eulerForwardUpdate.h:38-92):
 * q1 = S(C(q0) + r.flux (dt / volume)), S and C the solution-basis <-> conserved transforms.
 */
template <typename Policy, LawType Law, int nDim, floating_point Real, typename SolVarT>
void eulerForwardUpdateGlobal(const Policy policy, const MeshCellArray<nDim, Real>& cells, const Species<Law, Real>& species,
                              const Real dt, const par::DualArray<FluxResult<Law, nDim, Real>, nDim>& r,
                              const par::DualArray<SolVarT, nDim>& q0, par::DualArray<SolVarT, nDim>& q1)
{
    (void)policy;
    par::transform(
        [&](const SolVarT& v0, const geom::Volume<nDim, Real>& c, const FluxResult<Law, nDim, Real>& res) {
            return hcamd_detail::forward<Law, varset_traits<SolVarT>::basis>(species, v0, res.flux.elems.data(),
                                                                            dt / c.volume);
        },
        q1, q0, cells, r);
}

/*
 * eulerForwardUpdateLocal(policy, species, cfl, r, q0, q1) (spatial/eulerForwardUpdate.h:117-168):
 * q1 = S(C(q0) + r.flux (cfl / r.lambda)); the lambda of an accumulated residual is 0 in the
 * reference (rungeKutta.h:176-197), so callers pass the stage-0 spectral radius there.
 */
template <typename Policy, LawType Law, int nDim, floating_point Real, typename SolVarT>
void eulerForwardUpdateLocal(const Policy policy, const Species<Law, Real>& species, const Real cfl,
                             const par::DualArray<FluxResult<Law, nDim, Real>, nDim>& r,
                             const par::DualArray<SolVarT, nDim>& q0, par::DualArray<SolVarT, nDim>& q1)
{
    (void)policy;
    par::transform(
        [&](const SolVarT& v0, const FluxResult<Law, nDim, Real>& res) {
            return hcamd_detail::forward<Law, varset_traits<SolVarT>::basis>(species, v0, res.flux.elems.data(),
                                                                            cfl / res.lambda);
        },
        q1, q0, r);
}

/*
 * gradientCalc(policy, mesh, q[, dq]) (spatial/gradientCalc.h:16-313): per cell and direction the
 * sum of the neighbour differences q_R - q_L over its edges, the periodic wraps stopping one short
 * of the last row / column like the reference's loops; on the GPU.
 */
template <typename Policy, int nDim, floating_point Real, typename VarSetT>
void gradientCalc(const Policy policy, const Mesh<nDim, Real>& mesh, const SolutionField<VarSetT, nDim>& q,
                  par::DualArray<std::array<vardelta_t<VarSetT>, 2>, 2>& dq)
{
    static_assert(nDim == 2, "two-dimensional meshes");
    constexpr LawType Law = law_of_v<VarSetT>;
    assert(mesh.cells.shape() == q.interior.shape() && mesh.cells.shape() == dq.shape());
    const auto rk = ODE::Explicit::ssp11<Real>();
    double alpha[6][6], beta[6];
    hcamd_detail::rk_arrays(rk, alpha, beta);
    // the boundary kinds matter only through their periodicity here
    std::tuple<BoundaryCondition<Law, BoundaryType<Law>::Periodic>, BoundaryCondition<Law, BoundaryType<Law>::Fixed>> kinds;
    std::array<BoundaryType<Law>, 4> types;
    for (int b = 0; b < 4; ++b)
        types[b] = q.bcTypes[b] == BoundaryType<Law>::Periodic ? BoundaryType<Law>::Periodic : BoundaryType<Law>::Fixed;
    Species<Law, Real> sp{};
    if constexpr (Law == LawType::Euler) sp = get_air_species<Real>();
    const hc_case_desc d = hcamd_detail::describe<Law, varset_traits<VarSetT>::basis>(
        HC_FLUX_CENTRAL, 0, 0, HC_LIM_NOLIMIT1, kinds, types, sp, mesh, 1, alpha, beta, 1.0, HC_DT_GLOBAL);
    const auto nodes = hcamd_detail::flat_nodes(mesh);
    hc_context* ctx =
        hcamd_detail::session(hcamd_detail::hip_of(policy), nodes.data(), mesh.cells.shape(0), mesh.cells.shape(1), d);
    static_assert(sizeof(std::array<vardelta_t<VarSetT>, nDim>) == sizeof(double) * nDim * VarSetT::N, "packed deltas");
    hcamd_detail::check(hc_gradient(ctx, hcamd_detail::raw(q.interior), reinterpret_cast<double*>(dq.data())), ctx,
                        "hc_gradient");
}
template <typename Policy, int nDim, floating_point Real, typename VarSetT>
par::DualArray<std::array<vardelta_t<VarSetT>, 2>, 2> gradientCalc(const Policy policy, const Mesh<nDim, Real>& mesh,
                                                                  const SolutionField<VarSetT, nDim>& q)
{
    par::DualArray<std::array<vardelta_t<VarSetT>, 2>, 2> dq(q.interior.shape());
    gradientCalc(policy, mesh, q, dq);
    return dq;
}

// ---------------- time integration on the GPU ----------------

namespace hcamd_detail {
template <typename Policy, LawType Law, int nDim, floating_point Real, typename SolVarT, typename SecondOrderFlux,
          typename... BoundaryConds>
hc_context* integrate_session(const Policy policy, const ODE::Explicit::RungeKutta<Real>& rk, Real cfl,
                              const SecondOrderFlux& flux2, const std::tuple<BoundaryConds...>& bcs,
                              const Species<Law, Real>& species, const Mesh<nDim, Real>& mesh,
                              const SolutionField<SolVarT, nDim>& q0)
{
    static_assert(nDim == 2, "two-dimensional meshes");
    assert(q0.interior.shape() == mesh.cells.shape());
    (void)flux2;
    double alpha[6][6], beta[6];
    rk_arrays(rk, alpha, beta);
    const auto pol = hip_of(policy);
    const hc_case_desc d = describe<Law, varset_traits<SolVarT>::basis>(
        SecondOrderFlux::flux_id, SecondOrderFlux::vscale, SecondOrderFlux::pscale, SecondOrderFlux::limiter_id, bcs,
        q0.bcTypes, species, mesh, (int)rk.nstages, alpha, beta, cfl, pol.dt_mode);
    const auto nodes = flat_nodes(mesh);
    hc_context* ctx = session(pol, nodes.data(), mesh.cells.shape(0), mesh.cells.shape(1), d);
    const std::vector<double> bref = boundary_refs(q0);
    check(hc_set_boundary_refs(ctx, bref.data()), ctx, "hc_set_boundary_refs");
    check(hc_set_state(ctx, raw(q0.interior)), ctx, "hc_set_state");
    unsigned int status = 0;
    check(hc_get_status(ctx, &status), ctx, "hc_get_status");   // clear what earlier calls left
    check(hc_set_residual_smoothing(ctx, -1, 1.0), ctx, "hc_set_residual_smoothing");
    return ctx;
}
}   // namespace hcamd_detail

/*
 * integrate(policy, timeControls, rungeKutta, flux2, boundaryConds, species, mesh, q0)
 * (timestepping/rungeKutta.h:44-167): the whole time loop on the GPU (hc_step: fused stage
 * kernels replayed from HIP graphs); the final interior state is written back into q0.  Returns
 * the accumulated time of a global time step (0 for a local one).
 */
template <typename Policy, LawType Law, int nDim, floating_point Real, typename SolVarT, typename SecondOrderFlux,
          typename... BoundaryConds>
    requires par::execution_policy<Policy>
Real integrate(const Policy policy, const UnsteadyTimeControls<Real>& timeControls,
               const ODE::Explicit::RungeKutta<Real>& rungeKutta, const SecondOrderFlux& flux2,
               const std::tuple<BoundaryConds...> boundaryConds, const Species<Law, Real>& species,
               const Mesh<nDim, Real>& mesh, SolutionField<SolVarT, nDim>& q0)
{
    hc_context* ctx = hcamd_detail::integrate_session(policy, rungeKutta, timeControls.cfl, flux2, boundaryConds,
                                                      species, mesh, q0);
    hcamd_detail::check(hc_step(ctx, (int)timeControls.nTimesteps), ctx, "hc_step");
    hcamd_detail::check(hc_get_state(ctx, hcamd_detail::raw(q0.interior)), ctx, "hc_get_state");
    double t = 0.0;
    hcamd_detail::check(hc_get_time(ctx, &t), ctx, "hc_get_time");
    unsigned int status = 0;
    hcamd_detail::check(hc_get_status(ctx, &status), ctx, "hc_get_status");
    if (status) throw std::runtime_error("device raised status bits " + std::to_string(status));
    return Real(t);
}

// the reference drivers' call without a policy (script/euler/cylinder2D.cpp:184-190)
template <LawType Law, int nDim, floating_point Real, typename SolVarT, typename SecondOrderFlux, typename... BoundaryConds>
Real integrate(const UnsteadyTimeControls<Real>& timeControls, const ODE::Explicit::RungeKutta<Real>& rungeKutta,
               const SecondOrderFlux& flux2, const std::tuple<BoundaryConds...> boundaryConds,
               const Species<Law, Real>& species, const Mesh<nDim, Real>& mesh, SolutionField<SolVarT, nDim>& q0)
{
    return integrate(par::execution::hip, timeControls, rungeKutta, flux2, boundaryConds, species, mesh, q0);
}

/*
 * Steady-state integrate with SteadyTimeControls (controls.h:31-48, no integrate consumes it at the
 * reference's HEAD): marches until the mass residual's L2 norm has dropped by
 * timeControls.residual_drop (checked every 50 steps) or nTimesteps are taken, writes the state
 * back into q0 and returns the drop reached.  use_residual_smoothing smooths every stage's
 * accumulated residual with the implicit operator of OLD.fieldOperations/centralImplicitSmoothing.ipp
 * (along i, then along j) at alpha = smoothing_factor, which multiplies the stable time step.
 */
template <typename Policy, LawType Law, int nDim, floating_point Real, typename SolVarT, typename SecondOrderFlux,
          typename... BoundaryConds>
    requires par::execution_policy<Policy>
Real integrate(const Policy policy, const SteadyTimeControls<Real>& timeControls,
               const ODE::Explicit::RungeKutta<Real>& rungeKutta, const SecondOrderFlux& flux2,
               const std::tuple<BoundaryConds...> boundaryConds, const Species<Law, Real>& species,
               const Mesh<nDim, Real>& mesh, SolutionField<SolVarT, nDim>& q0)
{
    hc_context* ctx = hcamd_detail::integrate_session(policy, rungeKutta, timeControls.cfl, flux2, boundaryConds,
                                                      species, mesh, q0);
    if (timeControls.use_residual_smoothing)
        hcamd_detail::check(hc_set_residual_smoothing(ctx, 1, timeControls.smoothing_factor), ctx,
                            "hc_set_residual_smoothing");
    long long steps = 0;
    double drop = 1.0;
    hcamd_detail::check(hc_step_steady(ctx, (long long)timeControls.nTimesteps, timeControls.residual_drop, 50, &steps,
                                       &drop), ctx, "hc_step_steady");
    hcamd_detail::check(hc_get_state(ctx, hcamd_detail::raw(q0.interior)), ctx, "hc_get_state");
    return Real(drop);
}
