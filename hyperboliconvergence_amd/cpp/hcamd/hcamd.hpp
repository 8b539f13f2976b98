// hcamd.hpp -- C++20 drop-in for the reference's structured-mesh finite-volume hot path.
//
// Same names, template parameters and argument order as the reference's header-only API
// (lacrymose/hyperboliConvergence cpp/inc/): flux functors, limiters, make_muscl_flux,
// boundary-condition factories, Runge-Kutta tables, UnsteadyTimeControls, Mesh,
// SolutionField, residualCalc and integrate.  The point-wise functors evaluate on the host
// with the same code the HIP kernels run; residualCalc and integrate run on the GPU through
// the C-ABI of include/hcamd.h.  Link with -lhcamd (hyperboliconvergence_amd/libhcamd.so).
#pragma once

#include "hcamd.h"
#include "../../csrc/hc_point.hpp"

#include <array>
#include <cassert>
#include <cmath>
#include <cstddef>
#include <limits>
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

// ---------------- structured mesh (mesh/mesh.h:19-57, generate/twoD.h) ----------------

template <int nDim, floating_point Real> struct Mesh;

template <floating_point Real> struct Mesh<2, Real> {
    using Node = geom::Point<2, Real>;
    std::size_t ni = 0, nj = 0;          // cells
    std::vector<Node> nodes;             // (ni+1) x (nj+1), row-major
    const Node& node(std::size_t i, std::size_t j) const { return nodes[i * (nj + 1) + j]; }
    Node& node(std::size_t i, std::size_t j) { return nodes[i * (nj + 1) + j]; }
    std::array<Real, 2> period_i{0, 0};   // translation to the periodic image across boundary 1
    std::array<Real, 2> period_j{0, 0};   // ... across boundary 3
};

// mesh/generate/twoD.h:15-40
template <floating_point Real>
Mesh<2, Real> make_linspace_mesh(const std::array<std::size_t, 2>& cellDims, Real lox, Real hix, Real loy, Real hiy)
{
    Mesh<2, Real> m;
    m.ni = cellDims[0];
    m.nj = cellDims[1];
    m.nodes.resize((m.ni + 1) * (m.nj + 1));
    const Real dx = (hix - lox) / m.ni, dy = (hiy - loy) / m.nj;
    for (std::size_t i = 0; i <= m.ni; ++i)
        for (std::size_t j = 0; j <= m.nj; ++j) m.node(i, j) = {{lox + i * dx, loy + j * dy}};
    m.period_i = {hix - lox, 0};
    m.period_j = {0, hiy - loy};
    return m;
}

// mesh/generate/twoD.h:139-179 (uniform radial spacing, i = radius, j = theta)
template <floating_point Real>
Mesh<2, Real> make_cylinder_mesh(const std::array<Real, 2> lengths, const std::array<std::size_t, 2> ncells)
{
    Mesh<2, Real> m;
    m.ni = ncells[0];
    m.nj = ncells[1];
    m.nodes.resize((m.ni + 1) * (m.nj + 1));
    const Real dr = (lengths[1] - lengths[0]) / m.ni;
    const Real dth = (2. * M_PI) / m.nj;
    for (std::size_t i = 0; i <= m.ni; ++i)
        for (std::size_t j = 0; j <= m.nj; ++j) {
            const Real r = lengths[0] + i * dr, th = j * dth;
            m.node(i, j) = {{r * std::cos(th), r * std::sin(th)}};
        }
    return m;
}

// O-grid with geometric radial stretching (first spacing dr0) and node nth identical to node 0
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
    Mesh<2, Real> m;
    m.ni = nr;
    m.nj = nth;
    m.nodes.resize((nr + 1) * (nth + 1));
    for (std::size_t i = 0; i <= nr; ++i) {
        const Real r = (i == nr) ? lengths[1] : lengths[0] + dr0 * (std::pow(s, Real(i)) - 1.) / (s - 1.);
        for (std::size_t j = 0; j <= nth; ++j) {
            const Real th = (2. * M_PI * Real(j == nth ? 0 : j)) / nth;
            m.node(i, j) = {{r * std::cos(th), r * std::sin(th)}};
        }
    }
    return m;
}

// ---------------- solution field (solutionField/solutionField.h:56-91) ----------------

template <typename VarSet, int nDim> struct SolutionField;

template <typename VarSet> struct SolutionField<VarSet, 2> {
    static constexpr unsigned int nBoundaries = 4;
    std::size_t ni, nj;
    std::vector<VarSet> interior;                   // ni x nj, row-major
    std::array<std::vector<VarSet>, 4> boundary;    // per boundary: n_along x 2 ([.,0] ghost, [.,1] reference)
    std::array<int, 4> bcTypes{};                   // BoundaryType<Law> values
    explicit SolutionField(const std::array<std::size_t, 2>& s) : ni(s[0]), nj(s[1]), interior(s[0] * s[1])
    {
        boundary[0].resize(2 * nj);
        boundary[1].resize(2 * nj);
        boundary[2].resize(2 * ni);
        boundary[3].resize(2 * ni);
    }
    VarSet& operator()(std::size_t i, std::size_t j) { return interior[i * nj + j]; }
    const VarSet& operator()(std::size_t i, std::size_t j) const { return interior[i * nj + j]; }
};

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

// execution policies (parallalg/parallalg.h:65-87); the work runs on the GPU whichever is passed
namespace par {
namespace execution {
struct sequenced_policy {};
struct openmp_policy {};
struct hip_policy {
    int device = 0;
    int precision = HC_PRECISION_FP64;   // or HC_PRECISION_MIXED
    int dt_mode = HC_DT_GLOBAL;          // HC_DT_LOCAL: local time stepping with the stage-0 spectral radius
    int graph_steps = 50;
};
inline constexpr sequenced_policy seq{};
inline constexpr openmp_policy omp{};
inline constexpr hip_policy hip{};
}   // namespace execution
}   // namespace par

// ---------------- GPU forwarding ----------------

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

template <typename BC> constexpr int bc_id(const BC& bc)
{
    return static_cast<int>(bc.type());   // enumerators share the numbering of HC_BC_* for Periodic..Fixed
}

template <LawType Law, typename... BCs> int lookup_bc(int type, const std::tuple<BCs...>& bcs)
{
    int found = -1;
    std::apply([&](const auto&... bc) { ((found = (bc_id(bc) == type ? type : found)), ...); }, bcs);
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
inline int scalar_bc_code(int t)   // ScalarAdvectionBCs -> HC_BC_*
{
    return t == 0 ? HC_BC_PERIODIC : t == 1 ? HC_BC_RIEMANN : HC_BC_FIXED;
}

// One GPU context set up from the reference-style arguments.
template <LawType Law, int nDim, auto Basis, floating_point Real, typename HoFlux, typename... BCs>
hc_context* setup(const par::execution::hip_policy& pol, const HoFlux& hoflux, const std::tuple<BCs...>& bcs,
                  const Species<Law, Real>& species, const Mesh<2, Real>& mesh,
                  const SolutionField<VariableSet<Law, 2, Basis, Real>, 2>& q, int nstages, const double (*alpha)[6],
                  const double* beta, double cfl)
{
    static_assert(std::is_same_v<Real, double>, "the GPU path takes FP64 state");
    hc_context* ctx = nullptr;
    if (hc_create(&ctx, pol.device) != HC_OK) throw std::runtime_error("hc_create failed");
    std::vector<double> nodes(2 * mesh.nodes.size());
    for (std::size_t k = 0; k < mesh.nodes.size(); ++k) {
        nodes[2 * k] = mesh.nodes[k][0];
        nodes[2 * k + 1] = mesh.nodes[k][1];
    }
    check(hc_set_mesh(ctx, (int)mesh.ni, (int)mesh.nj, nodes.data()), ctx, "hc_set_mesh");
    hc_case_desc d{};
    d.law = Law == LawType::Euler ? HC_LAW_EULER : HC_LAW_SCALAR;
    d.basis = basis_id<Basis>();
    d.flux = HoFlux::flux_id;
    d.vscale = HoFlux::vscale;
    d.pscale = HoFlux::pscale;
    d.limiter = HoFlux::limiter_id;
    for (int b = 0; b < 4; ++b) {
        const int t = q.bcTypes[b];
        if (lookup_bc<Law>(t, bcs) < 0) throw std::runtime_error("no boundary condition in the tuple matches bcTypes");
        d.bc[b] = Law == LawType::Euler ? euler_bc_code(t) : scalar_bc_code(t);
    }
    d.nstages = nstages;
    for (int s = 0; s < 6; ++s) {
        d.beta[s] = beta[s];
        for (int k = 0; k < 6; ++k) d.alpha[s][k] = alpha[s][k];
    }
    d.cfl = cfl;
    d.dt_mode = pol.dt_mode;
    d.period_i[0] = mesh.period_i[0];
    d.period_i[1] = mesh.period_i[1];
    d.period_j[0] = mesh.period_j[0];
    d.period_j[1] = mesh.period_j[1];
    if constexpr (Law == LawType::Euler)
        d.sp = hc_species{species.gamma, species.minf, species.lref, species.nu, species.pr, species.dt, species.R,
                          species.gamma1};
    check(hc_set_case(ctx, &d), ctx, "hc_set_case");
    check(hc_set_precision(ctx, pol.precision), ctx, "hc_set_precision");
    constexpr int nv = nVar<Law, 2>;
    std::vector<double> bref;
    for (int b = 0; b < 4; ++b) {
        const std::size_t n_along = b < 2 ? q.nj : q.ni;
        for (std::size_t k = 0; k < n_along; ++k)
            for (int v = 0; v < nv; ++v) bref.push_back(q.boundary[b][2 * k + 1][v]);
    }
    check(hc_set_boundary_refs(ctx, bref.data()), ctx, "hc_set_boundary_refs");
    check(hc_set_state(ctx, reinterpret_cast<const double*>(q.interior.data())), ctx, "hc_set_state");
    check(hc_set_graph_steps(ctx, pol.graph_steps), ctx, "hc_set_graph_steps");
    return ctx;
}
}   // namespace hcamd_detail

/*
 * integrate(policy, timeControls, rungeKutta, flux2, boundaryConds, species, mesh, q0)
 * (timestepping/rungeKutta.h:37-167): runs the time loop on the GPU and writes the final
 * interior state back into q0.  Returns the accumulated time.
 */
template <typename Policy, LawType Law, int nDim, floating_point Real, auto Basis, typename SecondOrderFlux,
          typename... BoundaryConds>
Real integrate(const Policy policy, const UnsteadyTimeControls<Real>& timeControls,
               const ODE::Explicit::RungeKutta<Real>& rungeKutta, const SecondOrderFlux& flux2,
               const std::tuple<BoundaryConds...> boundaryConds, const Species<Law, Real>& species,
               const Mesh<nDim, Real>& mesh, SolutionField<VariableSet<Law, nDim, Basis, Real>, nDim>& q0)
{
    static_assert(sizeof(VariableSet<Law, nDim, Basis, Real>) == sizeof(Real) * nVar<Law, nDim>, "packed variable sets");
    double alpha[6][6], beta[6];
    for (int s = 0; s < 6; ++s) {
        beta[s] = rungeKutta.beta[s];
        for (int k = 0; k < 6; ++k) alpha[s][k] = rungeKutta.alpha[s][k];
    }
    hc_context* ctx = hcamd_detail::setup<Law, nDim, Basis>(hcamd_detail::hip_of(policy), flux2, boundaryConds, species,
                                                            mesh, q0, (int)rungeKutta.nstages, alpha, beta,
                                                            timeControls.cfl);
    hcamd_detail::check(hc_step(ctx, (int)timeControls.nTimesteps), ctx, "hc_step");
    hcamd_detail::check(hc_get_state(ctx, reinterpret_cast<double*>(q0.interior.data())), ctx, "hc_get_state");
    double t = 0.0;
    hcamd_detail::check(hc_get_time(ctx, &t), ctx, "hc_get_time");
    unsigned int status = 0;
    hc_get_status(ctx, &status);
    hc_destroy(ctx);
    if (status) throw std::runtime_error("device raised status bits " + std::to_string(status));
    return Real(t);
}

/*
 * Steady-state integrate with SteadyTimeControls (controls.h:31-48, for which the reference has no
 * integrate at HEAD): marches until the mass residual has dropped by timeControls.residual_drop
 * (hc_step_steady, checked every 50 steps) or nTimesteps are taken, writes the state back into q0
 * and returns the drop reached.  Residual smoothing is not built.
 */
template <typename Policy, LawType Law, int nDim, floating_point Real, auto Basis, typename SecondOrderFlux,
          typename... BoundaryConds>
Real integrate(const Policy policy, const SteadyTimeControls<Real>& timeControls,
               const ODE::Explicit::RungeKutta<Real>& rungeKutta, const SecondOrderFlux& flux2,
               const std::tuple<BoundaryConds...> boundaryConds, const Species<Law, Real>& species,
               const Mesh<nDim, Real>& mesh, SolutionField<VariableSet<Law, nDim, Basis, Real>, nDim>& q0)
{
    if (timeControls.use_residual_smoothing) throw std::runtime_error("residual smoothing is not built");
    double alpha[6][6], beta[6];
    for (int s = 0; s < 6; ++s) {
        beta[s] = rungeKutta.beta[s];
        for (int k = 0; k < 6; ++k) alpha[s][k] = rungeKutta.alpha[s][k];
    }
    hc_context* ctx = hcamd_detail::setup<Law, nDim, Basis>(hcamd_detail::hip_of(policy), flux2, boundaryConds, species,
                                                            mesh, q0, (int)rungeKutta.nstages, alpha, beta,
                                                            timeControls.cfl);
    long long steps = 0;
    double drop = 1.0;
    hcamd_detail::check(hc_step_steady(ctx, (long long)timeControls.nTimesteps, timeControls.residual_drop, 50, &steps,
                                       &drop), ctx, "hc_step_steady");
    hcamd_detail::check(hc_get_state(ctx, reinterpret_cast<double*>(q0.interior.data())), ctx, "hc_get_state");
    hc_destroy(ctx);
    return Real(drop);
}

/*
 * residualCalc(policy, hoflux, bcs, species, mesh, q, res) (spatial/residualCalc.h:45-69);
 * the LSQ metric arrays of the reference signature are computed on the device.
 */
template <typename Policy, LawType Law, int nDim, floating_point Real, auto Basis, typename HighOrderFlux,
          typename... BoundaryConds>
void residualCalc(const Policy policy, const HighOrderFlux& hoflux, const std::tuple<BoundaryConds...> bcs,
                  const Species<Law, Real>& species, const Mesh<nDim, Real>& mesh,
                  const SolutionField<VariableSet<Law, nDim, Basis, Real>, nDim>& q,
                  std::vector<FluxResult<Law, nDim, Real>>& res)
{
    const auto rk = ODE::Explicit::ssp11<Real>();
    double alpha[6][6], beta[6];
    for (int s = 0; s < 6; ++s) {
        beta[s] = rk.beta[s];
        for (int k = 0; k < 6; ++k) alpha[s][k] = rk.alpha[s][k];
    }
    hc_context* ctx = hcamd_detail::setup<Law, nDim, Basis>(hcamd_detail::hip_of(policy), hoflux, bcs, species, mesh, q,
                                                            1, alpha, beta, 1.0);
    constexpr int nv = nVar<Law, nDim>;
    std::vector<double> r((nv + 1) * q.ni * q.nj);
    hcamd_detail::check(hc_residual(ctx, r.data()), ctx, "hc_residual");
    hc_destroy(ctx);
    res.resize(q.ni * q.nj);
    for (std::size_t c = 0; c < res.size(); ++c) {
        for (int v = 0; v < nv; ++v) res[c].flux[v] = r[c * (nv + 1) + v];
        res[c].lambda = r[c * (nv + 1) + nv];
    }
}
