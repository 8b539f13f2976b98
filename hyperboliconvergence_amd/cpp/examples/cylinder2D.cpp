// cylinder2D.cpp -- the reference's cylinder driver pattern (cpp/script/euler/cylinder2D.cpp:73-190)
// on the drop-in API: Euler on an O-grid around a unit-diameter cylinder, slip wall, Riemann far
// field, periodic theta, Roe flux with MinMod2 MUSCL, classical RK4 with local time stepping.
//   usage: cylinder2D [nth nr nsteps [mixed]]   -- prints the state checksum and timing
#include <hcamd/hcamd.hpp>

#include <chrono>
#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv)
{
    constexpr LawType Law = LawType::Euler;
    constexpr int nDim = 2;
    using BasisT = BasisType<Law>;
    using BCType = BoundaryType<Law>;
    using Real = double;

    const std::size_t nth = argc > 1 ? std::atoi(argv[1]) : 128;
    const std::size_t nr = argc > 2 ? std::atoi(argv[2]) : 64;
    const std::size_t nt = argc > 3 ? std::atoi(argv[3]) : 100;
    const bool mixed = argc > 4 && std::atoi(argv[4]) != 0;

    constexpr Real mach = 0.3, pressure = 1. / 1.4, density = 1.;
    constexpr BasisT SolutionBasis = BasisT::Primitive;
    using Flux = RoeFlux<Law>;
    using Limiter = Limiters::MinMod2;
    using SolVarSet = VariableSet<Law, nDim, SolutionBasis, Real>;
    using SolField = SolutionField<SolVarSet, nDim>;

    const auto rk = ODE::Explicit::rk4<Real>();
    const UnsteadyTimeControls<Real> timeControls{.nTimesteps = nt, .cfl = 1.0};
    Species<Law, Real> species = get_air_species<Real>();
    species.minf = mach;

    const auto mesh = make_ogrid_mesh<Real>({0.5, 20.0}, {nr, nth}, 1e-3);
    const SolVarSet qref{{mach, 0., density, pressure}};
    SolField q({nr, nth});
    for (auto& c : q.interior) c = qref;
    for (auto& b : q.boundary)
        for (auto& v : b) v = qref;
    q.bcTypes = {int(BCType::InviscidWall), int(BCType::Riemann), int(BCType::Periodic), int(BCType::Periodic)};

    const std::tuple boundaryConditions{make_ghostCell_BCond<Law, BCType::Riemann>(),
                                        make_flux_BCond<Law, BCType::InviscidWall>(), make_periodic_BCond<Law>()};
    const auto hoflux = make_muscl_flux<Law>(Limiter{}, Flux{});

    par::execution::hip_policy policy;
    policy.dt_mode = HC_DT_LOCAL;
    policy.precision = mixed ? HC_PRECISION_MIXED : HC_PRECISION_FP64;

    const auto t0 = std::chrono::steady_clock::now();
    integrate(policy, timeControls, rk, hoflux, boundaryConditions, species, mesh, q);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    double sum = 0.0, pmin = 1e300;
    for (const auto& c : q.interior) {
        for (int v = 0; v < 4; ++v) sum += c[v];
        pmin = c[3] < pmin ? c[3] : pmin;
    }
    std::printf("cylinder2D %zux%zu steps %zu: checksum %.15e pmin %.6f time %.3f s\n", nth, nr, nt, sum, pmin, s);
    return 0;
}
