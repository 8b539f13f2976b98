// cylinder2D.cpp -- the reference's cylinder driver pattern (cpp/script/euler/cylinder2D.cpp:73-190)
// on the drop-in API: Euler on an O-grid around a unit-diameter cylinder, slip wall, Riemann far
// field, periodic theta, Roe flux with MinMod2 MUSCL, classical RK4 with local time stepping.
//   usage: cylinder2D [nth nr nsteps [mixed [outdir [drop]]]]   -- prints the state checksum and
//   timing; with outdir (or "-"), writes outdir/result.dat and outdir/mesh.dat in the columns and
//   blocks of the reference driver (cylinder2D.cpp:33-63, 192-272), for its gnuplot scripts; with
//   drop, marches to a steady state (SteadyTimeControls, controls.h:31-48) for at most nsteps
#include <hcamd/hcamd.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

// one line of result.dat after the cell coordinates: density, velocity, pressure and entropy
// relative to the reference state, speed of sound over the reference one (s = a^2 / (gamma rho^(gamma-1)))
template <floating_point Real>
void writeState(std::ofstream& os, const Species<LawType::Euler, Real>& species, const State<LawType::Euler, 2, Real>& state,
                const State<LawType::Euler, 2, Real>& sref)
{
    const Real gamma = species.gamma;
    const Real a = std::sqrt(state.speedOfSound2());
    const Real r = state.density();
    const Real s = a * a / (gamma * std::pow(r, gamma - 1));
    const Real r0 = sref.density(), p0 = sref.pressure(), u0 = std::sqrt(sref.velocity2());
    const Real a0 = std::sqrt(sref.speedOfSound2());
    const Real s0 = a0 * a0 / (gamma * std::pow(r0, gamma - 1));
    os << (r - r0) / r0 << " " << state.velocity(0) / u0 << " " << state.velocity(1) / u0 << " "
       << (state.pressure() - p0) / p0 << " " << (s - s0) / s0 << " " << a / a0 << "\n";
}

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
    if (argc > 6) {
        const SteadyTimeControls<Real> steady{.nTimesteps = nt, .cfl = 1.0, .residual_drop = std::atof(argv[6]),
                                              .use_residual_smoothing = false, .smoothing_factor = 1.0};
        const Real drop = integrate(policy, steady, rk, hoflux, boundaryConditions, species, mesh, q);
        std::printf("steady: residual drop %.6e\n", drop);
    } else {
        integrate(policy, timeControls, rk, hoflux, boundaryConditions, species, mesh, q);
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    double sum = 0.0, pmin = 1e300;
    for (const auto& c : q.interior) {
        for (int v = 0; v < 4; ++v) sum += c[v];
        pmin = c[3] < pmin ? c[3] : pmin;
    }
    std::printf("cylinder2D %zux%zu steps %zu: checksum %.15e pmin %.6f time %.3f s\n", nth, nr, nt, sum, pmin, s);

    if (argc > 5 && std::string(argv[5]) != "-") {
        // cells in (r, theta) order, theta fastest, a blank line after each ring; the reference
        // driver's coordinates are (theta, z = 0, r) of the cell centre (its r is uniform, the
        // stretched O-grid's is the centre's distance from the axis)
        const std::string dir = argv[5];
        std::ofstream res(dir + "/result.dat"), msh(dir + "/mesh.dat");
        if (!res.is_open() || !msh.is_open()) {
            std::printf("cannot open %s/result.dat or mesh.dat for writing\n", dir.c_str());
            return 1;
        }
        res << std::scientific;
        res.precision(8);
        const auto sref = set2State(species, qref);
        const Real dth = (2. * M_PI) / nth;
        for (std::size_t i = 0; i < nr; ++i) {
            for (std::size_t j = 0; j < nth; ++j) {
                Real cx = 0, cy = 0;
                for (int di = 0; di < 2; ++di)
                    for (int dj = 0; dj < 2; ++dj) {
                        cx += 0.25 * mesh.node(i + di, j + dj)[0];
                        cy += 0.25 * mesh.node(i + di, j + dj)[1];
                    }
                res << (j + 0.5) * dth << " " << 0. << " " << std::sqrt(cx * cx + cy * cy) << " ";
                writeState(res, species, set2State(species, q(i, j)), sref);
                if (j == nth - 1) res << "\n";
            }
        }
        msh << std::scientific;
        msh.precision(8);
        for (std::size_t i = 0; i <= nr; ++i)
            for (std::size_t j = 0; j <= nth; ++j) {
                msh << mesh.node(i, j)[0] << " " << mesh.node(i, j)[1] << " " << "\n";
                if (j == nth) msh << "\n";
            }
    }
    return 0;
}
