// cylinder2D.cpp -- the reference's cylinder driver (cpp/script/euler/cylinder2D.cpp:73-272) on the
// drop-in API, written in its idiom: par::DualShape from mesh.cells, par::fill of the interior and
// boundary VarFields with a set2Set free stream, bcTypes of BoundaryType<Law>, make_muscl_flux,
// integrate, and par::for_each_idx writers of result.dat / mesh.dat.  Euler on an O-grid around a
// unit-diameter cylinder, slip wall, Riemann far field, periodic theta, Roe flux with MinMod2 MUSCL,
// classical RK4 with local time stepping on the GPU.
//   usage: cylinder2D [nth nr nsteps [mixed [outdir [drop [smoothing]]]]]   -- prints the state
//   checksum and timing; with outdir (or "-"), writes outdir/result.dat and outdir/mesh.dat in the
//   columns and blocks of the reference driver, for its gnuplot scripts; with drop, marches to a
//   steady state (SteadyTimeControls, controls.h:31-48) for at most nsteps, with implicit residual
//   smoothing at that factor when smoothing > 1
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
    const Real density = sref.density(), pressure = sref.pressure(), velocity = std::sqrt(sref.velocity2());
    const Real sonic = std::sqrt(sref.speedOfSound2());
    const Real entropy = sonic * sonic / (gamma * std::pow(density, gamma - 1));
    os << (r - density) / density << " " << state.velocity(0) / velocity << " " << state.velocity(1) / velocity << " "
       << (state.pressure() - pressure) / pressure << " " << (s - entropy) / entropy << " " << a / sonic << "\n";
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
    constexpr Real velocity = mach;   // a = 1 for p = 1/gamma, rho = 1

    constexpr BasisT SolutionBasis = BasisT::Primitive;
    constexpr BasisT PrimBasis = BasisT::Primitive;
    using Flux = RoeFlux<Law>;
    using Limiter = Limiters::MinMod2;

    using SolVarSet = VariableSet<Law, nDim, SolutionBasis, Real>;
    using PrimVarSet = VariableSet<Law, nDim, PrimBasis, Real>;
    using SolField = SolutionField<SolVarSet, nDim>;
    using MeshT = Mesh<nDim, Real>;

    const ODE::Explicit::RungeKutta<Real> rk = ODE::Explicit::rk4<Real>();
    const UnsteadyTimeControls<Real> timeControls{.nTimesteps = nt, .cfl = 1.0};

    const Species<Law, Real> species = []() -> Species<Law, Real> {
        Species<Law, Real> gas = get_air_species<Real>();
        gas.minf = mach;
        return gas;
    }();

    // initialise mesh: the stretched O-grid of the cylinder cases, r from 0.5 to 20
    const MeshT mesh = make_ogrid_mesh<Real>({0.5, 20.0}, {nr, nth}, 1e-3);
    const par::DualShape<nDim> cellShape(mesh.cells.shape());

    // initialise solution
    const SolVarSet qref = set2Set<SolVarSet>(species, PrimVarSet{velocity, Real(0.), density, pressure});
    SolField q(cellShape);
    par::fill(q.interior, qref);

    // boundary conditions
    const std::tuple boundaryConditions{make_ghostCell_BCond<Law, BCType::Riemann>(),
                                        make_flux_BCond<Law, BCType::InviscidWall>(), make_periodic_BCond<Law>()};
    for (SolField::VarField& qb : q.boundary) par::fill(qb, qref);
    q.bcTypes[0] = BCType::InviscidWall;
    q.bcTypes[1] = BCType::Riemann;
    q.bcTypes[2] = BCType::Periodic;
    q.bcTypes[3] = BCType::Periodic;

    // high order reconstruction and flux functions
    const auto hoflux = make_muscl_flux<Law>(Limiter{}, Flux{});

    par::execution::hip_policy policy;
    policy.dt_mode = HC_DT_LOCAL;
    policy.precision = mixed ? HC_PRECISION_MIXED : HC_PRECISION_FP64;

    // integrate forward in time
    const auto t0 = std::chrono::steady_clock::now();
    if (argc > 6) {
        const Real factor = argc > 7 ? std::atof(argv[7]) : 1.0;
        const SteadyTimeControls<Real> steady{.nTimesteps = nt, .cfl = 1.0, .residual_drop = std::atof(argv[6]),
                                              .use_residual_smoothing = factor > 1.0, .smoothing_factor = factor};
        const Real drop = integrate(policy, steady, rk, hoflux, boundaryConditions, species, mesh, q);
        std::printf("steady: residual drop %.6e\n", drop);
    } else {
        integrate(policy, timeControls, rk, hoflux, boundaryConditions, species, mesh, q);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    double sum = 0.0, pmin = 1e300;
    par::for_each(
        [&](const SolVarSet& qc) {
            for (int v = 0; v < 4; ++v) sum += qc[v];
            pmin = qc[3] < pmin ? qc[3] : pmin;
        },
        q.interior);
    std::printf("cylinder2D %zux%zu steps %zu: checksum %.15e pmin %.6f time %.3f s\n", nth, nr, nt, sum, pmin, secs);

    // write solution to file: cells in (r, theta) order, theta fastest, a blank line after each
    // ring, coordinates (theta, z = 0, r) of the cell centre
    if (argc > 5 && std::string(argv[5]) != "-") {
        const std::string dir = argv[5];
        std::ofstream solutionFile(dir + "/result.dat");
        if (!solutionFile.is_open()) {
            std::printf("cannot open %s/result.dat for writing\n", dir.c_str());
            return 1;
        }
        const Real dth = (2. * M_PI) / nth;
        solutionFile << std::scientific;
        solutionFile.precision(8);
        const auto sref = set2State(species, qref);
        auto writer = [&](const par::DualIdx2 idx, const MeshT::Cell& c0, const SolVarSet& qc) -> void {
            const Real theta = (idx[1] + 0.5) * dth;
            const Real r = std::sqrt(c0.centre[0] * c0.centre[0] + c0.centre[1] * c0.centre[1]);
            solutionFile << theta << " " << 0. << " " << r << " ";
            writeState(solutionFile, species, set2State(species, qc), sref);
            if (idx[1] == mesh.cells.shape(1) - 1) solutionFile << "\n";
        };
        par::for_each_idx(writer, mesh.cells, q.interior);

        std::ofstream meshFile(dir + "/mesh.dat");
        if (!meshFile.is_open()) {
            std::printf("cannot open %s/mesh.dat for writing\n", dir.c_str());
            return 1;
        }
        meshFile << std::scientific;
        meshFile.precision(8);
        auto meshWriter = [&](const par::PrimalIdx2 idx, const MeshT::Node& p) -> void {
            meshFile << p[0] << " " << p[1] << " " << "\n";
            if (idx[1] == mesh.cells.shape(1)) meshFile << "\n";
        };
        par::for_each_idx(meshWriter, mesh.nodes);
    }
    return 0;
}
