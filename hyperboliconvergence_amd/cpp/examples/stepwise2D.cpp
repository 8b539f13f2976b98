// stepwise2D.cpp -- the body of the reference's integrate() (timestepping/rungeKutta.h:44-167) written
// by a caller on the drop-in's spatial entry points: per stage boundaryUpdate, qmetrics, residualCalc
// (on the GPU), the stage-0 spectralRadius, rungeKuttaAccumulation and eulerForwardUpdateGlobal, then
// the swap and copy of the solution fields.  The same case is then run with integrate() (the fused
// GPU time loop); both final states are printed as checksums and their largest difference.
//   Gresho-like vortex on a periodic square: Euler, primitive basis, SLAU or AUSM+up with van
//   Albada, SSP-RK3, global time step.
//   usage: stepwise2D [n nsteps [flux]]   flux: ausm (default) or slau
#include <hcamd/hcamd.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

template <typename Flux> int run(std::size_t n, std::size_t nt)
{
    constexpr LawType Law = LawType::Euler;
    constexpr int nDim = 2;
    using BasisT = BasisType<Law>;
    using BCType = BoundaryType<Law>;
    using Real = double;
    using SolVarSet = VariableSet<Law, nDim, BasisT::Primitive, Real>;
    using SolField = SolutionField<SolVarSet, nDim>;
    using MeshT = Mesh<nDim, Real>;
    using FluxRes = FluxResult<Law, nDim, Real>;
    using ResidualArray = par::DualArray<FluxRes, nDim>;
    using QMetArray = par::DualArray<lsq::QMetric<SolVarSet>, nDim>;

    const auto rk = ODE::Explicit::ssp33<Real>();
    const UnsteadyTimeControls<Real> timeControls{.nTimesteps = nt, .cfl = 0.5};
    Species<Law, Real> species = get_air_species<Real>();
    species.minf = 0.3;
    const auto policy = par::execution::omp;

    const MeshT mesh = make_linspace_mesh<Real>(par::DualShape2{{n, n}}, -1., 1., -1., 1.);
    SolField q0(mesh.cells.shape());
    // a smooth vortex at Mach 0.3 in a uniform stream
    const auto init = [&](const MeshT::Cell& c) -> SolVarSet {
        const Real x = c.centre[0], y = c.centre[1];
        const Real g = std::exp(-4. * (x * x + y * y));
        return SolVarSet{0.2 - 0.3 * y * g, 0.1 + 0.3 * x * g, 1.0 - 0.05 * g, 1. / 1.4 - 0.05 * g};
    };
    par::transform(init, q0.interior, mesh.cells);
    const SolVarSet qref{0.2, 0.1, 1.0, 1. / 1.4};
    for (auto& qb : q0.boundary) par::fill(qb, qref);
    for (auto& bc : q0.bcTypes) bc = BCType::Periodic;
    const std::tuple bcs{make_periodic_BCond<Law>()};
    const auto flux2 = make_muscl_flux<Law>(Limiters::VanAlbada2{}, Flux{});

    SolField qfused = copy(q0);

    // ---- the reference's integrate() body on the spatial entry points ----
    SolField q1 = copy(q0);
    SolField q2 = copy(q0);
    ResidualArray resTotal(q0.interior.shape());
    std::vector<ResidualArray> resStage = par::vec_of_Arrays<FluxRes, nDim>(rk.nstages, q0.interior.shape());
    const auto dxdx = xmetrics(policy, mesh.cells);
    QMetArray dqdx(mesh.cells.shape());
    Real t = 0;
    for (std::size_t tstep = 0; tstep < timeControls.nTimesteps; tstep++) {
        Real dt{};
        for (unsigned int stg = 0; stg < rk.nstages; stg++) {
            boundaryUpdate(mesh, bcs, species, q1);
            qmetrics(policy, mesh.cells, q1.interior, dqdx);
            residualCalc(policy, flux2, bcs, species, mesh, q1, dxdx, dqdx, resStage[stg]);
            if (stg == 0) dt = timeControls.cfl / spectralRadius(policy, mesh.cells, resStage[stg]);
            rungeKuttaAccumulation(policy, rk, stg, resStage, resTotal);
            eulerForwardUpdateGlobal(policy, mesh.cells, species, rk.beta[stg] * dt, resTotal, q0.interior, q2.interior);
            std::swap(q1, q2);
        }
        copy(policy, q0, q1);
        t += dt;
    }

    // ---- the same case through the fused GPU time loop ----
    const Real tf = integrate(par::execution::hip, timeControls, rk, flux2, bcs, species, mesh, qfused);

    double s0 = 0, s1 = 0, dmax = 0;
    for (std::size_t k = 0; k < q0.interior.flattened_length(); ++k)
        for (int v = 0; v < 4; ++v) {
            const double a = q0.interior.flatten(k)[v], b = qfused.interior.flatten(k)[v];
            s0 += a;
            s1 += b;
            dmax = std::fmax(dmax, std::fabs(a - b));
        }
    std::printf("stepwise2D %zux%zu steps %zu: checksum stepwise %.15e fused %.15e maxdiff %.3e time %.15e %.15e\n", n, n,
                nt, s0, s1, dmax, t, tf);
    return 0;
}

int main(int argc, char** argv)
{
    const std::size_t n = argc > 1 ? std::atoi(argv[1]) : 48;
    const std::size_t nt = argc > 2 ? std::atoi(argv[2]) : 10;
    const std::string flux = argc > 3 ? argv[3] : "ausm";
    if (flux == "slau") return run<Slau<>>(n, nt);
    return run<AusmPlusUP<>>(n, nt);
}
