"""The C++20 drop-in layer (hyperboliconvergence_amd/cpp/hcamd/hcamd.hpp): its point functors
evaluated on the host against the oracle, and the example driver built against libhcamd.so."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "hyperboliconvergence_amd", "cpp")


def _build_points(sanitized=False):
    # the sanitized build checks the functors' host code with AddressSanitizer + UBSan
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_points" + ("_asan" if sanitized else ""))
    flags = (["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"] if sanitized else ["-O2"])
    subprocess.run(["g++", "-std=c++20", *flags, "-ffp-contract=off", "-I", CPP, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dropin_points.cpp"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("sanitized", [False, True], ids=["plain", "asan_ubsan"])
def test_dropin_point_functors_bit_identical_to_oracle(sanitized):
    exe = _build_points(sanitized)
    rng = np.random.default_rng(21)
    lines, cases = [], []
    for k in range(70):
        fid = k % 7
        lid = (4, 3, 7)[k % 3]
        n = rng.standard_normal(2)
        n /= np.linalg.norm(n)
        area = rng.uniform(0.1, 2.0)
        ql = np.array([*rng.uniform(-0.5, 0.5, 2), rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.0)])
        qr = np.array([*rng.uniform(-0.5, 0.5, 2), rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.0)])
        lines.append(" ".join(repr(float(x)) for x in [fid, lid, *n, area, *ql, *qr]).replace(".0 ", " ", 0))
        cases.append((fid, lid, n, area, ql, qr))
    text = "\n".join(f"{c[0]} {c[1]} " + " ".join(repr(float(x)) for x in [*c[2], c[3], *c[4], *c[5]]) for c in cases)
    r = subprocess.run([exe], input=text, capture_output=True, text=True, check=True,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    out = r.stdout.split("\n")
    for k, (fid, lid, n, area, ql, qr) in enumerate(cases):
        got = np.array(list(map(float, out[3 * k].split())))
        cs = po.make_case(law=po.LAW_EULER, flux=fid, minf=0.3)
        f = po.face_from(area, (0.0, 0.0), n)
        assert np.array_equal(got, po.flux(cs, f, ql, qr)), (k, fid)
        lim = np.array(list(map(float, out[3 * k + 1].split())))
        assert np.array_equal(lim, [po.limit(lid, a, b) for a, b in zip(ql, qr)])
        st = np.array(list(map(float, out[3 * k + 2].split())))
        assert np.array_equal(st, po.set2state(cs, ql)), k


def test_example_drivers_build():
    """Both drivers, written in the reference's idiom (par::fill / par::transform / par::for_each_idx on
    par::DualArray fields, mesh.cells, SolField::VarField, set2Set, bcTypes of BoundaryType<Law>, the
    spatial entry points), compile against the drop-in header without warnings."""
    r = subprocess.run(["make", "-C", CPP, "-B", "all"], check=True, capture_output=True, text=True)
    assert "warning" not in r.stderr, r.stderr
    for exe in ("cylinder2D", "stepwise2D"):
        assert os.path.exists(os.path.join(CPP, "examples", exe))


def test_reference_driver_idiom_compiles():
    """A translation unit using the reference API surface that the drivers and integrate() use
    (cpp/script/euler/cylinder2D.cpp:159-177, timestepping/rungeKutta.h:44-167), including the
    spatial entry points with their reference argument lists, compiles against the drop-in."""
    src = r'''
#include <hcamd/hcamd.hpp>
constexpr LawType Law = LawType::Euler;
using Real = double;
using BasisT = BasisType<Law>;
using BCType = BoundaryType<Law>;
using SolVarSet = VariableSet<Law, 2, BasisT::Conserved, Real>;
using PrimVarSet = VariableSet<Law, 2, BasisT::Primitive, Real>;
using SolField = SolutionField<SolVarSet, 2>;
using MeshT = Mesh<2, Real>;
using FluxRes = FluxResult<Law, 2, Real>;
int main() {
    const auto species = get_air_species<Real>();
    const MeshT mesh = make_cylinder_mesh<Real>({1.0, 8.0}, {16, 32});
    const par::DualShape<2> cellShape(mesh.cells.shape());
    const SolVarSet qref = set2Set<SolVarSet>(species, PrimVarSet{Real(0.1), Real(0.), Real(1.), Real(1.)});
    SolField q(cellShape);
    par::fill(q.interior, qref);
    for (SolField::VarField& qb : q.boundary) par::fill(qb, qref);
    q.bcTypes = {BCType::InviscidWall, BCType::Riemann, BCType::Periodic, BCType::Periodic};
    const std::tuple bcs{make_ghostCell_BCond<Law, BCType::Riemann>(), make_flux_BCond<Law, BCType::InviscidWall>(),
                         make_periodic_BCond<Law>()};
    const auto hoflux = make_muscl_flux<Law>(Limiters::MinMod2{}, RoeFlux<Law>{});
    const auto policy = par::execution::omp;
    const auto dxdx = xmetrics(policy, mesh.cells);
    par::DualArray<lsq::QMetric<SolVarSet>, 2> dqdx(mesh.cells.shape());
    par::DualArray<FluxRes, 2> res(mesh.cells.shape());
    std::vector<par::DualArray<FluxRes, 2>> resStage = par::vec_of_Arrays<FluxRes, 2>(4, q.interior.shape());
    SolField q1 = copy(q);
    if (false) {
        boundaryUpdate(mesh, bcs, species, q);
        qmetrics(policy, mesh.cells, q.interior, dqdx);
        residualCalc(policy, hoflux, bcs, species, mesh, q, dxdx, dqdx, res);
        const Real dt = 0.5 / spectralRadius(policy, mesh.cells, res);
        rungeKuttaAccumulation(policy, ODE::Explicit::ssp34<Real>(), 0, resStage, res);
        eulerForwardUpdateGlobal(policy, mesh.cells, species, dt, res, q.interior, q1.interior);
        eulerForwardUpdateLocal(policy, species, Real(0.5), res, q.interior, q1.interior);
        const auto grad = gradientCalc(policy, mesh, q);
        (void)grad;
        integrate(policy, SteadyTimeControls<Real>{.nTimesteps = 10, .cfl = 0.5, .residual_drop = 1e-3,
                  .use_residual_smoothing = true, .smoothing_factor = 2.0}, ODE::Explicit::rk4<Real>(), hoflux, bcs,
                  species, mesh, q);
        integrate(UnsteadyTimeControls<Real>{.nTimesteps = 10, .cfl = 0.5}, ODE::Explicit::ssp34<Real>(), hoflux, bcs,
                  species, mesh, q);
        integrate(policy, UnsteadyTimeControls<Real>{.nTimesteps = 10, .cfl = 0.5}, ODE::Explicit::ssp34<Real>(), hoflux,
                  bcs, species, mesh, q);
    }
    auto writer = [&](const par::DualIdx2 idx, const MeshT::Cell& c, const SolVarSet& qc) {
        const auto st = set2State(species, qc);
        (void)idx; (void)c; (void)st;
    };
    par::for_each_idx(writer, mesh.cells, q.interior);
    return mesh.cells.shape(1) == 32 && q.boundary[2].shape(0) == 16 ? 0 : 1;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "idiom.cpp")
        open(c, "w").write(src)
        exe = os.path.join(d, "idiom")
        r = subprocess.run(["g++", "-std=c++20", "-fopenmp", "-Wall", "-I", CPP, "-I", os.path.join(ROOT, "include"), c,
                            "-o", exe, "-L", os.path.join(ROOT, "hyperboliconvergence_amd"), "-lhcamd",
                            "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + os.path.join(ROOT, "hyperboliconvergence_amd"),
                            "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        assert "warning" not in r.stderr, r.stderr[-4000:]
        # the host-only part (containers, set2Set, writers) runs without a GPU
        assert subprocess.run([exe]).returncode == 0


@pytest.mark.gpu
def test_stepwise_driver_equals_fused_integrate_and_oracle():
    """examples/stepwise2D: the reference's integrate() body on boundaryUpdate / qmetrics /
    residualCalc / spectralRadius / rungeKuttaAccumulation / eulerForwardUpdateGlobal gives the fused
    GPU integrate() bit for bit, for AUSM+up and SLAU."""
    exe = os.path.join(CPP, "examples", "stepwise2D")
    assert os.path.exists(exe), "build it with make -C hyperboliconvergence_amd/cpp"
    for flux in ("ausm", "slau"):
        out = subprocess.run([exe, "40", "8", flux], capture_output=True, text=True, check=True, timeout=120).stdout
        print(out)
        words = out.split()
        maxdiff = float(words[words.index("maxdiff") + 1])
        s_step = float(words[words.index("stepwise") + 1])
        s_fused = float(words[words.index("fused") + 1])
        assert maxdiff == 0.0 and s_step == s_fused


@pytest.mark.gpu
def test_example_driver_runs_and_matches_oracle(tmp_path):
    """The example driver's checksum against the oracle, and its result.dat / mesh.dat (written through the
    drop-in set2State) against the Python writers on the oracle's final state."""
    from hyperboliconvergence_amd import cases, results
    exe = os.path.join(CPP, "examples", "cylinder2D")
    assert os.path.exists(exe), "build it with make -C hyperboliconvergence_amd/cpp"
    out = subprocess.run([exe, "64", "32", "50", "0", str(tmp_path)], capture_output=True, text=True, check=True,
                         timeout=120).stdout
    checksum = float(out.split("checksum")[1].split()[0])
    # the C++ O-grid generator uses the same formula as cases.mesh_ogrid
    cfg = cases.config_cylinder_ogrid(nth=64, nr=32)
    qo, _, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 50, cfg["bref"])
    assert abs(checksum - qo.sum()) <= 1e-9 * abs(qo.sum())
    results.write_cylinder_result(tmp_path / "py_result.dat", cfg["nodes"], qo, cfg["q"][0, 0])
    results.write_mesh(tmp_path / "py_mesh.dat", cfg["nodes"])
    for name in ("result", "mesh"):
        cpp_text = (tmp_path / f"{name}.dat").read_text()
        py_text = (tmp_path / f"py_{name}.dat").read_text()
        assert cpp_text.count("\n\n") == py_text.count("\n\n") == (32 if name == "result" else 33)
        a, b = np.loadtxt(tmp_path / f"{name}.dat"), np.loadtxt(tmp_path / f"py_{name}.dat")
        assert a.shape == b.shape
        # 8 significant digits, GPU vs oracle within 1e-10, relative columns of small numbers
        assert np.allclose(a, b, rtol=1e-7, atol=1e-9), name
    # steady marching through the drop-in's SteadyTimeControls integrate: the example's reference Roe
    # takes the mass residual below 5e-2 of its start within 200 steps (the oracle: 4.7e-2 at
    # step 200), well before the frame-mixing divergence near step 400 on this grid
    out = subprocess.run([exe, "64", "32", "2000", "0", "-", "5e-2"], capture_output=True, text=True, check=True,
                         timeout=120).stdout
    drop = float(out.split("residual drop")[1].split()[0])
    assert 0.0 < drop <= 5e-2
