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


def test_example_driver_builds():
    subprocess.run(["make", "-C", CPP, "-B", "examples/cylinder2D"], check=True, capture_output=True)
    assert os.path.exists(os.path.join(CPP, "examples", "cylinder2D"))


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
