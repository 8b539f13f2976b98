"""The stage kernel's own code (hc_stage.hpp phases, hc_geom.cpp geometry), executed work-group by
work-group on the host by the test-only emulation, against the oracle on identical inputs.

The FP64 host execution is bit-identical to the oracle; this pins the kernel's arithmetic
order, LDS tile/halo indexing, periodic wrap and boundary handling without a GPU.  The GPU
tests (test_gpu_parity.py) then check the device execution.
"""
import numpy as np
import pytest

import pyemu
import pyoracle as po
from conftest import rel_err

G = 1.4


@pytest.fixture(autouse=True, params=[0, 1, 2], ids=["tiles0", "tiles1", "tiles2"])
def tile_config(request):
    """Every test runs with each tile shape of the kernels (the results must not depend on it)."""
    pyemu.set_tile_config(request.param)
    yield request.param
    pyemu.set_tile_config(0)


def _ogrid(nr, nth, basis, flux, limiter, bc=(po.BC_WALL, po.BC_RIEMANN, 0, 0), perturb=0.02, rk=44, dt=po.DT_LOCAL):
    from hyperboliconvergence_amd import cases
    nodes = po.mesh_ogrid(nr, nth, 0.5, 20.0, 1e-3)
    qinf = np.array([0.3, 0.0, 1.0, 1.0 / G])
    rng = np.random.default_rng(11)
    q = qinf * (1 + perturb * rng.standard_normal((nr, nth, 4)))
    bref = np.broadcast_to(qinf, (po.bref_size(nr, nth), 4)).copy()
    if basis == po.BASIS_CONSERVED:
        q, bref = cases.prim_to_cons(q), cases.prim_to_cons(bref)
    cs = po.make_case(law=po.LAW_EULER, basis=basis, flux=flux, limiter=limiter, bc=bc, rk=rk, cfl=1.0,
                      dt_mode=dt, minf=0.3)
    return cs, nodes, q, bref


@pytest.mark.parametrize("flux", range(7))
@pytest.mark.parametrize("basis", [0, 1])
def test_euler_residual_and_steps_bit_identical(flux, basis):
    cs, nodes, q, bref = _ogrid(13, 37, basis, flux, po.LIM_MINMOD2)   # ragged: not multiples of the tile
    r, st = pyemu.residual(cs, nodes, q, bref)
    assert np.array_equal(r, po.residual(cs, nodes, q, bref), equal_nan=True)
    qe, _, st = pyemu.integrate(cs, nodes, q, bref, 5)
    qo, _, _ = po.integrate(cs, nodes, q, 5, bref)
    assert np.array_equal(qe, qo, equal_nan=True)
    assert st == 0


@pytest.mark.parametrize("limiter", range(10))
def test_every_limiter_bit_identical(limiter):
    cs, nodes, q, bref = _ogrid(9, 20, 1, po.FLUX_ROE, limiter)
    qe, _, _ = pyemu.integrate(cs, nodes, q, bref, 3)
    qo, _, _ = po.integrate(cs, nodes, q, 3, bref)
    assert np.array_equal(qe, qo, equal_nan=True)


@pytest.mark.parametrize("bc", [(po.BC_FIXED, po.BC_ENTROPY, 0, 0), (po.BC_RIEMANN, po.BC_WALL, 0, 0),
                                (po.BC_WALL, po.BC_WALL, po.BC_RIEMANN, po.BC_FIXED)])
def test_boundary_kinds_bit_identical(bc):
    from hyperboliconvergence_amd import cases
    ni, nj = 11, 14
    nodes = po.mesh_linspace(ni, nj, 0.0, 1.0, 0.0, 1.3)
    qinf = np.array([0.3, 0.1, 1.0, 1.0 / G])
    rng = np.random.default_rng(12)
    q = qinf * (1 + 0.02 * rng.standard_normal((ni, nj, 4)))
    bref = np.broadcast_to(qinf, (po.bref_size(ni, nj), 4)).copy()
    cs = po.make_case(law=po.LAW_EULER, flux=po.FLUX_ROE, limiter=po.LIM_VANALBADA2, bc=bc, rk=33, cfl=0.5,
                      period_j=(0.0, 1.3), minf=0.3)
    r, _ = pyemu.residual(cs, nodes, q, bref)
    assert np.array_equal(r, po.residual(cs, nodes, q, bref), equal_nan=True)
    qe, te, _ = pyemu.integrate(cs, nodes, q, bref, 4)
    qo, to, _ = po.integrate(cs, nodes, q, 4, bref)
    assert np.array_equal(qe, qo) and te == to


@pytest.mark.parametrize("rk", [11, 22, 33, 34, 44])
@pytest.mark.parametrize("dt", [po.DT_GLOBAL, po.DT_LOCAL, po.DT_FIXED])
def test_rk_tables_bit_identical(rk, dt):
    cs, nodes, q, bref = _ogrid(8, 16, 1, po.FLUX_AUSMPUP, po.LIM_VANALBADA2, rk=rk, dt=dt)
    if dt == po.DT_FIXED:
        cs.cfl = 1e-4   # the time step itself
    qe, te, _ = pyemu.integrate(cs, nodes, q, bref, 3)
    qo, to, _ = po.integrate(cs, nodes, q, 3, bref)
    assert np.array_equal(qe, qo) and te == to


def test_scalar_c1_trajectory_bit_identical():
    from hyperboliconvergence_amd import cases
    cfg = cases.config_scalar_advection(40)
    qe, te, _ = pyemu.integrate(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"], 30)
    qo, to, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 30, cfg["bref"])
    assert np.array_equal(qe, qo) and te == to


def test_vortex_periodic_small_grid_smaller_than_tile():
    """nj smaller than the 32-wide tile and periodic in both directions (modular halo wrap)."""
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    cfg = cases.config_isentropic_vortex(12, hca.FLUX_SLAU)
    qe, _, _ = pyemu.integrate(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"], 4)
    qo, _, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 4, cfg["bref"])
    assert np.array_equal(qe, qo, equal_nan=True)


def test_mixed_precision_emulation_close_to_fp64():
    cs, nodes, q, bref = _ogrid(16, 32, 1, po.FLUX_ROE, po.LIM_MINMOD2)
    qm, _, _ = pyemu.integrate(cs, nodes, q, bref, 20, precision=1)
    qo, _, _ = po.integrate(cs, nodes, q, 20, bref)
    assert rel_err(qm, qo) < 1e-6


def test_product_geometry_matches_oracle():
    cs, nodes, q, bref = _ogrid(10, 24, 1, po.FLUX_ROE, po.LIM_MINMOD2)
    cells, X, fi, fj = pyemu.geometry(cs, nodes)
    assert np.array_equal(cells, po.dual(nodes), equal_nan=True)
    assert np.array_equal(X, po.xmetrics(cs, po.dual(nodes)), equal_nan=True)


@pytest.mark.parametrize("flux", [po.FLUX_SLAU, po.FLUX_AUSMPUP])
def test_adaptive_low_mach_scaling_bit_identical(flux):
    """LowMachScaling::Adaptive with a finite species dt (mu = lref / (a dt) ~ 0.4)."""
    cs, nodes, q, bref = _ogrid(9, 20, 1, flux, po.LIM_VANALBADA2)
    cs.vscale = cs.pscale = po.SCALE_ADAPTIVE
    cs.sp.dt = 2.0
    r, _ = pyemu.residual(cs, nodes, q, bref)
    assert np.array_equal(r, po.residual(cs, nodes, q, bref))
    qe, _, _ = pyemu.integrate(cs, nodes, q, bref, 5)
    qo, _, _ = po.integrate(cs, nodes, q, 5, bref)
    assert np.array_equal(qe, qo)


@pytest.mark.parametrize("flux", range(7))
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_uniform_geometry_kernels_bit_identical_to_the_general_ones(flux, precision):
    """On a periodic box whose spacing is exact in binary (h = 20/40) the host finds the geometry
    uniform; the uniform-geometry kernels (one slot set per face kind, one volume) then give the
    general kernels' bits, and the oracle's in FP64, over 4 SSP-RK3 steps with a global time step."""
    from hyperboliconvergence_amd import cases
    cfg = cases.config_isentropic_vortex(40, flux)
    out = {}
    for uni in (1, 0):
        pyemu.set_uniform_geometry(uni)
        try:
            out[uni] = pyemu.integrate(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"], 4, precision=precision)
            allowed = pyemu.set_uniform_geometry(0)
        finally:
            pyemu.set_uniform_geometry(0)
        assert allowed
    assert np.array_equal(out[1][0], out[0][0], equal_nan=True) and out[1][1] == out[0][1]
    if precision == 0:
        qo, to, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 4, cfg["bref"])
        assert np.array_equal(out[1][0], qo, equal_nan=True) and out[1][1] == to


def test_uniform_geometry_is_found_only_where_every_slot_is_equal():
    """A box whose spacing is not exact in binary (h = 20/24: volumes differ in the last bit), a
    periodic strip with a wall, and the cylinder O-grid are not uniform; the kernels then stay the
    general ones."""
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    pyemu.set_uniform_geometry(1)
    try:
        cfg = cases.config_isentropic_vortex(24, hca.FLUX_AUSMPUP)
        pyemu.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
        assert not pyemu.set_uniform_geometry(1)
        cfg = cases.config_isentropic_vortex(32, hca.FLUX_AUSMPUP)
        pyemu.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
        assert pyemu.set_uniform_geometry(1)
        cfg["case"].bc[2] = po.BC_WALL
        cfg["case"].bc[3] = po.BC_WALL
        pyemu.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
        assert not pyemu.set_uniform_geometry(1)
        cs, nodes, q, bref = _ogrid(8, 16, 1, po.FLUX_ROE, po.LIM_MINMOD2)
        pyemu.residual(cs, nodes, q, bref)
        assert not pyemu.set_uniform_geometry(1)
    finally:
        pyemu.set_uniform_geometry(0)
