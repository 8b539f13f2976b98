"""GPU: the uniform-geometry stage kernels (hc_set_uniform_geometry) on periodic boxes whose
spacing is exact in binary, against the general kernels bit for bit and against the oracle; every
tile configuration, FP64 and mixed, eager and in graphs; cases whose geometry is not uniform keep
the general kernels."""
import numpy as np
import pytest

import pyoracle as po
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _ctx(cfg, precision=0, tiles=0, graph_steps=0, uniform=1):
    import hyperboliconvergence_amd as hca
    ctx = hca.Context(0)
    ctx.set_mesh(cfg["nodes"])
    ctx.set_case(cfg["case"])
    ctx.set_precision(precision)
    ctx.set_tiles(tiles)
    ctx.set_graph_steps(graph_steps)
    ctx.set_uniform_geometry(uniform)
    ctx.set_boundary_refs(cfg["bref"])
    ctx.set_state(cfg["q"])
    return ctx


@pytest.mark.parametrize("flux", range(7))
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_uniform_kernels_equal_the_general_ones_on_a_256_box(flux, precision):
    from hyperboliconvergence_amd import cases
    cfg = cases.config_isentropic_vortex(256, flux)
    out = {}
    for uni in (1, 0):
        for tiles in range(3):
            with _ctx(cfg, precision, tiles, uniform=uni) as ctx:
                assert ctx.uniform_geometry() == bool(uni)
                r = ctx.residual()
                ctx.step(6)
                out[uni, tiles] = (r, ctx.get_state(), ctx.time())
    for key, (r, q, t) in out.items():
        assert np.array_equal(r, out[0, 0][0], equal_nan=True), key
        assert np.array_equal(q, out[0, 0][1], equal_nan=True) and t == out[0, 0][2], key
    if precision == 0:
        ro = po.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
        qo, to, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 6, cfg["bref"])
        assert rel_err(out[1, 0][0], ro) <= 1e-12
        if np.isfinite(qo).all():
            assert rel_err(out[1, 0][1], qo) <= 1e-10


def test_uniform_kernels_in_graphs_on_the_4096_box():
    """The 4096^2 AUSM+up vortex: graph-launched uniform kernels (autotuned) equal eager general
    ones over 4 steps, FP64 and mixed."""
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd import cases
    cfg = cases.config_isentropic_vortex(4096, hca.FLUX_AUSMPUP)
    for precision in (0, 1):
        with _ctx(cfg, precision, tiles=-1, graph_steps=2, uniform=1) as ctx:
            assert ctx.uniform_geometry()
            ctx.step(4)
            qu, tu = ctx.get_state(), ctx.time()
        with _ctx(cfg, precision, tiles=1, graph_steps=0, uniform=0) as ctx:
            assert not ctx.uniform_geometry()
            ctx.step(4)
            qg, tg = ctx.get_state(), ctx.time()
        assert np.array_equal(qu, qg) and tu == tg


def test_non_uniform_cases_keep_the_general_kernels():
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd import cases
    for cfg in (cases.config_cylinder(64, 32, 0.5, 10.0, flux=hca.FLUX_ROE),
                cases.config_isentropic_vortex(24, hca.FLUX_AUSMPUP)):
        with _ctx(cfg) as ctx:
            assert not ctx.uniform_geometry()
    b = hca.Batch(0)
    b.begin()
    for n in (32, 64):
        c = cases.config_isentropic_vortex(n, hca.FLUX_AUSMPUP)
        b.add_case(c["nodes"], c["case"])
    c = cases.config_isentropic_vortex(24, hca.FLUX_AUSMPUP)
    b.add_case(c["nodes"], c["case"])
    b.end()
    assert not b.uniform_geometry()   # one case of the batch is not uniform
    b.close()
