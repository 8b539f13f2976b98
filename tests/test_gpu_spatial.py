"""GPU: the spatial building blocks of a single-case context (hc_xmetrics, hc_qmetrics,
hc_boundary_update, hc_spectral_radius, hc_forward_update) against the oracle, and the
reference's integrate() body (timestepping/rungeKutta.h:97-164) written with them -- boundary
update, Q metric, residual, stage-0 spectral radius, Runge-Kutta accumulation in the reference's
order, forward-Euler update -- against the fused hc_step, bit for bit, for a global and a local
time step.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _cyl(basis=1, flux=None, dt_mode=None, bc1=None):
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd import cases
    cfg = cases.config_cylinder(64, 32, 0.5, 15.0, flux=hca.FLUX_ROE_FACEFRAME if flux is None else flux)
    rng = np.random.default_rng(5)
    cfg["q"] = cfg["q"] * (1.0 + 0.02 * rng.standard_normal(cfg["q"].shape))
    if bc1 is not None:
        cfg["case"].bc[1] = bc1
    if dt_mode is not None:
        cfg["case"].dt_mode = dt_mode
        cfg["case"].cfl = 0.8 if dt_mode == 0 else 1.0
    if basis == 0:
        cfg["case"].basis = 0
        cfg["q"], cfg["bref"] = cases.prim_to_cons(cfg["q"]), cases.prim_to_cons(cfg["bref"])
    return cfg


def _ctx(cfg):
    import hyperboliconvergence_amd as hca
    ctx = hca.Context(0)
    ctx.set_mesh(cfg["nodes"])
    ctx.set_case(cfg["case"])
    ctx.set_boundary_refs(cfg["bref"])
    ctx.set_state(cfg["q"])
    return ctx


def _rows(nodes):
    """(face, cell3) of every boundary row in the order of the reference arrays: boundary 0 and 1
    along j, then 2 and 3 along i; the face normal points into the domain."""
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    cells = po.dual(nodes)
    out = []
    for b in range(4):
        for t in range(nj if b < 2 else ni):
            if b == 0:
                p0, p1, ij = nodes[0, t], nodes[0, t + 1], (0, t)
            elif b == 1:
                p0, p1, ij = nodes[ni, t + 1], nodes[ni, t], (ni - 1, t)
            elif b == 2:
                p0, p1, ij = nodes[t + 1, 0], nodes[t, 0], (t, 0)
            else:
                p0, p1, ij = nodes[t, nj], nodes[t + 1, nj], (t, nj - 1)
            f = po.Face()
            po.lib().orc_surface(po._p(np.ascontiguousarray(p0)), po._p(np.ascontiguousarray(p1)), po._ref(f))
            out.append((b, f, cells[ij], ij))
    return out


@pytest.mark.parametrize("basis", [0, 1])
def test_metrics_and_boundary_update_match_the_oracle(basis):
    import hyperboliconvergence_amd as hca
    cfg = _cyl(basis, bc1=hca.BC_ENTROPY if basis == 0 else hca.BC_RIEMANN)
    cells = po.dual(cfg["nodes"])
    with _ctx(cfg) as ctx:
        X = ctx.xmetrics()
        Q = ctx.qmetrics(cfg["q"])
        ghost0 = np.full_like(cfg["bref"], -7.0)
        ghost = ctx.boundary_update(cfg["q"], cfg["bref"], ghost0)
    assert np.array_equal(X, po.xmetrics(cfg["case"], cells))
    assert np.array_equal(Q, po.qmetrics(cfg["case"], cells, cfg["q"]))
    for r, (b, f, c, (i, j)) in enumerate(_rows(cfg["nodes"])):
        kind = cfg["case"].bc[b]
        if kind == hca.BC_PERIODIC:
            assert np.all(ghost[r] == -7.0)
        elif kind == hca.BC_WALL:
            assert np.array_equal(ghost[r], cfg["bref"][r])
        else:
            qint = cfg["q"][i, j]
            exp = po.bc_update(cfg["case"], kind, f, c, qint, cfg["bref"][r])
            assert np.allclose(ghost[r], exp, rtol=1e-14, atol=1e-15), (r, ghost[r], exp)


@pytest.mark.parametrize("dt_mode", [0, 1], ids=["global", "local"])
@pytest.mark.parametrize("basis", [0, 1])
def test_reference_time_loop_from_building_blocks_equals_hc_step(dt_mode, basis):
    """The reference's integrate() body written with the building blocks equals hc_step bit for bit
    (local time step: the build's stage-0 spectral radius in place of resTotal.lambda = 0)."""
    cfg = _cyl(basis, dt_mode=dt_mode)
    cs = cfg["case"]
    ns = cs.nstages
    alpha = np.array([[cs.alpha[s][k] for k in range(6)] for s in range(6)])
    beta = np.array([cs.beta[s] for s in range(6)])
    nsteps = 5
    cells = po.dual(cfg["nodes"])
    with _ctx(cfg) as ctx:
        ctx.set_tiles(0)
        ctx.step(nsteps)
        q_fused = ctx.get_state()
        t_fused = ctx.time()
    with _ctx(cfg) as ctx:
        q0 = cfg["q"].copy()
        t = 0.0
        for _ in range(nsteps):
            q1 = q0.copy()
            res = []
            dt = None
            for s in range(ns):
                ctx.boundary_update(q1, cfg["bref"])     # ghost values (used inline by the residual)
                ctx.qmetrics(q1)
                ctx.set_state(q1)
                res.append(ctx.residual())
                if s == 0 and dt_mode == 0:
                    dt = cs.cfl / ctx.spectral_radius(res[0])
                rtot = np.zeros_like(res[0])
                for k in range(s + 1):
                    rtot[..., :4] = rtot[..., :4] + alpha[s, k] * res[k][..., :4]
                if dt_mode == 0:
                    q1 = ctx.forward_update(0, beta[s] * dt, rtot, q0)
                else:
                    rtot[..., 4] = res[0][..., 4]   # stage-0 spectral radius sums
                    q1 = ctx.forward_update(1, beta[s] * cs.cfl, rtot, q0)
            q0 = q1
            if dt_mode == 0:
                t += dt
    assert np.array_equal(q0, q_fused)
    if dt_mode == 0:
        assert t == t_fused
    qo, _, _ = po.integrate(cs, cfg["nodes"], cfg["q"], nsteps, cfg["bref"])
    assert rel_err(q0, qo) <= 1e-10


def test_spectral_radius_matches_the_oracle():
    cfg = _cyl(1, dt_mode=0)
    cells = po.dual(cfg["nodes"])
    with _ctx(cfg) as ctx:
        res = ctx.residual()
        sr = ctx.spectral_radius(res)
        # an all-zero spectral radius gives the reduction's initial value (spectralRadius.h:42)
        z = res.copy()
        z[..., 4] = 0.0
        sr0 = ctx.spectral_radius(z)
    assert sr == po.spectral_radius(cells, res)
    assert sr0 == np.finfo(np.float64).tiny


@pytest.mark.parametrize("periodic", [False, True])
def test_legacy_gradient_matches_the_oracle(periodic):
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    cfg = cases.config_isentropic_vortex(40, hca.FLUX_AUSMPUP) if periodic else _cyl(1)
    if not periodic:
        cfg["case"].bc[2] = cfg["case"].bc[3] = hca.BC_FIXED   # no periodic direction
    with _ctx(cfg) as ctx:
        g = ctx.gradient(cfg["q"])
    assert np.array_equal(g, po.gradient(cfg["case"], cfg["q"]))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_residual_smoothing_matches_the_oracle(mode):
    """On the cylinder (non-periodic along i, periodic along j) and the periodic vortex."""
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    for cfg in (_cyl(1), cases.config_isentropic_vortex(40, hca.FLUX_AUSMPUP)):
        with _ctx(cfg) as ctx:
            res = ctx.residual()
            out = ctx.smooth_residual(res, mode, 2.0)
        assert np.array_equal(out, po.smooth_residual(cfg["case"], res, mode, 2.0))


@pytest.mark.parametrize("dt_mode", [0, 1], ids=["global", "local"])
def test_smoothed_stepping_matches_the_oracle_and_identity_smoothing_is_the_fused_step(dt_mode):
    cfg = _cyl(1, dt_mode=dt_mode)
    with _ctx(cfg) as ctx:
        ctx.step(6)
        q_fused = ctx.get_state()
    with _ctx(cfg) as ctx:
        ctx.set_residual_smoothing(0, 1.0)   # explicit operator with gamma = 0
        ctx.step(6)
        q_ident = ctx.get_state()
    assert np.array_equal(q_ident, q_fused)
    with _ctx(cfg) as ctx:
        ctx.set_residual_smoothing(1, 2.0)
        ctx.step(6)
        q_sm = ctx.get_state()
        t = ctx.time()
    qo, to = po.integrate_smoothed(cfg["case"], cfg["nodes"], cfg["q"], 6, 1, 2.0, cfg["bref"])
    assert rel_err(q_sm, qo) <= 1e-12
    if dt_mode == 0:
        assert abs(t - to) <= 1e-14 * abs(to)


def test_steady_march_with_implicit_smoothing_converges_faster():
    """hc_step_steady with implicit smoothing (alpha = 2): the cylinder's mass residual drops
    100-fold in at most 250 steps, against more than 500 without."""
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    cfg = cases.config_cylinder(64, 32, 0.4, 0.0, flux=hca.FLUX_ROE_FACEFRAME)
    steps = {}
    for smooth in (False, True):
        with _ctx(cfg) as ctx:
            if smooth:
                ctx.set_residual_smoothing(1, 2.0)
            n, drop = ctx.step_steady(2000, 1e-2, 50)
            steps[smooth] = n
            assert drop <= 1e-2 and ctx.status() == 0
    print("steps to a 100-fold mass-residual drop:", steps)
    assert steps[True] <= 250 < 500 < steps[False]
