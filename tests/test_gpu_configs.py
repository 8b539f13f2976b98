"""GPU: the benchmark configurations at their full sizes, through the C-ABI, against the oracle.

* cylinder O-grid 512 x 256, reference Roe + Harten fix, MinMod2, LSQ, classical RK4, local
  time step at CFL 1, FP64: 300 steps (the reference algorithm's state leaves the finite numbers
  after about 380 steps, frame mixing of base/fluxes/roeFlux.ipp:40-54) within 1e-10 of the
  oracle;
* isentropic vortex 4096^2 with SLAU and with AUSM+up, van Albada, SSP-RK3, global time step,
  FP64: residual within 1e-12 and two steps within 1e-10 of the oracle;
* the 96-case batch with the reference Roe flux as configured: four steps of every case within
  1e-10 of that case's oracle trajectory (FP64), and the mixed path within 1e-6 of FP64;
* the adjoint case at 1024 x 512, AUSM+up, smooth van Albada, SSP-RK3, K = 400 steps with a
  checkpoint every 20 (sqrt K): the gradient of the wall force with respect to the initial
  state along a seeded direction against the FP64 tangent of the same 400 steps and against
  central differences of GPU FP64 trajectories; the gradient with respect to the free-stream
  Mach number (initial state, far-field references and the species' minf) against the tangent.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _threads():
    import os
    po.set_threads(min(16, os.cpu_count() or 1))


def _ctx(cfg, precision=0, tiles=-1, graph_steps=50):
    import hyperboliconvergence_amd as hca
    ctx = hca.Context(0)
    ctx.set_mesh(cfg["nodes"])
    ctx.set_case(cfg["case"])
    ctx.set_precision(precision)
    ctx.set_tiles(tiles)
    ctx.set_graph_steps(graph_steps)
    ctx.set_boundary_refs(cfg["bref"])
    ctx.set_state(cfg["q"])
    return ctx


def test_cylinder_512x256_reference_roe_300_steps_match_the_oracle():
    from hyperboliconvergence_amd import cases
    _threads()
    cfg = cases.config_cylinder_ogrid()
    with _ctx(cfg) as ctx:
        r = ctx.residual()
        ctx.step(300)
        q = ctx.get_state()
        st = ctx.status()
    ro = po.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
    qo, _, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 300, cfg["bref"])
    e_r, e_q = rel_err(r, ro), rel_err(q, qo)
    print(f"cylinder 512x256 Roe: residual {e_r:.2e}, 300 steps {e_q:.2e}")
    assert st == 0 and np.isfinite(q).all()
    assert e_r <= 1e-12 and e_q <= 1e-10


@pytest.mark.parametrize("flux", ["slau", "ausmpup"])
def test_vortex_4096_residual_and_two_steps_match_the_oracle(flux):
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd import cases
    _threads()
    cfg = cases.config_isentropic_vortex(4096, hca.FLUX_SLAU if flux == "slau" else hca.FLUX_AUSMPUP)
    with _ctx(cfg, graph_steps=0) as ctx:
        r = ctx.residual()
        ctx.step(2)
        q = ctx.get_state()
        st = ctx.status()
    ro = po.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
    e_r = rel_err(r, ro)
    del ro
    qo, _, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 2, cfg["bref"])
    e_q = rel_err(q, qo)
    print(f"vortex 4096^2 {flux}: residual {e_r:.2e}, 2 steps {e_q:.2e}")
    assert st == 0
    assert e_r <= 1e-12 and e_q <= 1e-10


def test_batch_of_96_with_the_reference_roe_matches_the_oracle_per_case():
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd import cases
    _threads()
    cfgs = [cases.config_cylinder(r["nth"], r["nr"], r["mach"], r["alpha_deg"], flux=hca.FLUX_ROE)
            for r in cases.batch_case_table()]
    out = {}
    for prec in (0, 1):
        b = hca.Batch(0)
        b.begin()
        for c in cfgs:
            b.add_case(c["nodes"], c["case"])
        b.end()
        b.set_precision(prec)
        for k, c in enumerate(cfgs):
            b.set_case_boundary_refs(k, c["bref"])
            b.set_case_state(k, c["q"])
        b.step(4)
        out[prec] = [b.get_case_state(k) for k in range(len(cfgs))]
        assert b.status() == 0
        b.close()
    worst, worst_mixed = 0.0, 0.0
    for k, c in enumerate(cfgs):
        qo, _, _ = po.integrate(c["case"], c["nodes"], c["q"], 4, c["bref"])
        worst = max(worst, rel_err(out[0][k], qo))
        worst_mixed = max(worst_mixed, rel_err(cases.prim_to_cons(out[1][k]), cases.prim_to_cons(out[0][k])))
    print(f"batch of 96, reference Roe, 4 steps: worst case vs oracle {worst:.2e}, mixed vs FP64 {worst_mixed:.2e}")
    assert worst <= 1e-10
    assert worst_mixed <= 1e-6


def test_adjoint_case_full_size_gradients():
    from hyperboliconvergence_amd import cases
    cfg = cases.config_adjoint()
    K, C = cfg["nsteps"], cfg["checkpoint_every"]
    assert (K, C) == (400, 20) and cfg["q"].shape == (512, 1024, 4)
    nodes, q0, bref, M = cfg["nodes"], cfg["q"], cfg["bref"], cfg["mach"]
    nb = bref.shape[0]
    with _ctx(cfg, tiles=0, graph_steps=0) as ctx:
        lam = ctx.residual()[..., 4]
        dt = 0.8 / float(np.max(lam / cases.cell_volumes(nodes)))
        cfg["case"].cfl = dt   # fixed step: CFL 0.8 of the initial state
        ctx.set_case(cfg["case"])
        ctx.set_tiles(0)

        def run(qs):
            ctx.set_boundary_refs(bref)
            ctx.set_state(qs)
            ctx.adjoint_forward(K, C)
            return ctx.get_state()

        qK = run(q0)
        g = cases.wall_force_x_grad(nodes, qK)
        q0b, bb, mb = ctx.adjoint_backward(g)
        # tangent of the same 400 steps along d (initial state) and along M (initial state, far field, minf)
        d = cases.smooth_perturbation(512, 1024, seed=7, amp=1e-3)
        ctx.set_boundary_refs(bref)
        ctx.set_state(q0)
        tan_d = float(np.sum(g * ctx.tangent_steps(K, d)))
        h = 1e-7
        dfs = (cases.freestream_cons(M + h) - cases.freestream_cons(M - h)) / (2 * h)
        ctx.set_boundary_refs(bref)
        ctx.set_state(q0)
        tan_M = float(np.sum(g * ctx.tangent_steps(K, np.broadcast_to(dfs, q0.shape).copy(),
                                                   np.broadcast_to(dfs, (nb, 4)).copy(), 1.0)))
        adj_d = float(np.sum(q0b * d))
        adj_M = float(np.sum(q0b * dfs[None, None, :]) + np.sum(bb * dfs[None, :]) + mb)
        fd_d = (cases.wall_force_x(nodes, run(q0 + 1e-6 * d)) - cases.wall_force_x(nodes, run(q0 - 1e-6 * d))) / 2e-6
        assert ctx.status() == 0
    print(f"adjoint 1024x512, K=400, C=20: dJ/dq0.d adjoint {adj_d:.12e} tangent {tan_d:.12e} fd {fd_d:.12e}; "
          f"dJ/dM adjoint {adj_M:.12e} tangent {tan_M:.12e}")
    assert abs(adj_d - tan_d) <= 1e-10 * abs(tan_d)
    assert abs(adj_d - fd_d) <= 1e-6 * abs(fd_d)
    assert abs(adj_M - tan_M) <= 1e-8 * abs(tan_M)
