"""GPU: tile shapes, the tile-shape autotuner and batched contexts, through the C-ABI.

Every tile shape must give the same results bit for bit (the shape changes which lanes
compute a face, never the arithmetic), so the autotuner may switch shapes freely; a batch of
cases stepped by one launch per stage must give, case by case, exactly the results of a
context holding that case alone.  Against the oracle the bars are the parity tolerances:
pow() in the far-field boundary is not correctly rounded on either side, so GPU and host
may differ in the last bit there and the difference grows slowly along a trajectory.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import rel_err

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
TRAJ_TOL = 1e-10


def _cyl(nth, nr, mach, alpha, flux=None, perturb=0.0, seed=3):
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    cfg = cases.config_cylinder(nth, nr, mach, alpha, flux=hca.FLUX_ROE if flux is None else flux)
    if perturb:
        rng = np.random.default_rng(seed)
        cfg["q"] = cfg["q"] * (1.0 + perturb * rng.standard_normal(cfg["q"].shape))
    return cfg


def _single(cfg, precision=0, tiles=None):
    import hyperboliconvergence_amd as hca
    ctx = hca.Context(0)
    ctx.set_mesh(cfg["nodes"])
    ctx.set_case(cfg["case"])
    ctx.set_precision(precision)
    if tiles is not None:
        ctx.set_tiles(tiles)
    ctx.set_boundary_refs(cfg["bref"])
    ctx.set_state(cfg["q"])
    return ctx


def _batch(cfgs, precision=0):
    import hyperboliconvergence_amd as hca
    b = hca.Batch(0)
    b.begin()
    for cfg in cfgs:
        b.add_case(cfg["nodes"], cfg["case"])
    b.end()
    b.set_precision(precision)
    for k, cfg in enumerate(cfgs):
        b.set_case_boundary_refs(k, cfg["bref"])
        b.set_case_state(k, cfg["q"])
    return b


def test_every_tile_shape_gives_identical_results():
    import hyperboliconvergence_amd as hca
    cfg = _cyl(40, 21, 0.3, 2.0, perturb=0.01)   # ragged in both directions for every shape
    out = []
    for tiles in range(hca.NTILECFG):
        with _single(cfg, tiles=tiles) as ctx:
            r = ctx.residual()
            ctx.step(6)
            out.append((r, ctx.get_state()))
            assert ctx.tiles()[0] == tiles
            assert ctx.status() == 0
    for k in range(1, len(out)):
        assert np.array_equal(out[0][0], out[k][0]) and np.array_equal(out[0][1], out[k][1])
    ro = po.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
    qo, _, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 6, cfg["bref"])
    assert rel_err(out[0][0], ro) <= RES_TOL
    assert rel_err(out[0][1], qo) <= TRAJ_TOL


def test_autotuner_times_every_shape_and_keeps_results():
    import hyperboliconvergence_amd as hca
    cfg = _cyl(64, 32, 0.4, -3.0, perturb=0.01)
    with _single(cfg, tiles=-1) as ctx:
        ctx.set_graph_steps(4)
        ctx.step(12)   # 2 trial steps per configuration, then graphs of 4 and eager steps
        q = ctx.get_state()
        chosen, ms = ctx.tiles()
    print("tile shape", chosen, "ms/step", ms)
    assert 0 <= chosen < hca.NTILECFG and np.all(ms > 0)
    assert ms[chosen] == ms.min()
    with _single(cfg, tiles=0) as ctx:
        ctx.step(12)
        q0 = ctx.get_state()
    assert np.array_equal(q, q0)
    qo, _, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 12, cfg["bref"])
    assert rel_err(q, qo) <= TRAJ_TOL


def _mixed_batch_cfgs():
    # different sizes, Mach numbers and angles of attack; perturbed so every face is active
    return [_cyl(64, 32, 0.3, 0.0, perturb=0.01, seed=1), _cyl(40, 21, 0.55, 4.0, perturb=0.01, seed=2),
            _cyl(128, 64, 0.25, -2.5, perturb=0.01, seed=3), _cyl(32, 16, 0.7, 1.0, perturb=0.01, seed=4)]


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_case_contexts(precision):
    cfgs = _mixed_batch_cfgs()
    with _batch(cfgs, precision) as b:
        assert len(b) == 4
        res = [b.case_residual(k) for k in range(4)]
        b.set_graph_steps(3)
        b.step(10)
        qb = [b.get_case_state(k) for k in range(4)]
        assert b.status() == 0
    for k, cfg in enumerate(cfgs):
        with _single(cfg, precision) as ctx:
            r = ctx.residual()
            ctx.step(10)
            q = ctx.get_state()
        assert np.array_equal(res[k], r), k
        assert np.array_equal(qb[k], q), k


@pytest.mark.parametrize("precision", [0, 1])
def test_shared_geometry_and_interleaved_tile_order_keep_every_bit(precision):
    """Cases on the same mesh share one device copy of the geometry (two meshes, one of them also
    ragged for every tile shape, with different Mach numbers and angles); interleaving their tiles
    in a step changes nothing: every tile shape, both orders and single-case contexts agree bit for
    bit, eager and in graphs."""
    import hyperboliconvergence_amd as hca
    cfgs = [_cyl(40, 21, 0.3, 0.0, perturb=0.01, seed=1), _cyl(64, 32, 0.55, 4.0, perturb=0.01, seed=2),
            _cyl(40, 21, 0.45, -3.0, perturb=0.01, seed=3), _cyl(64, 32, 0.25, 2.0, perturb=0.01, seed=4),
            _cyl(40, 21, 0.7, 1.5, perturb=0.01, seed=5)]
    ref = []
    for cfg in cfgs:
        with _single(cfg, precision) as ctx:
            r = ctx.residual()
            ctx.step(7)
            ref.append((r, ctx.get_state()))
    for tiles in range(hca.NTILECFG):
        for order in (1, 0, 1):   # switching the order back and forth rebuilds the table
            with _batch(cfgs, precision) as b:
                assert b.geometry_blocks() == 2
                b.set_tile_order(order)
                b.set_tiles(tiles)
                b.set_graph_steps(3 if order else 0)
                res = [b.case_residual(k) for k in range(len(cfgs))]
                b.step(7)
                assert b.status() == 0
                for k in range(len(cfgs)):
                    assert np.array_equal(res[k], ref[k][0]), (tiles, order, k)
                    assert np.array_equal(b.get_case_state(k), ref[k][1]), (tiles, order, k)


def test_batch_fp64_matches_oracle_per_case():
    cfgs = _mixed_batch_cfgs()[:3]
    with _batch(cfgs) as b:
        b.step(5)
        qb = [b.get_case_state(k) for k in range(3)]
        norms = [b.case_norms(k) for k in range(3)]
    for k, cfg in enumerate(cfgs):
        qo, _, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 5, cfg["bref"])
        assert rel_err(qb[k], qo) <= TRAJ_TOL, k
        ro = po.residual(cfg["case"], cfg["nodes"], qb[k], cfg["bref"])
        l2o, linfo = po.norms(po.dual(cfg["nodes"]), ro)
        assert np.allclose(norms[k][0], l2o, rtol=1e-10, atol=0), k
        assert np.allclose(norms[k][1], linfo, rtol=1e-10, atol=0), k


def test_batch_global_time_step_is_per_case():
    """Global dt (SSP-RK3): each case reduces its own max(lambda/V) and keeps its own time."""
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    cfgs = [cases.config_isentropic_vortex(n, flux=hca.FLUX_AUSMPUP) for n in (24, 40, 33)]
    with _batch(cfgs) as b:
        b.step(4)
        qb = [b.get_case_state(k) for k in range(3)]
        tb = [b.case_time(k) for k in range(3)]
    for k, cfg in enumerate(cfgs):
        qo, to, _ = po.integrate(cfg["case"], cfg["nodes"], cfg["q"], 4, cfg["bref"])
        assert rel_err(qb[k], qo) <= TRAJ_TOL, k
        assert abs(tb[k] - to) <= 1e-12 * to, k
    assert len(set(tb)) == 3
    for k, cfg in enumerate(cfgs):
        with _single(cfg) as ctx:
            ctx.step(4)
            assert np.array_equal(ctx.get_state(), qb[k]) and ctx.time() == tb[k], k


def test_batch_rejects_cases_with_different_schedules():
    import hyperboliconvergence_amd as hca
    a = _cyl(32, 16, 0.3, 0.0)
    c = _cyl(32, 16, 0.3, 0.0, flux=hca.FLUX_RUSANOV)
    b = hca.Batch(0)
    b.begin()
    b.add_case(a["nodes"], a["case"])
    with pytest.raises(hca.HCError):
        b.add_case(c["nodes"], c["case"])
    b.close()


def test_mixed_transonic_case_stays_finite_and_close_to_fp64():
    """Regression: a Mach 0.77 cylinder (supersonic pocket, sonic Roe states inside compressions)
    once produced 0/0 in the FP32 entropy fix of the mixed path after ~160 steps."""
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd import cases
    r = cases.batch_case_table()[12]
    cfg = cases.config_cylinder(r["nth"], r["nr"], r["mach"], r["alpha_deg"], flux=hca.FLUX_ROE_FACEFRAME)
    out = []
    for prec in (0, 1):
        with _single(cfg, prec) as ctx:
            ctx.set_graph_steps(50)
            ctx.step(300)
            out.append(ctx.get_state())
            assert ctx.status() == 0, prec
    qc = [cases.prim_to_cons(q) for q in out]
    err = max(np.max(np.abs(qc[1][..., v] - qc[0][..., v])) / np.max(np.abs(qc[0][..., v])) for v in range(4))
    print("mixed vs fp64 after 300 steps: %.3e" % err)
    assert err <= 1e-6
