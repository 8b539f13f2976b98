"""GPU: tangent and adjoint kernels through the C-ABI.

* residual level: the device kernels against their host execution (same code) and the
  dot-product identity between the device tangent and adjoint;
* trajectory level (fixed dt, SSP-RK3, AUSM+up with the smooth van Albada limiter on the
  cylinder O-grid): the adjoint of K steps against the tangent of K steps (dot-product
  identity), against central finite differences of the oracle trajectory (h = 1e-6), and for
  several checkpoint intervals (bit-identical gradients).
"""
import numpy as np
import pytest

import pyemu
import pyoracle as po
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _cfg(nth=32, nr=16, nsteps=12):
    from hyperboliconvergence_amd import cases
    cfg = cases.config_adjoint(nth, nr, nsteps=nsteps)
    res = po.residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"])
    vol = po.dual(cfg["nodes"])[..., 0]
    cfg["case"].cfl = 0.8 / np.max(res[..., 4] / vol)   # fixed step: CFL 0.8 of the initial state
    return cfg


def _ctx(cfg):
    import hyperboliconvergence_amd as hca
    ctx = hca.Context(0)
    ctx.set_mesh(cfg["nodes"])
    ctx.set_case(cfg["case"])
    ctx.set_tiles(0)
    ctx.set_boundary_refs(cfg["bref"])
    ctx.set_state(cfg["q"])
    return ctx


@pytest.mark.parametrize("flux", [0, 1, 2, 3, 4, 5, 6])
def test_tangent_and_adjoint_residual_match_host_execution(flux):
    import hyperboliconvergence_amd as hca
    cfg = _cfg()
    cfg["case"].flux = flux
    cfg["case"].pscale = hca.SCALE_ACOUSTIC if flux == hca.FLUX_SLAU else hca.SCALE_CONVECTIVE
    rng = np.random.default_rng(1)
    qd = rng.standard_normal(cfg["q"].shape)
    bd = rng.standard_normal(cfg["bref"].shape)
    yb = rng.standard_normal(cfg["q"].shape)
    with _ctx(cfg) as ctx:
        res, rd = ctx.tangent_residual(qd, bd, 0.3)
        qb, bb, mb = ctx.adjoint_residual(yb)
    res_e, rd_e = pyemu.tangent_residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"], qd, bd, 0.3)
    qb_e, bb_e, mb_e = pyemu.adjoint_residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"], yb)
    assert rel_err(res, res_e) <= 1e-12 and rel_err(rd, rd_e) <= 1e-12
    assert rel_err(qb, qb_e) <= 1e-12 and rel_err(bb, bb_e) <= 1e-12 and abs(mb - mb_e) <= 1e-12 * abs(mb_e)
    lhs = np.sum(yb * rd)
    rhs = np.sum(qb * qd) + np.sum(bb * bd) + mb * 0.3
    assert abs(lhs - rhs) <= 1e-12 * np.sum(np.abs(yb * rd))


def _box_case(kind, flux):
    """Meshes the adjoint tiles (7 x 16 cells) cover raggedly: a non-periodic box smaller than a tile
    with every far-field kind and a wall, and a doubly periodic box with ragged tiles in both
    directions."""
    from hyperboliconvergence_amd import cases
    import hyperboliconvergence_amd as hca
    g = 1.4
    rng = np.random.default_rng(4)
    if kind == "box":
        ni, nj = 9, 12
        nodes = po.mesh_linspace(ni, nj, 0.0, 1.0, 0.0, 1.3)
        bc = (po.BC_FIXED, po.BC_ENTROPY, po.BC_RIEMANN, po.BC_WALL)
        per_i = per_j = (0.0, 0.0)
    else:
        ni, nj = 10, 19
        nodes = po.mesh_linspace(ni, nj, -5.0, 5.0, -5.0, 5.0)
        bc = (0, 0, 0, 0)
        per_i, per_j = (10.0, 0.0), (0.0, 10.0)
    qinf = np.array([0.3, 0.1, 1.0, 1.0 / g])
    q = cases.prim_to_cons(qinf * (1 + 0.02 * rng.standard_normal((ni, nj, 4))))
    bref = cases.prim_to_cons(np.broadcast_to(qinf, (po.bref_size(ni, nj), 4)).copy())
    cs = po.make_case(law=po.LAW_EULER, basis=po.BASIS_CONSERVED, flux=flux, limiter=po.LIM_SMOOTHVANALBADA, bc=bc,
                      rk=33, cfl=1e-3, dt_mode=po.DT_FIXED, minf=0.3, period_i=per_i, period_j=per_j)
    return {"nodes": nodes, "case": cs, "q": q, "bref": bref}


@pytest.mark.parametrize("kind", ["box", "periodic"])
@pytest.mark.parametrize("flux", [2, 4])
def test_adjoint_tiles_on_ragged_and_small_grids_match_host_execution(kind, flux):
    """The tiled adjoint kernel (boundary faces of every kind, periodic wraps, tiles larger than
    the grid) against its host execution and the dot-product identity with the tangent."""
    cfg = _box_case(kind, flux)
    rng = np.random.default_rng(7)
    qd = rng.standard_normal(cfg["q"].shape)
    bd = rng.standard_normal(cfg["bref"].shape)
    yb = rng.standard_normal(cfg["q"].shape)
    with _ctx(cfg) as ctx:
        _, rd = ctx.tangent_residual(qd, bd, 0.3)
        qb, bb, mb = ctx.adjoint_residual(yb)
    qb_e, bb_e, mb_e = pyemu.adjoint_residual(cfg["case"], cfg["nodes"], cfg["q"], cfg["bref"], yb)
    assert rel_err(qb, qb_e) <= 1e-12 and rel_err(bb, bb_e) <= 1e-12
    assert abs(mb - mb_e) <= 1e-12 * max(abs(mb_e), 1e-300)
    lhs = np.sum(yb * rd)
    rhs = np.sum(qb * qd) + np.sum(bb * bd) + mb * 0.3
    assert abs(lhs - rhs) <= 1e-12 * np.sum(np.abs(yb * rd))


def _adjoint_grad(cfg, K, C):
    from hyperboliconvergence_amd import cases
    with _ctx(cfg) as ctx:
        ctx.adjoint_forward(K, C)
        qK = ctx.get_state()
        g = cases.wall_force_x_grad(cfg["nodes"], qK)
        q0b, bb, mb = ctx.adjoint_backward(g)
    return qK, g, q0b, bb, mb


def _set_rk(cfg, rk):
    n, a, b = po.rk_table(rk)
    cs = cfg["case"]
    cs.nstages = n
    for r in range(6):
        cs.beta[r] = b[r]
        for k in range(6):
            cs.alpha[r][k] = a[r, k]


@pytest.mark.parametrize("rk", [33, 11, 22, 34, 44])
def test_trajectory_adjoint_is_transpose_of_trajectory_tangent(rk):
    """Every Runge-Kutta table (SSP 1-4 stages and classical RK4, whose rows hold zeros): the
    backward sweep's fused stage adjoints against the trajectory tangent."""
    cfg = _cfg()
    _set_rk(cfg, rk)
    K = cfg["nsteps"]
    qK, g, q0b, bb, mb = _adjoint_grad(cfg, K, 4)
    rng = np.random.default_rng(2)
    qd = rng.standard_normal(cfg["q"].shape) * 1e-3
    bd = rng.standard_normal(cfg["bref"].shape) * 1e-3
    md = 0.7
    with _ctx(cfg) as ctx:
        qdK = ctx.tangent_steps(K, qd, bd, md)
        assert np.array_equal(ctx.get_state(), qK)
    lhs = np.sum(g * qdK)
    rhs = np.sum(q0b * qd) + np.sum(bb * bd) + mb * md
    print("tangent %.15e adjoint %.15e" % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-10 * (np.sum(np.abs(q0b * qd)) + np.sum(np.abs(bb * bd)) + abs(mb * md))


def test_checkpoint_interval_does_not_change_the_gradient():
    cfg = _cfg(nsteps=10)
    ref = _adjoint_grad(cfg, 10, 10)
    for C in (1, 3):
        out = _adjoint_grad(cfg, 10, C)
        assert np.array_equal(out[2], ref[2]) and np.array_equal(out[3], ref[3]) and out[4] == ref[4]


def test_trajectory_gradient_matches_oracle_finite_differences():
    """Wall force J(q_K) for the state and minf; a weighted sum of q_K (which sees the far field
    within K steps) for the boundary references.  The state direction is a smooth field of
    amplitude 1e-3 (the perturbation family of the adjoint case); at h = 1e-5 the central
    difference is within 2e-7 of the derivative (round-off of J's cancellations grows below)."""
    from hyperboliconvergence_amd import cases
    cfg = _cfg()
    K = cfg["nsteps"]
    _, _, q0b, bb, mb = _adjoint_grad(cfg, K, 4)

    def traj(q0, bref, minf):
        cs = type(cfg["case"]).from_buffer_copy(cfg["case"])
        cs.sp.minf = minf
        qK, _, _ = po.integrate(cs, cfg["nodes"], q0, K, bref)
        return qK

    def J(q0, bref, minf):
        return cases.wall_force_x(cfg["nodes"], traj(q0, bref, minf))

    d = cases.smooth_perturbation(16, 32, seed=7, amp=1e-3)
    h = 1e-5
    fd = (J(cfg["q"] + h * d, cfg["bref"], 0.4) - J(cfg["q"] - h * d, cfg["bref"], 0.4)) / (2 * h)
    ad = np.sum(q0b * d)
    print("dJ/dq0 . d: adjoint %.12e fd %.12e" % (ad, fd))
    assert abs(ad - fd) <= 1e-6 * abs(fd)
    h = 1e-6
    fd_m = (J(cfg["q"], cfg["bref"], 0.4 + h) - J(cfg["q"], cfg["bref"], 0.4 - h)) / (2 * h)
    print("dJ/dminf: adjoint %.12e fd %.12e" % (mb, fd_m))
    assert abs(mb - fd_m) <= 1e-6 * abs(fd_m)
    # weighted sum of the final state near the far field: the boundary references (the far-field
    # state reaches three rows in K steps; the difference quotient levels off near 7e-7 there)
    w = np.zeros_like(cfg["q"])
    w[-3:] = np.random.default_rng(3).standard_normal(w[-3:].shape)
    with _ctx(cfg) as ctx:
        ctx.adjoint_forward(K, 4)
        _, bbw, _ = ctx.adjoint_backward(w)
    e = np.random.default_rng(4).standard_normal(cfg["bref"].shape)
    h = 1e-6
    fd_b = (np.sum(w * traj(cfg["q"], cfg["bref"] + h * e, 0.4)) - np.sum(w * traj(cfg["q"], cfg["bref"] - h * e, 0.4))) / (2 * h)
    print("d(w.qK)/dbref . e: adjoint %.12e fd %.12e" % (np.sum(bbw * e), fd_b))
    assert abs(np.sum(bbw * e) - fd_b) <= 2e-6 * abs(fd_b)


def test_torch_autograd_gradient_through_mach_number():
    """J(M) = wall force after K steps, with M entering the initial state, the far-field
    references and the species' minf: torch.autograd through the adjoint against central
    differences of the oracle pipeline."""
    import torch
    from hyperboliconvergence_amd import cases
    from hyperboliconvergence_amd.adjoint import TrajectorySolver, trajectory
    cfg = _cfg()
    K, G = cfg["nsteps"], cases.GAMMA
    pert = torch.tensor(cases.smooth_perturbation(16, 32), dtype=torch.float64)
    nb = cfg["bref"].shape[0]
    area, nx, _ = cases.wall_faces(cfg["nodes"])
    w = torch.tensor(area * nx, dtype=torch.float64)

    def freestream(M):
        one = torch.ones_like(M)
        return torch.stack([M, 0.0 * M, one, one / (G * (G - 1.0)) + 0.5 * M * M])

    def J_of(qK):
        qw = qK[0]
        p = (G - 1.0) * (qw[:, 3] - 0.5 * (qw[:, 0] ** 2 + qw[:, 1] ** 2) / qw[:, 2])
        return -(p * w).sum()

    dev = "cuda" if torch.cuda.is_available() else "cpu"
    M = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
    fs = freestream(M)
    q0 = (fs[None, None, :] + pert).to(dev)
    bref = fs[None, :].expand(nb, 4).to(dev)
    solver = TrajectorySolver(cfg["nodes"], cfg["case"], K, 4)
    qK = trajectory(solver, q0, bref, M)
    J = J_of(qK.cpu())
    J.backward()
    solver.close()

    def J_oracle(m):
        cs = type(cfg["case"]).from_buffer_copy(cfg["case"])
        cs.sp.minf = m
        f = cases.freestream_cons(m)
        qK, _, _ = po.integrate(cs, cfg["nodes"], f[None, None, :] + pert.numpy(), K, np.broadcast_to(f, (nb, 4)).copy())
        return cases.wall_force_x(cfg["nodes"], qK)

    h = 1e-6
    fd = (J_oracle(0.4 + h) - J_oracle(0.4 - h)) / (2 * h)
    print("dJ/dM autograd %.12e fd %.12e" % (M.grad.item(), fd))
    assert abs(J.item() - J_oracle(0.4)) <= 1e-10 * abs(J_oracle(0.4)) + 1e-15
    assert abs(M.grad.item() - fd) <= 1e-6 * abs(fd)


def test_backward_after_a_second_forward_on_the_same_solver_raises():
    """The context keeps the checkpoints of its latest forward pass only: the backward of an
    earlier output must refuse rather than use the later pass's checkpoints."""
    import torch
    from hyperboliconvergence_amd.adjoint import TrajectorySolver, trajectory
    cfg = _cfg(nsteps=4)
    q0 = torch.tensor(cfg["q"], dtype=torch.float64, requires_grad=True)
    bref = torch.tensor(cfg["bref"], dtype=torch.float64)
    M = torch.tensor(cfg["case"].sp.minf, dtype=torch.float64)
    solver = TrajectorySolver(cfg["nodes"], cfg["case"], 4, 2)
    try:
        first = trajectory(solver, q0, bref, M)
        second = trajectory(solver, q0 * 1.0001, bref, M)
        with pytest.raises(RuntimeError, match="another forward pass"):
            first.sum().backward()
        second.sum().backward()   # the latest output still differentiates
        assert torch.isfinite(q0.grad).all()
    finally:
        solver.close()


def test_cuda_tensors_stay_on_the_gpu_and_give_the_host_path_bits():
    """The device-array entry points: torch tensors on the GPU (the library shares torch's HIP
    runtime) give the same forward state and gradients as CPU tensors through the host-array entry
    points; set/get state round-trips on the device; a host pointer is refused, not dereferenced."""
    import torch
    import hyperboliconvergence_amd as hca
    from hyperboliconvergence_amd.adjoint import TrajectorySolver, trajectory
    hca.library()
    assert torch.cuda.is_available(), "torch and libhcamd must share one HIP runtime"
    cfg = _cfg(nsteps=8)
    out = {}
    for dev in ("cpu", "cuda"):
        q0 = torch.tensor(cfg["q"], dtype=torch.float64, device=dev, requires_grad=True)
        bref = torch.tensor(cfg["bref"], dtype=torch.float64, device=dev, requires_grad=True)
        M = torch.tensor(cfg["case"].sp.minf, dtype=torch.float64, requires_grad=True)
        solver = TrajectorySolver(cfg["nodes"], cfg["case"], 8, 3)
        try:
            qK = trajectory(solver, q0, bref, M)
            assert qK.device.type == dev
            (qK * qK).sum().backward()
        finally:
            solver.close()
        out[dev] = [t.detach().cpu().numpy() for t in (qK, q0.grad, bref.grad, M.grad)]
    for a, b in zip(out["cpu"], out["cuda"]):
        assert np.array_equal(a, b)
    with _ctx(cfg) as ctx:
        q = torch.tensor(cfg["q"] * 1.01, dtype=torch.float64, device="cuda")
        ctx.set_state_device(q.data_ptr())
        assert np.array_equal(ctx.get_state(), cfg["q"] * 1.01)
        back = torch.empty_like(q)
        ctx.get_state_device(back.data_ptr())
        assert torch.equal(back, q)
        host = np.ascontiguousarray(cfg["q"])
        with pytest.raises(hca.HCError, match="device memory"):
            ctx.set_state_device(host.ctypes.data)
