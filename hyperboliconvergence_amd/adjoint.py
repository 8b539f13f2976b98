"""Differentiable K-step time integration of one case for PyTorch.

``trajectory(solver, q0, bref, minf)`` returns the state after K explicit Runge-Kutta steps
(fixed time step, conserved basis) and is differentiable with respect to the initial state
q0 [ni][nj][nv], the boundary reference states bref [rows][nv] and the species' free-stream
Mach number minf (which AUSM+up and SLAU use for their low-Mach scaling).  The forward pass
runs the HIP stage kernels through the C-ABI and keeps a checkpoint every C steps; the
backward pass runs the hand-written adjoint kernels through the stages in reverse,
recomputing each segment's states from its checkpoint (about sqrt(K) states in memory for
C ~ sqrt(K)).  Gradients are exact for the discrete scheme (up to round-off) wherever it is
differentiable.

The wrapper moves tensors through host memory (the C-ABI takes host arrays); the work stays
on the GPU.  Only this module needs torch.
"""
import numpy as np
import torch

from . import Context


class TrajectorySolver:
    """A single-case context set up for K steps with checkpoints every C steps."""

    def __init__(self, nodes, case, nsteps, checkpoint_every, device=0):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.case = type(case).from_buffer_copy(case)
        self.nsteps = int(nsteps)
        self.checkpoint_every = int(checkpoint_every)
        self.ctx = Context(device)
        self.ctx.set_mesh(self.nodes)
        self.ctx.set_case(self.case)
        self.ctx.set_tiles(0)
        # the context keeps the checkpoints of its latest forward pass only
        self.generation = 0

    def set_minf(self, minf):
        minf = float(minf)
        if minf != self.case.sp.minf:
            self.case.sp.minf = minf
            self.ctx.set_case(self.case)   # keeps the state and the boundary references
            self.ctx.set_tiles(0)

    def forward(self, q0, bref, minf):
        self.set_minf(minf)
        self.ctx.set_boundary_refs(bref)
        self.ctx.set_state(q0)
        self.ctx.adjoint_forward(self.nsteps, self.checkpoint_every)
        self.generation += 1
        return self.ctx.get_state()

    def backward(self, qbar):
        return self.ctx.adjoint_backward(qbar)

    def close(self):
        self.ctx.close()


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, bref, minf, solver):
        qK = solver.forward(q0.detach().cpu().double().numpy(), bref.detach().cpu().double().numpy(),
                            float(minf.detach().cpu()))
        ctx.solver = solver
        ctx.generation = solver.generation
        return torch.from_numpy(qK).to(device=q0.device, dtype=q0.dtype)

    @staticmethod
    def backward(ctx, grad_qK):
        if ctx.solver.generation != ctx.generation:
            raise RuntimeError("trajectory: the solver ran another forward pass since this output was computed; "
                               "its checkpoints belong to that pass (use one solver per pending backward)")
        q0b, bb, mb = ctx.solver.backward(grad_qK.detach().cpu().double().numpy())
        dev, dt = grad_qK.device, grad_qK.dtype
        return (torch.from_numpy(q0b).to(device=dev, dtype=dt), torch.from_numpy(bb).to(device=dev, dtype=dt),
                torch.tensor(mb, device=dev, dtype=dt), None)


def trajectory(solver, q0, bref, minf):
    """State after solver.nsteps steps from q0 (differentiable in q0, bref and minf)."""
    return _Trajectory.apply(q0, bref, minf, solver)
