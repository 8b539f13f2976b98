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

Tensors on the GPU stay there: the device-array entry points of the C-ABI read and write them
in place (torch and the library share one HIP runtime, see ``library()``); CPU tensors go through
the host-array entry points.  Only this module needs torch.
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
        self.device = int(device)
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

    # ---- float64 torch tensors on the context's GPU, read and written in place ----
    def _check_device(self, t):
        if t.device.type != "cuda" or (t.device.index or 0) != self.device:
            raise ValueError(f"trajectory: tensors must be on cuda:{self.device}, got {t.device}")

    def forward_device(self, q0, bref, minf):
        for t in (q0, bref):
            self._check_device(t)
        self.set_minf(minf)
        torch.cuda.current_stream(q0.device).synchronize()   # the inputs are written
        self.ctx.set_boundary_refs_device(bref.data_ptr())
        self.ctx.set_state_device(q0.data_ptr())
        self.ctx.adjoint_forward(self.nsteps, self.checkpoint_every)
        self.generation += 1
        out = torch.empty_like(q0)
        torch.cuda.current_stream(q0.device).synchronize()   # out's block is free
        self.ctx.get_state_device(out.data_ptr())
        return out

    def backward_device(self, qbar):
        self._check_device(qbar)
        q0b = torch.empty_like(qbar)
        bb = torch.empty((2 * self.ctx.nj + 2 * self.ctx.ni, self.ctx.nv), dtype=torch.float64, device=qbar.device)
        torch.cuda.current_stream(qbar.device).synchronize()
        mb = self.ctx.adjoint_backward_device(qbar.data_ptr(), q0b.data_ptr(), bb.data_ptr())
        return q0b, bb, mb

    def close(self):
        self.ctx.close()


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, bref, minf, solver):
        m = float(minf.detach().cpu())
        if q0.is_cuda:
            qK = solver.forward_device(q0.detach().to(torch.float64).contiguous(),
                                       bref.detach().to(device=q0.device, dtype=torch.float64).contiguous(), m)
        else:
            qK = torch.from_numpy(solver.forward(q0.detach().double().numpy(), bref.detach().cpu().double().numpy(), m))
        ctx.solver = solver
        ctx.generation = solver.generation
        ctx.bref_device, ctx.minf_device = bref.device, minf.device
        return qK.to(dtype=q0.dtype)

    @staticmethod
    def backward(ctx, grad_qK):
        if ctx.solver.generation != ctx.generation:
            raise RuntimeError("trajectory: the solver ran another forward pass since this output was computed; "
                               "its checkpoints belong to that pass (use one solver per pending backward)")
        dev, dt = grad_qK.device, grad_qK.dtype
        if grad_qK.is_cuda:
            q0b, bb, mb = ctx.solver.backward_device(grad_qK.detach().to(torch.float64).contiguous())
        else:
            q0b, bb, mb = ctx.solver.backward(grad_qK.detach().double().numpy())
            q0b, bb = torch.from_numpy(q0b), torch.from_numpy(bb)
        return (q0b.to(device=dev, dtype=dt), bb.to(device=ctx.bref_device, dtype=dt),
                torch.tensor(mb, device=ctx.minf_device, dtype=dt), None)


def trajectory(solver, q0, bref, minf):
    """State after solver.nsteps steps from q0 (differentiable in q0, bref and minf)."""
    return _Trajectory.apply(q0, bref, minf, solver)
