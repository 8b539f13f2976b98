"""MI355X-native structured-mesh finite-volume residual + Runge-Kutta hot path.

The compute path is the in-tree HIP library ``libhcamd.so`` (include/hcamd.h).
This module is a thin ctypes binding of that C-ABI for Python callers (tests,
bench, the adjoint wrapper); it contains no numerical fallback: importing
``Context`` without the built library raises immediately.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("HCAMD_LIBRARY", os.path.join(PKG, "libhcamd.so"))

LAW_SCALAR, LAW_EULER = 0, 1
BASIS_CONSERVED, BASIS_PRIMITIVE = 0, 1
FLUX_CENTRAL, FLUX_RUSANOV, FLUX_ROE, FLUX_SLAU, FLUX_AUSMPUP, FLUX_ROE_FACEFRAME, FLUX_ROE_UNPREC_WS = range(7)
SCALE_CONVECTIVE, SCALE_ACOUSTIC, SCALE_ADAPTIVE = range(3)
(LIM_NOLIMIT1, LIM_NOLIMIT2, LIM_NOLIMIT3, LIM_VANALBADA2, LIM_MINMOD2, LIM_SUPERBEE2,
 LIM_MC2, LIM_VANLEER2, LIM_CADA3, LIM_SMOOTHVANALBADA) = range(10)
BC_PERIODIC, BC_RIEMANN, BC_ENTROPY, BC_FIXED, BC_WALL = range(5)
DT_GLOBAL, DT_LOCAL, DT_FIXED = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
STATUS_NONFINITE, STATUS_NONPHYSICAL, STATUS_BCNORMAL = 1, 2, 4

C_ABI_SYMBOLS = (
    "hc_create", "hc_destroy", "hc_last_error", "hc_set_mesh", "hc_set_case", "hc_set_precision",
    "hc_set_boundary_refs", "hc_set_state", "hc_get_state", "hc_residual", "hc_step", "hc_step_steady", "hc_get_time",
    "hc_norms", "hc_get_status", "hc_synchronize", "hc_set_graph_steps", "hc_set_profiling",
    "hc_get_profile", "hc_device_state", "hc_stream_triad", "hc_version",
    "hc_batch_begin", "hc_batch_add_case", "hc_batch_end", "hc_batch_size", "hc_batch_set_boundary_refs",
    "hc_batch_set_state", "hc_batch_get_state", "hc_batch_residual", "hc_batch_norms", "hc_batch_get_time",
    "hc_set_tiles", "hc_get_tiles", "hc_set_uniform_geometry", "hc_get_uniform_geometry",
    "hc_set_tile_order", "hc_get_geometry_blocks",
    "hc_tangent_residual", "hc_adjoint_residual", "hc_adjoint_forward", "hc_adjoint_backward", "hc_tangent_steps",
    "hc_set_state_device", "hc_get_state_device", "hc_set_boundary_refs_device", "hc_adjoint_backward_device",
    "hc_step_norms",
    "hc_xmetrics", "hc_qmetrics", "hc_boundary_update", "hc_spectral_radius", "hc_forward_update",
    "hc_gradient", "hc_smooth_residual", "hc_set_residual_smoothing", "hc_spectral_radius_noalloc",
)
NTILECFG = 3


class Species(C.Structure):
    """Species<Euler,Real> (conservationLaws/euler/euler.h:48-60)."""
    _fields_ = [(n, C.c_double) for n in ("gamma", "minf", "lref", "nu", "pr", "dt", "R", "gamma1")]


class CaseDesc(C.Structure):
    """hc_case_desc of include/hcamd.h."""
    _fields_ = [("law", C.c_int), ("basis", C.c_int), ("flux", C.c_int), ("vscale", C.c_int),
                ("pscale", C.c_int), ("limiter", C.c_int), ("bc", C.c_int * 4),
                ("nstages", C.c_int), ("dt_mode", C.c_int),
                ("alpha", (C.c_double * 6) * 6), ("beta", C.c_double * 6), ("cfl", C.c_double),
                ("period_i", C.c_double * 2), ("period_j", C.c_double * 2), ("sp", Species)]


class HCError(RuntimeError):
    pass


_LIB = None


def library():
    """The loaded libhcamd.so; raises if it has not been built (no fallback path)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise HCError(f"{LIBPATH} is missing: build it with `python -m hyperboliconvergence_amd.build`")
    # One HIP runtime per process: torch's wheel carries its own libamdhip64.so.7.  Loaded first, it
    # also serves this library (same soname), so device pointers of torch tensors are valid here;
    # loaded after this library, it would be a second runtime in which torch finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIBPATH)
    P, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "hc_create": (i, [C.POINTER(P), i]),
        "hc_destroy": (i, [P]),
        "hc_last_error": (C.c_char_p, [P]),
        "hc_set_mesh": (i, [P, i, i, P]),
        "hc_set_case": (i, [P, P]),
        "hc_set_precision": (i, [P, i]),
        "hc_set_boundary_refs": (i, [P, P]),
        "hc_set_state": (i, [P, P]),
        "hc_get_state": (i, [P, P]),
        "hc_residual": (i, [P, P]),
        "hc_step": (i, [P, i]),
        "hc_step_steady": (i, [P, C.c_longlong, d, i, P, P]),
        "hc_get_time": (i, [P, P]),
        "hc_norms": (i, [P, P, P]),
        "hc_get_status": (i, [P, P]),
        "hc_synchronize": (i, [P]),
        "hc_set_graph_steps": (i, [P, i]),
        "hc_set_profiling": (i, [P, i]),
        "hc_get_profile": (i, [P, P]),
        "hc_device_state": (i, [P, P]),
        "hc_stream_triad": (i, [P, C.c_longlong, i, P]),
        "hc_version": (C.c_char_p, []),
        "hc_batch_begin": (i, [P]),
        "hc_batch_add_case": (i, [P, i, i, P, P, P]),
        "hc_batch_end": (i, [P]),
        "hc_batch_size": (i, [P, P]),
        "hc_batch_set_boundary_refs": (i, [P, i, P]),
        "hc_batch_set_state": (i, [P, i, P]),
        "hc_batch_get_state": (i, [P, i, P]),
        "hc_batch_residual": (i, [P, i, P]),
        "hc_batch_norms": (i, [P, i, P, P]),
        "hc_batch_get_time": (i, [P, i, P]),
        "hc_set_tiles": (i, [P, i]),
        "hc_get_tiles": (i, [P, P, P]),
        "hc_set_uniform_geometry": (i, [P, i]),
        "hc_get_uniform_geometry": (i, [P, P]),
        "hc_set_tile_order": (i, [P, i]),
        "hc_get_geometry_blocks": (i, [P, P]),
        "hc_tangent_residual": (i, [P, P, P, d, P, P]),
        "hc_adjoint_residual": (i, [P, P, P, P, P]),
        "hc_adjoint_forward": (i, [P, i, i]),
        "hc_adjoint_backward": (i, [P, P, P, P, P]),
        "hc_set_state_device": (i, [P, P]),
        "hc_step_norms": (i, [P, i, P, P]),
        "hc_get_state_device": (i, [P, P]),
        "hc_set_boundary_refs_device": (i, [P, P]),
        "hc_adjoint_backward_device": (i, [P, P, P, P, P]),
        "hc_tangent_steps": (i, [P, i, P, P, d, P]),
        "hc_xmetrics": (i, [P, P]),
        "hc_qmetrics": (i, [P, P, P]),
        "hc_boundary_update": (i, [P, P, P, P]),
        "hc_spectral_radius": (i, [P, P, P]),
        "hc_forward_update": (i, [P, i, d, P, P, P]),
        "hc_gradient": (i, [P, P, P]),
        "hc_smooth_residual": (i, [P, i, d, P]),
        "hc_set_residual_smoothing": (i, [P, i, d]),
        "hc_spectral_radius_noalloc": (i, [P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context of the C-ABI (hc_create ... hc_destroy)."""

    def __init__(self, device=0):
        self._L = library()
        self._h = C.c_void_p()
        rc = self._L.hc_create(C.byref(self._h), device)
        if rc != 0:
            raise HCError(f"hc_create failed ({rc})")
        self.ni = self.nj = self.nv = 0

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.hc_last_error(self._h)
            raise HCError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self._h:
            self._L.hc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_mesh(self, nodes):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
        self._check(self._L.hc_set_mesh(self._h, ni, nj, _p(nodes)), "hc_set_mesh")
        self.ni, self.nj = ni, nj

    def set_case(self, desc):
        self._check(self._L.hc_set_case(self._h, C.cast(C.pointer(desc), C.c_void_p)), "hc_set_case")
        self.nv = 4 if desc.law == LAW_EULER else 3

    def set_precision(self, precision):
        self._check(self._L.hc_set_precision(self._h, int(precision)), "hc_set_precision")

    def set_boundary_refs(self, bref):
        bref = np.ascontiguousarray(bref, dtype=np.float64)
        assert bref.size == (2 * self.nj + 2 * self.ni) * self.nv
        self._check(self._L.hc_set_boundary_refs(self._h, _p(bref)), "hc_set_boundary_refs")

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == self.ni * self.nj * self.nv
        self._check(self._L.hc_set_state(self._h, _p(q)), "hc_set_state")

    def get_state(self):
        q = np.zeros((self.ni, self.nj, self.nv))
        self._check(self._L.hc_get_state(self._h, _p(q)), "hc_get_state")
        return q

    def residual(self):
        r = np.zeros((self.ni, self.nj, self.nv + 1))
        self._check(self._L.hc_residual(self._h, _p(r)), "hc_residual")
        return r

    def step(self, nsteps):
        self._check(self._L.hc_step(self._h, int(nsteps)), "hc_step")

    def step_steady(self, max_steps, residual_drop, check_every=50):
        """Steady-state marching (hc_step_steady): (steps taken, residual drop reached)."""
        n, drop = C.c_longlong(), C.c_double()
        self._check(self._L.hc_step_steady(self._h, int(max_steps), float(residual_drop), int(check_every),
                                           C.byref(n), C.byref(drop)), "hc_step_steady")
        return n.value, drop.value

    def synchronize(self):
        self._check(self._L.hc_synchronize(self._h), "hc_synchronize")

    def time(self):
        t = C.c_double()
        self._check(self._L.hc_get_time(self._h, C.byref(t)), "hc_get_time")
        return t.value

    def norms(self):
        l2, linf = np.zeros(self.nv), np.zeros(self.nv)
        self._check(self._L.hc_norms(self._h, _p(l2), _p(linf)), "hc_norms")
        return l2, linf

    def step_norms(self, nsteps):
        """nsteps steps; the norms of the residual of the last step's start state (fused into it)."""
        l2, linf = np.zeros(self.nv), np.zeros(self.nv)
        self._check(self._L.hc_step_norms(self._h, int(nsteps), _p(l2), _p(linf)), "hc_step_norms")
        return l2, linf

    def status(self):
        s = C.c_uint()
        self._check(self._L.hc_get_status(self._h, C.byref(s)), "hc_get_status")
        return s.value

    def set_graph_steps(self, k):
        self._check(self._L.hc_set_graph_steps(self._h, int(k)), "hc_set_graph_steps")

    def set_profiling(self, mode):
        """0 off; 1 (or True) events around every launch, graphs off; 2 events around each step() call."""
        m = int(mode) if not isinstance(mode, bool) else (1 if mode else 0)
        self._check(self._L.hc_set_profiling(self._h, m), "hc_set_profiling")

    def profile(self):
        out = np.zeros(4)
        self._check(self._L.hc_get_profile(self._h, _p(out)), "hc_get_profile")
        return out

    def set_tiles(self, config):
        """-1: time every tile configuration on the first steps and keep the fastest; 0 .. NTILECFG-1:
        fixed configuration (shape, centre staging)."""
        self._check(self._L.hc_set_tiles(self._h, int(config)), "hc_set_tiles")

    def tiles(self):
        """(shape in use, milliseconds per step of each shape once timed)."""
        cfg = C.c_int()
        ms = np.zeros(NTILECFG)
        self._check(self._L.hc_get_tiles(self._h, C.byref(cfg), _p(ms)), "hc_get_tiles")
        return cfg.value, ms

    def set_uniform_geometry(self, enable):
        """1 (default): the uniform-geometry stage kernels when every case allows them; 0: never."""
        self._check(self._L.hc_set_uniform_geometry(self._h, int(enable)), "hc_set_uniform_geometry")

    def uniform_geometry(self):
        """Whether the uniform-geometry stage kernels are in use."""
        act = C.c_int()
        self._check(self._L.hc_get_uniform_geometry(self._h, C.byref(act)), "hc_get_uniform_geometry")
        return bool(act.value)

    def set_tile_order(self, interleave):
        """1 (default): a step runs the cases that share a geometry copy tile by tile together;
        0: case after case.  Bit-identical results."""
        self._check(self._L.hc_set_tile_order(self._h, int(interleave)), "hc_set_tile_order")

    def geometry_blocks(self):
        """Number of geometry copies on the device (cases with bitwise-equal geometry share one)."""
        n = C.c_int()
        self._check(self._L.hc_get_geometry_blocks(self._h, C.byref(n)), "hc_get_geometry_blocks")
        return n.value

    # ---- derivatives (single case, FP64) ----
    def _brows(self):
        return 2 * self.nj + 2 * self.ni

    def tangent_residual(self, qdot, bref_dot=None, minf_dot=0.0):
        """(R(q) [ni][nj][nv+1], Rdot [ni][nj][nv]) at the current state."""
        qdot = np.ascontiguousarray(qdot, dtype=np.float64)
        assert qdot.size == self.ni * self.nj * self.nv
        res = np.zeros((self.ni, self.nj, self.nv + 1))
        resd = np.zeros((self.ni, self.nj, self.nv))
        bd = None
        if bref_dot is not None:
            bref_dot = np.ascontiguousarray(bref_dot, dtype=np.float64)
            assert bref_dot.size == self._brows() * self.nv
            bd = _p(bref_dot)
        self._check(self._L.hc_tangent_residual(self._h, _p(qdot), bd, float(minf_dot), _p(res), _p(resd)),
                    "hc_tangent_residual")
        return res, resd

    def adjoint_residual(self, ybar):
        """(qbar [ni][nj][nv], bref_bar [rows][nv], minf_bar) at the current state."""
        ybar = np.ascontiguousarray(ybar, dtype=np.float64)
        assert ybar.size == self.ni * self.nj * self.nv
        qbar = np.zeros((self.ni, self.nj, self.nv))
        bb = np.zeros((self._brows(), self.nv))
        mb = C.c_double()
        self._check(self._L.hc_adjoint_residual(self._h, _p(ybar), _p(qbar), _p(bb), C.byref(mb)),
                    "hc_adjoint_residual")
        return qbar, bb, mb.value

    # ---- device arrays (integer device addresses, e.g. torch.Tensor.data_ptr() of float64 tensors) ----
    def set_state_device(self, q_ptr):
        self._check(self._L.hc_set_state_device(self._h, C.c_void_p(int(q_ptr))), "hc_set_state_device")

    def get_state_device(self, q_ptr):
        self._check(self._L.hc_get_state_device(self._h, C.c_void_p(int(q_ptr))), "hc_get_state_device")

    def set_boundary_refs_device(self, bref_ptr):
        self._check(self._L.hc_set_boundary_refs_device(self._h, C.c_void_p(int(bref_ptr))),
                    "hc_set_boundary_refs_device")

    def adjoint_backward_device(self, qbar_final_ptr, qbar_initial_ptr, bref_bar_ptr):
        """Device arrays in, device arrays out; returns the minf adjoint."""
        mb = C.c_double()
        self._check(self._L.hc_adjoint_backward_device(self._h, C.c_void_p(int(qbar_final_ptr)),
                                                       C.c_void_p(int(qbar_initial_ptr)),
                                                       C.c_void_p(int(bref_bar_ptr)), C.byref(mb)),
                    "hc_adjoint_backward_device")
        return mb.value

    def adjoint_forward(self, nsteps, checkpoint_every):
        self._check(self._L.hc_adjoint_forward(self._h, int(nsteps), int(checkpoint_every)), "hc_adjoint_forward")

    def adjoint_backward(self, qbar_final):
        qbar_final = np.ascontiguousarray(qbar_final, dtype=np.float64)
        assert qbar_final.size == self.ni * self.nj * self.nv
        q0b = np.zeros((self.ni, self.nj, self.nv))
        bb = np.zeros((self._brows(), self.nv))
        mb = C.c_double()
        self._check(self._L.hc_adjoint_backward(self._h, _p(qbar_final), _p(q0b), _p(bb), C.byref(mb)),
                    "hc_adjoint_backward")
        return q0b, bb, mb.value

    def tangent_steps(self, nsteps, qdot0, bref_dot=None, minf_dot=0.0):
        qdot0 = np.ascontiguousarray(qdot0, dtype=np.float64)
        assert qdot0.size == self.ni * self.nj * self.nv
        out = np.zeros((self.ni, self.nj, self.nv))
        bd = None
        if bref_dot is not None:
            bref_dot = np.ascontiguousarray(bref_dot, dtype=np.float64)
            bd = _p(bref_dot)
        self._check(self._L.hc_tangent_steps(self._h, int(nsteps), _p(qdot0), bd, float(minf_dot), _p(out)),
                    "hc_tangent_steps")
        return out

    # ---- spatial building blocks (single-case context) ----
    def xmetrics(self):
        """LSQ spatial metric per cell [ni][nj][3] = {xx, xy, yy}."""
        out = np.zeros((self.ni, self.nj, 3))
        self._check(self._L.hc_xmetrics(self._h, _p(out)), "hc_xmetrics")
        return out

    def qmetrics(self, q):
        """LSQ solution metric [ni][nj][nv][2] of the state q (solution basis)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros((self.ni, self.nj, self.nv, 2))
        self._check(self._L.hc_qmetrics(self._h, _p(q), _p(out)), "hc_qmetrics")
        return out

    def boundary_update(self, q, bref, ghost=None):
        """Ghost values [rows][nv] of the non-periodic boundaries (periodic rows keep `ghost`)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        bref = np.ascontiguousarray(bref, dtype=np.float64)
        out = np.zeros_like(bref) if ghost is None else np.array(ghost, dtype=np.float64, copy=True)
        self._check(self._L.hc_boundary_update(self._h, _p(q), _p(bref), _p(out)), "hc_boundary_update")
        return out

    def spectral_radius(self, res):
        """max over cells of lambda / volume of a residual [ni][nj][nv+1]."""
        res = np.ascontiguousarray(res, dtype=np.float64)
        out = C.c_double()
        self._check(self._L.hc_spectral_radius(self._h, _p(res), C.byref(out)), "hc_spectral_radius")
        return out.value

    def forward_update(self, mode, coef, res, q0):
        """mode 0: S(C(q0) + R coef / vol); mode 1: S(C(q0) + R coef / lambda)."""
        res = np.ascontiguousarray(res, dtype=np.float64)
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        q1 = np.zeros_like(q0)
        self._check(self._L.hc_forward_update(self._h, int(mode), float(coef), _p(res), _p(q0), _p(q1)),
                    "hc_forward_update")
        return q1

    def gradient(self, q):
        """Legacy neighbour-difference gradient [ni][nj][2][nv] of the state q."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.zeros((self.ni, self.nj, 2, self.nv))
        self._check(self._L.hc_gradient(self._h, _p(q), _p(out)), "hc_gradient")
        return out

    def smooth_residual(self, res, mode, alpha):
        """Residual smoothing of res [ni][nj][nv+1] (flux part): mode 0 explicit, 1 implicit, 2 both."""
        out = np.array(res, dtype=np.float64, copy=True, order="C")
        self._check(self._L.hc_smooth_residual(self._h, int(mode), float(alpha), _p(out)), "hc_smooth_residual")
        return out

    def set_residual_smoothing(self, mode, alpha=1.0):
        """Smooth every stage's accumulated residual in the time loop (mode -1 off, 0..2 as smooth_residual)."""
        self._check(self._L.hc_set_residual_smoothing(self._h, int(mode), float(alpha)), "hc_set_residual_smoothing")

    def spectral_radius_noalloc(self, res_dev_ptr, out_dev_ptr):
        """max lambda / volume of a device residual [ni][nj][nv+1] into a device double, allocating
        nothing (asynchronous on the context's stream)."""
        self._check(self._L.hc_spectral_radius_noalloc(self._h, C.c_void_p(int(res_dev_ptr)),
                                                       C.c_void_p(int(out_dev_ptr))), "hc_spectral_radius_noalloc")

    def stream_triad(self, n, reps):
        g = C.c_double()
        self._check(self._L.hc_stream_triad(self._h, int(n), int(reps), C.byref(g)), "hc_stream_triad")
        return g.value

    def device_state_ptr(self):
        p = C.c_void_p()
        self._check(self._L.hc_device_state(self._h, C.byref(p)), "hc_device_state")
        return p.value


def version():
    return library().hc_version().decode()


class Batch(Context):
    """A context holding several cases that share flux, limiter, basis and Runge-Kutta table;
    step() advances all of them with one launch per stage (hc_batch_* of include/hcamd.h)."""

    def __init__(self, device=0):
        super().__init__(device)
        self.shapes = []

    def begin(self):
        self._check(self._L.hc_batch_begin(self._h), "hc_batch_begin")
        self.shapes = []

    def add_case(self, nodes, desc):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
        cid = C.c_int()
        self._check(self._L.hc_batch_add_case(self._h, ni, nj, _p(nodes), C.cast(C.pointer(desc), C.c_void_p),
                                              C.byref(cid)), "hc_batch_add_case")
        self.shapes.append((ni, nj))
        self.nv = 4 if desc.law == LAW_EULER else 3
        return cid.value

    def end(self):
        self._check(self._L.hc_batch_end(self._h), "hc_batch_end")

    def __len__(self):
        return len(self.shapes)

    def set_case_boundary_refs(self, k, bref):
        ni, nj = self.shapes[k]
        bref = np.ascontiguousarray(bref, dtype=np.float64)
        assert bref.size == (2 * nj + 2 * ni) * self.nv
        self._check(self._L.hc_batch_set_boundary_refs(self._h, int(k), _p(bref)), "hc_batch_set_boundary_refs")

    def set_case_state(self, k, q):
        ni, nj = self.shapes[k]
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == ni * nj * self.nv
        self._check(self._L.hc_batch_set_state(self._h, int(k), _p(q)), "hc_batch_set_state")

    def get_case_state(self, k):
        ni, nj = self.shapes[k]
        q = np.zeros((ni, nj, self.nv))
        self._check(self._L.hc_batch_get_state(self._h, int(k), _p(q)), "hc_batch_get_state")
        return q

    def case_residual(self, k):
        ni, nj = self.shapes[k]
        r = np.zeros((ni, nj, self.nv + 1))
        self._check(self._L.hc_batch_residual(self._h, int(k), _p(r)), "hc_batch_residual")
        return r

    def case_norms(self, k):
        l2, linf = np.zeros(self.nv), np.zeros(self.nv)
        self._check(self._L.hc_batch_norms(self._h, int(k), _p(l2), _p(linf)), "hc_batch_norms")
        return l2, linf

    def case_time(self, k):
        t = C.c_double()
        self._check(self._L.hc_batch_get_time(self._h, int(k), C.byref(t)), "hc_batch_get_time")
        return t.value

