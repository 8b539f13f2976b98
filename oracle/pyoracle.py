"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle (oracle.cpp) restates the reference's residual + Runge-Kutta path on
the CPU.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py import this module; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

LAW_SCALAR, LAW_EULER = 0, 1
BASIS_CONSERVED, BASIS_PRIMITIVE = 0, 1
FLUX_CENTRAL, FLUX_RUSANOV, FLUX_ROE, FLUX_SLAU, FLUX_AUSMPUP, FLUX_ROE_FACEFRAME, FLUX_ROE_UNPREC_WS = range(7)
SCALE_CONVECTIVE, SCALE_ACOUSTIC, SCALE_ADAPTIVE = range(3)
(LIM_NOLIMIT1, LIM_NOLIMIT2, LIM_NOLIMIT3, LIM_VANALBADA2, LIM_MINMOD2, LIM_SUPERBEE2,
 LIM_MC2, LIM_VANLEER2, LIM_CADA3, LIM_SMOOTHVANALBADA) = range(10)
BC_PERIODIC, BC_RIEMANN, BC_ENTROPY, BC_FIXED, BC_WALL = range(5)
DT_GLOBAL, DT_LOCAL, DT_FIXED = 0, 1, 2


class Species(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "minf", "lref", "nu", "pr", "dt", "R", "gamma1")]


class Case(C.Structure):
    _fields_ = [("law", C.c_int), ("basis", C.c_int), ("flux", C.c_int), ("vscale", C.c_int),
                ("pscale", C.c_int), ("limiter", C.c_int), ("bc", C.c_int * 4),
                ("nstages", C.c_int), ("dt_mode", C.c_int),
                ("alpha", (C.c_double * 6) * 6), ("beta", C.c_double * 6), ("cfl", C.c_double),
                ("period_i", C.c_double * 2), ("period_j", C.c_double * 2), ("sp", Species)]


class Face(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("area", "cx", "cy", "nx", "ny", "tx", "ty")]


_lib = {}


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _lib:
        return _lib[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-C", HERE, name], check=True, capture_output=True)
    L = C.CDLL(path)
    d, i, P = C.c_double, C.c_int, C.c_void_p
    sig = {
        "orc_nvar": (i, [i]),
        "orc_surface": (None, [P, P, P]),
        "orc_volume": (None, [P, P, P, P, P]),
        "orc_dual": (None, [i, i, P, P]),
        "orc_mesh_linspace": (None, [i, i, d, d, d, d, P]),
        "orc_mesh_cylinder": (None, [i, i, d, d, P]),
        "orc_ogrid_ratio": (d, [i, d, d, d]),
        "orc_mesh_ogrid": (None, [i, i, d, d, d, P]),
        "orc_air_species": (Species, []),
        "orc_set2state": (None, [P, P, P]),
        "orc_state2set": (None, [P, i, P, P]),
        "orc_delta": (None, [P, i, i, P, P, P]),
        "orc_roe_average": (None, [P, P, P, P]),
        "orc_flux": (i, [P, P, P, P, P]),
        "orc_limit": (d, [i, d, d]),
        "orc_machsplit": (d, [i, i, d, d]),
        "orc_bc_update": (None, [P, i, P, P, P, P, P]),
        "orc_wall_flux": (None, [P, P, P, P]),
        "orc_xmetrics": (None, [P, i, i, P, P]),
        "orc_qmetrics": (None, [P, i, i, P, P, P]),
        "orc_lsq_solve": (None, [P, P, i, P, P]),
        "orc_muscl_face": (None, [P, P, P, P, P, P, P, P, P, P, P]),
        "orc_residual": (i, [P, i, i, P, P, P, P]),
        "orc_spectral_radius": (d, [i, i, P, P, i]),
        "orc_integrate": (i, [P, i, i, P, P, P, i, P, P]),
        "orc_norms": (None, [i, i, i, P, P, P, P]),
        "orc_set_threads": (None, [i]),
        "orc_get_threads": (i, []),
        "orc_rk_table": (None, [i, P, P, P]),
        "orc_gradient": (None, [P, i, i, P, P]),
        "orc_smooth_residual": (None, [P, i, i, i, d, P]),
        "orc_integrate_smoothed": (i, [P, i, i, P, P, P, i, i, d, P]),
    }
    for fn, (res, args) in sig.items():
        f = getattr(L, fn)
        f.restype = res
        f.argtypes = args
    _lib[name] = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _ref(x):
    return C.cast(C.pointer(x), C.c_void_p)


def air_species():
    return lib().orc_air_species()


def rk_table(which):
    n = C.c_int()
    a = np.zeros(36)
    b = np.zeros(6)
    lib().orc_rk_table(which, C.byref(n), _p(a), _p(b))
    return n.value, a.reshape(6, 6), b


def make_case(law=LAW_EULER, basis=BASIS_PRIMITIVE, flux=FLUX_ROE, limiter=LIM_MINMOD2,
              bc=(BC_PERIODIC,) * 4, rk=33, cfl=0.5, dt_mode=DT_GLOBAL, vscale=None, pscale=None,
              period_i=(0.0, 0.0), period_j=(0.0, 0.0), species=None, minf=None):
    cs = Case()
    cs.law, cs.basis, cs.flux, cs.limiter = law, basis, flux, limiter
    if vscale is None:
        vscale = SCALE_CONVECTIVE
    if pscale is None:
        pscale = SCALE_ACOUSTIC if flux == FLUX_SLAU else SCALE_CONVECTIVE
    cs.vscale, cs.pscale = vscale, pscale
    for k in range(4):
        cs.bc[k] = bc[k]
    n, a, b = rk_table(rk)
    cs.nstages = n
    for s in range(6):
        cs.beta[s] = b[s]
        for k in range(6):
            cs.alpha[s][k] = a[s, k]
    cs.cfl = cfl
    cs.dt_mode = dt_mode
    cs.period_i[0], cs.period_i[1] = period_i
    cs.period_j[0], cs.period_j[1] = period_j
    sp = species if species is not None else air_species()
    if minf is not None:
        sp.minf = minf
    cs.sp = sp
    return cs


def nvar(cs):
    return 4 if cs.law == LAW_EULER else 3


def surface(p0, p1):
    f = Face()
    lib().orc_surface(_p(np.ascontiguousarray(p0, float)), _p(np.ascontiguousarray(p1, float)), _ref(f))
    return f


def volume(p0, p1, p2, p3):
    out = np.zeros(3)
    a = [np.ascontiguousarray(x, float) for x in (p0, p1, p2, p3)]
    lib().orc_volume(*[_p(x) for x in a], _p(out))
    return out


def dual(nodes):
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    cells = np.zeros((ni, nj, 3))
    lib().orc_dual(ni, nj, _p(np.ascontiguousarray(nodes)), _p(cells))
    return cells


def mesh_linspace(ni, nj, lox, hix, loy, hiy):
    nodes = np.zeros((ni + 1, nj + 1, 2))
    lib().orc_mesh_linspace(ni, nj, lox, hix, loy, hiy, _p(nodes))
    return nodes


def mesh_cylinder(nr, nth, rc, rb):
    nodes = np.zeros((nr + 1, nth + 1, 2))
    lib().orc_mesh_cylinder(nr, nth, rc, rb, _p(nodes))
    return nodes


def mesh_ogrid(nr, nth, rc=0.5, rb=20.0, dr0=1e-3):
    nodes = np.zeros((nr + 1, nth + 1, 2))
    lib().orc_mesh_ogrid(nr, nth, rc, rb, dr0, _p(nodes))
    return nodes


def ogrid_ratio(nr, rc=0.5, rb=20.0, dr0=1e-3):
    return lib().orc_ogrid_ratio(nr, rc, rb, dr0)


def set2state(cs, q):
    st = np.zeros(8)
    lib().orc_set2state(_ref(cs), _p(np.ascontiguousarray(q, float)), _p(st))
    return st


def state2set(cs, basis, st):
    q = np.zeros(nvar(cs))
    lib().orc_state2set(_ref(cs), basis, _p(np.ascontiguousarray(st, float)), _p(q))
    return q


def delta(cs, frm, to, st, dq):
    out = np.zeros(4)
    lib().orc_delta(_ref(cs), frm, to, _p(np.ascontiguousarray(st, float)), _p(np.ascontiguousarray(dq, float)), _p(out))
    return out


def roe_average(cs, sl, sr):
    out = np.zeros(8)
    lib().orc_roe_average(_ref(cs), _p(np.ascontiguousarray(sl, float)), _p(np.ascontiguousarray(sr, float)), _p(out))
    return out


def face_from(area, c, n):
    f = Face()
    f.area, f.cx, f.cy = area, c[0], c[1]
    f.nx, f.ny = n[0], n[1]
    f.tx, f.ty = -n[1], n[0]
    return f


def flux(cs, face, ql, qr):
    out = np.zeros(nvar(cs) + 1)
    lib().orc_flux(_ref(cs), _ref(face), _p(np.ascontiguousarray(ql, float)), _p(np.ascontiguousarray(qr, float)), _p(out))
    return out


def limit(lim, a, b):
    return lib().orc_limit(lim, float(a), float(b))


def machsplit(which, sign, m, param=0.0):
    return lib().orc_machsplit(which, sign, float(m), float(param))


def bc_update(cs, bctype, face, cell3, qint, qref):
    out = np.zeros(nvar(cs))
    lib().orc_bc_update(_ref(cs), bctype, _ref(face), _p(np.ascontiguousarray(cell3, float)),
                        _p(np.ascontiguousarray(qint, float)), _p(np.ascontiguousarray(qref, float)), _p(out))
    return out


def wall_flux(cs, face, qint):
    out = np.zeros(5)
    lib().orc_wall_flux(_ref(cs), _ref(face), _p(np.ascontiguousarray(qint, float)), _p(out))
    return out


def xmetrics(cs, cells):
    ni, nj = cells.shape[:2]
    xm = np.zeros((ni, nj, 3))
    lib().orc_xmetrics(_ref(cs), ni, nj, _p(np.ascontiguousarray(cells)), _p(xm))
    return xm


def qmetrics(cs, cells, q):
    ni, nj = cells.shape[:2]
    qm = np.zeros((ni, nj, nvar(cs), 2))
    lib().orc_qmetrics(_ref(cs), ni, nj, _p(np.ascontiguousarray(cells)), _p(np.ascontiguousarray(q, float)), _p(qm))
    return qm


def lsq_solve(xm3, qm, direction):
    qm = np.ascontiguousarray(qm, float)
    nv = qm.shape[0]
    out = np.zeros(nv)
    lib().orc_lsq_solve(_p(np.ascontiguousarray(xm3, float)), _p(qm), nv, _p(np.ascontiguousarray(direction, float)), _p(out))
    return out


def muscl_face(cs, face, cl3, cr3, ql, qr, xl, xr, qml, qmr):
    out = np.zeros(nvar(cs) + 1)
    a = [np.ascontiguousarray(x, float) for x in (cl3, cr3, ql, qr, xl, xr, qml, qmr)]
    lib().orc_muscl_face(_ref(cs), _ref(face), *[_p(x) for x in a], _p(out))
    return out


def bref_size(ni, nj):
    return 2 * nj + 2 * ni


def residual(cs, nodes, q, bref=None, fast=False):
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    nv = nvar(cs)
    if bref is None:
        bref = np.zeros((bref_size(ni, nj), nv))
    res = np.zeros((ni, nj, nv + 1))
    rc = lib(fast).orc_residual(_ref(cs), ni, nj, _p(np.ascontiguousarray(nodes)), _p(np.ascontiguousarray(q, float)),
                                _p(np.ascontiguousarray(bref, float)), _p(res))
    if rc == -1:
        raise ValueError("invalid case for the oracle residual")
    return res


def integrate(cs, nodes, q, nsteps, bref=None, fast=False):
    """Returns (q_final, t, timers_ms[7]); q is copied, not modified."""
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    nv = nvar(cs)
    if bref is None:
        bref = np.zeros((bref_size(ni, nj), nv))
    qq = np.ascontiguousarray(q, float).copy()
    t = C.c_double()
    tm = np.zeros(7)
    rc = lib(fast).orc_integrate(_ref(cs), ni, nj, _p(np.ascontiguousarray(nodes)), _p(qq),
                                 _p(np.ascontiguousarray(bref, float)), nsteps, C.byref(t), _p(tm))
    if rc == -1:
        raise ValueError("invalid case for the oracle integrate")
    return qq, t.value, tm


def integrate_smoothed(cs, nodes, q, nsteps, mode, alpha, bref=None):
    """integrate with the accumulated stage residual smoothed before each update; returns (q, t)."""
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    if bref is None:
        bref = np.zeros((bref_size(ni, nj), nvar(cs)))
    qq = np.ascontiguousarray(q, float).copy()
    t = C.c_double()
    rc = lib().orc_integrate_smoothed(_ref(cs), ni, nj, _p(np.ascontiguousarray(nodes)), _p(qq),
                                      _p(np.ascontiguousarray(bref, float)), nsteps, int(mode), float(alpha), C.byref(t))
    if rc == -1:
        raise ValueError("invalid case for the oracle integrate")
    return qq, t.value


def spectral_radius(cells, res):
    ni, nj = cells.shape[:2]
    nv = res.shape[2] - 1
    return lib().orc_spectral_radius(ni, nj, _p(np.ascontiguousarray(cells)), _p(np.ascontiguousarray(res)), nv)


def norms(cells, res):
    ni, nj = cells.shape[:2]
    nv = res.shape[2] - 1
    l2, linf = np.zeros(nv), np.zeros(nv)
    lib().orc_norms(ni, nj, nv, _p(np.ascontiguousarray(cells)), _p(np.ascontiguousarray(res)), _p(l2), _p(linf))
    return l2, linf


def gradient(cs, q):
    """Legacy gradient (spatial/gradientCalc.h): [ni][nj][2][nvar] of neighbour differences."""
    ni, nj = q.shape[:2]
    out = np.zeros((ni, nj, 2, nvar(cs)))
    lib().orc_gradient(_ref(cs), ni, nj, _p(np.ascontiguousarray(q, float)), _p(out))
    return out


def smooth_residual(cs, res, mode, alpha):
    """Residual smoothing of res [ni][nj][nvar+1] (flux part): mode 0 explicit, 1 implicit, 2 both."""
    out = np.array(res, dtype=float, copy=True, order="C")
    ni, nj = out.shape[:2]
    lib().orc_smooth_residual(_ref(cs), ni, nj, int(mode), float(alpha), _p(out))
    return out


def set_threads(n, fast=False):
    lib(fast).orc_set_threads(int(n))


def get_threads(fast=False):
    return lib(fast).orc_get_threads()
