"""ctypes binding of the test-only host emulation of the stage kernels."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None


def make_locked(targets):
    """make the given targets under an exclusive file lock: parallel test workers must not load a
    library that another worker is rewriting."""
    import fcntl
    with open(os.path.join(HERE, ".make.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        return subprocess.run(["make", "-C", HERE, "-s", *targets], capture_output=True, text=True, timeout=900)


def lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "libhostemu.so")
        r = make_locked(["libhostemu.so"])
        assert r.returncode == 0, r.stderr[-2000:]
        _L = C.CDLL(path)
        P, i = C.c_void_p, C.c_int
        _L.emu_residual.argtypes = [P, i, i, P, P, P, i, P, P]
        _L.emu_integrate.argtypes = [P, i, i, P, P, P, i, i, P, P]
        _L.emu_geometry.argtypes = [P, i, i, P, P, P, P, P]
        _L.emu_set_tile_config.argtypes = [i]
        _L.emu_set_uniform_geometry.argtypes = [i]
        _L.emu_tangent_residual.argtypes = [P, i, i, P, P, P, P, P, C.c_double, P, P]
        _L.emu_adjoint_residual.argtypes = [P, i, i, P, P, P, P, P, P, P]
    return _L


def set_tile_config(cfg):
    """Tile shape of the emulated launches (index into the kernel's shape table)."""
    assert lib().emu_set_tile_config(int(cfg)) == 0


def set_uniform_geometry(enable):
    """1: emulate the uniform-geometry kernels on cases that allow them; returns whether the last
    case set up allowed them."""
    return bool(lib().emu_set_uniform_geometry(int(enable)))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def residual(cs, nodes, q, bref, precision=0):
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    nv = 4 if cs.law == 1 else 3
    res = np.zeros((ni, nj, nv + 1))
    st = C.c_uint()
    rc = lib().emu_residual(C.cast(C.pointer(cs), C.c_void_p), ni, nj, _p(np.ascontiguousarray(nodes)),
                            _p(np.ascontiguousarray(q, float)), _p(np.ascontiguousarray(bref, float)), precision,
                            _p(res), C.byref(st))
    assert rc == 0
    return res, st.value


def integrate(cs, nodes, q, bref, nsteps, precision=0):
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    qq = np.ascontiguousarray(q, float).copy()
    t = C.c_double()
    st = C.c_uint()
    rc = lib().emu_integrate(C.cast(C.pointer(cs), C.c_void_p), ni, nj, _p(np.ascontiguousarray(nodes)), _p(qq),
                             _p(np.ascontiguousarray(bref, float)), precision, nsteps, C.byref(t), C.byref(st))
    assert rc == 0
    return qq, t.value, st.value


def geometry(cs, nodes):
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    cells = np.zeros((ni, nj, 3))
    X = np.zeros((ni, nj, 3))
    fi = np.zeros((14, (ni + 1) * nj))
    fj = np.zeros((14, ni * (nj + 1)))
    rc = lib().emu_geometry(C.cast(C.pointer(cs), C.c_void_p), ni, nj, _p(np.ascontiguousarray(nodes)), _p(cells),
                            _p(X), _p(fi), _p(fj))
    assert rc == 0
    return cells, X, fi, fj


def tangent_residual(cs, nodes, q, bref, qd, brefd=None, minfd=0.0):
    """(R(q) [ni][nj][nv+1], Rdot [ni][nj][nv]) of the tangent kernels."""
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    nv = 4 if cs.law == 1 else 3
    res = np.zeros((ni, nj, nv + 1))
    resd = np.zeros((ni, nj, nv))
    bd = None if brefd is None else _p(np.ascontiguousarray(brefd, float))
    rc = lib().emu_tangent_residual(C.cast(C.pointer(cs), C.c_void_p), ni, nj, _p(np.ascontiguousarray(nodes)),
                                    _p(np.ascontiguousarray(q, float)), _p(np.ascontiguousarray(qd, float)),
                                    _p(np.ascontiguousarray(bref, float)), bd, float(minfd), _p(res), _p(resd))
    assert rc == 0, rc
    return res, resd


def adjoint_residual(cs, nodes, q, bref, ybar):
    """(qbar [ni][nj][nv], brefbar [rows][nv], minfbar) of the adjoint kernels."""
    ni, nj = nodes.shape[0] - 1, nodes.shape[1] - 1
    nv = 4 if cs.law == 1 else 3
    qbar = np.zeros((ni, nj, nv))
    brefbar = np.zeros_like(np.ascontiguousarray(bref, float))
    mb = C.c_double()
    rc = lib().emu_adjoint_residual(C.cast(C.pointer(cs), C.c_void_p), ni, nj, _p(np.ascontiguousarray(nodes)),
                                    _p(np.ascontiguousarray(q, float)), _p(np.ascontiguousarray(bref, float)),
                                    _p(np.ascontiguousarray(ybar, float)), _p(qbar), _p(brefbar), C.byref(mb))
    assert rc == 0, rc
    return qbar, brefbar, mb.value

