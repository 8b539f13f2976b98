"""ASCII result writers with the columns, number format and block layout of the reference's
drivers, so that their gnuplot scripts (cpp/cyl.gnp and the like) read this package's results.

  write_cylinder_result  script/euler/cylinder2D.cpp:33-63, 192-236   (theta z r + 6 state columns)
  write_mesh             script/euler/cylinder2D.cpp:238-272          (node x y, a block per row)
  write_gresho_result    script/euler/gresho2D.cpp:34-74, 226-265     (x y + 7 state columns)
  write_shocktube_result script/euler/shocktube.cpp:60-77, 108-133    (5 columns per cell)
  write_scalar_result    script/scalar/cartesian2D.cpp:76-93, 147-170 (x y  u v q)

States are primitive Euler sets {u, v, rho, p} (or scalar {u, v, q}) laid out [ni][nj][nv] as
the Context returns them; every quantity is formed on the host from those values.
"""
import math

import numpy as np

from .cases import cell_centres


def _state(q, gamma, R):
    """p, rho, a^2, u, v and |u|^2 of primitive sets (euler/transforms/primitive.ipp)."""
    u, v, r, p = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    t = p / (r * R)
    a2 = gamma * R * t
    return p, r, a2, u, v, u * u + v * v


def _entropy(a2, r, gamma):
    return a2 / (gamma * np.power(r, gamma - 1.0))


def _fmt(x, digits):
    return f"{x:.{digits}e}"


def write_cylinder_result(path, nodes, q, qref, gamma=1.4, R=287.058):
    """Cell by cell, rings (i) outer and theta (j) inner, a blank line after each ring:
    theta z r  (rho-rho0)/rho0  u/|u0|  v/|u0|  (p-p0)/p0  (s-s0)/s0  a/a0  with
    s = a^2 / (gamma rho^(gamma-1)); theta = (j + 1/2) 2 pi / n_theta, z = 0 and r the distance
    of the cell centre from the axis (the reference's uniform radial spacing does not apply to
    the stretched O-grid)."""
    ni, nj = q.shape[:2]
    p, r, a2, u, v, _ = _state(q, gamma, R)
    p0, r0, a20, u0, v0, vv0 = _state(np.asarray(qref, float), gamma, R)
    s, s0 = _entropy(a2, r, gamma), _entropy(a20, r0, gamma)
    speed0, sonic0 = math.sqrt(vv0), math.sqrt(a20)
    c = cell_centres(nodes)
    rad = np.sqrt(c[..., 0] ** 2 + c[..., 1] ** 2)
    dth = 2.0 * math.pi / nj
    with open(path, "w") as f:
        for i in range(ni):
            for j in range(nj):
                cols = [(j + 0.5) * dth, 0.0, rad[i, j], (r[i, j] - r0) / r0, u[i, j] / speed0, v[i, j] / speed0,
                        (p[i, j] - p0) / p0, (s[i, j] - s0) / s0, math.sqrt(a2[i, j]) / sonic0]
                f.write(" ".join(_fmt(x, 8) for x in cols) + "\n")
                if j == nj - 1:
                    f.write("\n")


def write_mesh(path, nodes):
    """Nodes row by row (x y per line), a blank line after each row of n_j + 1 nodes."""
    with open(path, "w") as f:
        for i in range(nodes.shape[0]):
            for j in range(nodes.shape[1]):
                f.write(f"{_fmt(nodes[i, j, 0], 8)} {_fmt(nodes[i, j, 1], 8)} \n")
            f.write("\n")


def write_gresho_result(path, nodes, q, qref, gamma=1.4, R=287.058):
    """Cell centre then rho/rho0  u  v  p-p0  s/s0  a/a0  |u|/a (the reference's velocity scale is
    1), 16 digits, a blank line after each row of cells (the last column index)."""
    ni, nj = q.shape[:2]
    p, r, a2, u, v, vv = _state(q, gamma, R)
    p0, r0, a20, _, _, _ = _state(np.asarray(qref, float), gamma, R)
    s, s0 = _entropy(a2, r, gamma), _entropy(a20, r0, gamma)
    a, a0 = np.sqrt(a2), math.sqrt(a20)
    c = cell_centres(nodes)
    with open(path, "w") as f:
        for i in range(ni):
            for j in range(nj):
                cols = [c[i, j, 0], c[i, j, 1], r[i, j] / r0, u[i, j], v[i, j], p[i, j] - p0, s[i, j] / s0, a[i, j] / a0,
                        math.sqrt(vv[i, j]) / a[i, j]]
                f.write(" ".join(_fmt(x, 16) for x in cols) + "\n")
                if j == nj - 1:
                    f.write("\n")


def write_shocktube_result(path, q, gamma=1.4, R=287.058):
    """One line per cell along the tube (the first cell of each strip row):
    rho  u/a  u  p  s."""
    p, r, a2, u, _, _ = _state(q[:, 0], gamma, R)
    a = np.sqrt(a2)
    s = _entropy(a2, r, gamma)
    with open(path, "w") as f:
        for k in range(q.shape[0]):
            f.write(" ".join(_fmt(x, 8) for x in (r[k], u[k] / a[k], u[k], p[k], s[k])) + "\n")


def write_scalar_result(path, nodes, q):
    """x y  u v q per cell (two spaces after y), a blank line after each row i."""
    c = cell_centres(nodes)
    ni, nj = q.shape[:2]
    with open(path, "w") as f:
        for i in range(ni):
            for j in range(nj):
                f.write(f"{_fmt(c[i, j, 0], 8)} {_fmt(c[i, j, 1], 8)}  {_fmt(q[i, j, 0], 8)} {_fmt(q[i, j, 1], 8)} "
                        f"{_fmt(q[i, j, 2], 8)}\n")
            f.write("\n")
