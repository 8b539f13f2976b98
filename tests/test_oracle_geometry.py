"""Pins of the oracle's geometry, mesh and Runge-Kutta tables by closed forms.

geom::surface (geometry/surface.ipp:12-29), geom::volume (geometry/volume.ipp:25-66),
dual (mesh/dual.ipp:42-83), make_linspace_mesh (mesh/generate/twoD.h:15-40),
the stretched O-grid, and ODE::Explicit tables (ode.h:107-197) + classical RK4.
"""
import math

import numpy as np
import pytest

import pyoracle as po


def test_surface_area_centre_and_right_handed_normal():
    f = po.surface([0.0, 0.0], [0.0, 2.0])
    assert f.area == 2.0
    assert (f.cx, f.cy) == (0.0, 1.0)
    # normal = orthog(tangent1/l): a face traversed along +y has its normal along +x
    assert (f.nx, f.ny) == (1.0, 0.0)
    assert (f.tx, f.ty) == (0.0, 1.0)
    rng = np.random.default_rng(0)
    for _ in range(50):
        p0, p1 = rng.standard_normal(2), rng.standard_normal(2)
        f = po.surface(p0, p1)
        t = p1 - p0
        assert math.isclose(f.area, np.linalg.norm(t), rel_tol=1e-15)
        assert np.allclose([f.cx, f.cy], 0.5 * (p0 + p1), rtol=0, atol=1e-15)
        assert math.isclose(f.nx ** 2 + f.ny ** 2, 1.0, rel_tol=1e-15)
        assert abs(f.nx * t[0] + f.ny * t[1]) < 1e-14 * f.area
        # metric[1] = (-n1, n0) exactly, right-handed (cross(n, t) = +1)
        assert (f.tx, f.ty) == (-f.ny, f.nx)


def test_volume_is_quarter_area_and_vertex_centroid():
    # unit square: |J(0,0)| of the bilinear map on [-1,1]^2 is area/4 (reference quirk)
    v = po.volume([0, 0], [1, 0], [0, 1], [1, 1])
    assert v[0] == 0.25
    assert (v[1], v[2]) == (0.5, 0.5)
    rng = np.random.default_rng(1)
    for _ in range(50):
        # random convex quad with corners ordered 0:(-,-) 1:(+,-) 2:(-,+) 3:(+,+)
        p = np.array([[-1, -1], [1, -1], [-1, 1], [1, 1]], float) + 0.3 * rng.uniform(-1, 1, (4, 2))
        v = po.volume(*p)
        poly = p[[0, 1, 3, 2]]
        shoelace = 0.5 * abs(np.dot(poly[:, 0], np.roll(poly[:, 1], -1)) - np.dot(poly[:, 1], np.roll(poly[:, 0], -1)))
        assert math.isclose(4 * v[0], shoelace, rel_tol=1e-14)
        assert np.allclose(v[1:], p.mean(axis=0), rtol=0, atol=1e-15)


def test_linspace_mesh_nodes_and_dual():
    nodes = po.mesh_linspace(8, 5, -1.0, 3.0, 2.0, 4.5)
    i = np.arange(9)[:, None]
    j = np.arange(6)[None, :]
    assert np.array_equal(nodes[..., 0], -1.0 + i * (4.0 / 8) + 0 * j)
    assert np.array_equal(nodes[..., 1], 2.0 + j * (2.5 / 5) + 0 * i)
    cells = po.dual(nodes)
    assert np.allclose(cells[..., 0], 0.25 * 0.5 * 0.5, rtol=1e-15)
    assert math.isclose(4 * cells[..., 0].sum(), 4.0 * 2.5, rel_tol=1e-14)


def test_ogrid_stretching():
    nr, nth = 64, 48
    nodes = po.mesh_ogrid(nr, nth, 0.5, 20.0, 1e-3)
    r = np.hypot(nodes[..., 0], nodes[..., 1])
    assert math.isclose(r[1, 0] - r[0, 0], 1e-3, rel_tol=1e-12)
    assert r[0, 0] == 0.5 and r[-1, 0] == 20.0
    s = po.ogrid_ratio(nr, 0.5, 20.0, 1e-3)
    dr = np.diff(r[:, 0])
    assert np.allclose(dr[1:] / dr[:-1], s, rtol=1e-9)
    assert np.array_equal(nodes[:, nth], nodes[:, 0])   # node nth identical to node 0
    cells = po.dual(nodes)
    assert (cells[..., 0] > 0).all()
    # polygonal annulus area: nth trapezoids per ring
    ring = 0.5 * nth * math.sin(2 * math.pi / nth) * (r[1:, 0] ** 2 - r[:-1, 0] ** 2)
    assert np.allclose(4 * cells[..., 0].sum(axis=1), ring, rtol=1e-12)


def test_cylinder_mesh_reference_generator():
    nodes = po.mesh_cylinder(4, 8, 1.0, 8.0)
    r = np.hypot(nodes[..., 0], nodes[..., 1])
    assert np.allclose(r, (1.0 + np.arange(5) * 7.0 / 4)[:, None], rtol=1e-15)


def _butcher(name):
    n, a, b = po.rk_table(name)
    # q_{s+1} = q0 + beta_s dt sum_k alpha[s][k] R_k  ->  Butcher A[s+1][k] = beta_s alpha[s][k]
    A = np.zeros((n, n))
    for s in range(n - 1):
        A[s + 1, : s + 1] = b[s] * a[s, : s + 1]
    w = b[n - 1] * a[n - 1, :n]
    c = A.sum(axis=1)
    return n, A, w, c


@pytest.mark.parametrize("name,order", [(11, 1), (22, 2), (33, 3), (34, 3), (44, 4)])
def test_rk_tables_order_conditions(name, order):
    n, A, w, c = _butcher(name)
    assert math.isclose(w.sum(), 1.0, rel_tol=1e-15)
    if order >= 2:
        assert math.isclose(w @ c, 0.5, rel_tol=1e-14)
    if order >= 3:
        assert math.isclose(w @ c ** 2, 1 / 3, rel_tol=1e-14)
        assert math.isclose(w @ A @ c, 1 / 6, rel_tol=1e-14)
    if order >= 4:
        assert math.isclose(w @ c ** 3, 1 / 4, rel_tol=1e-14)
        assert math.isclose(w @ (c * (A @ c)), 1 / 8, rel_tol=1e-14)
        assert math.isclose(w @ A @ c ** 2, 1 / 12, rel_tol=1e-14)
        assert math.isclose(w @ A @ A @ c, 1 / 24, rel_tol=1e-14)


def test_rk_tables_literal_values():
    # ode.h:143-197 literals
    n, a, b = po.rk_table(33)
    assert n == 3 and list(b[:3]) == [1.0, 0.5, 1.0]
    assert list(a[2, :3]) == [1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0]
    n, a, b = po.rk_table(34)
    assert n == 4 and list(b[:4]) == [0.5, 1.0, 0.5, 1.0]
    assert list(a[3, :4]) == [1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.5]


def test_product_mesh_generators_agree_with_oracle():
    from hyperboliconvergence_amd import cases
    assert np.array_equal(cases.mesh_linspace(16, 12, -10.0, 10.0, -3.0, 7.0),
                          po.mesh_linspace(16, 12, -10.0, 10.0, -3.0, 7.0))
    a = cases.mesh_ogrid(32, 40)
    b = po.mesh_ogrid(32, 40, 0.5, 20.0, 1e-3)
    assert np.allclose(a, b, rtol=0, atol=1e-14)
    assert cases.ogrid_ratio(256) == po.ogrid_ratio(256)
