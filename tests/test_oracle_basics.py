"""Pins the oracle's quadrature rules and bases against closed forms and the
reference's own known-answer tests (element/src/quadrature.rs:429-473,
element/src/lagrange/hex.rs:72-113)."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_gauss_legendre_exact(fr, n):
    x, w = fr.gauss_legendre(n)
    assert abs(w.sum() - 2.0) < 1e-14
    for k in range(2 * n):  # exact up to degree 2n-1
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-14, k
    # not exact at degree 2n (guards against a rule of the wrong order)
    assert abs(np.dot(w, x ** (2 * n)) - 2.0 / (2 * n + 1)) > 1e-6
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, xr, atol=1e-15, rtol=0) and np.allclose(w, wr, atol=1e-15, rtol=0)


def test_gl5_closed_form(fr):
    x, w = fr.gauss_legendre(5)
    a = math.sqrt(5 - 2 * math.sqrt(10 / 7)) / 3
    b = math.sqrt(5 + 2 * math.sqrt(10 / 7)) / 3
    assert np.allclose(x, [-b, -a, 0, a, b], atol=1e-16)
    assert np.allclose(w, [(322 - 13 * math.sqrt(70)) / 900, (322 + 13 * math.sqrt(70)) / 900,
                           128 / 225, (322 + 13 * math.sqrt(70)) / 900,
                           (322 - 13 * math.sqrt(70)) / 900], atol=1e-16)


def _tri_monomial(a, b):
    # integral of x^a y^b over the reference triangle
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


def _tet_monomial(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


@pytest.mark.parametrize("order,deg", [(1, 1), (3, 2), (5, 5), (6, 6)])
def test_tri_rule_exact(fr, order, deg):
    p, w = fr.tri_rule(order)
    # the reference's 12-point constants sum to 0.5 only within ~6e-12
    # (quadrature.rs:444-448 tests 1e-10 for orders >= 6)
    tol = 1e-14 if order < 6 else 1e-10
    assert abs(w.sum() - 0.5) < tol
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            got = np.dot(w, p[:, 0] ** a * p[:, 1] ** b)
            assert abs(got - _tri_monomial(a, b)) < max(tol, 1e-13), (a, b)


@pytest.mark.parametrize("order,deg,npts", [(1, 1, 1), (2, 2, 4), (5, 5, 15), (7, 7, 35)])
def test_tet_rule_exact(fr, order, deg, npts):
    p, w = fr.tet_rule(order)
    assert len(w) == npts
    assert abs(w.sum() - 1.0 / 6.0) < 1e-14
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                got = np.dot(w, p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
                assert abs(got - _tet_monomial(a, b, c)) < 1e-13, (a, b, c)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_lagrange_1d(fr, p):
    nodes = -1 + 2 * np.arange(p + 1) / p
    B, G = fr.lagrange_1d(p, nodes)
    assert np.allclose(B, np.eye(p + 1), atol=1e-14)  # nodal
    xs = np.linspace(-1, 1, 13)
    B, G = fr.lagrange_1d(p, xs)
    assert np.allclose(B.sum(1), 1.0, atol=1e-13)  # partition of unity
    assert np.allclose(G.sum(1), 0.0, atol=1e-12)
    # reproduces polynomials of degree <= p and their derivatives
    for k in range(p + 1):
        assert np.allclose(B @ nodes ** k, xs ** k, atol=1e-12)
        dk = k * xs ** (k - 1) if k else 0 * xs
        assert np.allclose(G @ nodes ** k, dk, atol=1e-11)


def _simplex_nodes(dim, order):
    """Reference-element DOF coordinates in the TriP*/TetP* local order."""
    verts = np.vstack([np.zeros(dim), np.eye(dim)])
    pts = list(verts)
    if order >= 2:
        edges = [(0, 1), (1, 2), (0, 2)] if dim == 2 else [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
        for a, b in edges:
            if order == 2:
                pts.append((verts[a] + verts[b]) / 2)
            else:
                pts.append((2 * verts[a] + verts[b]) / 3)
                pts.append((verts[a] + 2 * verts[b]) / 3)
    if order == 3:
        faces = [(0, 1, 2)] if dim == 2 else [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
        for f in faces:
            pts.append(sum(verts[i] for i in f) / 3)
    return np.array(pts)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_simplex_basis(fr, dim, order):
    nodes = _simplex_nodes(dim, order)
    for i, xi in enumerate(nodes):
        phi, _ = fr.simplex_basis(dim, order, xi)
        e = np.zeros(len(nodes)); e[i] = 1
        assert np.allclose(phi, e, atol=1e-13), (i, phi)
    rng = np.random.default_rng(1)
    for _ in range(5):
        xi = rng.random(dim); xi /= (xi.sum() + rng.random())
        phi, grad = fr.simplex_basis(dim, order, xi)
        assert abs(phi.sum() - 1) < 1e-13
        assert np.allclose(grad.sum(0), 0, atol=1e-12)
        # gradient matches a central finite difference
        h = 1e-6
        for k in range(dim):
            d = np.zeros(dim); d[k] = h
            fd = (fr.simplex_basis(dim, order, xi + d)[0] - fr.simplex_basis(dim, order, xi - d)[0]) / (2 * h)
            assert np.allclose(fd, grad[:, k], atol=1e-7)


def test_tri_p2_matches_reference_formula(fr):
    # lagrange/tri.rs:62-73 written out directly
    x, y = 0.2, 0.3
    l1, l2, l3 = 1 - x - y, x, y
    ref = [l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), l3 * (2 * l3 - 1), 4 * l1 * l2, 4 * l2 * l3, 4 * l1 * l3]
    phi, grad = fr.simplex_basis(2, 2, np.array([x, y]))
    assert np.allclose(phi, ref, atol=1e-15)
    assert np.allclose(grad[3], [4 * (1 - 2 * x - y), -4 * x], atol=1e-15)


def test_philox_known_answer(fr):
    # Philox4x32-10 known-answer vectors (Random123 kat_vectors: ctr = 0, key = 0
    # and ctr = key = all ones) checked through the uniform helper's first word.
    import ctypes as C
    L = fr.lib()
    f = L.fr_philox4x32_10
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = None
    out = (C.c_uint32 * 4)()
    ctr = (C.c_uint32 * 4)(0, 0, 0, 0); key = (C.c_uint32 * 2)(0, 0)
    f(ctr, key, out)
    assert list(out) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    ctr = (C.c_uint32 * 4)(*[0xffffffff] * 4); key = (C.c_uint32 * 2)(0xffffffff, 0xffffffff)
    f(ctr, key, out)
    assert list(out) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
