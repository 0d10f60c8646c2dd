"""Pins for the oracle's quadrature, Duffy rules, curls and mass matrices (closed forms, not restatements)."""

import math

import numpy as np
import pytest

from oracle import bem_oracle as O


def _tri_monomial(a, b):
    # int_T x^a y^b over the unit reference triangle = a! b! / (a + b + 2)!
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


def test_triangle_rule_order4_exactness():
    pts, w = O.triangle_rule(4)
    assert len(w) == 6
    assert abs(w.sum() - 0.5) < 1e-15
    for a in range(5):
        for b in range(5 - a):
            got = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b)
            assert abs(got - _tri_monomial(a, b)) < 2e-15, (a, b)
    # degree 5 is NOT integrated exactly (guards against a silently different, higher-order rule)
    assert abs(np.sum(w * pts[:, 0] ** 5) - _tri_monomial(5, 0)) > 1e-6


def test_low_order_rules():
    pts, w = O.triangle_rule(1)
    assert w.sum() == 0.5 and np.allclose(pts, [[1 / 3, 1 / 3]])
    pts, w = O.triangle_rule(2)
    for a, b in [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]:
        assert abs(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b) - _tri_monomial(a, b)) < 1e-15


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_gauss01_exactness(order):
    x, w = O.gauss_rule_1d(order)
    for k in range(2 * order):
        assert abs(np.sum(w * x**k) - 1.0 / (k + 1)) < 1e-15
    assert abs(np.sum(w * x ** (2 * order)) - 1.0 / (2 * order + 1)) > 1e-8


@pytest.mark.parametrize("kind,count", [("coincident", 1536), ("edge", 1280), ("vertex", 512)])
def test_duffy_rules_measure_and_smooth_integrands(kind, count):
    tp, sp, w = O.duffy_rule(4, kind)
    assert len(w) == count
    assert abs(w.sum() - 0.25) < 1e-14  # |T x T| = 1/4
    assert np.all(tp >= -1e-15) and np.all(tp.sum(1) <= 1 + 1e-15)
    assert np.all(sp >= -1e-15) and np.all(sp.sum(1) <= 1 + 1e-15)

    # smooth separable integrand f(x) g(y): exact = int_T f * int_T g
    def f(p):
        return 1.0 + p[:, 0] + 2.0 * p[:, 1]

    def g(p):
        return 1.0 + 3.0 * p[:, 0] - p[:, 1]

    exact = (0.5 + _tri_monomial(1, 0) + 2 * _tri_monomial(0, 1)) * (0.5 + 3 * _tri_monomial(1, 0) - _tri_monomial(0, 1))
    assert abs(np.sum(w * f(tp) * g(sp)) - exact) < 1e-13
    # non-separable polynomial in |x - y|^2 (the integrand family the singular maps must preserve)
    d2 = (tp[:, 0] - sp[:, 0]) ** 2 + (tp[:, 1] - sp[:, 1]) ** 2
    # reference value from a tensor product of the order-4 triangle rule (exact for degree 2 in each point)
    pts, ww = O.triangle_rule(4)
    ref = sum(ww[i] * ww[j] * ((pts[i, 0] - pts[j, 0]) ** 2 + (pts[i, 1] - pts[j, 1]) ** 2)
              for i in range(6) for j in range(6))
    assert abs(np.sum(w * d2) - ref) < 1e-13


@pytest.mark.parametrize("key", [(2, 0, 1, 0, 1), (2, 1, 2, 0, 2), (2, 2, 0, 1, 2), (3, 0, -1, 2, -1), (3, 2, -1, 1, -1)])
def test_remapped_rules_preserve_measure(key):
    base = O.duffy_rule(4, "edge" if key[0] == 2 else "vertex")
    tp, sp, w = O.remapped_rule(base, key)
    assert abs(w.sum() - 0.25) < 1e-14
    exact = _tri_monomial(1, 0) * _tri_monomial(0, 1)
    assert abs(np.sum(w * tp[:, 0] * sp[:, 1]) - exact) < 1e-13
    # singular locus: the shared vertex/edge of the remap is where test and trial points coincide
    va = key[1]
    corner = O._REF_VERTS[va]
    dmin = np.min(np.linalg.norm(tp - corner, axis=1))
    assert dmin < 0.05


def test_adjacency_info_orientation():
    assert O.adjacency_info((0, 1, 2), (0, 1, 2))[0] == "coincident"
    kind, tv, sv = O.adjacency_info((0, 1, 2), (5, 2, 1))
    # shared vertices 1 and 2: test locals (1,2), trial locals (2,1) -> sorted by trial index
    assert kind == "edge" and sv == (1, 2) and tv == (2, 1)
    kind, tv, sv = O.adjacency_info((7, 8, 9), (3, 4, 8))
    assert kind == "vertex" and tv == (1,) and sv == (2,)
    assert O.adjacency_info((0, 1, 2), (3, 4, 5))[0] == "regular"


def test_curls_closed_form_and_identities():
    fv = np.array([[[0.0, 0, 0], [1, 0, 0], [0, 1, 0]]])
    n = np.array([[0.0, 0, 1]])
    c = O.surface_curls(fv, n)[0]
    assert np.allclose(c, [[1, -1, 0], [0, 1, 0], [-1, 0, 0]], atol=1e-15)
    rng = np.random.default_rng(3)
    fv = rng.standard_normal((20, 3, 3))
    cr = np.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0])
    nn = cr / np.linalg.norm(cr, axis=1)[:, None]
    c = O.surface_curls(fv, nn)
    assert np.abs(c.sum(axis=1)).max() < 1e-10
    assert np.abs(np.einsum("fic,fc->fi", c, nn)).max() < 1e-10
    area2 = np.linalg.norm(cr, axis=1)
    for i in range(3):
        opp = np.linalg.norm(fv[:, (i + 2) % 3] - fv[:, (i + 1) % 3], axis=1)
        assert np.allclose(np.linalg.norm(c[:, i], axis=1), opp / area2, rtol=1e-10)


def test_mass_matrices_closed_form():
    fv = np.array([[0.0, 0, 0], [2, 0, 0], [0, 3, 0], [1, 1, 1]])
    faces = np.array([[0, 1, 2], [1, 3, 2]])
    m = O.Mesh(fv, faces, np.ones(2, int), np.ones(2, int))
    M11, M10 = O.identity_matrices(m)
    exact = np.zeros((4, 4))
    for t in range(2):
        blk = m.areas[t] / 12.0 * (np.ones((3, 3)) + np.eye(3))
        exact[np.ix_(faces[t], faces[t])] += blk
    assert np.abs(M11 - exact).max() < 1e-14
    assert np.allclose(M10[faces[0], 0], m.areas[0] / 3, rtol=1e-14)
    assert np.allclose(M10[faces[1], 1], m.areas[1] / 3, rtol=1e-14)


def test_kernel_conventions():
    # dG/dn_y of G = e^{ikr}/(4 pi r) against a central finite difference
    x = np.array([0.1, -0.2, 0.3])
    y = np.array([0.5, 0.4, -0.1])
    n = np.array([0.2, 0.3, 0.9])
    n /= np.linalg.norm(n)
    k = 7.0

    def G(y_):
        r = np.linalg.norm(y_ - x)
        return np.exp(1j * k * r) / (4 * np.pi * r)

    h = 1e-6
    fd = (G(y + h * n) - G(y - h * n)) / (2 * h)
    rv = y - x
    r = np.linalg.norm(rv)
    an = G(y) * (1j * k - 1 / r) * np.dot(rv, n) / r
    assert abs(fd - an) / abs(an) < 1e-8
