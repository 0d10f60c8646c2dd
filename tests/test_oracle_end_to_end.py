"""End-to-end pins of the FP64 oracle: analytic pulsating sphere, LAPACK residual, direct field sum."""

import math

import numpy as np
import pytest

from beat_hip import synthetic as syn
from oracle import bem_oracle as O


@pytest.fixture(scope="module")
def sphere320():
    V, F, T = syn.icosphere(2, 0.1)
    return O.Mesh(V, F, T, np.ones(len(F), int))


def test_pulsating_sphere_analytic(sphere320):
    f, c, rho, a = 1000.0, 343.0, 1.205, 0.1
    pts = O.fibonacci_observation({"spherical_sampling_enabled": True, "spherical_sampling_points": 200})[0]
    out = O.solve_frequency(sphere320, f, np.ones((sphere320.n_faces, 1)), c, rho, points=pts.astype(float))
    k, om = 2 * math.pi * f / c, 2 * math.pi * f
    pa = O.pulsating_sphere_pressure(np.linalg.norm(pts, axis=1), a, k, rho, om)
    u = out["field"][:, 0]
    assert np.abs(20 * np.log10(np.abs(u) / np.abs(pa))).max() < 0.1
    assert np.degrees(np.abs(np.angle(u / pa))).max() < 1.0
    # surface pressure vs p(a) = i rho omega v a / (ika - 1)
    p_surf = 1j * rho * om * a / (1j * k * a - 1)
    assert np.abs(out["pressure"][:, 0] / p_surf - 1).max() < 0.01
    # LAPACK residual of the dense solve
    res = np.linalg.norm(out["lhs"] @ out["pressure"] - out["rhs"]) / np.linalg.norm(out["rhs"])
    assert res < 1e-12


def test_pulsating_sphere_converges_with_refinement():
    # the discretisation error must fall with mesh refinement (a wrong sign/term would plateau)
    f, c, rho, a = 800.0, 343.0, 1.21, 0.1
    k, om = 2 * math.pi * f / c, 2 * math.pi * f
    errs = []
    for n in (1, 2):
        V, F, T = syn.icosphere(n, a)
        m = O.Mesh(V, F, T, np.ones(len(F), int))
        pts = np.array([[0.0, 0.0, 2.0], [1.0, 1.0, 1.0]])
        out = O.solve_frequency(m, f, np.ones((len(F), 1)), c, rho, points=pts)
        pa = O.pulsating_sphere_pressure(np.linalg.norm(pts, axis=1), a, k, rho, om)
        errs.append(np.abs(out["field"][:, 0] / pa - 1).max())
    assert errs[1] < 0.5 * errs[0]


def test_field_matches_direct_double_loop():
    V, F, T = syn.icosphere(0, 0.1)
    m = O.Mesh(V, F, T, np.ones(len(F), int))
    rng = np.random.default_rng(1)
    p = rng.standard_normal(len(V)) + 1j * rng.standard_normal(len(V))
    q = rng.standard_normal(len(F)) + 1j * rng.standard_normal(len(F))
    k = 5.0
    x = np.array([[0.3, -0.1, 0.5], [0.0, 0.0, 2.0]])
    got = O.evaluate_field(x, m, p, q, k)
    pts, w = O.triangle_rule(4)
    ref = np.zeros(len(x), complex)
    for oi, xo in enumerate(x):
        for e in range(len(F)):
            v = m.face_vertices[e]
            for qi in range(6):
                xi, eta = pts[qi]
                phi = np.array([1 - xi - eta, xi, eta])
                y = phi @ v
                r = np.linalg.norm(y - xo)
                G = np.exp(1j * k * r) / (4 * np.pi * r)
                dG = G * (1j * k - 1 / r) * np.dot(y - xo, m.normals[e]) / r
                ww = w[qi] * 2 * m.areas[e]
                ref[oi] += dG * (phi @ p[F[e]]) * ww - G * q[e] * ww
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


def test_regular_operator_symmetries(sphere320):
    # H (hypersingular) is complex symmetric: exactly for the regular part (same tensor rule on both sides),
    # to Duffy-quadrature accuracy once the (non-symmetric) singular maps are added
    k = 10.0
    H0 = O.assemble_operators(sphere320, k, with_singular=False).hypersingular
    assert np.abs(H0 - H0.T).max() / np.abs(H0).max() < 1e-14
    ops = O.assemble_operators(sphere320, k)
    H = ops.hypersingular
    assert np.abs(H - H.T).max() / np.abs(H).max() < 1e-6
    # rigid translation invariance of the assembled LHS
    V2 = sphere320.vertices + np.array([0.3, -1.0, 2.0])
    m2 = O.Mesh(V2, sphere320.faces, sphere320.tags, sphere320.mesh_ids)
    ops2 = O.assemble_operators(m2, k)
    assert np.abs(ops2.double_layer - ops.double_layer).max() / np.abs(ops.double_layer).max() < 1e-9
