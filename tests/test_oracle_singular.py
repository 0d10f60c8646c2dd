"""Pin the oracle's singular (Duffy) integration against an independent route.

Independent route: the closed-form potential of a uniform flat triangle, int_T 1/|x-y| dS_y
(Wilton et al. 1984 / Graglia 1993 edge formula), integrated over the test triangle with a
subdivided high-order rule. This fixes the k -> 0 single-layer blocks S (and through them the
curl term of H) of coincident, edge- and vertex-adjacent pairs. The closed form itself is checked
against brute-force quadrature at a distant point.
"""

import numpy as np
import pytest

from oracle import bem_oracle as O


def triangle_potential(x, tri):
    """int_T 1/|x - y| dS_y for points x (n,3) and triangle tri (3,3)."""
    v0, v1, v2 = tri
    n = np.cross(v1 - v0, v2 - v0)
    n /= np.linalg.norm(n)
    d = (x - v0) @ n  # signed height
    rho = x - d[:, None] * n[None, :]
    ad = np.abs(d)
    total = np.zeros(len(x))
    for a, b in ((v0, v1), (v1, v2), (v2, v0)):
        L = b - a
        lh = L / np.linalg.norm(L)
        uh = np.cross(lh, n)  # outward in-plane edge normal for a counter-clockwise triangle
        lp = (b - rho) @ lh
        lm = (a - rho) @ lh
        p0 = (a - rho) @ uh
        rp = np.linalg.norm(x - b, axis=1)
        rm = np.linalg.norm(x - a, axis=1)
        r0sq = p0 * p0 + d * d
        with np.errstate(divide="ignore", invalid="ignore"):
            log_term = np.where(np.abs(p0) > 1e-14, p0 * np.log((rp + lp) / (rm + lm)), 0.0)
            at = np.arctan2(p0 * lp, r0sq + ad * rp) - np.arctan2(p0 * lm, r0sq + ad * rm)
        total += log_term - ad * at
    return total


def _subdivided_rule(tri, levels):
    pts, w = O.triangle_rule(4)
    tris = [np.asarray(tri, float)]
    for _ in range(levels):
        new = []
        for a, b, c in tris:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            new += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = new
    X, W, B = [], [], []
    tri = np.asarray(tri, float)
    e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
    M = np.stack([e1, e2], axis=1)
    for t in tris:
        area2 = np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))
        x = (1 - pts[:, :1] - pts[:, 1:]) * t[0] + pts[:, :1] * t[1] + pts[:, 1:] * t[2]
        # barycentrics of x in the parent triangle -> P1 basis values
        uv = np.linalg.lstsq(M, (x - tri[0]).T, rcond=None)[0].T
        B.append(np.stack([1 - uv[:, 0] - uv[:, 1], uv[:, 0], uv[:, 1]], axis=1))
        X.append(x)
        W.append(w * area2)
    return np.concatenate(X), np.concatenate(W), np.concatenate(B)


def test_closed_form_potential_far_point():
    tri = np.array([[0.0, 0, 0], [1.0, 0.2, 0], [0.3, 0.9, 0.1]])
    x = np.array([[0.4, 0.3, 2.0], [-1.0, 0.5, 0.7]])
    Xs, Ws, _ = _subdivided_rule(tri, 4)
    brute = np.array([np.sum(Ws / np.linalg.norm(Xs - xi, axis=1)) for xi in x])
    assert np.allclose(triangle_potential(x, tri), brute, rtol=1e-9)


def _pair_mesh(kind):
    if kind == "coincident":
        V = np.array([[0.0, 0, 0], [0.01, 0, 0], [0.003, 0.008, 0]])
        F = np.array([[0, 1, 2]])
        return V, F, (0, 0)
    if kind == "edge":
        V = np.array([[0.0, 0, 0], [0.01, 0, 0], [0.003, 0.008, 0], [0.008, -0.007, 0.004]])
        F = np.array([[0, 1, 2], [1, 0, 3]])
        return V, F, (0, 1)
    V = np.array([[0.0, 0, 0], [0.01, 0, 0], [0.003, 0.008, 0], [-0.009, -0.002, 0.003], [-0.004, -0.009, -0.001]])
    F = np.array([[0, 1, 2], [0, 3, 4]])
    return V, F, (0, 1)


@pytest.mark.parametrize("kind", ["coincident", "edge", "vertex"])
def test_singular_single_layer_vs_closed_form(kind):
    V, F, (t, s) = _pair_mesh(kind)
    mesh = O.Mesh(V, F, np.ones(len(F), int), np.ones(len(F), int))
    cache = O.build_singular_cache(mesh, 4)
    sel = [i for i in range(len(cache.test)) if cache.test[i] == t and cache.trial[i] == s][0]
    one = O.SingularCache(cache.test[[sel]], cache.trial[[sel]], cache.rule[[sel]], cache.rules, cache.keys)
    Sb, Dsb, Db, Hb = O.singular_pair_blocks(mesh, 1e-9, one)
    Xs, Ws, Bs = _subdivided_rule(mesh.face_vertices[t], 6)
    pot = triangle_potential(Xs, mesh.face_vertices[s]) / (4 * np.pi)
    ref = (Bs * (Ws * pot)[:, None]).sum(0)
    err = np.abs(Sb[0].real - ref) / np.abs(ref)
    assert err.max() < 5e-5, (kind, err)
    # k -> 0: hypersingular reduces to curl_t . curl_s * int int G
    curls = O.surface_curls(mesh.face_vertices, mesh.normals)
    cp = curls[t] @ curls[s].T
    h_ref = cp * ref.sum()
    assert np.abs(Hb[0].real - h_ref).max() / np.abs(h_ref).max() < 5e-5
