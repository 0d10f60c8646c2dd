"""Oracle pins for x / xy plane symmetry.

The symmetric formulation (image trial elements, geometric image-singular pairs with Duffy rules, P1 orbit row
weights, image field sources) applied to a fundamental domain must reproduce the plain solve of the full
mirror-symmetric model. The two discretisations coincide except for the Duffy orientation of pairs that touch
the symmetry plane (the full model's mirrored elements have the opposite local vertex order), so agreement is
at the level of the Duffy quadrature error (~1e-5), while a wrong sign or a missing image term shows up at 1e-1.
"""

import numpy as np
import pytest

from beat_hip import synthetic as syn
from oracle import bem_oracle as O

FREQ, C0, RHO = 1200.0, 343.0, 1.205
PTS = np.array([[0.3, 0.2, 1.5], [1.0, 0.5, -0.3], [0.2, 1.1, 0.4], [-0.7, -0.4, 0.9]])


@pytest.fixture(scope="module")
def full_solution():
    V, F, T = syn.uv_sphere(10, 20, 0.1)
    m = O.Mesh(V, F, T, np.ones(len(F), int))
    return V, F, T, O.solve_frequency(m, FREQ, np.ones((len(F), 1)), C0, RHO, points=PTS)


def test_uv_sphere_is_mirror_symmetric():
    V, F, T = syn.uv_sphere(6, 12)
    assert syn.check_closed_oriented(F) and syn.signed_volume(V, F) > 0
    ref = {tuple(sorted(tuple(np.round(V[i], 12)) for i in f)) for f in F}
    for sg in ([-1, 1, 1], [1, -1, 1]):
        W = V * np.array(sg)
        assert {tuple(sorted(tuple(np.round(W[i], 12)) for i in f)) for f in F} == ref


@pytest.mark.parametrize("mode", ["x", "xy"])
def test_symmetric_solve_reproduces_full_model(full_solution, mode):
    V, F, T, full = full_solution
    v, f, t = syn.symmetric_part(V, F, T, mode)
    m = O.Mesh(v, f, t, np.ones(len(f), int))
    out = O.solve_frequency_symmetric(m, FREQ, np.ones((len(f), 1)), mode, C0, RHO, points=PTS)
    idx = [int(np.argmin(np.linalg.norm(V - p, axis=1))) for p in v]
    pf, ph = full["pressure"][idx, 0], out["pressure"][:, 0]
    assert np.linalg.norm(ph - pf) / np.linalg.norm(pf) < 5e-5
    assert np.linalg.norm(out["field"] - full["field"]) / np.linalg.norm(full["field"]) < 1e-5
    # the pulsating sphere closed form holds on the reduced model as well
    k = 2 * np.pi * FREQ / C0
    pa = O.pulsating_sphere_pressure(np.linalg.norm(PTS, axis=1), 0.1, k, RHO, 2 * np.pi * FREQ)
    assert np.abs(20 * np.log10(np.abs(out["field"][:, 0]) / np.abs(pa))).max() < 0.2


def test_image_singular_pairs_are_the_plane_crossing_pairs_of_the_full_model():
    """Image pairs (t, s) with R s touching t = vertex-sharing pairs (t, s') of the full model whose trial s'
    lies in the mirrored half (same kinds)."""
    V, F, T = syn.uv_sphere(6, 12)
    v, f, t = syn.symmetric_part(V, F, T, "x")
    m = O.Mesh(v, f, t, np.ones(len(f), int))
    img = O.image_singular_cache(m, 4, np.array([-1.0, 1.0, 1.0]))
    vid = [int(np.argmin(np.linalg.norm(V - p, axis=1))) for p in v]
    ff = np.asarray(vid)[f]  # half faces in full vertex ids
    cen = V[F].mean(1)
    mirror_faces = F[cen[:, 0] < 0]
    expect = {}
    for ti, ft in enumerate(ff):
        for fs in mirror_faces:
            kind, _, _ = O.adjacency_info(ft, fs)
            if kind != "regular":
                # the mirrored element is the image of the half face with the reflected centroid
                c = V[fs].mean(0) * np.array([-1.0, 1.0, 1.0])
                si = int(np.argmin(np.linalg.norm(m.face_vertices.mean(1) - c, axis=1)))
                expect[(ti, si)] = kind
    got = {(int(a), int(b)): img.keys[r][0] for a, b, r in zip(img.test, img.trial, img.rule)}
    assert set(got) == set(expect)
    code = {"coincident": 1, "edge": 2, "vertex": 3}
    assert all(got[p] == code[expect[p]] for p in got)


def test_snap_validate_and_orbit_weights():
    v = np.array([[1e-12, 0.5, 0.1], [0.2, -3e-13, 0.0], [0.3, 0.4, 0.5], [0.0, 0.0, 1.0]])
    tol = O.symmetry_plane_tolerance(v)
    assert tol == max(1e-9, np.linalg.norm(v.max(0) - v.min(0)) * 1e-6)
    s = O.snap_symmetry_planes(v, "xy")
    assert s[0, 0] == 0.0 and s[1, 1] == 0.0 and s[2, 0] == 0.3
    assert np.array_equal(O.p1_orbit_weights(s, "x"), [2, 1, 1, 2])
    assert np.array_equal(O.p1_orbit_weights(s, "xy"), [2, 2, 1, 4])
    O.validate_fundamental_domain(s, "xy")
    with pytest.raises(ValueError, match="positive Y fundamental domain"):
        O.validate_fundamental_domain(np.array([[0.1, -0.2, 0.0], [0.2, 0.3, 0.1]]), "xy")
    assert [t[0] for t in O.symmetry_transforms("xy")] == ["identity", "x", "y", "xy"]
    assert [t[2] for t in O.symmetry_transforms("xy")] == [1.0, -1.0, -1.0, 1.0]
