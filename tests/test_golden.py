"""The oracle (NumPy and C restatements) reproduces the committed golden fixtures."""

from pathlib import Path

import numpy as np
import pytest

from oracle import bem_oracle as O

GOLD = Path(__file__).resolve().parent / "golden"


def test_tetra_pair_operators_numpy_and_c():
    g = np.load(GOLD / "tetra_pair_ops.npz")
    m = O.Mesh(g["vertices"], g["faces"], np.ones(len(g["faces"]), int), np.ones(len(g["faces"]), int))
    ops = O.assemble_operators(m, float(g["k"]))
    for name, key in (("single_layer", "S"), ("double_layer", "D"), ("adjoint_double_layer", "Ds"),
                      ("hypersingular", "H")):
        a = getattr(ops, name)
        assert np.abs(a - g[key]).max() <= 1e-13 * np.abs(g[key]).max()
    from oracle import beat_cpu as B

    c = B.CpuMesh(m, "double").assemble(float(g["k"]))
    for name, key in (("single_layer", "S"), ("double_layer", "D"), ("adjoint_double_layer", "Ds"),
                      ("hypersingular", "H")):
        assert np.abs(getattr(c, name) - g[key]).max() <= 1e-12 * np.abs(g[key]).max()


def test_sphere_fixture_rows():
    g = np.load(GOLD / "sphere1280_1khz.npz")
    m = O.Mesh(g["vertices"], g["faces"], np.ones(len(g["faces"]), int), np.ones(len(g["faces"]), int))
    k = float(g["k"])
    ops = O.assemble_operators(m, k)
    M11, M10 = O.identity_matrices(m)
    lhs, rhs_op = O.burton_miller(ops, M11, M10, k)
    assert np.abs(lhs[:8] - g["lhs_rows"]).max() <= 1e-12 * np.abs(g["lhs_rows"]).max()
    # solution satisfies the fixture system (rows) and the fixture field is consistent with the pressure
    q = 1j * 1.205 * 2 * np.pi * 1000.0 * np.ones(len(g["faces"]))
    u = O.evaluate_field(g["points"], m, g["pressure"], q, k)
    assert np.allclose(u, g["field"], rtol=1e-12)


def test_c_fp64_build_reproduces_numpy_oracle_fixture():
    # the FP64 C restatement (fast oracle of the larger GPU parity fixtures) against the NumPy oracle's 1,280-
    # triangle sphere fixture: boundary pressure and field at its 100 points
    from oracle import beat_cpu as B

    g = np.load(GOLD / "sphere1280_1khz.npz")
    m = O.Mesh(g["vertices"], g["faces"], np.ones(len(g["faces"]), int), np.ones(len(g["faces"]), int))
    out = B.CpuMesh(m, "double").solve_frequency(1000.0, np.ones((len(g["faces"]), 1)), 343.0, 1.205, g["points"])
    assert np.linalg.norm(out["pressure"][:, 0] - g["pressure"]) <= 1e-12 * np.linalg.norm(g["pressure"])
    assert np.linalg.norm(out["field"][:, 0] - g["field"]) <= 1e-12 * np.linalg.norm(g["field"])


@pytest.mark.parametrize("freq", [1000.0, 8000.0])
def test_fp32_restatement_within_stated_bound(freq):
    # the FP32 build (the reference's production precision) stays within 0.02 dB / 0.1 degrees and 5e-4 relative
    # boundary pressure of the FP64 build on the 1,280-triangle sphere (measured: <= 1e-4 dB, 1e-3 degrees)
    from beat_hip import synthetic as syn
    from oracle import beat_cpu as B

    V, F, T = syn.icosphere(3, 0.1)
    m = O.Mesh(V, F, T, np.ones(len(F), int))
    pts = O.fibonacci_observation({"spherical_sampling_enabled": True, "spherical_sampling_points": 50})[0]
    q = np.ones((len(F), 1))
    a = B.CpuMesh(m, "double").solve_frequency(freq, q, 343.0, 1.205, pts.astype(float))
    b = B.CpuMesh(m, "single").solve_frequency(freq, q, 343.0, 1.205, pts.astype(float))
    fa, fb = a["field"][:, 0], b["field"][:, 0].astype(complex)
    assert np.abs(20 * np.log10(np.abs(fb) / np.abs(fa))).max() <= 0.02
    assert np.degrees(np.abs(np.angle(fb / fa))).max() <= 0.1
    assert np.linalg.norm(b["pressure"][:, 0] - a["pressure"][:, 0]) <= 5e-4 * np.linalg.norm(a["pressure"][:, 0])
    # and it is not the FP64 result relabelled: FP32 rounding is visible
    assert np.abs(fb - fa).max() > 0


def test_horn6k_fixture_mesh_is_the_generator_output():
    from beat_hip import synthetic as syn

    g = np.load(GOLD / "horn6k_fp64.npz")
    V, F, T = syn.horn_in_box()
    assert np.array_equal(F, g["faces"]) and np.array_equal(T, g["tags"])
    assert np.allclose(V, g["vertices"], rtol=0, atol=1e-12)


def test_horn6k_fixture_field_follows_from_its_pressure():
    # internal consistency of the fixture: the stored field is the field of the stored boundary pressure
    g = np.load(GOLD / "horn6k_fp64.npz")
    m = O.Mesh(g["vertices"], g["faces"], g["tags"], np.ones(len(g["faces"]), int))
    for i, f in enumerate(g["freqs"]):
        k = 2 * np.pi * f / 343.0
        q = 1j * 1.205 * 2 * np.pi * f * (g["tags"] == 2)
        u = O.evaluate_field(g["points"], m, g["pressure"][i], q, k)
        assert np.linalg.norm(u - g["field"][i]) <= 1e-11 * np.linalg.norm(g["field"][i])


def test_sphere40_fixture_spot_check():
    from beat_hip import synthetic as syn
    from oracle import beat_cpu as B

    g = np.load(GOLD / "sphere1280_40f.npz")
    V, F, T = syn.icosphere(3, 0.1)
    assert np.array_equal(F, g["faces"]) and np.allclose(V, g["vertices"], rtol=0, atol=1e-14)
    m = O.Mesh(V, F, T, np.ones(len(F), int))
    i = 33
    out = B.CpuMesh(m, "double").solve_frequency(float(g["freqs"][i]), np.ones((len(F), 1)), 343.0, 1.205, g["points"])
    assert np.linalg.norm(out["field"][:, 0] - g["field"][i]) <= 1e-6 * np.linalg.norm(g["field"][i])
