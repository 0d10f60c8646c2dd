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
