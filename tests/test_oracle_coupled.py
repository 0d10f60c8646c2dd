"""Pins of the coupled FEM-BEM oracle (oracle/coupled_oracle.py): closed forms, the reference's own fixture
expectations (src/blab/solvers/julia_local/tests/coupled_solver_tests.jl), and physics that a wrong sign or a
dropped term in any coupling block would break."""

import math
from pathlib import Path

import numpy as np
import pytest

from beat_hip import synthetic as syn
from oracle import bem_oracle as O
from oracle import coupled_oracle as CO

GOLDEN = Path(__file__).parent / "golden"
RHO, C0 = 1.21, 343.0


def test_unit_tetrahedron_closed_form():
    x = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    K, M, det = CO.tet_matrices(x)
    Kx = np.array([[3.0, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]) / 6.0
    Mx = (np.ones((4, 4)) + np.eye(4)) / 120.0
    assert det == pytest.approx(1.0)
    assert np.allclose(K, Kx, atol=1e-15) and np.allclose(M, Mx, atol=1e-15)
    # a reflected, scaled copy: K scales with the edge length, M with the volume
    K2, M2, _ = CO.tet_matrices(2.0 * x[[0, 2, 1, 3]])
    assert np.allclose(K2, 2.0 * Kx[np.ix_([0, 2, 1, 3], [0, 2, 1, 3])], atol=1e-14)
    assert np.allclose(M2, 8.0 * Mx, atol=1e-15)


def test_sealed_cube_modes():
    """First modes of the rigid unit cube: c / 2 (three-fold), within the 8 % of the reference's own test on its
    4 x 4 x 4 structured cube (coupled_solver_tests.jl:280-287)."""
    import scipy.linalg as sla

    cube = syn.structured_cube(4)
    K, M = CO.fem_matrices(cube.vertices, cube.tetrahedra)
    assert np.abs(K @ np.ones(len(K))).max() < 1e-12 and M.sum() == pytest.approx(1.0)
    w = sla.eigh(K, M, eigvals_only=True)
    modes = C0 * np.sqrt(np.clip(w[w > 1e-8 * w.max()], 0, None)) / (2 * math.pi)
    assert modes[0] == pytest.approx(C0 / 2, rel=0.08)
    assert modes[:3].max() / modes[:3].min() < 1.08


def test_reference_fixture_expectations():
    """femvolume.msh / exterior_conforming.msh (the reference's tests/fixtures, committed under tests/golden):
    the counts and operator properties its tests assert (coupled_solver_tests.jl:176-200, 289-311, 397-452)."""
    from beat_hip import fem as FE
    from beat_hip import geometry as G

    fm = FE.read_gmsh41_volume(GOLDEN / "femvolume.msh", 0.001)
    assert fm.n_vertices == 842 and len(fm.tetrahedra) == 2925
    assert fm.physical_tag(3, "Volume") == 1 and fm.physical_tag(2, "Radiator") == 2
    assert fm.physical_tag(2, "Interface") == 3
    V, F, T = G.read_gmsh22(str(GOLDEN / "exterior_conforming.msh"), scale=0.001)
    bem = O.Mesh(V, F, T, np.zeros(len(F), np.int64))
    im = CO.interface_map(fm.vertices, fm.boundary_faces, fm.boundary_tags, bem, 3, 2)
    assert len(im["fem_vertices"]) == 106 and len(im["fem_faces"]) == 180
    assert set(im["sign"].tolist()) <= {-1, 1}
    fl, bf, ft, bt = CO.interface_ops(fm.vertices, fm.n_vertices, fm.boundary_faces, bem, im)
    assert fl.shape == (842, 106) and bf.shape == (2424, 106) and ft.shape == (106, 842) and bt.shape == (106, 1214)
    assert np.allclose(bf[im["bem_faces"]].sum(axis=1), im["sign"])
    # rigid piston on the radiator (FEM, sign +1) and on BEM tag 1 (sign -1), completion factor 2
    p = dict(fem_tags=[2], fem_signs=[1.0], bem_tags=[1], bem_signs=[-1.0], axis=(0, 0, 1), completion=2.0)
    po = CO.piston_ops(fm.vertices, fm.tetrahedra, fm.boundary_faces, fm.boundary_tags, bem, [p])
    n = CO.outward_normals(fm.vertices, fm.tetrahedra, fm.boundary_faces)
    fa = np.array([CO.tri_area(fm.vertices[f]) for f in fm.boundary_faces])
    sel = fm.boundary_tags == 2
    fem_area = float(np.sum(n[sel, 2] * fa[sel]))
    bsel = bem.tags == 1
    bem_area = float(np.sum(-bem.normals[bsel, 2] * bem.areas[bsel]))
    assert po["fem_surface"].sum() == pytest.approx(fem_area) and po["fem_force"].sum() == pytest.approx(2 * fem_area)
    assert po["bem_surface"].sum() == pytest.approx(bem_area) and po["bem_force"].sum() == pytest.approx(2 * bem_area)
    assert np.allclose(po["bem_normal_velocity"][bsel, 0], -bem.normals[bsel, 2])
    assert np.all(po["bem_normal_velocity"][~bsel, 0] == 0.0)


def test_prescribed_load_hand_case():
    """The reference's mixed-excitation unit test (coupled_solver_tests.jl:326-395): velocity 2 x weight 0.5 on
    the tagged face of one unit tetrahedron, rho = 1, omega = 2 gives i/3 on its three vertices and 0 on the
    fourth."""
    v = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    load = CO.velocity_load(v, 4, [(0, 1, 2)], [9], 9, 1.0, 2.0, 2.0 * 0.5)
    assert np.allclose(load[:3], 1j / 3.0) and load[3] == 0


def test_miki_bounds():
    """Passive lining with 1-5 % absorption at 100 Hz for 30 mm of 5,000 Pa s/m^2 (coupled_solver_tests.jl:212-238)."""
    y = CO.miki_admittance(100.0, 343.0, 1.21, 0.03, 5000.0)
    assert y.real > 0 and y.imag < 0
    z, air = 1.0 / y, 1.21 * 343.0
    assert 0.01 < 1.0 - abs((z - air) / (z + air)) ** 2 < 0.05


@pytest.fixture(scope="module")
def shell():
    a, b = 0.1, 0.15
    vm, (V, F, T) = syn.spherical_shell(2, a, b, 2)
    bem = O.Mesh(V, F, T, np.zeros(len(F), np.int64))
    im = CO.interface_map(vm.vertices, vm.boundary_faces, vm.boundary_tags, bem, syn.TAG_OUTER, syn.TAG_OUTER)
    rng = np.random.default_rng(0)
    pts = rng.normal(size=(24, 3))
    pts = 2.0 * pts / np.linalg.norm(pts, axis=1)[:, None]
    return dict(a=a, vm=vm, bem=bem, im=im, pts=pts)


def test_pulsating_sphere_through_fem_shell(shell):
    """Inner sphere of the FEM shell moving with unit velocity along the FEM outward normal (towards the centre):
    the field beyond the BEM surface is the pulsating sphere's with radial velocity -1 (closed form) to the
    discretisation error; the opposite sign is off by ~200 %."""
    vm, f = shell["vm"], 500.0
    k, w = 2 * math.pi * f / C0, 2 * math.pi * f
    out = CO.solve_coupled(vm.vertices, vm.tetrahedra, vm.boundary_faces, vm.boundary_tags, shell["bem"], shell["im"],
                           f, C0, RHO, [{"kind": "normal_velocity", "fem_tags": [syn.TAG_INNER], "fem_weights": [1.0]}],
                           points=shell["pts"])
    r = np.linalg.norm(shell["pts"], axis=1)
    good = O.pulsating_sphere_pressure(r, shell["a"], k, RHO, w, -1.0)
    bad = O.pulsating_sphere_pressure(r, shell["a"], k, RHO, w, 1.0)
    err = np.linalg.norm(out["field"][:, 0] - good) / np.linalg.norm(good)
    assert err < 0.03
    assert np.linalg.norm(out["field"][:, 0] - bad) / np.linalg.norm(bad) > 1.0
    # interface continuity of the monolithic solution
    im = shell["im"]
    pf, pb = out["fem_pressure"][im["fem_vertices"], 0], out["bem_pressure"][im["bem_vertices"], 0]
    assert np.linalg.norm(pf - pb) / np.linalg.norm(pf) < 1e-12


PISTON = dict(fem_tags=[2], fem_signs=[1.0], bem_tags=[], bem_signs=[], axis=(0, 0, 1), completion=1.0,
              re=6.0, le=0.0005, bl=7.0, mmd=0.015, cms=0.0005, rms=1.0)


def _piston_checks(shell, piston):
    vm, f = shell["vm"], 500.0
    k, w = 2 * math.pi * f / C0, 2 * math.pi * f
    out = CO.solve_coupled(vm.vertices, vm.tetrahedra, vm.boundary_faces, vm.boundary_tags, shell["bem"], shell["im"],
                           f, C0, RHO, [{"kind": "voltage", "transducer": 1, "amplitude": 2.83}], pistons=[piston],
                           points=shell["pts"])
    u, i = out["diaphragm_velocity"][0, 0], out["voice_coil_current"][0, 0]
    a = shell["a"]

    def h1(x):
        return -np.exp(1j * x) * (x + 1j) / x ** 2

    def dh1(x):
        return np.exp(1j * x) * (2j + 2 * x - 1j * x * x) / x ** 3

    zr = 4 * math.pi * a * a / 3 * 1j * RHO * C0 * h1(k * a) / dh1(k * a)  # radiation impedance, closed form
    zm = complex(piston["rms"], -w * piston["mmd"] + 1 / (w * piston["cms"]))
    ze = complex(piston["re"], -w * piston["le"])
    u_exact = piston["bl"] * 2.83 / (ze * (zm + zr) + piston["bl"] ** 2)
    p_in = 0.5 * (2.83 * np.conj(i)).real
    p_out = 0.5 * piston["rms"] * abs(u) ** 2 + 0.5 * piston["re"] * abs(i) ** 2 + 0.5 * zr.real * abs(u) ** 2
    r = np.linalg.norm(shell["pts"], axis=1)
    field = CO.oscillating_sphere_pressure(r, shell["pts"][:, 2] / r, a, k, RHO, C0, u)
    return (abs(u - u_exact) / abs(u_exact), abs(p_in - p_out) / p_in,
            np.linalg.norm(out["field"][:, 0] - field) / np.linalg.norm(field))


def test_oscillating_sphere_piston_power_balance(shell):
    """The inner sphere as an electrodynamic rigid piston along z driven at 2.83 V: its velocity matches
    Bl e / (Z_e (Z_m + Z_r) + Bl^2) with the oscillating sphere's closed-form radiation impedance, the far field
    the oscillating sphere's, and electrical input power = mechanical + electrical losses + radiated power."""
    du, dp, df = _piston_checks(shell, PISTON)
    assert du < 0.04 and dp < 0.01 and df < 0.1


def test_planted_sign_error_breaks_the_power_balance(shell, monkeypatch):
    """A sign error planted in the oracle's piston force coupling (on the CPU) is caught by the same checks."""
    real = CO.piston_ops

    def wrong(*args, **kw):
        out = real(*args, **kw)
        out["fem_force"] = -out["fem_force"]
        return out

    monkeypatch.setattr(CO, "piston_ops", wrong)
    du, dp, _ = _piston_checks(shell, PISTON)
    assert du > 0.2 or dp > 0.2


def test_schur_complement_matches_dense_elimination():
    """The product's host condensation (SuperLU on the interior) against dense elimination, the reference's
    "blocked exact FEM Schur complement" check (coupled_solver_tests.jl:58-89), on a lossy cube system."""
    from beat_hip.coupled import schur_complement
    import scipy.sparse as sp

    cube = syn.structured_cube(3)
    K, M = CO.fem_matrices(cube.vertices, cube.tetrahedra)
    k = 2 * math.pi * 300.0 / C0
    A = K - k * k * M - 0.01j * k * k * M
    ret = np.nonzero(np.abs(cube.vertices[:, 2] - 1.0) < 1e-12)[0]
    S, parts = schur_complement(sp.csr_matrix(A), ret, block=5)
    I = np.setdiff1d(np.arange(len(A)), ret)
    ref = A[np.ix_(ret, ret)] - A[np.ix_(ret, I)] @ np.linalg.solve(A[np.ix_(I, I)], A[np.ix_(I, ret)])
    assert np.linalg.norm(S - ref) / np.linalg.norm(ref) < 1e-12
