"""GPU parity of the coupled FEM-BEM solve against the FP64 oracle (oracle/coupled_oracle.py, monolithic):

* the compiled-system protocol on the shell cases (moving inner sphere + BEM source with bulk loss; electrodynamic
  piston; Miki-lined wall), condensed and monolithic, FP64 and float32 (FP32 LU + FP64 refinement);
* the reference's own fixture (femvolume.msh + exterior_conforming.msh at 500 Hz, bulk loss 0.01, FEM radiator and
  a BEM-only source) with the bounds its tests assert (coupled_solver_tests.jl:454-568): relative residual,
  pressure continuity and BEM replay below 1e-8, flux conservation below 1e-10, FP32 against FP64 below 1e-4.

Tolerances: FP64 results within 1e-9 (relative, per block) of the oracle -- the same discretisation, only the
summation order and the condensation differ; float32 within 1e-5 (FP32 BEM assembly; the refined solve is exact
for the assembled FP32 matrix; the wire rounds to complex64).
"""

import math
from pathlib import Path

import numpy as np
import pytest

import coupled_cases as CC

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden"


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _factory(dev):
    from beat_hip.coupled import CoupledSolver

    return lambda model, **opts: CoupledSolver(model, dev=dev, **opts)


@pytest.mark.parametrize("case,precision,static,validation", [
    ("velocity", "float64", True, False), ("velocity", "float64", False, True), ("velocity", "float32", True, False),
    ("piston", "float64", True, False), ("piston", "float32", True, False), ("wall", "float64", True, False),
    ("wall", "float32", False, False),
])
def test_coupled_request_matches_oracle(gpu_device, tmp_path, case, precision, static, validation):
    from beat_hip import system as SY

    print(f"coupled {case} {precision} {'condensed' if static else 'monolithic'}", flush=True)
    req = CC.coupled_request(tmp_path, case, precision=precision, static=static, validation=validation,
                             freqs=(400.0, 900.0))
    events = []
    cancelled, solved = SY.solve_system_request(req, None, events.append, coupled_factory=_factory(gpu_device))
    assert (cancelled, solved) == (False, 2)
    tol = 1e-9 if precision == "float64" else 1e-5
    for ev, f in zip(events, (400.0, 900.0)):
        res = ev["result"]
        ref = CC.oracle_case(req, f)
        q = {x["id"]: CC.wire_values(x) for x in res["quantities"]}
        errs = {
            "fem": _rel(q["fem_nodal_pressure"], ref["fem_pressure"].T),
            "bem": _rel(q["bem_boundary_pressure"], ref["bem_pressure"].T),
            "neumann": _rel(q["bem_boundary_neumann"], ref["bem_neumann"].T),
            "flux": _rel(q["interface_normal_derivative"], ref["interface_flux"].T),
            "field": _rel(q["far"], ref["field"].T),
        }
        if case == "piston":
            errs["u"] = _rel(q["u"], ref["diaphragm_velocity"].T)
            errs["i"] = _rel(q["i"], ref["voice_coil_current"].T)
        d = res["diagnostics"]
        print(f"  {f:g} Hz", {k: f"{v:.1e}" for k, v in errs.items()}, "residual", d["relative_residual"],
              "continuity", d["pressure_continuity_error"])
        assert max(errs.values()) < tol, errs
        assert d["formulation"] == ("fem_interface_condensed" if static else "monolithic")
        assert d["relative_residual"] < (1e-10 if precision == "float32" else 1e-8)
        if validation:
            assert d["all_bem_replay_error"] < 1e-8
        if static:
            assert d["fem_interior_residual"] < 1e-10


@pytest.fixture(scope="module")
def fixture_model():
    """The reference's baseline interface case: femvolume.msh (842 vertices) coupled through its Interface group
    to exterior_conforming.msh (1,214 vertices; tag 1 a BEM-only moving surface), bulk loss 0.01."""
    import scipy.sparse as sp

    from beat_hip import fem as FE
    from beat_hip import geometry as G
    from beat_hip.coupled import CoupledModel

    fm = FE.read_gmsh41_volume(GOLDEN / "femvolume.msh", 0.001)
    bm = G.BoundaryMesh(*G.read_gmsh22(str(GOLDEN / "exterior_conforming.msh"), scale=0.001))
    im = FE.build_conforming_interface_map(fm, bm, fm.physical_tag(2, "Interface"), 2)
    sel = np.nonzero(bm.tags == 1)[0]
    pv = sp.csr_matrix((np.ones(len(sel)), (sel, np.zeros(len(sel), np.int64))), shape=(bm.n_faces, 1))
    model = CoupledModel(fm, bm, im, 1.21, 343.0, [], pv, np.full(fm.n_vertices, 0.01), [])
    return model, fm.physical_tag(2, "Radiator")


def _fixture_solve(dev, model, radiator, precision, static, validation):
    from beat_hip.coupled import CoupledSolver, Excitation

    exc = [Excitation("normal_velocity", (radiator,), (1.0,), 0, 0, 1.0),
           Excitation("normal_velocity", (), (), 1, 0, 1.0)]
    s = CoupledSolver(model, precision=precision, static_condensation=static, validation_diagnostics=validation,
                      regular_order=1, singular_order=1, dev=dev)
    try:
        sols, _, _, _ = s.solve(500.0, exc)
    finally:
        s.close()
    return sols


def test_reference_fixture_bounds(gpu_device, fixture_model):
    model, radiator = fixture_model
    print("reference fixture, monolithic FP64 with validation", flush=True)
    sols = _fixture_solve(gpu_device, model, radiator, "float64", False, True)
    for s in sols:
        print(f"  residual {s.relative_residual:.2e} continuity {s.pressure_continuity_error:.2e} "
              f"flux {s.flux_conservation_error:.2e} replay {s.all_bem_replay_error:.2e}")
        assert s.relative_residual < 1e-8 and s.pressure_continuity_error < 1e-8
        assert s.flux_conservation_error < 1e-10 and s.all_bem_replay_error < 1e-8
    # the BEM-only source: its Neumann data is i rho omega on tag 1 plus the interface flux
    omega = 2 * math.pi * 500.0
    expect = 1j * 1.21 * omega * model.prescribed_velocity.toarray()[:, 0] + model_flux(model, sols[1])
    assert np.allclose(sols[1].bem_neumann, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
    print("reference fixture, condensed FP64 and float32", flush=True)
    cond = _fixture_solve(gpu_device, model, radiator, "float64", True, False)
    f32 = _fixture_solve(gpu_device, model, radiator, "float32", True, False)
    for a, b, c in zip(sols, cond, f32):
        for key in ("fem_pressure", "bem_pressure", "interface_flux"):
            e64 = _rel(getattr(b, key), getattr(a, key))
            e32 = _rel(getattr(c, key), getattr(a, key))
            print(f"  {key}: condensed {e64:.1e} float32 {e32:.1e}")
            assert e64 < 1e-9 and e32 < 1e-4


def model_flux(model, sol):
    from beat_hip import fem as FE

    ops = FE.interface_operators(model.fem_mesh, model.bem_mesh, model.interface_map)
    return ops.bem_flux @ sol.interface_flux


def test_reference_fixture_matches_oracle(gpu_device, fixture_model):
    """The same fixture against the FP64 oracle (monolithic, dense LAPACK)."""
    from oracle import bem_oracle as O
    from oracle import coupled_oracle as CO

    model, radiator = fixture_model
    fm, bm, im = model.fem_mesh, model.bem_mesh, model.interface_map
    om = O.Mesh(bm.vertices, bm.faces, bm.tags, np.zeros(bm.n_faces, np.int64))
    imo = dict(fem_vertices=im.fem_vertex_indices, bem_vertices=im.fem_to_bem_vertex_indices,
               fem_faces=im.fem_face_indices, bem_faces=im.bem_face_indices, sign=im.normal_sign)
    ref = CO.solve_coupled(fm.vertices, fm.tetrahedra, fm.boundary_faces, fm.boundary_tags, om, imo, 500.0, 343.0,
                           1.21, [{"kind": "normal_velocity", "fem_tags": [radiator], "fem_weights": [1.0]},
                                  {"kind": "normal_velocity", "bem_source": 1}],
                           prescribed=model.prescribed_velocity.toarray(), loss_by_vertex=model.bulk_loss_by_vertex,
                           regular_order=1, singular_order=1)
    print("reference fixture, condensed FP64 against the oracle", flush=True)
    sols = _fixture_solve(gpu_device, model, radiator, "float64", True, False)
    for e, s in enumerate(sols):
        for key in ("fem_pressure", "bem_pressure", "interface_flux"):
            err = _rel(getattr(s, key), ref[key][:, e])
            print(f"  excitation {e} {key}: {err:.1e}")
            assert err < 1e-9


def test_batched_frequencies_match_single(gpu_device, tmp_path):
    """Three frequencies solved as one device batch equal the same frequencies solved one at a time (the batched
    LU may take other launch shapes: agreement to the solve's accuracy, 1e-10 FP64 / refinement tolerance)."""
    from beat_hip import coupled_system as CS
    from beat_hip.coupled import CoupledSolver

    req = CC.coupled_request(tmp_path, "piston", precision="float64", freqs=(300.0,))
    job = CS.parse_coupled_request(req)
    for precision, tol in (("float64", 1e-10), ("float32", 1e-8)):
        print(f"batched vs single, {precision}", flush=True)
        s = CoupledSolver(job["model"], precision=precision, dev=gpu_device)
        try:
            freqs = [250.0, 700.0, 1300.0]
            batch = s.solve_batch(freqs, job["excitations"], CC.observation_points())
            for f, (sols, field, _, _) in zip(freqs, batch):
                one, field1, _, _ = s.solve(f, job["excitations"], CC.observation_points())
                for a, b in zip(sols, one):
                    assert _rel(a.fem_pressure, b.fem_pressure) < tol
                    assert _rel(a.diaphragm_velocity, b.diaphragm_velocity) < tol
                assert _rel(field, field1) < tol
        finally:
            s.close()
