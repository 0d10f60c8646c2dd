"""Compiled-system exterior protocol (coupled_solver.jl solve_exterior_request and worker loop) with a test double
for the device: request checks, tag aggregation, excitations, output wire and worker events."""

import io
import json
import math

import numpy as np
import pytest

from beat_hip import system as SY
from beat_hip import system_worker as SW
from beat_hip import synthetic as syn
from test_worker_protocol import FakeEngine


def _system_request(tmp_path, precision="float32", outputs=None, **opts):
    # mesh "a": icosphere with the lower half tagged 7 (moving) and the rest 3; mesh "b": a second sphere
    # (tag 5 only) translated along x. Tags are renumbered per mesh in order of appearance.
    V, F, T = syn.icosphere(1, 0.1)
    T = np.where(V[F].mean(axis=1)[:, 2] < 0, 7, 3)
    pa = tmp_path / "a.msh"
    syn.write_gmsh22(pa, V, F, T, scale=0.001)
    V2, F2, T2 = syn.icosphere(1, 0.05)
    pb = tmp_path / "b.msh"
    syn.write_gmsh22(pb, V2, F2, np.full(len(F2), 5), scale=0.001)
    system = {
        "contract_version": 1, "id": "sys", "name": "sys",
        "meshes": [
            {"id": "a", "name": "a", "file": str(pa), "purpose": "bem_surface", "scale_to_m": 0.001,
             "translation_m": [0.0, 0.0, 0.0]},
            {"id": "b", "name": "b", "file": str(pb), "purpose": "bem_surface", "scale_to_m": 0.001,
             "translation_m": [0.5, 0.0, 0.0]},
        ],
        "regions": [{"id": "air", "name": "air", "kind": "unbounded_air", "mesh_ids": ["a", "b"],
                     "sound_speed_m_per_s": 343.0, "density_kg_per_m3": 1.205}],
        "boundaries": [
            {"id": "cone", "name": "cone", "region_id": "air", "kind": "moving",
             "group": {"mesh_id": "a", "dimension": 2, "tag": 7}},
            {"id": "box", "name": "box", "region_id": "air", "kind": "rigid",
             "group": {"mesh_id": "a", "dimension": 2, "tag": 3}},
            {"id": "port2", "name": "port2", "region_id": "air", "kind": "moving",
             "group": {"mesh_id": "b", "dimension": 2, "tag": 5}},
        ],
        "interfaces": [],
        "components": [
            {"id": "woofer", "name": "woofer", "kind": "ideal_velocity_source", "boundary_ids": ["cone", "box"],
             "parameters": {"boundary_motion_weights": {"cone": 2.0}}},
            {"id": "tweeter", "name": "tweeter", "kind": "ideal_velocity_source", "boundary_ids": ["port2"]},
        ],
        "excitation_ports": [
            {"id": "p_w", "name": "p_w", "component_id": "woofer", "kind": "normal_velocity"},
            {"id": "p_t", "name": "p_t", "component_id": "tweeter", "kind": "normal_velocity"},
        ],
        "assumptions": [],
    }
    if outputs is None:
        outputs = [
            {"id": "far", "quantity": "exterior_pressure", "target_ids": [],
             "options": {"points_m": [[0.0, 0.0, 2.0], [2.0, 0.0, 0.0], [0.0, 1.0, 1.0]]}},
            {"id": "p", "quantity": "bem_boundary_pressure", "target_ids": ["air"], "options": {}},
            {"id": "q", "quantity": "bem_boundary_neumann", "target_ids": [], "options": {}},
            {"id": "z", "quantity": "radiation_impedance", "target_ids": [], "options": {}},
        ]
    options = {"precision": precision, "bem_backend": "rocm"}
    options.update(opts)
    return {"schema_version": 1, "compiled_system": system, "frequencies_hz": [250.0, 1000.0, 3000.0],
            "excitation_port_ids": ["p_w", "p_t"], "outputs": outputs, "solver_options": options}


def _wire_array(w):
    a = np.asarray(w["real"]) + 1j * np.asarray(w["imag"])
    return a.reshape(w["shape"])


def test_aggregation_renumbers_tags_per_mesh(tmp_path):
    req = _system_request(tmp_path)
    sysd = req["compiled_system"]
    dom = SY.aggregate_bem_region(sysd["meshes"], sysd["regions"][0], sysd["boundaries"])
    mesh = dom["mesh"]
    # mesh a's two tags become solver tags 1 and 2 (order of first appearance); mesh b's tag 5 becomes 3
    V, F, _ = syn.icosphere(1, 0.1)
    lower = V[F].mean(axis=1)[:, 2] < 0
    first_is_lower = bool(lower[0])
    assert np.all(mesh.tags[:80][lower] == (1 if first_is_lower else 2))
    assert np.all(mesh.tags[:80][~lower] == (2 if first_is_lower else 1))
    assert set(np.unique(mesh.tags[80:]).tolist()) == {3}
    assert dom["boundary_tag_by_id"]["port2"] == 3
    assert dom["vertex_offset_by_mesh_id"] == {"a": 0, "b": 42}
    # mesh b translated by 0.5 m along x, no welding
    assert np.isclose(mesh.vertices[42:, 0].mean(), 0.5, atol=1e-12)


def test_excitations_drives_and_impedance(tmp_path):
    req = _system_request(tmp_path)
    job = SY.parse_exterior_request(req)
    exc = job["excitations"]
    assert [e["component_id"] for e in exc] == ["woofer", "tweeter"]
    assert exc[0]["amplitudes"] == [2.0] and len(exc[0]["tags"]) == 1  # rigid "box" is not moving
    q = SY.unit_drives(job["mesh"], exc)
    cone = job["mesh"].tags == exc[0]["tags"][0]
    assert np.all(q[cone, 0] == 2.0) and np.all(q[~cone, 0] == 0.0)
    assert np.all(q[80:, 1] == 1.0) and np.all(q[:80, 1] == 0.0)
    # impedance: conj(10 * sum(face-average pressure * area)) / 2, times the symmetry factor
    p = np.linspace(1, 2, job["mesh"].n_vertices) * (1 + 0.5j)
    z = SY.component_impedance(job["mesh"], p, exc[0], 1)
    m = job["mesh"]
    force = 10 * sum(p[m.faces[f]].mean() * m.areas[f] for f in np.nonzero(cone)[0])
    assert np.isclose(z, complex(force.real / 2, -force.imag / 2), rtol=1e-14)
    assert np.isclose(SY.component_impedance(m, p, exc[0], 4), 4 * z, rtol=1e-14)


def test_complex_array_wire_is_row_major():
    a = np.arange(6).reshape(2, 3) + 1j * np.arange(6, 12).reshape(2, 3)
    w = SY.complex_array_wire(a, np.complex64)
    assert w["dtype"] == "complex64" and w["shape"] == [2, 3]
    assert w["real"] == [0, 1, 2, 3, 4, 5] and w["imag"] == [6, 7, 8, 9, 10, 11]
    assert SY.complex_array_wire(a, np.complex128)["dtype"] == "complex128"


@pytest.mark.parametrize("precision,cdt", [("float32", "complex64"), ("float64", "complex128")])
def test_results_carry_every_output(tmp_path, precision, cdt):
    req = _system_request(tmp_path, precision=precision)
    events = []
    cancelled, solved = SY.solve_system_request(req, FakeEngine, events.append, batch=2)
    assert not cancelled and solved == 3 and len(events) == 3
    job = SY.parse_exterior_request(req)
    V, F = job["mesh"].n_vertices, job["mesh"].n_faces
    for i, ev in enumerate(events):
        assert ev["type"] == "result"
        r = ev["result"]
        assert r["schema_version"] == 1 and r["excitation_port_ids"] == ["p_w", "p_t"]
        assert r["freq_hz"] == req["frequencies_hz"][i]
        qs = {q["id"]: q for q in r["quantities"]}
        assert [q["id"] for q in r["quantities"]] == ["far", "p", "q", "z"]
        assert qs["far"]["axes"] == ["excitation", "observation"] and qs["far"]["unit"] == "Pa"
        assert qs["far"]["values"]["shape"] == [2, 3] and qs["far"]["values"]["dtype"] == cdt
        assert qs["far"]["metadata"]["points_m"] == req["outputs"][0]["options"]["points_m"]
        assert qs["p"]["values"]["shape"] == [2, V] and qs["p"]["target_id"] == "air"
        assert qs["q"]["unit"] == "Pa/m" and qs["q"]["values"]["shape"] == [2, F]
        omega = 2 * math.pi * r["freq_hz"]
        neu = _wire_array(qs["q"]["values"])
        np.testing.assert_allclose(neu, (1j * 1.205 * omega * SY.unit_drives(job["mesh"], job["excitations"])).T,
                                   rtol=1e-6)
        assert qs["z"]["axes"] == ["radiator"] and qs["z"]["unit"] == "Pa*s/m^3"
        assert qs["z"]["values"]["shape"] == [2]
        # the test double returns unit pressure: impedance = (5 * moving area, 0) per component
        m = job["mesh"]
        z = _wire_array(qs["z"]["values"])
        for e, ex in enumerate(job["excitations"]):
            area = m.areas[np.isin(m.tags, ex["tags"])].sum()
            assert np.isclose(z[e], 5 * area, rtol=1e-6)
        d = r["diagnostics"]
        assert d["precision"] == precision and d["bem_backend"] == "rocm"
        assert d["formulation"] == "exterior_burton_miller_neumann" and d["bounded_region_count"] == 0
        assert set(d["timings"]) == {"assembly_s", "solve_s", "field_s", "mesh_setup_s", "cache_setup_s"}
        assert (d["timings"]["mesh_setup_s"] > 0) == (i == 0)
        json.dumps(r, allow_nan=False)


@pytest.mark.parametrize("mutate,message", [
    (lambda r: r["compiled_system"]["regions"].append(
        {"id": "in", "name": "in", "kind": "bounded_air", "mesh_ids": [], "sound_speed_m_per_s": 343.0,
         "density_kg_per_m3": 1.2}), "shared sound speed and density"),
    (lambda r: r["solver_options"].update(precision="float16"), "float32 or float64"),
    (lambda r: r["solver_options"].update(bem_backend="cuda"), "serves bem_backend 'rocm'"),
    (lambda r: r["solver_options"].update(regular_quadrature_mode="wavelength"), "only on the CPU backend"),
    (lambda r: r["solver_options"].update(singular_order=5), "between 1 and 4"),
    (lambda r: r["outputs"].append({"id": "x", "quantity": "interior_pressure"}), "Unsupported exterior output"),
    (lambda r: r["outputs"].append({"id": "y", "quantity": "exterior_pressure", "options": {}}), "points_m"),
    (lambda r: r["compiled_system"]["components"][1].update(boundary_ids=[]), "no moving boundaries"),
    (lambda r: r["compiled_system"]["excitation_ports"][0].update(kind="pressure"), "normal-velocity"),
    (lambda r: r.update(schema_version=2), "schema"),
])
def test_request_errors(tmp_path, mutate, message):
    req = _system_request(tmp_path)
    mutate(req)
    with pytest.raises(ValueError, match=message):
        SY.solve_system_request(req, FakeEngine, lambda e: None)


def test_worker_loop_events(tmp_path, monkeypatch):
    monkeypatch.setattr(SW, "engine_factory", FakeEngine)
    req = _system_request(tmp_path)
    path = tmp_path / "request.json"
    path.write_text(json.dumps(req))
    cancel = tmp_path / "cancel"
    req_c = dict(req, cancel_path=str(cancel))
    path_c = tmp_path / "request_c.json"
    path_c.write_text(json.dumps(req_c))
    cancel.write_text("cancel")
    lines = [json.dumps({"request": str(path), "operation": "solve"}),
             json.dumps({"request": str(path), "operation": "bem_field"}),
             json.dumps({"request": str(path_c), "operation": "solve"})]
    out = []
    SW.worker_loop(io.StringIO("\n".join(lines) + "\n"), out.append)
    types = [e["type"] for e in out]
    assert types == ["ready", "result", "result", "result", "completed", "failed", "cancelled"]
    assert out[4]["solved_count"] == 3 and "bem_field" in out[5]["error"] and out[6]["solved_count"] == 0
