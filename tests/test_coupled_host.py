"""Host side of the coupled FEM-BEM path (no device): the FEM module against the oracle, the Gmsh 4.1 reader, and
the compiled-system protocol (parse, excitations, result wire) with an oracle-backed stand-in for the device
solver."""

import json
import math

import numpy as np
import pytest
import scipy.sparse as sp

import coupled_cases as CC
from beat_hip import coupled as CP
from beat_hip import coupled_system as CS
from beat_hip import fem as FE
from beat_hip import synthetic as syn
from beat_hip import system as SY
from oracle import bem_oracle as O
from oracle import coupled_oracle as CO


def test_gmsh41_round_trip_and_restriction(tmp_path):
    vm, _ = syn.spherical_shell(1, 0.1, 0.15, 2)
    path = tmp_path / "v.msh"
    FE.write_gmsh41_volume(path, vm)
    back = FE.read_gmsh41_volume(path, scale=2.0, translation=(1.0, 0.0, 0.0))
    assert np.allclose(back.vertices, 2.0 * vm.vertices + [1.0, 0, 0])
    assert np.array_equal(back.tetrahedra, vm.tetrahedra) and np.array_equal(back.boundary_tags, vm.boundary_tags)
    assert back.physical_tag(2, "Interface") == syn.TAG_OUTER
    with pytest.raises(ValueError, match="not found in dimension 2"):
        back.physical_tag(2, "Nope")
    # two disjoint tetrahedra, the second selected (coupled_solver_tests.jl:240-278)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [2, 0, 0], [3, 0, 0], [2, 1, 0], [2, 0, 1.0]])
    faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3), (4, 5, 6), (4, 5, 7), (4, 6, 7), (5, 6, 7)]
    m = FE.VolumeMesh(v, [(0, 1, 2, 3), (4, 5, 6, 7)], [10, 20], faces, [101] * 4 + [202] * 4)
    sel, vmap, fmap = FE.restrict_volume_mesh(m, [20])
    assert sel.n_vertices == 4 and sel.tetrahedra.tolist() == [[0, 1, 2, 3]] and sel.tetra_tags.tolist() == [20]
    assert sel.boundary_tags.tolist() == [202] * 4
    assert vmap == {4: 0, 5: 1, 6: 2, 7: 3} and fmap == {4: 0, 5: 1, 6: 2, 7: 3}
    with pytest.raises(ValueError, match="contain no tetrahedra"):
        FE.restrict_volume_mesh(m, [99])


def test_gmsh41_reader_rejects(tmp_path):
    p = tmp_path / "bad.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError, match="Gmsh 4.1"):
        FE.read_gmsh41_volume(p)
    p.write_text("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError, match="ASCII"):
        FE.read_gmsh41_volume(p)


def test_fem_host_matches_oracle():
    vm, (V, F, T) = syn.spherical_shell(1, 0.1, 0.15, 2)
    K, M = FE.p1_fem_matrices(vm)
    Ko, Mo = CO.fem_matrices(vm.vertices, vm.tetrahedra)
    assert np.abs(K.toarray() - Ko).max() < 1e-13 * np.abs(Ko).max()
    assert np.abs(M.toarray() - Mo).max() < 1e-15
    n = FE.outward_boundary_normals(vm)
    assert np.allclose(n, CO.outward_normals(vm.vertices, vm.tetrahedra, vm.boundary_faces))
    faces = np.nonzero(vm.boundary_tags == syn.TAG_INNER)[0]
    B = FE.boundary_mass_matrix(vm, faces, np.arange(vm.n_vertices)).toarray()
    Bo = CO.boundary_mass(vm.vertices, vm.n_vertices, vm.boundary_faces[faces], np.arange(vm.n_vertices))
    assert np.allclose(B, Bo, atol=1e-16)
    load = FE.prescribed_velocity_load(vm, syn.TAG_INNER, 1.2, 300.0, 0.5 + 0.1j)
    assert np.allclose(load, CO.velocity_load(vm.vertices, vm.n_vertices, vm.boundary_faces, vm.boundary_tags,
                                              syn.TAG_INNER, 1.2, 300.0, 0.5 + 0.1j))
    from beat_hip import geometry as G

    bm = G.BoundaryMesh(V, F, T)
    im = FE.build_conforming_interface_map(vm, bm, syn.TAG_OUTER, syn.TAG_OUTER)
    om = O.Mesh(V, F, T, np.zeros(len(F), np.int64))
    imo = CO.interface_map(vm.vertices, vm.boundary_faces, vm.boundary_tags, om, syn.TAG_OUTER, syn.TAG_OUTER)
    assert np.array_equal(im.fem_vertex_indices, imo["fem_vertices"])
    assert np.array_equal(im.fem_to_bem_vertex_indices, imo["bem_vertices"])
    assert np.array_equal(im.bem_face_indices, imo["bem_faces"]) and np.array_equal(im.normal_sign, imo["sign"])
    ops = FE.interface_operators(vm, bm, im)
    for a, b in zip((ops.fem_load, ops.bem_flux, ops.fem_trace, ops.bem_trace),
                    CO.interface_ops(vm.vertices, vm.n_vertices, vm.boundary_faces, om, imo)):
        assert np.allclose(a.toarray(), b, atol=1e-16)
    td = FE.Transducer("t", [syn.TAG_INNER], [1.0], [], [], np.array([0.0, 0.3, 1.0]), 2.0, 1, 6, 1e-4, 7, 0.01,
                       5e-4, 1.0)
    to = FE.transducer_operators(vm, bm, [td])
    po = CO.piston_ops(vm.vertices, vm.tetrahedra, vm.boundary_faces, vm.boundary_tags, om,
                       [dict(fem_tags=[syn.TAG_INNER], fem_signs=[1.0], bem_tags=[], bem_signs=[],
                             axis=(0.0, 0.3, 1.0), completion=2.0)])
    for key in ("fem_surface", "fem_force", "bem_surface", "bem_force", "bem_normal_velocity"):
        assert np.allclose(to[key].toarray(), po[key], atol=1e-16), key


def test_dynamic_stiffness_loss_forms_agree():
    """Scalar loss factor and per-vertex loss mass give the same system (coupled_solver_tests.jl:212-226)."""
    vm, _ = syn.spherical_shell(1, 0.1, 0.15, 1)
    K, M = FE.p1_fem_matrices(vm)
    k = 2 * math.pi * 200.0 / 343.0
    a = FE.dynamic_stiffness(K, M, k, bulk_loss_factor=0.02)
    b = FE.dynamic_stiffness(K, M, k, bulk_loss_mass=0.02 * M)
    assert abs(a - b).max() < 1e-12 * abs(a).max()


class OracleSolver:
    """Stand-in for the device solver: the oracle's monolithic solve of the model the protocol built."""

    def __init__(self, model, **opts):
        self.model, self.opts = model, opts
        fm, bm = model.fem_mesh, model.bem_mesh
        self.n_fem, self.V, self.n_if, self.nt = fm.n_vertices, bm.n_vertices, model.interface_map.count, len(
            model.transducers)
        self.n = self.n_fem + self.V + self.n_if + 2 * self.nt
        self.timings = {"fem_matrix_cache_s": 0.0}
        self.closed = False

    def solve(self, freq, excitations, points=None):
        m = self.model
        bm = m.bem_mesh
        om = O.Mesh(bm.vertices, bm.faces, bm.tags, np.zeros(bm.n_faces, np.int64))
        im = m.interface_map
        imo = dict(fem_vertices=im.fem_vertex_indices, bem_vertices=im.fem_to_bem_vertex_indices,
                   fem_faces=im.fem_face_indices, bem_faces=im.bem_face_indices, sign=im.normal_sign)
        pist = [dict(fem_tags=t.fem_tags, fem_signs=t.fem_signs, bem_tags=t.bem_tags, bem_signs=t.bem_signs,
                     axis=t.motion_axis, completion=t.surface_completion_factor, re=t.re_ohm, le=t.le_h,
                     bl=t.bl_n_per_a, mmd=t.mmd_kg, cms=t.cms_m_per_n, rms=t.rms_n_s_per_m) for t in m.transducers]
        exc = [{"kind": e.kind, "fem_tags": list(e.fem_tags), "fem_weights": list(e.fem_weights),
                "bem_source": e.bem_source_index, "transducer": e.transducer_index, "amplitude": e.amplitude}
               for e in excitations]
        walls = [(w["tag"], w["thickness_m"], w["flow_resistivity_pa_s_per_m2"]) for w in m.wall_impedances]
        fm = m.fem_mesh
        out = CO.solve_coupled(fm.vertices, fm.tetrahedra, fm.boundary_faces, fm.boundary_tags, om, imo, freq,
                               m.sound_speed, m.density, exc, pistons=pist,
                               prescribed=m.prescribed_velocity.toarray(), loss_by_vertex=m.bulk_loss_by_vertex,
                               walls=walls, regular_order=self.opts["regular_order"],
                               singular_order=self.opts["singular_order"], points=points)
        sols = [CP.CoupledSolution(out["fem_pressure"][:, e], out["bem_pressure"][:, e], out["interface_flux"][:, e],
                                   out["bem_neumann"][:, e], out["diaphragm_velocity"][:, e],
                                   out["voice_coil_current"][:, e], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
                for e in range(len(excitations))]
        return sols, out.get("field"), [], {}

    def close(self):
        self.closed = True


@pytest.mark.parametrize("case", ["velocity", "piston", "wall"])
def test_protocol_builds_the_oracle_system(tmp_path, case):
    """Parse the compiled-system request, build the model, run the frequency loop through the oracle stand-in:
    every output equals the oracle solution of the same request read independently from its files."""
    req = CC.coupled_request(tmp_path, case, freqs=(500.0,))
    events = []
    cancelled, solved = SY.solve_system_request(req, None, events.append, coupled_factory=OracleSolver)
    assert (cancelled, solved) == (False, 1)
    res = events[0]["result"]
    ref = CC.oracle_case(req, 500.0)
    q = {x["id"]: CC.wire_values(x) for x in res["quantities"]}
    assert np.allclose(q["fem_nodal_pressure"], ref["fem_pressure"].T, rtol=1e-12, atol=1e-12 * np.abs(ref["fem_pressure"]).max())
    assert np.allclose(q["bem_boundary_pressure"], ref["bem_pressure"].T, rtol=1e-12,
                       atol=1e-12 * np.abs(ref["bem_pressure"]).max())
    assert np.allclose(q["far"], ref["field"].T, rtol=1e-10, atol=1e-10 * np.abs(ref["field"]).max())
    if case == "piston":
        assert np.allclose(q["u"], ref["diaphragm_velocity"].T)
    d = res["diagnostics"]
    assert d["formulation"] == "fem_interface_condensed" and d["interface_ids"] == ["opening"]
    assert d["bounded_region_count"] == 1 and d["transducer_count"] == (1 if case == "piston" else 0)
    assert d["pressure_continuity_error"] < 1e-10
    json.dumps(res, allow_nan=False)


@pytest.mark.parametrize("mutate,message", [
    (lambda r: r["solver_options"].update(precision="float16"), "Unsupported coupled precision"),
    (lambda r: r["solver_options"].update(bem_backend="cuda"), "serves bem_backend 'rocm'"),
    (lambda r: r["solver_options"].update(validation_diagnostics=True, static_condensation=True), "cannot be combined"),
    (lambda r: r["compiled_system"]["regions"][0].update(density_kg_per_m3=1.0), "shared sound speed"),
    (lambda r: r["compiled_system"]["regions"][0]["loss_model"].update(bulk_loss_factor=2.0), "between 0 and 1"),
    (lambda r: r["compiled_system"]["boundaries"].pop(1), "unassigned boundary tag"),
    (lambda r: r["compiled_system"]["interfaces"][0].update(bounded_boundary_id="ext_open"),
     "does not reference a bounded FEM boundary"),
    (lambda r: r["compiled_system"]["interfaces"][0]["topology"]["fem_vertex_indices"].__setitem__(0, 10 ** 6),
     "outside the selected FEM volume groups"),
    (lambda r: r["outputs"].append({"id": "x", "quantity": "interior_pressure"}), "Unsupported coupled output"),
    (lambda r: r["outputs"].append({"id": "y", "quantity": "exterior_pressure", "options": {}}), "points_m"),
    (lambda r: r["compiled_system"]["excitation_ports"][0].update(kind="voltage"), "incompatible"),
    (lambda r: r.update(excitation_port_ids=[]), "at least one excitation port"),
    (lambda r: r["compiled_system"]["meshes"][0].update(purpose="bem_surface"), "must be a FEM volume"),
])
def test_protocol_errors(tmp_path, mutate, message):
    req = CC.coupled_request(tmp_path, "velocity", freqs=(500.0,))
    mutate(req)
    with pytest.raises(ValueError, match=message):
        SY.solve_system_request(req, None, lambda e: None, coupled_factory=OracleSolver)


@pytest.mark.parametrize("mutate,message", [
    (lambda p: p.pop("bl_n_per_a"), "is missing: bl_n_per_a"),
    (lambda p: p.update(extra=1), "unsupported parameters: extra"),
    (lambda p: p.update(motion_axis=[0, 0, 0]), "nonzero length"),
    (lambda p: p.update(re_ohm=0.0), "re_ohm must be greater than zero"),
    (lambda p: p.update(surface_completion_factor=2), "symmetry_role is inconsistent"),
    (lambda p: p.update(motion_profile="cone"), "rigid-translation"),
])
def test_transducer_parameter_errors(tmp_path, mutate, message):
    req = CC.coupled_request(tmp_path, "piston", freqs=(500.0,))
    mutate(req["compiled_system"]["components"][0]["parameters"])
    with pytest.raises(ValueError, match=message):
        SY.solve_system_request(req, None, lambda e: None, coupled_factory=OracleSolver)


def test_cancel_before_and_between_frequencies(tmp_path):
    req = CC.coupled_request(tmp_path, "velocity", freqs=(500.0, 600.0))
    flag = tmp_path / "cancel"
    req["cancel_path"] = str(flag)
    flag.write_text("x")
    assert SY.solve_system_request(req, None, lambda e: None, coupled_factory=OracleSolver) == (True, 0)
    flag.unlink()
    events = []

    def emit(e):
        events.append(e)
        flag.write_text("x")  # cancel after the first result

    assert SY.solve_system_request(req, None, emit, coupled_factory=OracleSolver) == (True, 1)
