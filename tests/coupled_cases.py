"""Coupled FEM-BEM test cases shared by the CPU protocol tests and the GPU parity tests: an air shell between two
spheres (FEM, Gmsh 4.1 file) radiating through its outer sphere (BEM, Gmsh 2.2 file, with an optional small
separate sphere), as compiled-system requests, and the oracle's solution of the same system."""

import json
import math

import numpy as np

from beat_hip import fem as FE
from beat_hip import synthetic as syn

RHO, C0 = 1.21, 343.0
EXTRA_CENTER = np.array([0.32, 0.0, 0.05])
PISTON = dict(re_ohm=6.0, le_h=0.0005, bl_n_per_a=7.0, mmd_kg=0.015, cms_m_per_n=0.0005, rms_n_s_per_m=1.0)


def shell_meshes(subdivisions=2, extra=True):
    """(VolumeMesh, (V, F, T) of the BEM surface): the outer sphere (tag 3) and, with `extra`, a sphere of radius
    0.05 at EXTRA_CENTER (tag 5)."""
    vm, (V, F, T) = syn.spherical_shell(subdivisions, 0.1, 0.15, 2)
    if extra:
        V2, F2, _ = syn.icosphere(1, 0.05)
        V = np.concatenate([V, V2 + EXTRA_CENTER[None, :]])
        F = np.concatenate([F, F2 + (len(V) - len(V2))])
        T = np.concatenate([T, np.full(len(F2), 5)])
    return vm, (V, F, T)


def observation_points(n=16, radius=1.5, seed=3):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(n, 3))
    return radius * p / np.linalg.norm(p, axis=1)[:, None]


def coupled_request(tmp_path, case="velocity", precision="float64", static=True, validation=False,
                    freqs=(400.0, 900.0), extra=True, outputs=None):
    """A compiled-system request for `case`:
    velocity  -- the inner sphere moves (ideal velocity source, weight 1) and, with `extra`, a second port drives
                 the small BEM sphere (weight 0.5); FEM bulk loss 0.02
    piston    -- the inner sphere is an electrodynamic rigid piston along z driven at the reference voltage
    wall      -- the inner sphere is a rigid Miki-lined wall; the small BEM sphere is the source
    """
    from beat_hip import geometry as G

    vm, (V, F, T) = shell_meshes(extra=extra or case == "wall")
    fem_path, bem_path = tmp_path / "shell.msh", tmp_path / "exterior.msh"
    FE.write_gmsh41_volume(fem_path, vm)
    syn.write_gmsh22(bem_path, V, F, T, scale=1.0)
    bm = G.BoundaryMesh(V, F, T)
    top = FE.build_conforming_interface_map(vm, bm, syn.TAG_OUTER, syn.TAG_OUTER)
    inner_kind = "rigid" if case == "wall" else "moving"
    inner = {"id": "inner", "region_id": "cavity", "kind": inner_kind, "group": {"mesh_id": "fem", "tag": syn.TAG_INNER}}
    if case == "wall":
        inner["parameters"] = {"wall_impedance": {"model": "miki", "thickness_m": 0.04,
                                                  "flow_resistivity_pa_s_per_m2": 8000.0}}
    boundaries = [inner,
                  {"id": "cav_open", "region_id": "cavity", "kind": "interface",
                   "group": {"mesh_id": "fem", "tag": syn.TAG_OUTER}},
                  {"id": "ext_open", "region_id": "air", "kind": "interface", "group": {"mesh_id": "bem", "tag": 3}}]
    components, ports = [], []
    if extra or case == "wall":
        boundaries.append({"id": "ext_src", "region_id": "air", "kind": "moving", "group": {"mesh_id": "bem", "tag": 5}})
    if case == "velocity":
        components.append({"id": "src", "kind": "ideal_velocity_source", "boundary_ids": ["inner"],
                           "parameters": {"boundary_motion_weights": {"inner": 1.0}}})
        ports.append({"id": "p_inner", "kind": "normal_velocity", "component_id": "src"})
    if case == "piston":
        components.append({"id": "drv", "kind": "electrodynamic_transducer", "boundary_ids": ["inner"],
                           "parameters": dict(PISTON, motion_axis=[0.0, 0.0, 1.0])})
        ports.append({"id": "p_drv", "kind": "voltage", "component_id": "drv"})
    if (extra and case != "piston") or case == "wall":
        components.append({"id": "ext", "kind": "ideal_velocity_source", "boundary_ids": ["ext_src"],
                           "parameters": {"boundary_motion_weights": {"ext_src": 0.5}}})
        ports.append({"id": "p_ext", "kind": "normal_velocity", "component_id": "ext"})
    pts = observation_points()
    if outputs is None:
        outputs = [{"id": q, "quantity": q} for q in ("fem_nodal_pressure", "bem_boundary_pressure",
                                                       "bem_boundary_neumann", "interface_normal_derivative")]
        if case == "piston":
            outputs += [{"id": "u", "quantity": "diaphragm_velocity"}, {"id": "i", "quantity": "voice_coil_current"}]
        outputs.append({"id": "far", "quantity": "exterior_pressure", "options": {"points_m": pts.tolist()}})
    return {
        "schema_version": 1,
        "frequencies_hz": list(freqs),
        "excitation_port_ids": [p["id"] for p in ports],
        "outputs": outputs,
        "solver_options": {"precision": precision, "bem_backend": "rocm", "quadrature_order": 2,
                           "singular_order": 2, "validation_diagnostics": validation, "static_condensation": static},
        "compiled_system": {
            "meshes": [{"id": "fem", "purpose": "fem_volume", "file": str(fem_path), "scale_to_m": 1.0,
                        "translation_m": [0.0, 0.0, 0.0]},
                       {"id": "bem", "purpose": "bem_surface", "file": str(bem_path), "scale_to_m": 1.0,
                        "translation_m": [0.0, 0.0, 0.0]}],
            "regions": [{"id": "cavity", "name": "cavity", "kind": "bounded_air", "mesh_ids": ["fem"],
                         "volume_groups": [{"mesh_id": "fem", "tag": syn.TAG_VOLUME}], "sound_speed_m_per_s": C0,
                         "density_kg_per_m3": RHO,
                         "loss_model": {"bulk_loss_factor": 0.02 if case == "velocity" else 0.0}},
                        {"id": "air", "name": "air", "kind": "unbounded_air", "mesh_ids": ["bem"],
                         "sound_speed_m_per_s": C0, "density_kg_per_m3": RHO}],
            "boundaries": boundaries,
            "components": components,
            "excitation_ports": ports,
            "interfaces": [{"id": "opening", "bounded_boundary_id": "cav_open", "unbounded_boundary_id": "ext_open",
                            "topology": {"fem_vertex_indices": top.fem_vertex_indices.tolist(),
                                         "fem_to_bem_vertex_indices": top.fem_to_bem_vertex_indices.tolist(),
                                         "fem_face_indices": top.fem_face_indices.tolist(),
                                         "bem_face_indices": top.bem_face_indices.tolist(),
                                         "normal_sign": top.normal_sign.tolist()}}],
        },
    }


def oracle_case(request, freq):
    """The oracle's monolithic FP64 solution of a coupled_request at one frequency: dict of blocks with one
    column per excitation port in request order, and the field at the exterior_pressure points."""
    from beat_hip import geometry as G
    from oracle import bem_oracle as O
    from oracle import coupled_oracle as CO

    sysd = request["compiled_system"]
    vm = FE.read_gmsh41_volume(sysd["meshes"][0]["file"])
    V, F, T = G.read_gmsh22(sysd["meshes"][1]["file"])
    # solver tags: 3 -> 1, 5 -> 2 (first appearance), as the unbounded region's aggregation numbers them
    tagmap = {}
    for t in T.tolist():
        tagmap.setdefault(t, len(tagmap) + 1)
    Ts = np.array([tagmap[t] for t in T.tolist()])
    bem = O.Mesh(V, F, Ts, np.zeros(len(F), np.int64))
    top = sysd["interfaces"][0]["topology"]
    im = dict(fem_vertices=np.array(top["fem_vertex_indices"]), bem_vertices=np.array(top["fem_to_bem_vertex_indices"]),
              fem_faces=np.array(top["fem_face_indices"]), bem_faces=np.array(top["bem_face_indices"]),
              sign=np.array(top["normal_sign"]))
    bounds = {b["id"]: b for b in sysd["boundaries"]}
    # FEM solver tags: the cavity's boundaries numbered in request order
    fem_tag = {}
    for b in sysd["boundaries"]:
        if b["region_id"] == "cavity":
            fem_tag[b["group"]["tag"]] = len(fem_tag) + 1
    btags = np.array([fem_tag[int(t)] for t in vm.boundary_tags])
    comps = {c["id"]: c for c in sysd["components"]}
    pistons, excs, pres_cols = [], [], []
    for c in sysd["components"]:
        if c["kind"] == "electrodynamic_transducer":
            p = c["parameters"]
            pistons.append(dict(fem_tags=[fem_tag[bounds[b]["group"]["tag"]] for b in c["boundary_ids"]],
                                fem_signs=[1.0] * len(c["boundary_ids"]), bem_tags=[], bem_signs=[],
                                axis=p["motion_axis"], completion=1.0, re=p["re_ohm"], le=p["le_h"],
                                bl=p["bl_n_per_a"], mmd=p["mmd_kg"], cms=p["cms_m_per_n"], rms=p["rms_n_s_per_m"]))
    for pid in request["excitation_port_ids"]:
        port = next(p for p in sysd["excitation_ports"] if p["id"] == pid)
        c = comps[port["component_id"]]
        if port["kind"] == "voltage":
            excs.append({"kind": "voltage", "transducer": 1, "amplitude": 2.83})
            continue
        w = c["parameters"]["boundary_motion_weights"]
        ftags, fw, src = [], [], 0
        for b in c["boundary_ids"]:
            g = bounds[b]["group"]
            if g["mesh_id"] == "fem":
                ftags.append(fem_tag[g["tag"]])
                fw.append(w[b])
            else:
                col = np.where(Ts == tagmap[g["tag"]], w[b], 0.0)
                pres_cols.append(col)
                src = len(pres_cols)
        excs.append({"kind": "normal_velocity", "fem_tags": ftags, "fem_weights": fw, "bem_source": src})
    prescribed = np.stack(pres_cols, axis=1) if pres_cols else None
    cav = sysd["regions"][0]
    loss = np.full(vm.n_vertices, cav["loss_model"]["bulk_loss_factor"])
    walls = []
    for b in sysd["boundaries"]:
        wi = (b.get("parameters") or {}).get("wall_impedance")
        if wi:
            walls.append((fem_tag[b["group"]["tag"]], wi["thickness_m"], wi["flow_resistivity_pa_s_per_m2"]))
    pts = None
    for o in request["outputs"]:
        if o["quantity"] == "exterior_pressure":
            pts = np.asarray(o["options"]["points_m"], float)
    return CO.solve_coupled(vm.vertices, vm.tetrahedra, vm.boundary_faces, btags, bem, im, freq, C0, RHO, excs,
                            pistons=pistons, prescribed=prescribed, loss_by_vertex=loss, walls=walls,
                            regular_order=request["solver_options"]["quadrature_order"],
                            singular_order=request["solver_options"]["singular_order"], points=pts)


def wire_values(q):
    v = q["values"]
    a = np.asarray(v["real"]) + 1j * np.asarray(v["imag"])
    return a.reshape(v["shape"])


def dump(request, path):
    with open(path, "w") as fh:
        json.dump(request, fh)
