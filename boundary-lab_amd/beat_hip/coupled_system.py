"""Coupled FEM-BEM system requests: the compiled-system wire of the reference's coupled worker for a system with
bounded (FEM) regions and one unbounded (BEM) region, solved by ``beat_hip.coupled``.

Reference: ``solve_request`` and its helpers (src/blab/solvers/julia_local/coupled_solver.jl:571-1070,
1465-2240): FEM domains aggregated from Gmsh 4.1 volumes with per-region bulk loss and Miki wall linings,
interface maps from the wire, electrodynamic transducers (rigid pistons with Thiele-Small parameters),
normal-velocity and voltage excitation ports, and the outputs ``fem_nodal_pressure``, ``bem_boundary_pressure``,
``bem_boundary_neumann``, ``interface_normal_derivative``, ``diaphragm_velocity``, ``voice_coil_current`` and
``exterior_pressure``; per-interface pressure-continuity and flux-conservation errors in the diagnostics.

Served with ``bem_backend="rocm"``: the formulation is the reference's ROCm one -- FEM interior condensed on the
host (``static_condensation``, the production default) or the monolithic system (``validation_diagnostics``,
which adds the full-system residual and the BEM-only replay error) -- with the dense system on the GPU.
"""

from __future__ import annotations

import math
import os
import time

import numpy as np
import scipy.sparse as sp

from . import fem as FE
from . import geometry as G
from .system import SCHEMA_VERSION, aggregate_bem_region, object_by_id, quantity_wire

DEFAULT_TRANSDUCER_REFERENCE_VOLTAGE_V = 2.83


def _close(a, b):
    return math.isclose(float(a), float(b), rel_tol=1e-9, abs_tol=0.0)


def aggregate_fem_domains(meshes, bounded_regions, boundaries):
    """FEM volumes of the bounded regions, restricted to their volume groups and concatenated, with solver
    boundary tags numbered 1, 2, ... over the regions' boundaries, per-vertex bulk loss and wall linings
    (coupled_solver.jl:626-803)."""
    verts, tets, ttags, bfaces, btags, loss = [], [], [], [], [], []
    domains, walls, terminations = [], [], []
    tag_by_id, domain_by_boundary = {}, {}
    next_tag = 1
    vo = fo = 0
    quadratic = None
    for region in bounded_regions:
        mids = [str(m) for m in region.get("mesh_ids", [])]
        if len(mids) != 1:
            raise ValueError("Each bounded region must currently contain exactly one FEM mesh.")
        mid = mids[0]
        res = object_by_id(meshes, mid, "mesh")
        if str(res["purpose"]) != "fem_volume":
            raise ValueError("Bounded region mesh must be a FEM volume.")
        full = FE.read_gmsh41_volume(str(res["file"]), scale=float(res["scale_to_m"]),
                                     translation=res.get("translation_m", (0.0, 0.0, 0.0)))
        sel_tags = [int(g["tag"]) for g in region.get("volume_groups", []) if str(g["mesh_id"]) == mid]
        mesh, vmap, fmap = FE.restrict_volume_mesh(full, sel_tags)
        if quadratic is None:
            quadratic = mesh.is_quadratic
        elif quadratic != mesh.is_quadratic:
            raise ValueError("All bounded FEM regions must use the same tetrahedron order.")
        bl = float((region.get("loss_model") or {}).get("bulk_loss_factor", 0.0))
        if not (math.isfinite(bl) and 0.0 <= bl <= 1.0):
            raise ValueError("Bounded-region FEM bulk loss factor must be finite and between 0 and 1.")
        local = {}
        for b in boundaries:
            if str(b["region_id"]) != str(region["id"]) or str(b["group"]["mesh_id"]) != mid:
                continue
            src = int(b["group"]["tag"])
            if src in local:
                raise ValueError(f"FEM physical tag {src} on mesh {mid!r} is assigned more than once.")
            local[src] = next_tag
            bid = str(b["id"])
            tag_by_id[bid] = next_tag
            domain_by_boundary[bid] = len(domains)
            params = b.get("parameters") or {}
            kind = str(b["kind"])
            if kind == "plane_wave_tube_termination":
                if params:
                    raise ValueError(f"Plane-wave tube termination {bid!r} does not accept parameters.")
                terminations.append({"boundary_id": bid, "tag": next_tag})
            if "wall_impedance" in params:
                if kind != "rigid":
                    raise ValueError(f"Wall impedance boundary {bid!r} must be rigid.")
                tr = params["wall_impedance"]
                if str(tr.get("model", "miki")) != "miki":
                    raise ValueError(f"Unsupported wall impedance model on {bid!r}.")
                th = float(tr.get("thickness_m", 0.03))
                sig = float(tr.get("flow_resistivity_pa_s_per_m2", 5000.0))
                if not (math.isfinite(th) and th > 0.0):
                    raise ValueError("Wall-lining thickness must be finite and positive.")
                if not (math.isfinite(sig) and sig > 0.0):
                    raise ValueError("Wall-lining airflow resistivity must be finite and positive.")
                walls.append({"boundary_id": bid, "tag": next_tag, "thickness_m": th,
                              "flow_resistivity_pa_s_per_m2": sig})
            next_tag += 1
        remapped = np.array([local.get(int(t), 0) for t in mesh.boundary_tags], dtype=np.int64)
        if np.any(remapped == 0):
            raise ValueError(f"Selected FEM region {str(region['id'])!r} contains an unassigned boundary tag.")
        verts.append(mesh.vertices)
        tets.append(mesh.tetrahedra + vo)
        ttags.append(mesh.tetra_tags)
        bfaces.append(mesh.boundary_faces + vo)
        btags.append(remapped)
        loss.append(np.full(mesh.n_vertices, bl))
        domains.append({"id": str(region["id"]), "mesh_id": mid, "vertex_map": vmap, "face_map": fmap,
                        "vertex_offset": vo, "face_offset": fo, "vertex_count": mesh.n_vertices})
        vo += mesh.n_vertices
        fo += len(mesh.boundary_faces)
    if quadratic:
        raise ValueError("Quadratic tetrahedra are currently supported only by the pure interior FEM solve.")
    vm = FE.VolumeMesh(np.concatenate(verts), np.concatenate(tets), np.concatenate(ttags), np.concatenate(bfaces),
                       np.concatenate(btags), {})
    return {"mesh": vm, "domains": domains, "fem_boundary_tag_by_id": tag_by_id,
            "domain_by_boundary_id": domain_by_boundary, "bulk_loss_by_vertex": np.concatenate(loss),
            "wall_impedances": walls, "plane_wave_terminations": terminations}


def _remapped(index_map, wire_index, label):
    v = index_map.get(int(wire_index), -1)
    if v < 0:
        raise ValueError(f"{label} lies outside the selected FEM volume groups.")
    return v


def combined_interface_map(interfaces, fem_dom, boundaries, bem_dom):
    """Per-interface maps from the wire topology (0-based indices into the original FEM mesh and the BEM mesh of
    the unbounded boundary), offset into the aggregated meshes and concatenated; the flux dofs of interface j
    are ranges[j] (coupled_solver.jl:102-123, 805-869)."""
    maps, ranges = [], []
    nxt = 0
    for itf in interfaces:
        iid = str(itf["id"])
        d = fem_dom["domain_by_boundary_id"].get(str(itf["bounded_boundary_id"]))
        if d is None:
            raise ValueError(f"Interface {iid!r} does not reference a bounded FEM boundary.")
        dom = fem_dom["domains"][d]
        top = itf["topology"]
        local = FE.InterfaceMap(
            [_remapped(dom["vertex_map"], i, "FEM interface vertex") for i in top["fem_vertex_indices"]],
            [int(i) for i in top["fem_to_bem_vertex_indices"]],
            [_remapped(dom["face_map"], i, "FEM interface face") for i in top["fem_face_indices"]],
            [int(i) for i in top["bem_face_indices"]],
            [int(s) for s in top["normal_sign"]])
        ub = object_by_id(boundaries, str(itf["unbounded_boundary_id"]), "unbounded boundary")
        bmid = str(ub["group"]["mesh_id"])
        bvo = bem_dom["vertex_offset_by_mesh_id"].get(bmid, -1)
        bfo = bem_dom["face_offset_by_mesh_id"].get(bmid, -1)
        if bvo < 0 or bfo < 0:
            raise ValueError(f"Interface {iid!r} references BEM mesh {bmid!r} outside the unbounded region.")
        m = local.offset(dom["vertex_offset"], dom["face_offset"], bvo, bfo)
        maps.append(m)
        ranges.append(range(nxt, nxt + m.count))
        nxt += m.count
    return FE.InterfaceMap.concatenate(maps), maps, ranges


def electrodynamic_transducers(components, boundaries, fem_tag_by_id, bem_tag_by_id, fem_mesh, bem_mesh, symmetry):
    """Electrodynamic components as rigid pistons with their parameters checked as the reference checks them
    (coupled_solver.jl:871-1069). Returns (transducers, index_by_component_id (1-based))."""
    out, index = [], {}
    scalar = ("re_ohm", "le_h", "bl_n_per_a", "mmd_kg", "cms_m_per_n", "rms_n_s_per_m")
    required = scalar + ("motion_axis",)
    optional = ("motion_profile", "boundary_motion_signs", "boundary_motion_weights", "symmetry_role",
                "surface_completion_factor", "physical_driver_orbit_count", "fractional_symmetry_axes")
    for comp in components:
        if str(comp["kind"]) != "electrodynamic_transducer":
            continue
        cid = str(comp["id"])
        p = comp.get("parameters") or {}
        missing = [n for n in required if n not in p]
        if missing:
            raise ValueError(f"Electrodynamic component {cid!r} is missing: " + ", ".join(missing))
        unsupported = sorted(str(n) for n in p if str(n) not in required and str(n) not in optional)
        if unsupported:
            raise ValueError(f"Electrodynamic component {cid!r} has unsupported parameters: " + ", ".join(unsupported))
        if str(p.get("motion_profile", "rigid_translation")) != "rigid_translation":
            raise ValueError("Electrodynamic components currently require a rigid-translation piston motion profile.")
        axis = np.asarray(p["motion_axis"], dtype=np.float64).reshape(-1)
        if axis.size != 3:
            raise ValueError("motion_axis must contain exactly three values.")
        if not np.all(np.isfinite(axis)):
            raise ValueError("motion_axis must contain finite values.")
        na = float(np.linalg.norm(axis))
        if not na > 0.0:
            raise ValueError("motion_axis must have nonzero length.")
        axis = axis / na
        completion = int(p.get("surface_completion_factor", 1))
        orbits = int(p.get("physical_driver_orbit_count", 1))
        if completion not in (1, 2, 4):
            raise ValueError("surface_completion_factor must be 1, 2, or 4.")
        if orbits not in (1, 2, 4):
            raise ValueError("physical_driver_orbit_count must be 1, 2, or 4.")
        role = str(p.get("symmetry_role", "complete_representative"))
        if role != ("fractional_driver" if completion > 1 else "complete_representative"):
            raise ValueError("symmetry_role is inconsistent with surface_completion_factor.")
        frac = [str(a) for a in p.get("fractional_symmetry_axes", [])]
        if len(frac) != len(set(frac)):
            raise ValueError("fractional_symmetry_axes must not contain duplicates.")
        if not all(a in ("x", "y") for a in frac):
            raise ValueError("fractional_symmetry_axes may contain only x and y.")
        active = {"off": [], "x": ["x"], "xy": ["x", "y"]}[symmetry]
        if not all(a in active for a in frac):
            raise ValueError("fractional_symmetry_axes must be active in the selected symmetry mode.")
        if completion != 2 ** len(frac):
            raise ValueError("surface_completion_factor must equal 2 raised to the number of fractional_symmetry_axes.")
        for a in frac:
            if abs(axis[0 if a == "x" else 1]) > 1e-8:
                raise ValueError("motion_axis must lie in every symmetry plane that cuts the physical driver.")
        represented = completion * orbits
        expected = G.symmetry_reduction_factor(symmetry)
        if represented != expected:
            raise ValueError(f"Electrodynamic component {cid!r} represents {represented} symmetry images, but "
                             f"{symmetry} symmetry requires {expected}.")
        signs = p.get("boundary_motion_signs") or {}
        weights = p.get("boundary_motion_weights") or {}
        own = set(str(b) for b in comp.get("boundary_ids", []))
        bad = sorted(str(b) for b in signs if str(b) not in own)
        if bad:
            raise ValueError(f"Electrodynamic component {cid!r} has motion signs for unrelated boundaries: "
                             + ", ".join(bad))
        bad = sorted(str(b) for b in weights if str(b) not in own)
        if bad:
            raise ValueError(f"Electrodynamic component {cid!r} has motion weights for unrelated boundaries: "
                             + ", ".join(bad))
        ft, fs, bt, bs = [], [], [], []
        for bid in [str(b) for b in comp.get("boundary_ids", [])]:
            b = object_by_id(boundaries, bid, "boundary")
            if str(b["kind"]) != "moving":
                raise ValueError(f"Electrodynamic component {cid!r} boundary {bid!r} is not moving.")
            sign = float(signs.get(bid, 1))
            if sign not in (-1.0, 1.0):
                raise ValueError("Electrodynamic boundary motion signs must be -1 or +1.")
            w = float(weights.get(bid, 1))
            if not (math.isfinite(w) and w > 0.0):
                raise ValueError("Electrodynamic boundary motion weights must be finite and greater than zero.")
            if bid in fem_tag_by_id:
                tag = fem_tag_by_id[bid]
                if not np.any(fem_mesh.boundary_tags == tag):
                    raise ValueError(f"Electrodynamic FEM boundary tag {tag} is outside the selected volume groups.")
                ft.append(tag)
                fs.append(sign * w)
            elif bid in bem_tag_by_id:
                tag = bem_tag_by_id[bid]
                if not np.any(bem_mesh.tags == tag):
                    raise ValueError(f"Electrodynamic BEM boundary tag {tag} is not present in the exterior mesh.")
                bt.append(tag)
                bs.append(sign * w)
            else:
                raise ValueError(f"Electrodynamic component {cid!r} references a boundary outside the active "
                                 "acoustic regions.")
        if not ft and not bt:
            raise ValueError(f"Electrodynamic component {cid!r} has no moving acoustic boundary.")
        vals = {n: float(p[n]) for n in scalar}
        if not all(math.isfinite(v) for v in vals.values()):
            raise ValueError("Electrodynamic parameters must be finite numbers.")
        for n, ok, msg in (("re_ohm", vals["re_ohm"] > 0, "re_ohm must be greater than zero."),
                           ("le_h", vals["le_h"] >= 0, "le_h must not be negative."),
                           ("bl_n_per_a", vals["bl_n_per_a"] > 0, "bl_n_per_a must be greater than zero."),
                           ("mmd_kg", vals["mmd_kg"] > 0, "mmd_kg must be greater than zero."),
                           ("cms_m_per_n", vals["cms_m_per_n"] > 0, "cms_m_per_n must be greater than zero."),
                           ("rms_n_s_per_m", vals["rms_n_s_per_m"] >= 0, "rms_n_s_per_m must not be negative.")):
            if not ok:
                raise ValueError(msg)
        out.append(FE.Transducer(cid, ft, fs, bt, bs, axis, float(completion), orbits, vals["re_ohm"], vals["le_h"],
                                 vals["bl_n_per_a"], vals["mmd_kg"], vals["cms_m_per_n"], vals["rms_n_s_per_m"]))
        index[cid] = len(out)
    return out, index


def parse_coupled_request(request):
    """Everything solve_request prepares before its frequency loop (coupled_solver.jl:1465-1795), for a system
    with bounded regions and one unbounded region. Returns the job dict."""
    from .coupled import CoupledModel, Excitation

    if int(request.get("schema_version", 0)) != SCHEMA_VERSION:
        raise ValueError("Unsupported system solve request schema.")
    system = request["compiled_system"]
    meshes, regions = system.get("meshes", []), system.get("regions", [])
    boundaries, components = system.get("boundaries", []), system.get("components", [])
    ports, interfaces = system.get("excitation_ports", []), system.get("interfaces", [])
    bounded = [r for r in regions if str(r["kind"]) == "bounded_air"]
    unbounded = [r for r in regions if str(r["kind"]) == "unbounded_air"]
    if not unbounded:
        raise ValueError("Interior (FEM-only) system solves are not served by the HIP backend.")
    if len(unbounded) != 1:
        raise ValueError("Coupled backend currently requires exactly one unbounded region.")
    ub = unbounded[0]
    c0, rho0 = float(bounded[0]["sound_speed_m_per_s"]), float(bounded[0]["density_kg_per_m3"])
    for r in regions:
        if not (_close(r["sound_speed_m_per_s"], c0) and _close(r["density_kg_per_m3"], rho0)):
            raise ValueError("Coupled backend currently requires one shared sound speed and density across all "
                             "regions.")
    opts = request.get("solver_options") or {}
    pname = str(opts.get("precision", "float64")).lower()
    if pname in ("float32", "complex64"):
        precision = "float32"
    elif pname in ("float64", "complex128"):
        precision = "float64"
    else:
        raise ValueError(f"Unsupported coupled precision: {pname}. Expected float32 or float64.")
    backend = str(opts.get("bem_backend", "cpu")).lower()
    if backend not in ("cpu", "cuda", "rocm"):
        raise ValueError(f"Unsupported coupled BEM backend: {backend}. Expected cpu, cuda, or rocm.")
    if backend != "rocm":
        raise ValueError(f"The HIP coupled backend serves bem_backend 'rocm', not {backend!r}.")
    symmetry = G.normalized_symmetry_mode(opts.get("symmetry", "off"))
    vref = float(opts.get("transducer_reference_voltage_v", DEFAULT_TRANSDUCER_REFERENCE_VOLTAGE_V))
    if not vref > 0.0:
        raise ValueError("transducer_reference_voltage_v must be greater than zero.")
    fdt = np.float32 if precision == "float32" else np.float64
    t0 = time.perf_counter()
    fem_dom = aggregate_fem_domains(meshes, bounded, boundaries)
    bem_dom = aggregate_bem_region(meshes, ub, boundaries)
    fm, bm = fem_dom["mesh"], bem_dom["mesh"]
    tol = max(G.symmetry_plane_tolerance(fm.vertices), G.symmetry_plane_tolerance(bm.vertices))
    axes = {"off": (), "x": (0,), "xy": (0, 1)}[symmetry]
    if axes:
        fv = fm.vertices.copy()
        bv = np.array(bm.vertices, dtype=np.float64)
        for a in axes:
            fv[np.abs(fv[:, a]) <= tol, a] = 0.0
            bv[np.abs(bv[:, a]) <= tol, a] = 0.0
        fm.vertices = fv
        bm = G.BoundaryMesh(bv, bm.faces, bm.tags, bm.mesh_ids)
        for a in axes:
            i = int(np.argmin(fv[:, a]))
            if fv[i, a] < -tol:
                raise ValueError(f"FEM mesh is not in the positive {'XY'[a]} fundamental domain for "
                                 f"{symmetry.upper()} symmetry. Vertex {i + 1} has coordinate {fv[i, a]} m.")
        G.validate_symmetry_domain(bm, symmetry)
    imap, maps, ranges = combined_interface_map(interfaces, fem_dom, boundaries, bem_dom)
    mesh_setup_s = time.perf_counter() - t0
    tds, td_index = electrodynamic_transducers(components, boundaries, fem_dom["fem_boundary_tag_by_id"],
                                               bem_dom["boundary_tag_by_id"], fm, bm, symmetry)
    port_ids = [str(p) for p in request.get("excitation_port_ids", [])]
    if not port_ids:
        raise ValueError("Coupled solve requires at least one excitation port.")
    excitations = []
    pr, pc, pv = [], [], []
    n_pres = 0
    for port_id in port_ids:
        port = object_by_id(ports, port_id, "excitation port")
        comp = object_by_id(components, str(port["component_id"]), "component")
        pk, ck = str(port["kind"]), str(comp["kind"])
        if pk == "normal_velocity" and ck == "ideal_velocity_source":
            weights = (comp.get("parameters") or {}).get("boundary_motion_weights", {}) or {}
            ftags, fw, btags, bw = [], [], [], []
            for bid in [str(b) for b in comp.get("boundary_ids", [])]:
                b = object_by_id(boundaries, bid, "boundary")
                if str(b["kind"]) != "moving":
                    raise ValueError(f"Prescribed-velocity component {str(comp['id'])!r} references non-moving "
                                     f"boundary {bid!r}.")
                w = float(weights.get(bid, 1))
                if not (math.isfinite(w) and w > 0.0):
                    raise ValueError("Prescribed-velocity boundary motion weights must be finite and greater than "
                                     "zero.")
                if bid in fem_dom["fem_boundary_tag_by_id"]:
                    tag = fem_dom["fem_boundary_tag_by_id"][bid]
                    if not np.any(fm.boundary_tags == tag):
                        raise ValueError(f"Moving boundary tag {tag} is not on the selected FEM volume groups.")
                    ftags.append(tag)
                    fw.append(w)
                elif bid in bem_dom["boundary_tag_by_id"]:
                    tag = bem_dom["boundary_tag_by_id"][bid]
                    if not np.any(bm.tags == tag):
                        raise ValueError(f"Moving boundary tag {tag} is not present in the exterior BEM mesh.")
                    btags.append(tag)
                    bw.append(w)
                else:
                    raise ValueError(f"Prescribed-velocity component {str(comp['id'])!r} references boundary "
                                     f"{bid!r} outside the active acoustic regions.")
            if not ftags and not btags:
                raise ValueError(f"Prescribed-velocity component {str(comp['id'])!r} must own at least one moving "
                                 "FEM or BEM boundary.")
            src = 0
            if btags:
                n_pres += 1
                src = n_pres
                wmap = dict(zip(btags, bw))
                for f in range(bm.n_faces):
                    w = wmap.get(int(bm.tags[f]), 0.0)
                    if w != 0.0:
                        pr.append(f)
                        pc.append(src - 1)
                        pv.append(w)
            excitations.append(Excitation("normal_velocity", tuple(ftags), tuple(fw), src, 0, 1.0))
        elif pk == "voltage" and ck == "electrodynamic_transducer":
            ti = td_index.get(str(comp["id"]), 0)
            if ti <= 0:
                raise ValueError(f"Voltage port {port_id} references an unresolved electrodynamic component.")
            excitations.append(Excitation("voltage", (), (), 0, ti, complex(fdt(vref))))
        else:
            raise ValueError(f"Excitation port {port_id} kind {pk!r} is incompatible with component kind {ck!r}.")
    prescribed = sp.csr_matrix((np.asarray(pv, float), (np.asarray(pr, np.int64), np.asarray(pc, np.int64))),
                               shape=(bm.n_faces, n_pres))
    validation = bool(opts.get("validation_diagnostics", True))
    static = bool(opts.get("static_condensation", not validation))
    if static and validation:
        raise ValueError("FEM static condensation cannot be combined with full-matrix validation diagnostics.")
    outputs = list(request.get("outputs", []))
    for out in outputs:
        q = str(out["quantity"])
        if q == "exterior_pressure":
            if not (out.get("options") or {}).get("points_m"):
                raise ValueError("exterior_pressure output requires options.points_m.")
        elif q not in ("fem_nodal_pressure", "bem_boundary_pressure", "bem_boundary_neumann",
                       "interface_normal_derivative", "diaphragm_velocity", "voice_coil_current"):
            raise ValueError(f"Unsupported coupled output quantity: {q}")
    model = CoupledModel(fm, bm, imap, float(fdt(rho0)), float(fdt(c0)), tds, prescribed,
                         fem_dom["bulk_loss_by_vertex"], fem_dom["wall_impedances"])
    mesh_ids = [str(m) for m in ub.get("mesh_ids", [])]
    nxt_v = sorted(set(bem_dom["vertex_offset_by_mesh_id"].values())) + [bm.n_vertices]
    nxt_f = sorted(set(bem_dom["face_offset_by_mesh_id"].values())) + [bm.n_faces]

    def count_after(offs, table, off):
        return table[table.index(off) + 1] - off

    return {
        "model": model, "precision": precision, "symmetry": symmetry, "static_condensation": static,
        "validation_diagnostics": validation, "static_condensation_requested": static,
        "regular_order": int(opts.get("quadrature_order", 2)), "singular_order": int(opts.get("singular_order", 2)),
        "excitations": excitations, "port_ids": port_ids, "outputs": outputs, "interfaces": interfaces,
        "interface_maps": maps, "interface_ranges": ranges, "fem_domains": fem_dom, "bem_domain": bem_dom,
        "unbounded_region": ub, "bounded_regions": bounded, "transducers": tds, "reference_voltage": vref,
        "mesh_setup_s": mesh_setup_s, "complex_dtype": np.complex64 if precision == "float32" else np.complex128,
        "frequencies": [float(fdt(f)) for f in request.get("frequencies_hz", [])],
        "cancel_path": request.get("cancel_path"), "bem_mesh_ids": mesh_ids,
        "bem_vertex_counts": [count_after(None, nxt_v, bem_dom["vertex_offset_by_mesh_id"][m]) for m in mesh_ids],
        "bem_face_counts": [count_after(None, nxt_f, bem_dom["face_offset_by_mesh_id"][m]) for m in mesh_ids],
    }


def per_interface_errors(sol, fem_mesh, bem_mesh, maps, ranges):
    """Pressure continuity and integrated flux conservation per interface (coupled_solver.jl:571-624)."""
    perr, ferr = [], []
    for im, rg in zip(maps, ranges):
        tf = sol.fem_pressure[im.fem_vertex_indices]
        tb = sol.bem_pressure[im.fem_to_bem_vertex_indices]
        scale = max(np.linalg.norm(tf), np.linalg.norm(tb), np.finfo(float).eps)
        perr.append(float(np.linalg.norm(tf - tb) / scale))
        flux = sol.interface_flux[list(rg)]
        dof = {int(v): i for i, v in enumerate(im.fem_vertex_indices)}
        fi = bi = 0.0
        for r, f in enumerate(im.fem_face_indices):
            face = fem_mesh.boundary_faces[f]
            avg = sum(flux[dof[int(v)]] for v in face) / 3.0
            fi += FE.triangle_areas(fem_mesh.vertices, face[None, :])[0] * avg
            b = im.bem_face_indices[r]
            bi += im.normal_sign[r] * bem_mesh.areas[b] * sol.bem_neumann[b]
        fs = max(abs(fi), abs(bi), np.finfo(float).eps)
        ferr.append(float(abs(fi - bi) / fs))
    return perr, ferr


def _rows(vectors, cdt):
    return np.stack([np.asarray(v) for v in vectors]).astype(cdt) if vectors else np.zeros((0, 0), cdt)


def frequency_result(job, freq, sols, field, ys, tim, first, solver) -> dict:
    """The result object of one frequency (coupled_solver.jl:1876-2222)."""
    cdt = job["complex_dtype"]
    model = job["model"]
    fd, bd = job["fem_domains"], job["bem_domain"]
    sets = [per_interface_errors(s, model.fem_mesh, model.bem_mesh, job["interface_maps"], job["interface_ranges"])
            for s in sols]
    n_if = len(job["interfaces"])
    pe = [max(e[0][i] for e in sets) for i in range(n_if)]
    fe = [max(e[1][i] for e in sets) for i in range(n_if)]
    quantities = []
    offset = 0
    tds = job["transducers"]
    td_meta = {"component_ids": [t.id for t in tds],
               "surface_completion_factors": [float(t.surface_completion_factor) for t in tds],
               "physical_driver_orbit_counts": [int(t.physical_driver_orbit_count) for t in tds]}
    for out in job["outputs"]:
        q = str(out["quantity"])
        if q == "fem_nodal_pressure":
            quantities.append(quantity_wire(out, _rows([s.fem_pressure for s in sols], cdt), "Pa",
                                            ["excitation", "fem_node"], cdt, metadata={
                                                "mesh_ids": [d["mesh_id"] for d in fd["domains"]],
                                                "region_ids": [d["id"] for d in fd["domains"]],
                                                "node_offsets": [d["vertex_offset"] for d in fd["domains"]],
                                                "node_counts": [d["vertex_count"] for d in fd["domains"]]}))
        elif q == "bem_boundary_pressure":
            quantities.append(quantity_wire(out, _rows([s.bem_pressure for s in sols], cdt), "Pa",
                                            ["excitation", "bem_node"], cdt, metadata={
                                                "mesh_ids": job["bem_mesh_ids"],
                                                "vertex_offsets": [bd["vertex_offset_by_mesh_id"][m]
                                                                   for m in job["bem_mesh_ids"]],
                                                "vertex_counts": job["bem_vertex_counts"]}))
        elif q == "bem_boundary_neumann":
            quantities.append(quantity_wire(out, _rows([s.bem_neumann for s in sols], cdt), "Pa/m",
                                            ["excitation", "bem_face"], cdt, metadata={
                                                "mesh_ids": job["bem_mesh_ids"],
                                                "face_offsets": [bd["face_offset_by_mesh_id"][m]
                                                                 for m in job["bem_mesh_ids"]],
                                                "face_counts": job["bem_face_counts"]}))
        elif q == "interface_normal_derivative":
            quantities.append(quantity_wire(out, _rows([s.interface_flux for s in sols], cdt), "Pa/m",
                                            ["excitation", "interface_node"], cdt, metadata={
                                                "interface_ids": [str(i["id"]) for i in job["interfaces"]],
                                                "interface_offsets": [r.start for r in job["interface_ranges"]],
                                                "interface_counts": [len(r) for r in job["interface_ranges"]]}))
        elif q == "diaphragm_velocity":
            quantities.append(quantity_wire(out, _rows([s.diaphragm_velocity for s in sols], cdt), "m/s",
                                            ["excitation", "transducer"], cdt, metadata=dict(td_meta)))
        elif q == "voice_coil_current":
            quantities.append(quantity_wire(out, _rows([s.voice_coil_current for s in sols], cdt), "A",
                                            ["excitation", "transducer"], cdt, metadata=dict(td_meta)))
        elif q == "exterior_pressure":
            opts = out.get("options") or {}
            pts = opts.get("points_m", [])
            n = len(pts)
            quantities.append(quantity_wire(out, field[offset : offset + n].T, "Pa", ["excitation", "observation"],
                                            cdt, metadata={"points_m": pts,
                                                           "observation_domains": opts.get("observation_domains", [])}))
            offset += n
    static = bool(getattr(solver, "condensed", job["static_condensation"]))
    t = {
        "assembly_s": tim.get("assembly_s", 0.0), "solve_s": tim.get("solve_s", 0.0), "field_s": tim.get("field_s", 0.0),
        "mesh_setup_s": job["mesh_setup_s"] if first else 0.0,
        "cache_setup_s": sum(solver.timings.values()) if first else 0.0,
        "fem_matrix_cache_s": solver.timings.get("fem_matrix_cache_s", 0.0) if first else 0.0,
        "interface_operator_cache_s": solver.timings.get("interface_operator_cache_s", 0.0) if first else 0.0,
        "bem_device_regular_cache_s": solver.timings.get("bem_device_cache_s", 0.0) if first else 0.0,
        "fem_system_s": tim.get("fem_system_s", 0.0), "bem_operator_s": tim.get("bem_operator_s", 0.0),
        "bem_matrix_s": tim.get("bem_matrix_s", 0.0), "fem_condensation_s": tim.get("fem_condensation_s", 0.0),
        "fem_condensation_analysis_s": 0.0,
        "fem_condensation_factorization_s": tim.get("fem_condensation_factorization_s", 0.0),
        "fem_condensation_partition_s": tim.get("fem_condensation_partition_s", 0.0),
        "fem_schur_extraction_s": tim.get("fem_condensation_schur_extraction_s", 0.0),
        "fem_rhs_condensation_s": max((s.fem_rhs_condensation_s or 0.0) for s in sols),
        "fem_reconstruction_s": max((s.fem_reconstruction_s or 0.0) for s in sols),
        "block_assembly_s": tim.get("block_assembly_s", 0.0),
        "coupled_factorization_s": tim.get("coupled_factorization_s", 0.0), "replay_factorization_s": 0.0,
    }
    diagnostics = {
        "precision": job["precision"], "bem_backend": "rocm", "linear_backend": "rocm",
        "symmetry": job["symmetry"],
        "formulation": "fem_interface_condensed" if static else "monolithic",
        "linear_solver": ("hip_cpu_superlu_schur_plus_hip_dense_lu" if static else "hip_dense_lu")
        + ("_fp64_refinement" if job["precision"] == "float32" else ""),
        "full_system_order": solver.n_fem + solver.V + solver.n_if + 2 * solver.nt,
        "solved_system_order": solver.n,
        "bounded_region_count": len(job["bounded_regions"]), "interface_count": n_if,
        "transducer_count": len(tds), "transducer_reference_voltage_v": job["reference_voltage"],
        "fem_bulk_loss_factors_by_region": {str(r["id"]): float((r.get("loss_model") or {}).get("bulk_loss_factor", 0.0))
                                            for r in job["bounded_regions"]},
        "wall_impedance_boundary_ids": [w["boundary_id"] for w in job["fem_domains"]["wall_impedances"]],
        "static_condensation_requested": job["static_condensation_requested"],
        "static_condensation_active": static,
        "fem_condensation_backend": "cpu_superlu" if static else None,
        "fem_schur_block_size": 256 if static else 0, "fem_schur_thread_count": 1 if static else 0,
        "pressure_continuity_error": max(pe) if pe else None,
        "flux_conservation_error": max(fe) if fe else None,
        "interface_ids": [str(i["id"]) for i in job["interfaces"]],
        "interface_pressure_continuity_errors": pe, "interface_flux_conservation_errors": fe,
        "regular_quadrature_order": job["regular_order"], "singular_quadrature_order": job["singular_order"],
        "relative_residual": max(s.relative_residual for s in sols),
        "timings": t,
    }
    if job["validation_diagnostics"]:
        diagnostics["all_bem_replay_error"] = max(s.all_bem_replay_error for s in sols)
    diagnostics["fem_interior_residual"] = (max(s.fem_interior_residual for s in sols)
                                            if static and sols[0].fem_interior_residual is not None else None)
    return {"schema_version": 1, "freq_hz": float(freq), "excitation_port_ids": list(job["port_ids"]),
            "quantities": quantities, "diagnostics": diagnostics}


def _cancelled(job) -> bool:
    return job["cancel_path"] is not None and os.path.isfile(str(job["cancel_path"]))


def solve_coupled_request(request, emit, event_mode: bool = True, dev=None, solver_factory=None):
    """Run a coupled FEM-BEM compiled-system request: one ``result`` per frequency (wrapped as an event in worker
    mode), returns (cancelled, solved_count) (coupled_solver.jl:1465-2240). solver_factory(model, **options)
    builds the CoupledSolver (tests inject their own); dev is the device it uses."""
    from .coupled import CoupledSolver

    if _cancelled({"cancel_path": request.get("cancel_path")}):
        return True, 0
    job = parse_coupled_request(request)
    if _cancelled(job):
        return True, 0
    points = []
    for out in job["outputs"]:
        if str(out["quantity"]) == "exterior_pressure":
            points.append(np.asarray(out["options"]["points_m"], dtype=np.float64).reshape(-1, 3))
    pts = np.concatenate(points) if points else None
    opts = dict(precision=job["precision"], static_condensation=job["static_condensation"],
                validation_diagnostics=job["validation_diagnostics"], regular_order=job["regular_order"],
                singular_order=job["singular_order"], symmetry=job["symmetry"])
    if solver_factory is None:
        solver = CoupledSolver(job["model"], dev=dev, **opts)
    else:
        solver = solver_factory(job["model"], **opts)
    solved = 0
    freqs = job["frequencies"]
    # frequencies go to the device in batches (one batched assembly, LU and refinement per batch)
    nb = max(1, int(solver.batch_capacity())) if hasattr(solver, "batch_capacity") else 1
    try:
        i = 0
        while i < len(freqs):
            if _cancelled(job):
                return True, solved
            chunk = freqs[i : i + nb]
            t0 = time.perf_counter()
            if hasattr(solver, "solve_batch"):
                results = solver.solve_batch(chunk, job["excitations"], pts)
            else:
                results = [solver.solve(f, job["excitations"], pts) for f in chunk]
            wall = (time.perf_counter() - t0) / len(chunk)
            for j, (sols, field, ys, tim) in enumerate(results):
                if _cancelled(job):
                    return True, solved
                tim.setdefault("assembly_s", tim.get("fem_system_s", 0.0) + tim.get("fem_condensation_s", 0.0)
                               + tim.get("bem_operator_s", 0.0) + tim.get("bem_matrix_s", 0.0)
                               + tim.get("block_assembly_s", 0.0))
                tim.setdefault("solve_s", max(0.0, wall - tim["assembly_s"] - tim.get("field_s", 0.0)))
                res = frequency_result(job, chunk[j], sols, field, ys, tim, first=(i + j == 0), solver=solver)
                emit({"type": "result", "result": res} if event_mode else res)
                solved = i + j + 1
            i += len(chunk)
    finally:
        solver.close()
    return _cancelled(job), solved
