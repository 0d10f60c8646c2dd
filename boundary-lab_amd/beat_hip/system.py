"""Physical-system exterior solve: the compiled-system request/response wire of the coupled worker, served by the
HIP engine for a system with one unbounded region and no bounded (FEM) region.

Reference: ``solve_request`` / ``solve_exterior_request`` and their helpers
(src/blab/solvers/julia_local/coupled_solver.jl:17-87,125-265,267-569,1465-1492,2515-2575), reached from Boundary
Lab through ``PhysicalSystemProductionBackend(bem_backend="rocm")`` (src/blab/solvers/coupled_backend.py:815-843,
237-305) with the wire of src/blab/system_contract.py. Per frequency: the Burton-Miller system of the aggregated
BEM surface is solved once for every excitation port (one channel each), then the requested outputs are formed:
``exterior_pressure`` at given points, ``bem_boundary_pressure`` (P1), ``bem_boundary_neumann`` (DP0) and
``radiation_impedance`` per component.

Precision ``float32`` runs the mixed path (FP32 LU with FP64 refinement; results are rounded to complex64 on the
wire as the reference's Float32 solve reports them), ``float64`` the plain FP64 path. Systems with bounded regions
and one unbounded region are handed to the coupled FEM-BEM solve (``beat_hip.coupled_system``); pure interior
systems (no unbounded region) and the retained-cache ``bem_field`` operation are not served: they fail with an
error naming the unsupported part.
"""

from __future__ import annotations

import math
import os
import time

import numpy as np

from . import geometry as G
from . import driver as D
from .driver import DEFAULT_FREQUENCY_BATCH

SCHEMA_VERSION = 1


def object_by_id(items, object_id, label):
    """coupled_solver.jl:17-22."""
    for item in items:
        if str(item["id"]) == object_id:
            return item
    raise ValueError(f"Unknown {label} id: {object_id}")


def combine_tagged(meshes):
    """Concatenate boundary meshes without welding and renumber physical tags to solver tags 1, 2, ... in order
    of first appearance, per mesh (BeatEngineCore.jl:266-311). Returns (mesh, vertex_offsets, face_offsets,
    tag_maps)."""
    parts, tag_maps, voffs, foffs = [], [], [], []
    next_tag, vo, fo = 1, 0, 0
    for V, F, T in meshes:
        local = {}
        solver_tags = np.empty(len(T), dtype=np.int64)
        for i, t in enumerate(np.asarray(T).tolist()):
            if t not in local:
                local[t] = next_tag
                next_tag += 1
            solver_tags[i] = local[t]
        parts.append((V, F, solver_tags))
        tag_maps.append(local)
        voffs.append(vo)
        foffs.append(fo)
        vo += len(V)
        fo += len(F)
    if not parts:
        raise ValueError("An unbounded region must contain at least one BEM mesh.")
    return G.combine(parts), voffs, foffs, tag_maps


def aggregate_bem_region(meshes, region, boundaries):
    """The unbounded region's BEM surfaces, scaled and translated, combined without welding, and the solver tag
    of each of the region's boundaries (coupled_solver.jl:40-87)."""
    mesh_ids = [str(m) for m in region.get("mesh_ids", [])]
    if not mesh_ids:
        raise ValueError("Unbounded region must contain at least one BEM mesh.")
    resources = [object_by_id(meshes, mid, "mesh") for mid in mesh_ids]
    if not all(str(r["purpose"]) == "bem_surface" for r in resources):
        raise ValueError("Unbounded region meshes must be BEM surfaces.")
    loaded = []
    for r in resources:
        loaded.append(G.read_gmsh22(str(r["file"]), scale=float(r["scale_to_m"]),
                                    translation=r.get("translation_m", (0.0, 0.0, 0.0))))
    mesh, voffs, foffs, tag_maps = combine_tagged(loaded)
    tag_map_by_mesh = dict(zip(mesh_ids, tag_maps))
    boundary_tag_by_id = {}
    for b in boundaries:
        if str(b["region_id"]) != str(region["id"]):
            continue
        bid = str(b["id"])
        group = b["group"]
        mid = str(group["mesh_id"])
        tmap = tag_map_by_mesh.get(mid)
        if tmap is None:
            raise ValueError(f"Exterior boundary {bid!r} references mesh {mid!r} outside its unbounded region.")
        src = int(group["tag"])
        solver_tag = tmap.get(src, 0)
        if solver_tag <= 0:
            raise ValueError(f"Exterior boundary {bid!r} tag {src} is not present in BEM mesh {mid!r}.")
        boundary_tag_by_id[bid] = solver_tag
    return {
        "mesh": mesh,
        "vertex_offset_by_mesh_id": dict(zip(mesh_ids, voffs)),
        "face_offset_by_mesh_id": dict(zip(mesh_ids, foffs)),
        "boundary_tag_by_id": boundary_tag_by_id,
    }


def exterior_excitations(port_ids, ports, components, boundaries, boundary_tag_by_id, region):
    """One excitation per port: a normal-velocity port of an ideal velocity source, moving with amplitude
    ``boundary_motion_weights[boundary]`` (default 1) on each of its moving boundaries
    (coupled_solver.jl:196-232)."""
    out = []
    for port_id in port_ids:
        port = object_by_id(ports, port_id, "excitation port")
        if str(port["kind"]) != "normal_velocity":
            raise ValueError(f"Exterior excitation port {port_id} must be a normal-velocity port.")
        comp = object_by_id(components, str(port["component_id"]), "component")
        if str(comp["kind"]) != "ideal_velocity_source":
            raise ValueError(f"Exterior excitation port {port_id} must reference an ideal velocity source.")
        weights = (comp.get("parameters") or {}).get("boundary_motion_weights", {}) or {}
        tags, amps = [], []
        for bid in [str(x) for x in comp.get("boundary_ids", [])]:
            b = object_by_id(boundaries, bid, "boundary")
            if str(b["region_id"]) != str(region["id"]):
                raise ValueError(f"Exterior component {str(comp['id'])!r} references a boundary outside its region.")
            if str(b["kind"]) != "moving":
                continue
            tag = boundary_tag_by_id.get(bid, 0)
            if tag <= 0:
                raise ValueError(f"Exterior moving boundary {bid} did not resolve to a BEM tag.")
            amp = float(weights.get(bid, 1.0))
            if not (math.isfinite(amp) and amp > 0.0):
                raise ValueError("Exterior boundary motion weights must be finite and positive.")
            tags.append(tag)
            amps.append(amp)
        if not tags:
            raise ValueError(f"Exterior component {str(comp['id'])!r} has no moving boundaries.")
        out.append({"port_id": port_id, "component_id": str(comp["id"]), "tags": tags, "amplitudes": amps})
    return out


def unit_drives(mesh: G.BoundaryMesh, excitations) -> np.ndarray:
    """(F, n_exc) face amplitudes: the Neumann data of excitation e is i rho omega times column e
    (coupled_solver.jl:234-243; a face takes the amplitude of the last matching (tag, amplitude) entry)."""
    q = np.zeros((mesh.n_faces, len(excitations)), dtype=np.float64)
    for e, ex in enumerate(excitations):
        for tag, amp in zip(ex["tags"], ex["amplitudes"]):
            q[mesh.tags == tag, e] = amp
    return q


def component_impedance(mesh: G.BoundaryMesh, pressure: np.ndarray, excitation, symmetry_factor: int) -> complex:
    """Force = sum over the excitation's faces of the face-average P1 pressure times area, times the symmetry
    factor and 10; reported as conj(force) / 2 (coupled_solver.jl:254-265)."""
    sel = np.isin(mesh.tags, np.asarray(excitation["tags"]))
    avg = pressure[mesh.faces[sel]].mean(axis=1)
    force = complex(np.sum(avg * mesh.areas[sel])) * symmetry_factor * 10.0
    return complex(force.real / 2.0, -force.imag / 2.0)


def complex_array_wire(values, complex_dtype) -> dict:
    """{"dtype", "shape", "real", "imag"} with row-major flattening (coupled_solver.jl:125-142)."""
    a = np.asarray(values).astype(complex_dtype)
    flat = a.reshape(-1)
    return {
        "dtype": "complex64" if a.dtype == np.complex64 else "complex128",
        "shape": list(a.shape),
        "real": flat.real.tolist(),
        "imag": flat.imag.tolist(),
    }


def quantity_wire(output, values, unit, axes, complex_dtype, metadata=None) -> dict:
    """coupled_solver.jl:144-154."""
    targets = output.get("target_ids") or []
    return {
        "id": str(output["id"]),
        "quantity": str(output["quantity"]),
        "unit": unit,
        "target_id": str(targets[0]) if targets else None,
        "axes": list(axes),
        "values": complex_array_wire(values, complex_dtype),
        "metadata": dict(metadata or {}),
    }


def _options(request):
    return request.get("solver_options") or {}


def parse_exterior_request(request):
    """Checks of solve_request / solve_exterior_request up to the frequency loop (coupled_solver.jl:1465-1492,
    267-300). Returns a dict with the prepared inputs."""
    if int(request.get("schema_version", 0)) != SCHEMA_VERSION:
        raise ValueError("Unsupported system solve request schema.")
    system = request["compiled_system"]
    regions = system.get("regions", [])
    bounded = [r for r in regions if str(r["kind"]) == "bounded_air"]
    unbounded = [r for r in regions if str(r["kind"]) == "unbounded_air"]
    if not unbounded:
        raise ValueError("Interior (FEM-only) system solves are not served by the HIP exterior backend.")
    if len(unbounded) != 1:
        raise ValueError("Coupled backend currently requires exactly one unbounded region.")
    if bounded:
        raise ValueError("A system with bounded regions is a coupled FEM-BEM solve (beat_hip.coupled_system), not "
                         "an exterior one.")
    region = unbounded[0]
    opts = _options(request)
    precision = str(opts.get("precision", "float32")).lower()
    if precision not in ("float32", "float64"):
        raise ValueError("Exterior precision must be float32 or float64.")
    backend = str(opts.get("bem_backend", "cpu")).lower()
    if backend not in ("cpu", "cuda", "rocm"):
        raise ValueError("Exterior BEM backend must be cpu, cuda, or rocm.")
    if backend != "rocm":
        raise ValueError(f"The HIP exterior backend serves bem_backend 'rocm', not {backend!r}.")
    mode = str(opts.get("regular_quadrature_mode", "fixed")).lower()
    if mode not in ("fixed", "wavelength"):
        raise ValueError(f"Unsupported regular quadrature mode: {mode}. Expected fixed or wavelength.")
    if mode != "fixed":
        raise ValueError("Wavelength-driven regular quadrature is currently available only on the CPU backend.")
    regular_order = int(opts.get("quadrature_order", 4))
    singular_order = int(opts.get("singular_order", 4))
    if not 1 <= singular_order <= 4:
        raise ValueError("Duffy 1D Gauss order must be between 1 and 4.")
    symmetry = G.normalized_symmetry_mode(opts.get("symmetry", "off"))
    c = float(region["sound_speed_m_per_s"])
    rho = float(region["density_kg_per_m3"])
    if precision == "float32":
        c, rho = float(np.float32(c)), float(np.float32(rho))
    if not (c > 0.0 and rho > 0.0):
        raise ValueError("Exterior sound speed and density must be positive.")
    t0 = time.perf_counter()
    dom = aggregate_bem_region(system.get("meshes", []), region, system.get("boundaries", []))
    mesh = G.snap_symmetry_planes(dom["mesh"], symmetry)
    G.validate_symmetry_domain(mesh, symmetry)
    port_ids = [str(p) for p in request.get("excitation_port_ids", [])]
    excitations = exterior_excitations(port_ids, system.get("excitation_ports", []), system.get("components", []),
                                       system.get("boundaries", []), dom["boundary_tag_by_id"], region)
    mesh_setup_s = time.perf_counter() - t0
    outputs = list(request.get("outputs", []))
    point_sets = []
    for out in outputs:
        q = str(out["quantity"])
        if q == "exterior_pressure":
            pts = (out.get("options") or {}).get("points_m", [])
            if not pts:
                raise ValueError("exterior_pressure output requires options.points_m.")
            point_sets.append(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
        elif q not in ("bem_boundary_pressure", "bem_boundary_neumann", "radiation_impedance"):
            raise ValueError(f"Unsupported exterior output quantity: {q}")
    fdt = np.float32 if precision == "float32" else np.float64
    return {
        "precision": precision,
        "engine_precision": "mixed" if precision == "float32" else "fp64",
        "complex_dtype": np.complex64 if precision == "float32" else np.complex128,
        "float_dtype": fdt,
        "symmetry": symmetry,
        "sound_speed": c,
        "density": rho,
        "mesh": mesh,
        "excitations": excitations,
        "port_ids": port_ids,
        "components": system.get("components", []),
        "outputs": outputs,
        "point_sets": point_sets,
        "mesh_setup_s": mesh_setup_s,
        "frequencies": [float(fdt(f)) for f in request.get("frequencies_hz", [])],
        "cancel_path": request.get("cancel_path"),
        "regular_order": regular_order,
        "singular_order": singular_order,
    }


def _cancelled(job) -> bool:
    return job["cancel_path"] is not None and os.path.isfile(str(job["cancel_path"]))


def frequency_result(job, freq, pressure, field, timings, first) -> dict:
    """The result object of one frequency (coupled_solver.jl:420-548). pressure: (V, n_exc) complex in the
    mesh numbering; field: (n_points_total, n_exc) or None."""
    mesh, cdt = job["mesh"], job["complex_dtype"]
    omega = 2.0 * math.pi * freq
    qunit = unit_drives(mesh, job["excitations"])
    neumann = (1j * job["density"] * omega) * qunit  # (F, n_exc)
    quantities = []
    offset = 0
    field_s = 0.0
    for out in job["outputs"]:
        q = str(out["quantity"])
        if q == "exterior_pressure":
            pts = (out.get("options") or {}).get("points_m", [])
            n = len(pts)
            quantities.append(quantity_wire(out, field[offset : offset + n].T, "Pa", ["excitation", "observation"],
                                            cdt, metadata={"points_m": pts}))
            offset += n
            field_s = timings["field_s"]
        elif q == "bem_boundary_pressure":
            quantities.append(quantity_wire(out, pressure.T, "Pa", ["excitation", "bem_node"], cdt))
        elif q == "bem_boundary_neumann":
            quantities.append(quantity_wire(out, neumann.T, "Pa/m", ["excitation", "bem_face"], cdt))
        elif q == "radiation_impedance":
            by_comp = {ex["component_id"]: (e, ex) for e, ex in enumerate(job["excitations"])}
            values = []
            for comp in job["components"]:
                cid = str(comp["id"])
                if cid not in by_comp:
                    raise ValueError(f"radiation_impedance needs an excitation of component {cid!r}.")
                e, ex = by_comp[cid]
                values.append(component_impedance(mesh, pressure[:, e], ex, G.symmetry_reduction_factor(
                    job["symmetry"])))
            quantities.append(quantity_wire(out, np.asarray(values), "Pa*s/m^3", ["radiator"], cdt))
    diagnostics = {
        "precision": job["precision"],
        "bem_backend": "rocm",
        "symmetry": job["symmetry"],
        "formulation": "exterior_burton_miller_neumann",
        "linear_solver": "hip_dense_lu_fp64_refinement" if job["precision"] == "float32" else "hip_dense_lu",
        "bounded_region_count": 0,
        "interface_count": 0,
        "regular_quadrature_mode": "fixed",
        "regular_quadrature_order": job["regular_order"],
        "blas_threads": 1,
        "timings": {
            "assembly_s": timings["assembly_s"],
            "solve_s": timings["solve_s"],
            "field_s": field_s,
            "mesh_setup_s": job["mesh_setup_s"] if first else 0.0,
            "cache_setup_s": 0.0,
        },
    }
    return {
        "schema_version": 1,
        "freq_hz": freq,
        "excitation_port_ids": list(job["port_ids"]),
        "quantities": quantities,
        "diagnostics": diagnostics,
    }


def solve_system_request(request, engine_factory, emit, batch: int = DEFAULT_FREQUENCY_BATCH,
                         event_mode: bool = True, coupled_factory=None):
    """Run a compiled-system request; emits one ``result`` per frequency (wrapped as an event in worker mode,
    bare in one-shot mode) and returns (cancelled, solved_count). Exterior systems (coupled_solver.jl:267-569):
    engine_factory(prepared_mesh, qunit) -> beat_hip.engine.Engine. Systems with bounded regions and one
    unbounded region go to the coupled FEM-BEM solve (beat_hip.coupled_system; coupled_factory(model, **options)
    -> beat_hip.coupled.CoupledSolver, default a solver on device 0)."""
    regions = (request.get("compiled_system") or {}).get("regions", [])
    kinds = [str(r.get("kind")) for r in regions]
    if "bounded_air" in kinds and "unbounded_air" in kinds:
        from .coupled_system import solve_coupled_request

        return solve_coupled_request(request, emit, event_mode=event_mode, solver_factory=coupled_factory)
    job = parse_exterior_request(request)
    if _cancelled(job):
        return True, 0
    mesh, exc = job["mesh"], job["excitations"]
    prep = G.prepare(mesh, symmetry=job["symmetry"], regular_order=job["regular_order"],
                     singular_order=job["singular_order"])
    qunit = unit_drives(mesh, exc)
    obs = np.concatenate(job["point_sets"], axis=0) if job["point_sets"] else None
    freqs = job["frequencies"]
    nb = max(1, min(int(batch), len(freqs))) if freqs else 1
    # one engine for every excitation: the system is assembled and factored once per frequency, the
    # right-hand sides, solves and field run per group of at most four excitations on the same factors
    eng = engine_factory(prep, qunit)
    fit = getattr(eng, "fit_batch", None)
    if fit is not None:
        nb = fit(job["engine_precision"], nb, 0 if obs is None else len(obs))
    solved = 0
    b = None
    try:
        b = eng.batch(job["engine_precision"], nb, obs, job["density"], job["sound_speed"])
        for i0 in range(0, len(freqs), nb):
            if _cancelled(job):
                return True, solved
            fr = np.asarray(freqs[i0 : i0 + nb], dtype=np.float64)
            r = b.run(fr)
            flags = getattr(r, "flags", None)
            n = len(fr)
            for j in range(n):
                if flags is not None and int(flags[j]):
                    raise RuntimeError(f"Dense solve failed at {fr[j]:g} Hz: {D.status_text(int(flags[j]))}.")
                timings = {"assembly_s": r.t_assembly / n, "solve_s": r.t_solve / n, "field_s": r.t_field / n}
                res = frequency_result(job, freqs[i0 + j], r.pressure[j], None if obs is None else r.field[j],
                                       timings, first=(i0 + j == 0))
                emit({"type": "result", "result": res} if event_mode else res)
                solved = i0 + j + 1
    finally:
        if b is not None:
            b.close()
        eng.close()
    return _cancelled(job), solved
