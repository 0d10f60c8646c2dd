"""Solve-request driver: the job semantics of the reference worker, on the HIP engine.

Restates src/blab/solvers/julia_local/solver.jl (cited per function) in Python:
request parsing and validation, mesh loading, radiator/channel drives with crossovers, polar and Fibonacci
observation sets, per-frequency solve, channel-basis synthesis, SPL, impedance, and the NDJSON event payloads
(``initialized``, ``status``, ``result``, ``completed``/``cancelled``/``failed``).

Differences by design (documented in the result diagnostics):
* all channels of a frequency are solved from ONE factorization (the reference re-factors per channel,
  BeatEngineRocmOperators.jl:238-242), and frequencies are batched on the device;
* ``precision`` is "mixed" (default) or "fp64"; mixed results carry their FP64 refinement residual and
  pivot growth, and optionally (``fp64_check``) their error against the FP64 GPU path.
"""

from __future__ import annotations

import functools
import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import geometry as G
from .wire import Float32Row

# ---------------------------------------------------------------------------------------------
# config parsing (solver.jl:46-250)
# ---------------------------------------------------------------------------------------------


def get_value(raw, key, default=None):
    return raw[key] if isinstance(raw, dict) and key in raw and raw[key] is not None else default


def validate_crossover_config(owner, crossover):
    """solver.jl:54-69"""
    if crossover is None:
        return
    ctype = str(get_value(crossover, "type", "none")).lower()
    if ctype not in ("none", "lowpass", "highpass"):
        raise ValueError(f"{owner} crossover type must be none, lowpass, or highpass.")
    if ctype == "none":
        return
    freq = get_value(crossover, "frequency_hz")
    if freq is None or float(freq) <= 0.0:
        raise ValueError(f"{owner} crossover frequency_hz must be > 0.")
    filt = str(get_value(crossover, "filter", "butterworth")).lower()
    if filt not in ("butterworth", "linkwitz_riley"):
        raise ValueError(f"{owner} crossover filter must be butterworth or linkwitz_riley.")
    order = int(get_value(crossover, "order", 1))
    if order not in (1, 2, 4, 6):
        raise ValueError(f"{owner} crossover order must be 1, 2, 4, or 6.")
    if filt == "linkwitz_riley" and order not in (2, 4, 6):
        raise ValueError(f"{owner} Linkwitz-Riley order must be 2, 4, or 6.")


def symmetry_mode_from_config(config):
    """solver.jl:71-75"""
    mode = str(get_value(config, "symmetry", "off")).strip().lower()
    if mode not in ("off", "x", "xy"):
        raise ValueError(f"Unsupported symmetry mode: {mode}. Expected off, x, or xy.")
    return mode


BACKEND_ALIASES = {
    "beat_cuda": "cuda", "gpu": "cuda", "julia_local": "cuda", "local_julia": "cuda", "beat_cpu": "cpu",
    "beat_rocm": "rocm", "amd": "rocm", "amdgpu": "rocm",
}


def beat_backend_from_request(request):
    """solver.jl:76-91 (this engine runs every accepted backend id on the HIP device)."""
    backend = str(get_value(request, "beat_engine_backend", "rocm")).strip().lower()
    backend = BACKEND_ALIASES.get(backend, backend)
    if backend not in ("cuda", "cpu", "rocm"):
        raise ValueError(f"Unsupported BEAT Engine backend: {backend}. Expected cuda, cpu, or rocm.")
    return backend


def regular_quadrature_mode_from_config(config):
    """solver.jl:93-101 for a GPU backend: fixed quadrature (the rule of quadrature_order for every frequency);
    the wavelength-driven selection is the CPU backend's only, as in the reference."""
    mode = str(get_value(config, "regular_quadrature_mode", get_value(config, "quadrature_mode", "fixed")))
    mode = mode.strip().lower()
    if mode not in ("fixed", "wavelength"):
        raise ValueError(f"Unsupported regular quadrature mode: {mode}. Expected fixed or wavelength.")
    if mode == "wavelength":
        raise ValueError("Wavelength-driven regular quadrature is currently implemented only for the BEAT CPU backend.")
    return mode


@dataclass
class MeshInput:
    path: str
    scale: float
    name: str
    translation: tuple


def mesh_inputs_from_config(config):
    """solver.jl:152-179"""
    meshes = get_value(config, "meshes", []) or []
    if meshes:
        out, seen = [], set()
        for i, mesh in enumerate(meshes, start=1):
            name = str(get_value(mesh, "name", f"mesh_{i}"))
            if name in seen:
                raise ValueError(f"Duplicate mesh name: {name}")
            seen.add(name)
            tr = get_value(mesh, "translation_m", [0.0, 0.0, 0.0])
            if len(tr) != 3:
                raise ValueError(f"Mesh '{name}' translation_m must contain three values.")
            scale = get_value(mesh, "scale_factor", get_value(config, "scale_factor", 0.001))
            out.append(MeshInput(str(mesh["file"]), float(scale), name, tuple(float(v) for v in tr)))
        return out
    return [MeshInput(str(config["mesh_file"]), float(get_value(config, "scale_factor", 0.001)), "mesh",
                      (0.0, 0.0, 0.0))]


def radiator_inputs_from_config(config, mesh_inputs):
    """solver.jl:185-233"""
    raw = get_value(config, "radiators", []) or []
    lookup = {m.name: i for i, m in enumerate(mesh_inputs, start=1)}
    none = {"type": "none"}
    if not raw:
        return [{
            "name": "Radiator", "tag": int(get_value(config, "tag_throat", 2)), "mesh": mesh_inputs[0].name,
            "mesh_id": 1, "channel": "main", "velocity_offset_db": 0.0, "level_db": 0.0, "polarity": 1,
            "delay_ms": 0.0, "hpf": none, "lpf": none,
        }]
    out = []
    for r in raw:
        name = str(get_value(r, "name", "Radiator"))
        mesh_name = get_value(r, "mesh")
        if mesh_name is None:
            if len(mesh_inputs) != 1:
                raise ValueError(f"Radiator '{name}' must specify 'mesh' when multiple meshes are configured.")
            mesh_name = mesh_inputs[0].name
        mesh_name = str(mesh_name)
        if mesh_name not in lookup:
            raise ValueError(f"Radiator '{name}' references unknown mesh '{mesh_name}'.")
        validate_crossover_config(name, get_value(r, "hpf"))
        validate_crossover_config(name, get_value(r, "lpf"))
        out.append({
            "name": name, "tag": int(r["tag"]), "mesh": mesh_name, "mesh_id": lookup[mesh_name],
            "channel": str(get_value(r, "channel", "main")),
            "velocity_offset_db": float(get_value(r, "velocity_offset_db", 0.0)),
            "level_db": float(get_value(r, "level_db", 0.0)), "polarity": int(get_value(r, "polarity", 1)),
            "delay_ms": float(get_value(r, "delay_ms", 0.0)), "hpf": get_value(r, "hpf", none),
            "lpf": get_value(r, "lpf", none),
        })
    return out


def channel_inputs_from_config(config):
    """solver.jl:235-250"""
    channels = {}
    for ch in get_value(config, "channels", []) or []:
        name = str(ch["name"])
        validate_crossover_config(f"channel {name}", get_value(ch, "hpf"))
        validate_crossover_config(f"channel {name}", get_value(ch, "lpf"))
        channels[name] = {
            "level_db": float(get_value(ch, "level_db", 0.0)), "polarity": int(get_value(ch, "polarity", 1)),
            "delay_ms": float(get_value(ch, "delay_ms", 0.0)), "hpf": get_value(ch, "hpf", {"type": "none"}),
            "lpf": get_value(ch, "lpf", {"type": "none"}),
        }
    return channels


def load_combined_mesh(mesh_inputs):
    """solver.jl:252-269: raw physical tags are kept, element mesh ids are the 1-based input index."""
    parts = [G.read_gmsh22(m.path, m.scale, m.translation) for m in mesh_inputs]
    return G.combine(parts)


# ---------------------------------------------------------------------------------------------
# drives and crossovers (solver.jl:271-318, 390-410, 590-605)
# ---------------------------------------------------------------------------------------------


def butterworth_poles(order):
    return [np.exp(1j * (math.pi / 2 + (2 * k - 1) * math.pi / (2 * order))) for k in range(1, order + 1)]


def butterworth_response(ctype, order, cutoff_hz, freq):
    omega = 2 * math.pi * freq
    wc = 2 * math.pi * float(cutoff_hz)
    s = 1j * omega
    resp = 1.0 + 0j
    for pole in butterworth_poles(order):
        sp = wc * pole if ctype == "lowpass" else wc / pole
        resp *= (-sp) / (s - sp) if ctype == "lowpass" else s / (s - sp)
    return resp


def crossover_response(crossover, freq):
    if crossover is None:
        return 1.0 + 0j
    ctype = str(get_value(crossover, "type", "none")).lower()
    if ctype == "none":
        return 1.0 + 0j
    filt = str(get_value(crossover, "filter", "butterworth")).lower()
    order = int(get_value(crossover, "order", 1))
    cutoff = get_value(crossover, "frequency_hz")
    if cutoff is None:
        raise ValueError(f"Crossover frequency_hz must be set for {ctype}.")
    if filt == "linkwitz_riley":
        sec = butterworth_response(ctype, order // 2, cutoff, freq)
        return sec * sec
    return butterworth_response(ctype, order, cutoff, freq)


def channel_drive(channel, freq):
    omega = 2 * math.pi * freq
    level = 10.0 ** (float(channel["level_db"]) / 20.0)
    delay = np.exp(-1j * omega * float(channel["delay_ms"]) / 1000.0)
    xo = crossover_response(get_value(channel, "hpf"), freq) * crossover_response(get_value(channel, "lpf"), freq)
    return complex(float(channel["polarity"]) * level) * delay * xo


def drive_for_radiator(radiator, channels, freq):
    omega = 2 * math.pi * freq
    ch = channels.get(str(get_value(radiator, "channel", "main")))
    if ch is not None:
        return channel_drive(ch, freq) * 10.0 ** (float(radiator["velocity_offset_db"]) / 20.0)
    level = 10.0 ** ((float(radiator["level_db"]) + float(radiator["velocity_offset_db"])) / 20.0)
    delay = np.exp(-1j * omega * float(radiator["delay_ms"]) / 1000.0)
    xo = crossover_response(get_value(radiator, "hpf"), freq) * crossover_response(get_value(radiator, "lpf"), freq)
    return complex(float(radiator["polarity"]) * level) * delay * xo


def radiator_faces(mesh, radiator):
    return (mesh.tags == int(radiator["tag"])) & (mesh.mesh_ids == int(radiator["mesh_id"]))


def validate_radiator_channels(radiators, channels):
    if not channels:
        return
    for r in radiators:
        cname = str(get_value(r, "channel", "main"))
        if cname not in channels:
            raise ValueError(f"Radiator '{r['name']}' references unknown channel '{cname}'.")


def validate_radiator_elements(mesh, radiators):
    for r in radiators:
        if not np.any(radiator_faces(mesh, r)):
            raise ValueError(f"No elements found for radiator '{r['name']}' tag={r['tag']} on mesh '{r['mesh']}'.")


def channel_unit_qunit(mesh, radiators, channel_names):
    """Unit Neumann drives per channel, (F, C): q_e / (i rho omega) (solver.jl:434-458, 590-597).

    Later radiators overwrite earlier ones on shared faces, as in the reference's assignment.
    """
    q = np.zeros((mesh.n_faces, len(channel_names)), dtype=np.complex128)
    for ci, cname in enumerate(channel_names):
        for r in radiators:
            drive = 10.0 ** (float(r["velocity_offset_db"]) / 20.0) if str(get_value(r, "channel", "main")) == cname \
                else 0.0
            q[radiator_faces(mesh, r), ci] = drive
    return q


# ---------------------------------------------------------------------------------------------
# observation sets (solver.jl:320-388)
# ---------------------------------------------------------------------------------------------


def polar_observation_points(config):
    step = float(get_value(config, "step_size", 5.0))
    amin = float(get_value(config, "min_angle", -180.0))
    amax = float(get_value(config, "max_angle", 180.0))
    if step <= 0:
        raise ValueError("step_size must be positive.")
    if amin < -180 or amax > 180:
        raise ValueError("polar angle range must stay within [-180, 180] degrees.")
    if amax < amin:
        raise ValueError("max_angle must be >= min_angle.")
    if not (amin <= 0 <= amax):
        raise ValueError("polar angle range must include 0 degrees.")
    n = int(math.floor((amax - amin) / step + 1e-9)) + 1
    angles = (amin + step * np.arange(n, dtype=np.float64)).astype(np.float32)
    if angles.size == 0 or angles[-1] < np.float32(amax):
        angles = np.append(angles, np.float32(amax))
    angles = np.clip(angles, np.float32(amin), np.float32(amax)).astype(np.float32)
    d = float(get_value(config, "distance", 2.0))
    if d <= 0:
        raise ValueError("distance must be positive.")
    df = np.float32(d)
    off = np.float32(get_value(config, "axial_offset", 0.0))
    ang = np.float32(math.pi) * angles / np.float32(180.0)
    s, c = np.sin(ang).astype(np.float32), np.cos(ang).astype(np.float32)
    z = np.zeros_like(s)
    horizontal = np.stack([df * s, z, df * c + off], axis=1).astype(np.float32)
    vertical = np.stack([z, df * s, df * c + off], axis=1).astype(np.float32)
    return angles, horizontal, vertical, int(np.argmin(np.abs(angles.astype(np.float64))))


def spherical_observation(config):
    if not bool(get_value(config, "spherical_sampling_enabled", False)):
        return None
    n = int(get_value(config, "spherical_sampling_points", 6000))
    if n <= 0:
        raise ValueError("spherical_sampling_points must be positive.")
    d = float(get_value(config, "distance", 2.0))
    if d <= 0:
        raise ValueError("distance must be positive.")
    df = np.float32(d)
    off = np.float32(get_value(config, "axial_offset", 0.0))
    golden = np.float32(math.pi * (3.0 - math.sqrt(5.0)))
    i = np.arange(n)
    zu = (1.0 - (2.0 * i + 1.0) / n).astype(np.float32)
    xy = np.sqrt(np.maximum(np.float32(1.0) - zu * zu, np.float32(0.0))).astype(np.float32)
    az = (i.astype(np.float32) * golden).astype(np.float32)
    x = (df * xy * np.cos(az)).astype(np.float32)
    y = (df * xy * np.sin(az)).astype(np.float32)
    pts = np.stack([x, y, (df * zu + off).astype(np.float32)], axis=1)
    meta = {
        "r_distance_m": np.full(n, df, dtype=np.float32),
        "theta_polar_rad": np.arccos(np.clip(zu, -1, 1)).astype(np.float32),
        "phi_azimuth_rad": np.mod(np.arctan2(y, x), np.float32(2 * math.pi)).astype(np.float32),
    }
    return pts, meta


# ---------------------------------------------------------------------------------------------
# post-processing (solver.jl:488-631)
# ---------------------------------------------------------------------------------------------


def pressure_to_spl(p):
    with np.errstate(divide="ignore"):
        return (20.0 * np.log10(np.abs(p) / 20e-6)).astype(np.float32)


def complex_rows_to_wire(rows):
    return {"real": [Float32Row(r.real) for r in rows], "imag": [Float32Row(r.imag) for r in rows]}


@functools.lru_cache(maxsize=64)
def _reference_bracket(angles: tuple, n_values: int, ref_deg: float):
    """Where the flat-target reference angle falls among the polar angles: (i, j, frac) with the interpolated value
    v[i] (1 - frac) + v[j] frac (indices into the values), computed once per angle set."""
    ref = (float(ref_deg) + 180.0) % 360.0 - 180.0
    ang = np.asarray(angles, dtype=np.float64)
    n = n_values
    if math.isclose(ang[0], -180.0) and math.isclose(ang[-1], 180.0):
        ang, n = ang[:-1], n - 1
    ea = np.concatenate([ang - 360.0, ang, ang + 360.0])
    # the first bracket [ea[i], ea[i + 1]] holding ref in list order; with ascending angles it is found by bisection
    # (the last i with ea[i] <= ref, stepped back while the bracket before also holds ref)
    if np.all(np.diff(ea) > 0):
        i = int(np.searchsorted(ea, ref, side="right")) - 1
        while i > 0 and ea[i - 1] <= ref <= ea[i]:
            i -= 1
        candidates = range(max(i, 0), min(i + 1, len(ea) - 1))
    else:  # angles out of order: every bracket, in list order
        candidates = range(len(ea) - 1)
    for i in candidates:
        lo, hi = ea[i], ea[i + 1]
        if lo <= ref <= hi:
            if math.isclose(lo, hi):
                return i % n, i % n, None
            return i % n, (i + 1) % n, (ref - lo) / (hi - lo)
    k = int(np.argmin(np.abs(ea - ref))) % n
    return k, k, None


def interpolate_complex_reference(values, angles_deg, ref_deg):
    if len(values) == 0:
        return 0j
    if len(values) == 1:
        return complex(values[0])
    key = angles_deg if type(angles_deg) is tuple else tuple(np.asarray(angles_deg, np.float64).tolist())
    i, j, frac = _reference_bracket(key, len(values), float(ref_deg))
    val = np.asarray(values, dtype=np.complex128)
    if frac is None:
        return val[i]
    return val[i] * (1.0 - frac) + val[j] * frac


def flat_target_corrections(channel_names, horizontal_rows, angles_deg, ref_deg, flat_target):
    out = {}
    for ci, name in enumerate(channel_names):
        if not flat_target:
            out[name] = 1.0
            continue
        mag = abs(interpolate_complex_reference(horizontal_rows[ci], angles_deg, ref_deg))
        out[name] = 1.0 if mag <= 1e-12 else 1.0 / mag
    return out


DEFAULT_CHANNEL = {"level_db": 0.0, "polarity": 1, "delay_ms": 0.0, "hpf": {"type": "none"},
                   "lpf": {"type": "none"}}


def synthesize_channel_basis(channel_names, hrows, vrows, srows, channels, freq, angles_deg, ref_deg, flat):
    corr = flat_target_corrections(channel_names, hrows, angles_deg, ref_deg, flat)
    weights = [channel_drive(channels.get(n, DEFAULT_CHANNEL), freq) * corr[n] for n in channel_names]
    hp = sum(hrows[i] * weights[i] for i in range(len(channel_names)))
    vp = sum(vrows[i] * weights[i] for i in range(len(channel_names)))
    hs, vs = pressure_to_spl(hp), pressure_to_spl(vp)
    ref = hs[int(np.argmin(np.abs(np.asarray(angles_deg, np.float64))))]
    sn = None
    if srows is not None:
        sp = sum(srows[i] * weights[i] for i in range(len(channel_names)))
        sn = (pressure_to_spl(sp) - ref).astype(np.float32)
    return {"horizontal_spl": hs, "vertical_spl": vs, "horizontal_norm": (hs - ref).astype(np.float32),
            "vertical_norm": (vs - ref).astype(np.float32), "sphere_norm": sn, "corrections": corr,
            "weights": weights}


def radiator_selection(mesh, radiators):
    """Per radiator, the vertex ids (F_r, 3) and areas of its faces (for impedance_for_radiators)."""
    out = []
    for r in radiators:
        sel = radiator_faces(mesh, r)
        out.append((mesh.faces[sel], mesh.areas[sel]))
    return out


def impedance_for_radiators(mesh, pressure, radiators, drives, symmetry_factor=1, selection=None):
    """solver.jl:607-631: F = 10 * sum p_avg A * factor; [Re(F/drive)/2, -Im(F/drive)/2]. ``selection``: the
    radiators' faces from radiator_selection (formed here when not given)."""
    out = []
    if selection is None:
        selection = radiator_selection(mesh, radiators)
    for (fr, ar), drive in zip(selection, drives):
        if abs(drive) <= 0.0:
            out.append([float("nan"), float("nan")])
            continue
        pav = (pressure[fr[:, 0]] + pressure[fr[:, 1]] + pressure[fr[:, 2]]) / 3.0
        force = 10.0 * np.sum(pav * ar * symmetry_factor)
        z = force / drive
        out.append([float(np.float32(z.real / 2)), float(np.float32(-z.imag / 2))])
    return out


# ---------------------------------------------------------------------------------------------
# job
# ---------------------------------------------------------------------------------------------


@dataclass
class Job:
    """A parsed, mesh-loaded request ready for device work."""

    request: dict
    config: dict
    frequencies: np.ndarray
    mesh: G.BoundaryMesh
    radiators: list
    channels: dict
    channel_names: list
    angles: np.ndarray
    horizontal: np.ndarray
    vertical: np.ndarray
    sphere: tuple | None
    rho: float
    sound_speed: float
    flat_target: bool
    ref_angle: float
    symmetry: str
    precision: str
    qunit: np.ndarray
    obs: np.ndarray
    cancel_path: str | None
    fp64_check: bool
    regular_order: int = 4
    singular_order: int = 4

    @functools.cached_property
    def radiator_selection(self):
        """The radiators' faces (vertex ids and areas), formed once per job."""
        return radiator_selection(self.mesh, self.radiators)

    def prepare(self) -> G.PreparedMesh:
        """The device-ready tables of the job's mesh under its rules."""
        return G.prepare(self.mesh, symmetry=self.symmetry, regular_order=self.regular_order,
                         singular_order=self.singular_order)

    @property
    def n_horizontal(self):
        return len(self.horizontal)


def parse_request(request) -> tuple[dict, dict]:
    """Everything the reference validates before emitting ``initialized`` (solver.jl:680-708)."""
    schema = int(get_value(request, "schema_version", 1))
    if schema not in (1, 2):
        raise ValueError(f"Unsupported solve request schema_version {schema}.")
    config = request["config"]
    symmetry = symmetry_mode_from_config(config)
    beat_backend_from_request(request)
    freqs = np.asarray(request["frequencies_hz"], dtype=np.float32)
    if freqs.size == 0:
        raise ValueError("frequencies_hz must contain at least one frequency.")
    mesh_inputs = mesh_inputs_from_config(config)
    radiators = radiator_inputs_from_config(config, mesh_inputs)
    channels = channel_inputs_from_config(config)
    validate_radiator_channels(radiators, channels)
    angles, hor, ver, _ = polar_observation_points(config)
    sphere = spherical_observation(config)
    init = {
        "polar_angle_deg": angles.tolist(),
        "radiator_names": [r["name"] for r in radiators],
        "sphere_metadata": None if sphere is None else {k: v.tolist() for k, v in sphere[1].items()},
    }
    parsed = dict(config=config, symmetry=symmetry, freqs=freqs, mesh_inputs=mesh_inputs, radiators=radiators,
                  channels=channels, angles=angles, hor=hor, ver=ver, sphere=sphere)
    return init, parsed


def build_job(request, parsed) -> Job:
    config = parsed["config"]
    regular_quadrature_mode_from_config(config)
    regular_order = int(get_value(config, "quadrature_order", 4))
    singular_order = int(get_value(config, "singular_order", 4))
    if not 1 <= singular_order <= 4:
        raise ValueError("Duffy 1D Gauss order must be between 1 and 4.")
    mesh = load_combined_mesh(parsed["mesh_inputs"])
    # x / xy symmetry: the mesh is the fundamental domain (solver.jl:718-719)
    mesh = G.snap_symmetry_planes(mesh, parsed["symmetry"])
    G.validate_symmetry_domain(mesh, parsed["symmetry"])
    validate_radiator_elements(mesh, parsed["radiators"])
    names = sorted({str(get_value(r, "channel", "main")) for r in parsed["radiators"]})
    qunit = channel_unit_qunit(mesh, parsed["radiators"], names)
    pts = [parsed["hor"], parsed["ver"]] + ([parsed["sphere"][0]] if parsed["sphere"] is not None else [])
    precision = str(get_value(config, "precision", os.environ.get("BEAT_HIP_PRECISION", "mixed"))).lower()
    if precision not in ("mixed", "fp64"):
        raise ValueError("precision must be 'mixed' or 'fp64'.")
    return Job(
        request=request, config=config, frequencies=parsed["freqs"], mesh=mesh, radiators=parsed["radiators"],
        channels=parsed["channels"], channel_names=names, angles=parsed["angles"], horizontal=parsed["hor"],
        vertical=parsed["ver"], sphere=parsed["sphere"], rho=float(np.float32(get_value(config, "rho", 1.21))),
        sound_speed=float(np.float32(get_value(config, "sound_speed", 343.0))),
        flat_target=bool(get_value(config, "flat_target_normalization_enabled", True)),
        ref_angle=float(get_value(config, "flat_target_reference_angle_deg", 0.0)),
        symmetry=parsed["symmetry"], precision=precision, qunit=qunit,
        obs=np.concatenate(pts).astype(np.float64), cancel_path=get_value(request, "cancel_path"),
        fp64_check=bool(get_value(config, "fp64_check", False)), regular_order=regular_order,
        singular_order=singular_order,
    )


def result_payload(job: Job, freq: float, p_vc: np.ndarray, field_oc: np.ndarray, timings: dict,
                   diagnostics: dict) -> dict:
    """Per-frequency result dict (solver.jl:921-996). p_vc (V, C) boundary pressure, field_oc (n_obs, C)."""
    nh, nv = len(job.horizontal), len(job.vertical)
    C = len(job.channel_names)
    hrows = [field_oc[:nh, c] for c in range(C)]
    vrows = [field_oc[nh : nh + nv, c] for c in range(C)]
    srows = [field_oc[nh + nv :, c] for c in range(C)] if job.sphere is not None else None
    syn = synthesize_channel_basis(job.channel_names, hrows, vrows, srows, job.channels, freq, job.angles,
                                   job.ref_angle, job.flat_target)
    drives = [drive_for_radiator(r, job.channels, freq) * syn["corrections"].get(str(get_value(r, "channel", "main")),
                                                                                  1.0)
              for r in job.radiators]
    mixed = sum(p_vc[:, c] * syn["weights"][c] for c in range(C))
    imp = impedance_for_radiators(job.mesh, mixed, job.radiators, drives, G.symmetry_reduction_factor(job.symmetry),
                                  selection=job.radiator_selection)
    return {
        "freq_hz": float(np.float32(freq)),
        "horizontal_spl_norm_db": Float32Row(syn["horizontal_norm"]),
        "vertical_spl_norm_db": Float32Row(syn["vertical_norm"]),
        "impedance": imp,
        "horizontal_spl_db": Float32Row(syn["horizontal_spl"]),
        "vertical_spl_db": Float32Row(syn["vertical_spl"]),
        "sphere_spl_norm_db": None if syn["sphere_norm"] is None else Float32Row(syn["sphere_norm"]),
        "channel_names": list(job.channel_names),
        "horizontal_pressure": complex_rows_to_wire(hrows),
        "vertical_pressure": complex_rows_to_wire(vrows),
        "sphere_pressure": None if srows is None else complex_rows_to_wire(srows),
        "timings": {k: float(np.float32(v)) for k, v in timings.items()},
        "diagnostics": diagnostics,
    }


def diagnostics_for(job: Job, res, i, extra=None) -> dict:
    d = {
        "convergence_info": int(res.info[i]),
        "message": "HIP direct dense solve" + (" + FP64 iterative refinement" if job.precision == "mixed" else ""),
        "backend": "rocm",
        "symmetry": job.symmetry,
        "regular_assembly_mode": "hip_row_strip_slots",
        "blas_threads": 0,
        "regular_quadrature_mode": "fixed",
        "regular_quadrature_order": job.regular_order,
        "regular_quadrature_base_order": job.regular_order,
        "precision": job.precision,
        "relative_residual": float(res.residual[i]),
        "pivot_growth": float(res.growth[i]),
        "refinement_iterations": int(res.iters[i]),
    }
    if extra:
        d.update(extra)
    return d


def fp64_errors(p_mixed, f_mixed, p_ref, f_ref) -> dict:
    """Per-result error of a mixed-precision solve against the FP64 path."""
    with np.errstate(divide="ignore", invalid="ignore"):
        db = np.abs(20.0 * np.log10(np.abs(f_mixed) / np.abs(f_ref)))
        ph = np.degrees(np.abs(np.angle(f_mixed / f_ref)))
    return {
        "fp64_pressure_rel_error": float(np.linalg.norm(p_mixed - p_ref) / max(np.linalg.norm(p_ref), 1e-300)),
        "fp64_field_max_db_error": float(np.nanmax(db)),
        "fp64_field_max_phase_error_deg": float(np.nanmax(ph)),
    }


def cancelled(job: Job) -> bool:
    return job.cancel_path is not None and os.path.isfile(str(job.cancel_path))


# Frequencies per device batch: the LU's latency-bound panel chain is shared by the batch, so larger batches
# factor more matrices per chain (serving stream of 16 mixed-size jobs: p50 / p99 latency 0.36 / 1.80 s at 8,
# 0.26 / 1.33 s at 32); 40 holds the whole 40-frequency sweep in one batch.
DEFAULT_FREQUENCY_BATCH = 40


def balanced_batch(n: int, nb_max: int) -> int:
    """Batch size for n frequencies in batches of at most nb_max, as even as possible (a padded last batch
    costs as much device work as a full one): 40 frequencies at most 32 per batch run as 2 x 20, not 32 + 8."""
    n, nb_max = max(1, int(n)), max(1, int(nb_max))
    batches = -(-n // nb_max)
    return -(-n // batches)


class JobFailed(RuntimeError):
    """A solve the engine could not stand behind (LU breakdown, non-finite values, a residual or pivot growth
    above its limit after the FP64 retry). Reported as the job's `failed` event, as the reference reports a
    singular factorization (julia_local/solver.jl:1033)."""


def status_text(bits: int) -> str:
    from .engine import STATUS_BREAKDOWN, STATUS_GROWTH, STATUS_NONFINITE, STATUS_RESIDUAL

    parts = []
    if bits & STATUS_BREAKDOWN:
        parts.append("LU breakdown (zero pivot)")
    if bits & STATUS_NONFINITE:
        parts.append("non-finite solution")
    if bits & STATUS_RESIDUAL:
        parts.append("relative residual above the tolerance")
    if bits & STATUS_GROWTH:
        parts.append("pivot growth above the limit")
    return ", ".join(parts) or "ok"


def plan_batch(job: Job, eng, batch: int | None = None) -> int:
    """Frequencies per device batch: the request's `frequency_batch` (default DEFAULT_FREQUENCY_BATCH), at most the
    job's frequency count, shrunk until every batch the job holds at once fits the device memory (the FP64 check
    batch of `fp64_check` included)."""
    n = len(job.frequencies)
    nb = int(batch or get_value(job.config, "frequency_batch", DEFAULT_FREQUENCY_BATCH))
    nb = max(1, min(nb, n))
    fit = getattr(eng, "fit_batch", None)  # device engines shrink the batch to the memory available
    if fit is not None:
        companions = ("fp64",) if (job.fp64_check and job.precision == "mixed") else ()
        nb = fit(job.precision, nb, 0 if job.obs is None else len(job.obs), companions=companions)
    return balanced_batch(n, nb)


@dataclass
class RawResult:
    """One solved frequency before post-processing: boundary pressure (V, C), field (n_obs, C), the batch result
    and index its diagnostics come from, extra diagnostics, per-frequency timings."""
    index: int
    freq: float
    pressure: np.ndarray
    field: np.ndarray
    res: object
    j: int
    extra: dict
    timings: dict


def solve_raw(job: Job, eng, nb: int, freqs=None, graph: bool = True):
    """Generator over RawResult of every frequency (the job's, or `freqs`) in order (solver.jl:805-997).

    Frequencies run in device batches of nb. A job of several batches captures the batch's launch sequence as
    HIP graphs after its first (eager) batch and replays them for the rest. Each batch's solves are checked on
    the device (bem_solve_status): a mixed-precision frequency whose LU broke down, whose solution is not finite,
    whose refinement missed its tolerance or whose pivot growth is above the limit is solved again on the plain
    FP64 path (one bounded retry, as the reference retries a corrupt rocSOLVER result,
    JS/BeatEngineRocmOperators.jl:123-130); if that fails too, or an FP64 job fails its checks, JobFailed is
    raised. Closing the generator (e.g. on cancellation) releases the batches."""
    freqs = job.frequencies.astype(np.float64) if freqs is None else np.asarray(freqs, dtype=np.float64)
    n_obs = 0 if job.obs is None else len(job.obs)
    b = eng.batch(job.precision, nb, job.obs, job.rho, job.sound_speed)
    ref = eng.batch("fp64", nb, job.obs, job.rho, job.sound_speed) if (job.fp64_check and job.precision == "mixed") \
        else None
    retry = None
    try:
        for bi, i0 in enumerate(range(0, len(freqs), nb)):
            if graph and bi == 1 and getattr(b, "capture", None) is not None:
                b.capture(warm=False)
            fr = freqs[i0 : i0 + nb]
            res = b.run(fr)
            rref = ref.run(fr) if ref is not None else None
            n = len(fr)
            flags = getattr(res, "flags", None)
            for j in range(n):
                bits = 0 if flags is None else int(flags[j])
                p_j, f_j, r_j, jj, extra = res.pressure[j], res.field[j], res, j, {}
                if bits:
                    freq = float(np.float32(fr[j]))
                    if job.precision != "mixed":
                        raise JobFailed(f"Dense solve failed at {freq:g} Hz: {status_text(bits)}.")
                    if retry is None:
                        retry = _retry_batch(eng, b, job, n_obs)
                    rr = retry.run(fr[j : j + 1])
                    rbits = 0 if getattr(rr, "flags", None) is None else int(rr.flags[0])
                    if rbits:
                        raise JobFailed(f"Dense solve failed at {freq:g} Hz: {status_text(bits)} in the mixed "
                                        f"solve, {status_text(rbits)} in the FP64 retry.")
                    p_j, f_j, r_j, jj = rr.pressure[0], rr.field[0], rr, 0
                    extra = {"fp64_retry": True, "fp64_retry_reason": status_text(bits), "precision": "fp64"}
                if rref is not None:
                    extra.update(fp64_errors(p_j, f_j, rref.pressure[j], rref.field[j]))
                timings = {"assembly_s": res.t_assembly / n, "solve_s": res.t_solve / n, "field_s": res.t_field / n}
                yield RawResult(i0 + j, float(fr[j]), p_j, f_j, r_j, jj, extra, timings)
    finally:
        for x in (b, ref, retry):
            if x is not None:
                x.close()


def payload_for(job: Job, raw: RawResult, freq: float, extra_diagnostics=None) -> dict:
    """The `result` payload of one solved frequency for `job` (solver.jl:921-996)."""
    extra = dict(raw.extra)
    if extra_diagnostics:
        extra.update(extra_diagnostics)
    return result_payload(job, freq, raw.pressure, raw.field, raw.timings, diagnostics_for(job, raw.res, raw.j, extra))


def solve_results(job: Job, eng, nb: int, graph: bool = True, extra_diagnostics=None):
    """Generator over (index, payload) of every frequency of the job in request order (see solve_raw)."""
    gen = solve_raw(job, eng, nb, graph=graph)
    try:
        for raw in gen:
            yield raw.index, payload_for(job, raw, float(job.frequencies[raw.index]), extra_diagnostics)
    finally:
        gen.close()


def _retry_batch(eng, b, job: Job, n_obs: int):
    """FP64 batch of one frequency for retries. Its workspace may need to grow the mesh's, which captured graphs
    prevent: the main batch then drops its graphs and continues eagerly."""
    from . import _lib

    try:
        return eng.batch("fp64", 1, job.obs, job.rho, job.sound_speed)
    except _lib.BemError as exc:
        if exc.status != _lib.BEM_E_STATE:
            raise
        b.close()
        return eng.batch("fp64", 1, job.obs, job.rho, job.sound_speed)


def run_job(request, engine_factory, emit, batch: int | None = None, graph: bool = True):
    """Drive one request to completion, emitting event dicts through ``emit`` (solver.jl:680-1020).

    engine_factory(prepared_mesh, qunit) -> beat_hip.engine.Engine. Cancellation (the request's cancel_path) is
    checked before every frequency's result, as the reference checks it before every frequency
    (solver.jl:805-809); device work already enqueued for the current batch completes first. A failed solve
    ends the job with a `failed` event (solver.jl:1033)."""
    init, parsed = parse_request(request)
    emit({"type": "initialized", **init})
    emit({"type": "status", "message": f"BEAT Engine loading {len(parsed['mesh_inputs'])} mesh"
          f"{'' if len(parsed['mesh_inputs']) == 1 else 'es'} (HIP engine)"})
    job = build_job(request, parsed)
    prep = job.prepare()
    emit({"type": "status", "message": f"Initializing BEAT Engine using native HIP assembly and GPU solve "
          f"({job.precision}, {job.mesh.n_faces} triangles, {job.mesh.n_vertices} P1 dofs)..."})
    if cancelled(job):
        emit({"type": "cancelled", "solved_count": 0})
        return
    eng = engine_factory(prep, job.qunit)
    total = len(job.frequencies)
    solved = 0
    try:
        nb = plan_batch(job, eng, batch)
        gen = solve_results(job, eng, nb, graph=graph)
        try:
            for i, payload in gen:
                if cancelled(job):
                    emit({"type": "cancelled", "solved_count": solved})
                    return
                solved = i + 1
                emit({"type": "result", "solved_count": solved, "total_count": total, "result": payload})
        except JobFailed as exc:
            emit({"type": "failed", "error": str(exc)})
            return
        finally:
            gen.close()
    finally:
        eng.close()
    emit({"type": "completed", "solved_count": total})


def now():
    return time.perf_counter()
