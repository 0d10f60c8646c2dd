"""Solver-backend contract types and wire helpers.

When Boundary Lab itself is importable (``blab`` on Python >= 3.11), its own classes are used so that objects
cross the boundary unchanged. Otherwise this module provides field-for-field equivalents of
src/blab/solvers/base.py:13-110 and the wire helpers of src/blab/protocol.py:170-278,304-317.
"""

from __future__ import annotations

from dataclasses import dataclass, field, fields, is_dataclass, asdict
from typing import Any, Callable

import numpy as np

try:  # pragma: no cover - exercised only where Boundary Lab is installed
    from blab.solvers.base import (  # type: ignore
        FrequencyResult,
        FrequencySolveTimings,
        SolveMetadata,
        SolveRequest,
        SolverCapabilities,
        SolverDiagnostics,
    )
    from blab.protocol import frequency_result_from_dict, solve_request_from_config_and_frequencies  # type: ignore

    HAVE_BLAB = True
except Exception:  # Boundary Lab not importable: local equivalents
    HAVE_BLAB = False

    @dataclass(frozen=True)
    class FrequencySolveTimings:
        assembly_s: float = 0.0
        solve_s: float = 0.0
        field_s: float = 0.0

    @dataclass(frozen=True)
    class SolverDiagnostics:
        convergence_info: int | None = None
        message: str | None = None
        backend: str | None = None
        symmetry: str | None = None
        regular_assembly_mode: str | None = None

    @dataclass(frozen=True)
    class FrequencyResult:
        freq_hz: float
        horizontal_spl_norm_db: np.ndarray
        vertical_spl_norm_db: np.ndarray
        impedance: np.ndarray
        horizontal_spl_db: np.ndarray | None = None
        vertical_spl_db: np.ndarray | None = None
        sphere_spl_norm_db: np.ndarray | None = None
        channel_names: np.ndarray | None = None
        horizontal_pressure: np.ndarray | None = None
        vertical_pressure: np.ndarray | None = None
        sphere_pressure: np.ndarray | None = None
        timings: FrequencySolveTimings = field(default_factory=FrequencySolveTimings)
        diagnostics: SolverDiagnostics | None = None

        @property
        def has_channel_basis(self) -> bool:
            return (
                self.channel_names is not None
                and self.horizontal_pressure is not None
                and self.vertical_pressure is not None
            )

    @dataclass(frozen=True)
    class SolveRequest:
        config: Any
        frequencies_hz: np.ndarray
        assets: tuple = ()
        status_callback: Callable[[str], None] | None = field(default=None, compare=False, repr=False)

    @dataclass(frozen=True)
    class SolveMetadata:
        polar_angle_deg: np.ndarray
        radiator_names: np.ndarray
        sphere_metadata: dict | None = None

    @dataclass(frozen=True)
    class SolverCapabilities:
        supports_spherical_sampling: bool = True
        supports_impedance: bool = True
        supports_burton_miller: bool = True
        supports_flat_target_normalization: bool = True
        supports_channel_resynthesis: bool = False
        supports_cancellation: bool = True
        supports_streaming: bool = True
        supports_remote_assets: bool = False
        supports_parallel_workers: bool = False
        supports_symmetry: bool = False
        is_remote: bool = False

    def _f32(raw):
        return None if raw is None else np.asarray(raw, dtype=np.float32)

    def _c64(raw):
        if raw is None:
            return None
        if not isinstance(raw, dict):
            return np.asarray(raw, dtype=np.complex64)
        return (np.asarray(raw.get("real"), np.float32) + 1j * np.asarray(raw.get("imag"), np.float32)).astype(
            np.complex64)

    def frequency_result_from_dict(raw: dict) -> FrequencyResult:
        """src/blab/protocol.py:263-278 (diagnostics keep the five contract keys, :250-260)."""
        t = raw.get("timings") or {}
        d = raw.get("diagnostics")
        diag = None
        if d is not None:
            ci = d.get("convergence_info")
            diag = SolverDiagnostics(
                convergence_info=None if ci is None else int(ci),
                message=None if d.get("message") is None else str(d["message"]),
                backend=None if d.get("backend") is None else str(d["backend"]),
                symmetry=None if d.get("symmetry") is None else str(d["symmetry"]),
                regular_assembly_mode=None if d.get("regular_assembly_mode") is None else str(d["regular_assembly_mode"]),
            )
        return FrequencyResult(
            freq_hz=float(raw["freq_hz"]),
            horizontal_spl_norm_db=_f32(raw["horizontal_spl_norm_db"]),
            vertical_spl_norm_db=_f32(raw["vertical_spl_norm_db"]),
            impedance=_f32(raw["impedance"]),
            horizontal_spl_db=_f32(raw.get("horizontal_spl_db")),
            vertical_spl_db=_f32(raw.get("vertical_spl_db")),
            sphere_spl_norm_db=_f32(raw.get("sphere_spl_norm_db")),
            channel_names=None if raw.get("channel_names") is None else np.asarray(raw.get("channel_names")),
            horizontal_pressure=_c64(raw.get("horizontal_pressure")),
            vertical_pressure=_c64(raw.get("vertical_pressure")),
            sphere_pressure=_c64(raw.get("sphere_pressure")),
            timings=FrequencySolveTimings(float(t.get("assembly_s", 0.0)), float(t.get("solve_s", 0.0)),
                                          float(t.get("field_s", 0.0))),
            diagnostics=diag,
        )

    def solve_request_from_config_and_frequencies(config, frequencies_hz, *, include_assets=False) -> dict:
        """src/blab/protocol.py:304-317 for dict or dataclass configs (assets are staged by the caller)."""
        if is_dataclass(config):
            cfg = asdict(config)
        else:
            cfg = dict(config)
        return {"schema_version": 2, "config": cfg,
                "frequencies_hz": np.asarray(frequencies_hz, dtype=np.float32).tolist()}


BEAT_ROCM_CAPABILITIES = SolverCapabilities(
    supports_remote_assets=False,
    supports_parallel_workers=False,
    supports_symmetry=True,
    supports_channel_resynthesis=True,
    is_remote=False,
)


def config_to_request_dict(config) -> dict:
    """A SimulationConfig (dataclass) or plain dict as the request ``config`` object."""
    if isinstance(config, dict):
        return dict(config)
    if is_dataclass(config):
        out = {}
        for f in fields(config):
            v = getattr(config, f.name)
            out[f.name] = asdict(v) if is_dataclass(v) else (
                [asdict(x) if is_dataclass(x) else x for x in v] if isinstance(v, (list, tuple)) else v)
        return out
    raise TypeError("config must be a SimulationConfig-like dataclass or a dict")
