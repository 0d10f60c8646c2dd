"""In-process ``beat_rocm`` solver backend (the L5 contract of the reference).

Mirrors BeatEngineRocmBackend (src/blab/solvers/beat_engine_backend.py:465-585) and the registry entry
(src/blab/solvers/registry.py:59-71): ``backend_id="beat_rocm"``, label, capabilities, ``create_session``,
``warm_up``; sessions expose ``metadata``, ``solve_stream(stop_requested=...)`` and ``stop()``
(src/blab/solvers/base.py:92-110). Unlike the subprocess adapter, sessions from several server threads share one
device scheduler (continuous batching), so ``blab server --max-running-jobs N`` overlaps jobs on the GPU.

Register with Boundary Lab's registry (src/blab/solvers/registry.py:59-71) by replacing the ``beat_rocm`` factory with
``lambda **kw: BeatHipRocmBackend(**kw)``.
"""

from __future__ import annotations

import queue
from typing import Callable, Iterator

import numpy as np

from . import contract as K
from .driver import DEFAULT_FREQUENCY_BATCH
from .serving import shared_scheduler


class BeatHipSession:
    def __init__(self, request: "K.SolveRequest", scheduler):
        self.request_payload = request
        cfg = K.config_to_request_dict(request.config)
        payload = {"schema_version": 2, "config": cfg,
                   "frequencies_hz": np.asarray(request.frequencies_hz, dtype=np.float32).tolist(),
                   "beat_engine_backend": "rocm"}
        self._status("Initializing BEAT Engine")
        self.handle, init = scheduler.submit(payload)
        first = self.handle.events.get()
        assert first["type"] == "initialized"
        sm = init.get("sphere_metadata") or {}
        self._metadata = K.SolveMetadata(
            polar_angle_deg=np.asarray(init["polar_angle_deg"], dtype=np.float32),
            radiator_names=np.asarray(init.get("radiator_names", ["Radiator"])),
            sphere_metadata={k: np.asarray(v, dtype=np.float32) for k, v in sm.items()},
        )
        self._stopped = False

    @property
    def metadata(self):
        return self._metadata

    def _status(self, msg):
        cb = getattr(self.request_payload, "status_callback", None)
        if cb is not None:
            cb(msg)

    def solve_stream(self, *, stop_requested: Callable[[], bool] | None = None) -> Iterator["K.FrequencyResult"]:
        while True:
            if stop_requested is not None and stop_requested():
                self.stop()
            try:
                ev = self.handle.events.get(timeout=0.1)
            except queue.Empty:
                continue
            t = ev.get("type")
            if t == "result":
                if not self._stopped:
                    yield K.frequency_result_from_dict(ev["result"])
            elif t == "status":
                self._status(str(ev.get("message", "")))
            elif t in ("completed", "cancelled"):
                return
            elif t == "failed":
                raise RuntimeError(str(ev.get("error", "BEAT Engine solver failed.")))

    def stop(self) -> None:
        self._stopped = True
        self.handle.cancel()


class BeatHipRocmBackend:
    backend_id = "beat_rocm"
    label = "BEAT Engine (AMD ROCm)"
    beat_engine_backend = "rocm"
    capabilities = K.BEAT_ROCM_CAPABILITIES

    def __init__(self, *, slots: int = 4, batch: int = DEFAULT_FREQUENCY_BATCH, hbm_budget_bytes: int | None = None,
                 **_ignored):
        self._sched_kw = dict(slots=slots, batch=batch, hbm_budget_bytes=hbm_budget_bytes)

    @property
    def scheduler(self):
        return shared_scheduler(**self._sched_kw)

    def create_session(self, request: "K.SolveRequest") -> BeatHipSession:
        return BeatHipSession(request, self.scheduler)

    def warm_up(self, mode: str = "worker", *, status_callback=None) -> None:
        mode = str(mode or "off").strip().lower()
        if mode in ("off", "0", "false", "no", "none"):
            return
        if status_callback:
            status_callback("Initializing BEAT Engine (HIP)")
        from .worker import shared_device

        shared_device()
        if status_callback:
            status_callback("BEAT Engine ready")


def create_backend(**kwargs) -> BeatHipRocmBackend:
    """Factory with the signature the reference registry uses (registry.py:59-71,116-120)."""
    return BeatHipRocmBackend(**kwargs)
