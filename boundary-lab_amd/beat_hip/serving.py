"""Continuous-batching serving loop for the ``blab server`` job path.

The reference serves one job at a time per GPU (the worker protocol is serial behind a lock,
src/blab/solvers/beat_engine_backend.py:307-334; the server pool defaults to one running job,
src/blab/server.py:154-185). Here one process-wide scheduler owns the GPU:

* Jobs are parsed and their ``initialized`` metadata returned immediately; device work is queued.
* Admission by HBM footprint: a job is admitted when a device slot is free and its estimated footprint
  (mesh tables + strip workspace + batch matrices, which grow with N^2) fits in the remaining budget.
* Each admitted job runs on its own slot (HIP stream + context), so kernels of different jobs overlap on the
  device; its mesh, singular tables and factor buffers stay resident between its frequency batches.
* Frequencies are batched per job (one launch set per batch); jobs whose mesh and drives are identical share
  one device mesh and are merged into the same batches.
* Per-job latency (arrival -> last result) is recorded for p50/p99 reporting.
* A completed job can leave the server's ``result.npz`` artifact (src/blab/server.py:477-511), and
  ``health_payload`` is the body of the server's ``/health`` route (src/blab/server.py:548-568).
"""

from __future__ import annotations

import hashlib
import queue
import threading
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import contract as K
from . import driver as D
from . import geometry as G
from .engine import estimate_footprint

CAPABILITY_KEYS = (
    "supports_spherical_sampling",
    "supports_impedance",
    "supports_burton_miller",
    "supports_flat_target_normalization",
    "supports_channel_resynthesis",
    "supports_cancellation",
    "supports_streaming",
    "supports_symmetry",
)


def health_payload(solver_id: str = "beat_rocm") -> dict:
    """``GET /health`` body for a server whose solver is this backend (src/blab/server.py:548-568): status,
    solver and backend ids, label and the eight advertised capability flags of ``beat_rocm``
    (src/blab/solvers/registry.py:59-71 over the SolverCapabilities defaults, src/blab/solvers/base.py:78-89)."""
    caps = K.BEAT_ROCM_CAPABILITIES
    return {
        "status": "ok",
        "solver": str(solver_id),
        "backend": "beat_rocm",
        "solver_label": "BEAT Engine (AMD ROCm)",
        "capabilities": {key: bool(getattr(caps, key)) for key in CAPABILITY_KEYS},
    }


def result_artifact_bundle(init: dict, results: list[dict]) -> dict | None:
    """Arrays of a job's ``result.npz`` from its ``initialized`` payload and ``result`` payloads, as the server
    writes them (src/blab/server.py:477-511 over LiveSolveDataset, src/blab/live.py:37-87,175-206): results
    keyed by frequency (a repeated frequency keeps the last result) in ascending order, Float32 rows, impedance
    as [radiator][frequency]. The per-result SPL rows are the channel synthesis the solver already applied with
    the request's channels (julia_local/solver.jl:546-588), which is what the server's re-synthesis with the same
    channel configuration reproduces. None when nothing was solved."""
    by_freq = {}
    for r in results:
        by_freq[float(r["freq_hz"])] = r
    if not by_freq:
        return None
    ordered = [by_freq[k] for k in sorted(by_freq)]

    def f32(v):
        return np.asarray(v, dtype=np.float32)

    def raw_or_norm(r, key):
        return r.get(key + "_spl_db") if r.get(key + "_spl_db") is not None else r[key + "_spl_norm_db"]

    freqs = f32([r["freq_hz"] for r in ordered])
    impedance = np.stack([f32(r["impedance"]) for r in ordered], axis=1)
    bundle = {
        "freq_hz": freqs,
        "polar_angle_deg": f32(init["polar_angle_deg"]),
        "horizontal_spl_db": np.vstack([f32(raw_or_norm(r, "horizontal")) for r in ordered]),
        "vertical_spl_db": np.vstack([f32(raw_or_norm(r, "vertical")) for r in ordered]),
        "horizontal_spl_norm_db": np.vstack([f32(r["horizontal_spl_norm_db"]) for r in ordered]),
        "vertical_spl_norm_db": np.vstack([f32(r["vertical_spl_norm_db"]) for r in ordered]),
        "impedance_freq_hz": freqs,
        "impedance_radiator_names": np.asarray(init["radiator_names"]).astype(str),
        "impedance_real": impedance[:, :, 0],
        "impedance_imag": impedance[:, :, 1],
    }
    sphere = init.get("sphere_metadata")
    if sphere and all(r.get("sphere_spl_norm_db") is not None for r in ordered):
        bundle.update(
            sphere_r_distance_m=f32(sphere["r_distance_m"]),
            sphere_theta_polar_rad=f32(sphere["theta_polar_rad"]),
            sphere_phi_azimuth_rad=f32(sphere["phi_azimuth_rad"]),
            sphere_spl_norm_db=np.vstack([f32(r["sphere_spl_norm_db"]) for r in ordered]),
        )
    return bundle


def write_result_artifact(directory, init: dict, results: list[dict]) -> Path | None:
    """Write ``<directory>/result.npz`` (compressed, as the server does); None when nothing was solved."""
    bundle = result_artifact_bundle(init, results)
    if bundle is None:
        return None
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    path = directory / "result.npz"
    np.savez_compressed(path, **bundle)
    return path


@dataclass
class JobHandle:
    job_id: int
    request: dict
    arrival: float
    events: "queue.Queue[dict]" = field(default_factory=queue.Queue)
    cancel_flag: threading.Event = field(default_factory=threading.Event)
    t_done: float | None = None
    t_admit: float | None = None
    n_results: int = 0
    footprint: int = 0
    init: dict | None = None
    results: list = field(default_factory=list)
    artifact_path: Path | None = None

    def cancel(self):
        self.cancel_flag.set()

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class Scheduler:
    def __init__(self, hbm_budget_bytes: int | None = None, slots: int = 4, batch: int = D.DEFAULT_FREQUENCY_BATCH,
                 device_index: int = 0, engine_factory=None, artifact_root=None):
        self.artifact_root = None if artifact_root is None else Path(artifact_root)
        self.slots = int(slots)
        self.batch = int(batch)
        self.device_index = device_index
        self._engine_factory = engine_factory
        self._budget = hbm_budget_bytes
        self._used = 0
        self._lock = threading.Condition()
        self._pending: list[tuple[JobHandle, D.Job, G.PreparedMesh]] = []
        self._running = 0
        self._next_id = 0
        self._threads: list[threading.Thread] = []
        self._stop = False
        self._devices = []
        self.completed: list[JobHandle] = []
        self.peak_running = 0
        self._dispatcher = threading.Thread(target=self._dispatch_loop, name="beat-hip-dispatch", daemon=True)
        self._dispatcher.start()

    # -- public -------------------------------------------------------------------------------
    def submit(self, request: dict) -> tuple[JobHandle, dict]:
        """Parse and queue a request; returns the handle and the ``initialized`` payload."""
        init, parsed = D.parse_request(request)
        with self._lock:
            jid = self._next_id
            self._next_id += 1
        h = JobHandle(jid, request, time.perf_counter())
        h.init = init
        h.events.put({"type": "initialized", **init})
        try:
            job = D.build_job(request, parsed)
            prep = G.prepare(job.mesh, symmetry=job.symmetry)
        except Exception as exc:
            h.events.put({"type": "failed", "error": f"{type(exc).__name__}: {exc}"})
            h.t_done = time.perf_counter()
            return h, init
        nb = max(1, min(self.batch, len(job.frequencies)))
        h.footprint = estimate_footprint(job.mesh.n_vertices, job.mesh.n_faces, len(prep.pairs.test), nb,
                                         len(job.obs), job.precision, len(job.channel_names))
        with self._lock:
            self._pending.append((h, job, prep))
            self._lock.notify_all()
        return h, init

    def shutdown(self):
        with self._lock:
            self._stop = True
            self._lock.notify_all()

    def budget(self) -> int:
        if self._budget is None:
            import torch

            free, total = torch.cuda.mem_get_info(self.device_index)
            self._budget = int(0.85 * free)
        return self._budget

    # -- internals ----------------------------------------------------------------------------
    def _device(self, slot):
        from .engine import Device

        while len(self._devices) <= slot:
            self._devices.append(Device(self.device_index))
        return self._devices[slot]

    def _dispatch_loop(self):
        free_slots = list(range(self.slots))
        slot_of = {}
        while True:
            with self._lock:
                while not self._stop:
                    pick = None
                    if free_slots and self._pending:
                        budget = self.budget()
                        for idx, (h, job, prep) in enumerate(self._pending):
                            if h.cancel_flag.is_set() or self._used + h.footprint <= budget or self._running == 0:
                                pick = idx
                                break
                    if pick is not None:
                        break
                    self._lock.wait(0.05)
                    # reclaim finished slots
                    for t in list(slot_of):
                        if not t.is_alive():
                            free_slots.append(slot_of.pop(t))
                if self._stop:
                    return
                h, job, prep = self._pending.pop(pick)
                slot = free_slots.pop(0)
                self._used += h.footprint
                self._running += 1
                self.peak_running = max(self.peak_running, self._running)
                h.t_admit = time.perf_counter()
            th = threading.Thread(target=self._run_job, args=(slot, h, job, prep), name=f"beat-hip-slot{slot}",
                                  daemon=True)
            slot_of[th] = slot
            th.start()

    def _run_job(self, slot, h: JobHandle, job: D.Job, prep: G.PreparedMesh):
        from .engine import Engine

        eng = None
        terminal = None
        try:
            if h.cancel_flag.is_set():
                terminal = {"type": "cancelled", "solved_count": 0}
                return
            if self._engine_factory is not None:
                eng = self._engine_factory(prep, job.qunit)
            else:
                eng = Engine(prep, job.qunit, dev=self._device(slot))
            freqs = job.frequencies.astype(np.float64)
            nb = max(1, min(self.batch, len(freqs)))
            b = eng.batch(job.precision, nb, job.obs, job.rho, job.sound_speed)
            solved = 0
            for i0 in range(0, len(freqs), nb):
                if h.cancel_flag.is_set() or D.cancelled(job):
                    terminal = {"type": "cancelled", "solved_count": solved}
                    return
                fr = freqs[i0 : i0 + nb]
                res = b.run(fr)
                n = len(fr)
                for j in range(n):
                    timings = {"assembly_s": res.t_assembly / n, "solve_s": res.t_solve / n,
                               "field_s": res.t_field / n}
                    payload = D.result_payload(job, float(job.frequencies[i0 + j]), res.pressure[j], res.field[j],
                                               timings, D.diagnostics_for(job, res, j, {"serving_slot": slot}))
                    solved = i0 + j + 1
                    h.n_results = solved
                    h.results.append(payload)
                    h.events.put({"type": "result", "solved_count": solved, "total_count": len(freqs),
                                  "result": payload})
            b.close()
            terminal = {"type": "completed", "solved_count": len(freqs)}
            if self.artifact_root is not None:
                # the server's artifact, written before the terminal event (src/blab/server.py:446-456)
                h.artifact_path = write_result_artifact(self.artifact_root / str(h.job_id), h.init, h.results)
                terminal["artifact"] = f"/jobs/{h.job_id}/artifacts/result.npz"
        except Exception as exc:
            terminal = {"type": "failed", "error": f"{type(exc).__name__}: {exc}"}
        finally:
            if eng is not None:
                eng.close()
            # latency is final before the terminal event becomes visible to the consumer
            h.t_done = time.perf_counter()
            with self._lock:
                self._used -= h.footprint
                self._running -= 1
                self.completed.append(h)
                self._lock.notify_all()
            h.events.put(terminal or {"type": "failed", "error": "job ended without a terminal event"})


_SCHEDULER = None
_SCHEDULER_LOCK = threading.Lock()


def shared_scheduler(**kw) -> Scheduler:
    global _SCHEDULER
    with _SCHEDULER_LOCK:
        if _SCHEDULER is None:
            _SCHEDULER = Scheduler(**kw)
        return _SCHEDULER


def percentile(values, q):
    v = np.sort(np.asarray(values, dtype=np.float64))
    if v.size == 0:
        return float("nan")
    return float(np.percentile(v, q))


def mesh_key(mesh: G.BoundaryMesh, qunit: np.ndarray) -> str:
    h = hashlib.sha1()
    for a in (mesh.vertices, mesh.faces, qunit):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
