"""Continuous-batching serving loop for the ``blab server`` job path.

The reference serves one job at a time per GPU (the worker protocol is serial behind a lock,
src/blab/solvers/beat_engine_backend.py:307-334; the server pool defaults to one running job,
src/blab/server.py:154-185). Here one process-wide scheduler owns the GPU:

* ``submit`` validates the request and returns its ``initialized`` metadata at once; mesh loading and
  preparation run on a small host pool, so a request's arrival time is when it was submitted, and a job waiting
  for host preparation or for the device is counted in its latency.
* Admission by HBM footprint: a prepared job is admitted when a device slot is free and its estimated footprint
  (mesh tables + strip workspace + batch matrices, which grow with N^2) fits in the remaining budget. Jobs are
  considered in arrival order; a smaller job may pass one that does not fit, but once a job has been passed
  ``max_bypass`` times nothing behind it is admitted until it is (no starvation of large meshes).
* Each admitted job runs on its own slot (HIP stream + context, one stream per slot so that the slots map onto
  the device's hardware queues), so kernels of different jobs overlap on the device; its mesh, singular tables
  and factor buffers stay resident between its frequency batches. Slots launch their batches eagerly: a stream
  capture in one slot thread was invalidated by the allocations and copies the other slots make meanwhile
  (the worker path, one device thread, captures its multi-batch jobs).
* Jobs whose mesh, drives, observation points, medium and precision are identical (``mesh_key``) and that are
  waiting at the same time are merged: one device mesh, their frequencies concatenated into shared device
  batches (fuller batches amortise the LU's latency-bound panel chain), each job's results routed back to it.
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
    t_ready: float | None = None  # host preparation finished
    t_admit: float | None = None
    n_results: int = 0
    footprint: int = 0
    init: dict | None = None
    results: list = field(default_factory=list)
    artifact_path: Path | None = None
    key: str | None = None
    bypassed: int = 0
    group_size: int = 0  # jobs that shared its device batches (1: alone)

    def cancel(self):
        self.cancel_flag.set()

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


@dataclass
class _Pending:
    h: JobHandle
    job: D.Job
    prep: G.PreparedMesh
    nb: int


class Scheduler:
    def __init__(self, hbm_budget_bytes: int | None = None, slots: int = 4, batch: int = D.DEFAULT_FREQUENCY_BATCH,
                 device_index: int = 0, engine_factory=None, artifact_root=None, prep_workers: int = 2,
                 max_bypass: int = 4, merge: bool = True, fanout: int = 1, graph: bool = False):
        from concurrent.futures import ThreadPoolExecutor

        self.artifact_root = None if artifact_root is None else Path(artifact_root)
        self.slots = int(slots)
        self.batch = int(batch)
        self.device_index = device_index
        self.max_bypass = int(max_bypass)
        self.merge = bool(merge)
        self.fanout = int(fanout)
        self.graph = bool(graph)
        self._engine_factory = engine_factory
        self._budget = hbm_budget_bytes
        self._used = 0
        self._lock = threading.Condition()
        self._pending: list[_Pending] = []
        self._running = 0
        self._next_id = 0
        self._stop = False
        self._devices = []
        self._prep_pool = ThreadPoolExecutor(max_workers=max(1, int(prep_workers)), thread_name_prefix="beat-hip-prep")
        self.completed: list[JobHandle] = []
        self.peak_running = 0
        self.merged_jobs = 0
        self._dispatcher = threading.Thread(target=self._dispatch_loop, name="beat-hip-dispatch", daemon=True)
        self._dispatcher.start()

    # -- public -------------------------------------------------------------------------------
    def submit(self, request: dict, arrival: float | None = None) -> tuple[JobHandle, dict]:
        """Validate and queue a request; returns the handle and the ``initialized`` payload at once. `arrival`
        (time.perf_counter() clock) defaults to now."""
        init, parsed = D.parse_request(request)
        with self._lock:
            jid = self._next_id
            self._next_id += 1
        h = JobHandle(jid, request, time.perf_counter() if arrival is None else float(arrival))
        h.init = init
        h.events.put({"type": "initialized", **init})
        self._prep_pool.submit(self._prepare, h, request, parsed)
        return h, init

    def shutdown(self):
        with self._lock:
            self._stop = True
            self._lock.notify_all()
        self._prep_pool.shutdown(wait=False)

    def budget(self) -> int:
        if self._budget is None:
            import torch

            free, total = torch.cuda.mem_get_info(self.device_index)
            self._budget = int(0.85 * free)
        return self._budget

    # -- internals ----------------------------------------------------------------------------
    def _footprint(self, job: D.Job, prep: G.PreparedMesh, nb: int) -> int:
        precs = (job.precision, "fp64") if (job.fp64_check and job.precision == "mixed") else (job.precision,)
        return sum(estimate_footprint(job.mesh.n_vertices, job.mesh.n_faces, len(prep.pairs.test), nb, len(job.obs),
                                      pr, len(job.channel_names)) for pr in precs)

    def _prepare(self, h: JobHandle, request: dict, parsed: dict):
        try:
            job = D.build_job(request, parsed)
            prep = job.prepare()
            # frequencies per device batch: the configured batch, shrunk until the job alone fits the budget
            nb = max(1, min(self.batch, len(job.frequencies)))
            budget = self.budget() if self._engine_factory is None else (self._budget or 1 << 62)
            while nb > 1 and self._footprint(job, prep, nb) > budget:
                nb -= 1
            nb = D.balanced_batch(len(job.frequencies), nb)
            h.footprint = self._footprint(job, prep, nb)
            h.key = mesh_key(job, prep) if self.merge else None
        except Exception as exc:
            h.t_done = time.perf_counter()
            with self._lock:
                self.completed.append(h)
            h.events.put({"type": "failed", "error": f"{type(exc).__name__}: {exc}"})
            return
        h.t_ready = time.perf_counter()
        with self._lock:
            self._pending.append(_Pending(h, job, prep, nb))
            self._pending.sort(key=lambda e: (e.h.arrival, e.h.job_id))
            self._lock.notify_all()

    def _device(self, slot):
        from . import _lib
        from .engine import Device

        while len(self._devices) <= slot:
            dev = Device(self.device_index)
            _lib.check(dev.lib.bem_ctx_set_fanout(dev.ctx, max(1, self.fanout)), "bem_ctx_set_fanout")
            self._devices.append(dev)
        return self._devices[slot]

    def _pick(self) -> int | None:
        """Index of the next pending job to admit (arrival order, bounded bypass), or None."""
        budget = self.budget() if self._engine_factory is None else (self._budget or 1 << 62)
        for idx, e in enumerate(self._pending):
            if e.h.cancel_flag.is_set() or self._used + e.h.footprint <= budget or self._running == 0:
                for skipped in self._pending[:idx]:
                    skipped.h.bypassed += 1
                return idx
            if e.h.bypassed >= self.max_bypass:
                return None  # this job has waited long enough: nothing behind it goes first
        return None

    def _dispatch_loop(self):
        free_slots = list(range(self.slots))
        slot_of = {}
        while True:
            with self._lock:
                while not self._stop:
                    for t in list(slot_of):  # reclaim finished slots
                        if not t.is_alive():
                            free_slots.append(slot_of.pop(t))
                    pick = self._pick() if (free_slots and self._pending) else None
                    if pick is not None:
                        break
                    self._lock.wait(0.05)
                if self._stop:
                    return
                lead = self._pending.pop(pick)
                group = [lead]
                if lead.h.key is not None and not lead.h.cancel_flag.is_set():
                    for e in list(self._pending):
                        if e.h.key == lead.h.key and not e.h.cancel_flag.is_set():
                            self._pending.remove(e)
                            group.append(e)
                    self.merged_jobs += len(group) - 1
                if len(group) > 1:
                    # a merged group fills batches with the union of its frequencies: grow the lead's batch to
                    # the configured size while the footprint still fits beside what is running
                    budget = self.budget() if self._engine_factory is None else (self._budget or 1 << 62)
                    total = sum(len(e.job.frequencies) for e in group)
                    nb = max(lead.nb, min(self.batch, total))
                    while nb > lead.nb and (self._footprint(lead.job, lead.prep, nb) > budget or (
                            self._running and self._used + self._footprint(lead.job, lead.prep, nb) > budget)):
                        nb -= 1
                    lead.nb = D.balanced_batch(total, nb)
                    lead.h.footprint = self._footprint(lead.job, lead.prep, nb)
                slot = free_slots.pop(0)
                self._used += lead.h.footprint
                self._running += 1
                self.peak_running = max(self.peak_running, self._running)
                now = time.perf_counter()
                for e in group:
                    e.h.t_admit = now
                    e.h.group_size = len(group)
            th = threading.Thread(target=self._run_group, args=(slot, group), name=f"beat-hip-slot{slot}", daemon=True)
            slot_of[th] = slot
            th.start()

    def _finish(self, e: _Pending, terminal: dict):
        h = e.h
        if terminal["type"] == "completed" and self.artifact_root is not None:
            # the server's artifact, written before the terminal event (src/blab/server.py:446-456)
            h.artifact_path = write_result_artifact(self.artifact_root / str(h.job_id), h.init, h.results)
            terminal["artifact"] = f"/jobs/{h.job_id}/artifacts/result.npz"
        # latency is final before the terminal event becomes visible to the consumer
        h.t_done = time.perf_counter()
        with self._lock:
            self.completed.append(h)
        h.events.put(terminal)

    def _run_group(self, slot, group: list[_Pending]):
        from .engine import Engine

        lead = group[0]
        eng = None
        open_ = {id(e): e for e in group}
        gen = None
        try:
            for e in group:
                if e.h.cancel_flag.is_set() or D.cancelled(e.job):
                    self._finish(e, {"type": "cancelled", "solved_count": 0})
                    del open_[id(e)]
            if not open_:
                return
            members = [e for e in group if id(e) in open_]
            if self._engine_factory is not None:
                eng = self._engine_factory(lead.prep, lead.job.qunit)
            else:
                eng = Engine(lead.prep, lead.job.qunit, dev=self._device(slot))
            # the members' frequencies one after the other, in shared device batches
            owner = [(e, i) for e in members for i in range(len(e.job.frequencies))]
            freqs = np.concatenate([e.job.frequencies.astype(np.float64) for e in members])
            nb = max(1, min(lead.nb, len(freqs)))  # the footprint admitted is the lead's at this batch size
            gen = D.solve_raw(lead.job, eng, nb, freqs=freqs, graph=self.graph)
            extra = {"serving_slot": slot, "serving_group": len(members)}
            for raw in gen:
                e, i = owner[raw.index]
                if id(e) not in open_:
                    continue
                if e.h.cancel_flag.is_set() or D.cancelled(e.job):
                    self._finish(e, {"type": "cancelled", "solved_count": e.h.n_results})
                    del open_[id(e)]
                    if not open_:
                        break
                    continue
                payload = D.payload_for(e.job, raw, float(e.job.frequencies[i]), extra)
                e.h.n_results = i + 1
                e.h.results.append(payload)
                e.h.events.put({"type": "result", "solved_count": i + 1, "total_count": len(e.job.frequencies),
                                "result": payload})
                if i + 1 == len(e.job.frequencies):
                    self._finish(e, {"type": "completed", "solved_count": i + 1})
                    del open_[id(e)]
        except Exception as exc:
            msg = str(exc) if isinstance(exc, D.JobFailed) else f"{type(exc).__name__}: {exc}"
            for e in list(open_.values()):
                self._finish(e, {"type": "failed", "error": msg})
            open_.clear()
        finally:
            if gen is not None:
                gen.close()
            if eng is not None:
                eng.close()
            for e in list(open_.values()):
                self._finish(e, {"type": "failed", "error": "job ended without a terminal event"})
            with self._lock:
                self._used -= lead.h.footprint
                self._running -= 1
                self._lock.notify_all()


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


def mesh_key(job: D.Job, prep: G.PreparedMesh | None = None) -> str:
    """Identity of everything the device solve of a job depends on besides its frequencies: mesh geometry and
    faces, unit drives, symmetry, observation points, medium, precision, rules and the FP64 check. Jobs with equal keys
    can share one device mesh and one stream of frequency batches."""
    h = hashlib.sha1()
    for a in (job.mesh.vertices, job.mesh.faces, job.qunit, job.obs):
        a = np.ascontiguousarray(a)
        h.update(str((a.dtype, a.shape)).encode())
        h.update(a.tobytes())
    h.update(repr((job.symmetry, job.precision, job.rho, job.sound_speed, bool(job.fp64_check), job.regular_order,
                   job.singular_order)).encode())
    return h.hexdigest()
