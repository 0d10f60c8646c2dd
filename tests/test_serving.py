"""Serving scheduler and in-process backend with a test double for the device (host logic only)."""

import json
import threading
import time

import numpy as np
import pytest

from beat_hip import backend as BK
from beat_hip import contract as K
from beat_hip import serving as S
from beat_hip import synthetic as syn
from test_worker_protocol import FakeBatch, FakeEngine


class SlowEngine(FakeEngine):
    active = 0
    peak = 0
    lock = threading.Lock()

    def batch(self, precision, nf, obs, rho, c, **kw):
        eng = self

        class B(FakeBatch):
            def run(self, freqs):
                with SlowEngine.lock:
                    SlowEngine.active += 1
                    SlowEngine.peak = max(SlowEngine.peak, SlowEngine.active)
                time.sleep(0.02)
                try:
                    return super().run(freqs)
                finally:
                    with SlowEngine.lock:
                        SlowEngine.active -= 1

        return B(eng.prep, eng.qunit, obs, rho, c, nf)


def _mesh(tmp_path, n=1, name="m.msh"):
    V, F, T = syn.icosphere(n, 0.1)
    p = tmp_path / name
    syn.write_gmsh22(p, V, F, T, scale=1.0)
    return str(p)


def _req(path, nf=5):
    return {"schema_version": 2, "frequencies_hz": list(np.linspace(200, 2000, nf)),
            "config": {"mesh_file": path, "scale_factor": 1.0, "step_size": 45.0}}


def _drain(h):
    out = []
    while True:
        ev = h.events.get(timeout=30)
        out.append(ev)
        if ev["type"] in ("completed", "failed", "cancelled"):
            return out


def test_jobs_overlap_and_complete(tmp_path):
    SlowEngine.peak = 0
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=3, batch=2, engine_factory=SlowEngine)
    path = _mesh(tmp_path)
    handles = [sched.submit(_req(path, nf=6))[0] for _ in range(5)]
    results = [_drain(h) for h in handles]
    sched.shutdown()
    for ev in results:
        assert ev[0]["type"] == "initialized" and ev[-1]["type"] == "completed"
        assert [e["solved_count"] for e in ev if e["type"] == "result"] == list(range(1, 7))
    assert 2 <= sched.peak_running <= 3
    assert SlowEngine.peak >= 2
    assert all(h.latency is not None and h.latency > 0 for h in handles)


def test_admission_by_footprint_serialises_big_jobs(tmp_path):
    sched = S.Scheduler(hbm_budget_bytes=1, slots=4, batch=2, engine_factory=SlowEngine)
    path = _mesh(tmp_path)
    handles = [sched.submit(_req(path, nf=3))[0] for _ in range(3)]
    for h in handles:
        _drain(h)
    sched.shutdown()
    assert sched.peak_running == 1  # nothing fits beside a running job; one at a time, none starved


def test_same_results_alone_and_in_mix(tmp_path):
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=4, batch=3, engine_factory=FakeEngine)
    a = _mesh(tmp_path, 1, "a.msh")
    b = _mesh(tmp_path, 2, "b.msh")
    alone = _drain(sched.submit(_req(a, nf=7))[0])
    hs = [sched.submit(_req(b, nf=9))[0], sched.submit(_req(a, nf=7))[0], sched.submit(_req(b, nf=4))[0]]
    mixed = [_drain(h) for h in hs][1]
    sched.shutdown()
    ra = [e["result"] for e in alone if e["type"] == "result"]
    rm = [e["result"] for e in mixed if e["type"] == "result"]
    for x, y in zip(ra, rm):
        x = dict(x, diagnostics=None, timings=None)
        y = dict(y, diagnostics=None, timings=None)
        assert json.dumps(x, sort_keys=True) == json.dumps(y, sort_keys=True)


def test_cancel_and_failure(tmp_path):
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=1, batch=1, engine_factory=SlowEngine)
    path = _mesh(tmp_path)
    h1, _ = sched.submit(_req(path, nf=50))
    h1.cancel()
    ev = _drain(h1)
    assert ev[-1]["type"] == "cancelled"
    h2, init = sched.submit(_req(str(tmp_path / "missing.msh")))
    ev = _drain(h2)
    assert ev[-1]["type"] == "failed"
    sched.shutdown()


def test_backend_contract(tmp_path, monkeypatch):
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=2, batch=2, engine_factory=FakeEngine)
    monkeypatch.setattr(BK, "shared_scheduler", lambda **kw: sched)
    be = BK.create_backend()
    assert be.backend_id == "beat_rocm" and be.label == "BEAT Engine (AMD ROCm)"
    caps = be.capabilities
    assert caps.supports_symmetry and caps.supports_channel_resynthesis and not caps.supports_parallel_workers
    assert not caps.is_remote
    path = _mesh(tmp_path)
    statuses = []
    req = K.SolveRequest(config={"mesh_file": path, "scale_factor": 1.0, "step_size": 30.0},
                         frequencies_hz=np.array([300.0, 600.0, 900.0], np.float32),
                         status_callback=statuses.append)
    session = be.create_session(req)
    assert session.metadata.polar_angle_deg.shape == (13,)
    res = list(session.solve_stream())
    assert [r.freq_hz for r in res] == [300.0, 600.0, 900.0]
    assert res[0].has_channel_basis and res[0].diagnostics.backend == "rocm"
    sched.shutdown()


def test_footprint_grows_with_n_squared():
    a = S.estimate_footprint(1000, 2000, 27000, 8, 146)
    b = S.estimate_footprint(4000, 8000, 108000, 8, 146)
    assert 10 < b / a < 17
