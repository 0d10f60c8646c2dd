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
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=3, batch=2, engine_factory=SlowEngine, merge=False)
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
    sched = S.Scheduler(hbm_budget_bytes=1, slots=4, batch=2, engine_factory=SlowEngine, merge=False)
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
    a = S.estimate_footprint(1000, 2000, 27000, 4, 146)
    b = S.estimate_footprint(4000, 8000, 108000, 4, 146)
    assert 10 < b / a < 17
    # more than one frequency group (8 frequencies, strips of 4): a second strip set while a set stays within
    # 4 GiB, so doubling the batch doubles everything but the mesh; a larger set stays single
    small_set = 4 * 3 * 2000 * 1024 * 8 * 2
    assert S.estimate_footprint(1000, 2000, 27000, 8, 146) - 2 * a > -small_set / 2
    big_set = 4 * 3 * 8000 * 4032 * 8 * 2
    assert big_set > 4 << 30
    assert S.estimate_footprint(4000, 8000, 108000, 8, 146) - 2 * b < -big_set / 2


ARTIFACT_KEYS = {"freq_hz", "polar_angle_deg", "horizontal_spl_db", "vertical_spl_db", "horizontal_spl_norm_db",
                 "vertical_spl_norm_db", "impedance_freq_hz", "impedance_radiator_names", "impedance_real",
                 "impedance_imag"}
SPHERE_KEYS = {"sphere_r_distance_m", "sphere_theta_polar_rad", "sphere_phi_azimuth_rad", "sphere_spl_norm_db"}


def test_result_artifact_written_on_completion(tmp_path):
    # the server's result.npz (src/blab/server.py:477-511): ascending frequencies, Float32 rows taken from the
    # result events, impedance as [radiator][frequency], balloon arrays when the sphere is sampled
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=1, batch=2, engine_factory=FakeEngine,
                        artifact_root=tmp_path / "jobs")
    path = _mesh(tmp_path)
    req = _req(path)
    req["frequencies_hz"] = [900.0, 300.0, 1500.0, 600.0, 1200.0]
    req["config"].update(spherical_sampling_enabled=True, spherical_sampling_points=16)
    h, init = sched.submit(req)
    ev = _drain(h)
    assert ev[-1]["type"] == "completed"
    assert ev[-1]["artifact"] == f"/jobs/{h.job_id}/artifacts/result.npz"
    assert h.artifact_path == tmp_path / "jobs" / str(h.job_id) / "result.npz"
    results = sorted((e["result"] for e in ev if e["type"] == "result"), key=lambda r: r["freq_hz"])
    with np.load(h.artifact_path) as z:
        assert set(z.files) == ARTIFACT_KEYS | SPHERE_KEYS
        assert z["freq_hz"].dtype == np.float32
        assert z["freq_hz"].tolist() == [300.0, 600.0, 900.0, 1200.0, 1500.0]
        n_ang = len(init["polar_angle_deg"])
        for key in ("horizontal_spl_db", "vertical_spl_db", "horizontal_spl_norm_db", "vertical_spl_norm_db"):
            assert z[key].shape == (5, n_ang) and z[key].dtype == np.float32
            for i, r in enumerate(results):
                np.testing.assert_array_equal(z[key][i], np.asarray(r[key], np.float32))
        n_rad = len(init["radiator_names"])
        assert z["impedance_real"].shape == (n_rad, 5)
        for i, r in enumerate(results):
            imp = np.asarray(r["impedance"], np.float32)
            np.testing.assert_array_equal(z["impedance_real"][:, i], imp[:, 0])
            np.testing.assert_array_equal(z["impedance_imag"][:, i], imp[:, 1])
        assert z["impedance_radiator_names"].tolist() == list(init["radiator_names"])
        assert z["sphere_spl_norm_db"].shape == (5, 16)
        np.testing.assert_array_equal(z["sphere_theta_polar_rad"],
                                      np.asarray(init["sphere_metadata"]["theta_polar_rad"], np.float32))
    # a cancelled job leaves no artifact, as the server returns before writing it
    h2, _ = sched.submit(_req(path, nf=40))
    h2.cancel()
    assert _drain(h2)[-1]["type"] == "cancelled" and h2.artifact_path is None
    sched.shutdown()


def test_artifact_bundle_keeps_last_result_per_frequency_and_skips_balloon_without_sphere():
    init = {"polar_angle_deg": [0.0, 90.0], "radiator_names": ["a", "b"], "sphere_metadata": None}

    def res(f, v):
        return {"freq_hz": f, "horizontal_spl_norm_db": [v, v], "vertical_spl_norm_db": [v, v],
                "impedance": [[v, -v], [2 * v, 0.0]]}

    b = S.result_artifact_bundle(init, [res(200.0, 1.0), res(100.0, 2.0), res(200.0, 3.0)])
    assert set(b) == ARTIFACT_KEYS
    assert b["freq_hz"].tolist() == [100.0, 200.0]
    assert b["horizontal_spl_db"].tolist() == [[2.0, 2.0], [3.0, 3.0]]  # raw rows fall back to the normalised
    assert b["impedance_real"].tolist() == [[2.0, 3.0], [4.0, 6.0]]
    assert S.result_artifact_bundle(init, []) is None


def test_health_payload_matches_server_schema():
    h = S.health_payload()
    assert set(h) == {"status", "solver", "backend", "solver_label", "capabilities"}
    assert h["status"] == "ok" and h["backend"] == "beat_rocm" and h["solver_label"] == "BEAT Engine (AMD ROCm)"
    assert set(h["capabilities"]) == set(S.CAPABILITY_KEYS) and len(S.CAPABILITY_KEYS) == 8
    assert all(h["capabilities"].values())  # beat_rocm advertises every flag the route reports
    assert json.loads(json.dumps(h)) == h


class RecordingEngine(SlowEngine):
    """Slow double that records the frequency lists of its device batches."""

    batches = []

    def batch(self, precision, nf, obs, rho, c, **kw):
        inner = super().batch(precision, nf, obs, rho, c, **kw)
        rec = RecordingEngine.batches

        class B:
            def run(self, freqs):
                rec.append(list(np.round(freqs, 3)))
                return inner.run(freqs)

            def close(self):
                pass

        return B()


def test_identical_meshes_merge_into_shared_batches(tmp_path):
    # while one job holds the only slot, three requests for the same mesh and drives queue up; they are admitted
    # together and their frequencies share device batches, each job still receiving exactly its own results
    RecordingEngine.batches = []
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=1, batch=8, engine_factory=RecordingEngine)
    other = _mesh(tmp_path, 2, "other.msh")
    same = _mesh(tmp_path, 1, "same.msh")
    h0 = sched.submit(_req(other, nf=200))[0]  # 25 slow batches hold the slot for about half a second
    time.sleep(0.1)
    reqs = [_req(same, nf=3), _req(same, nf=4), _req(same, nf=2)]
    reqs[1]["frequencies_hz"] = [250.0, 750.0, 1250.0, 1750.0]
    hs = [sched.submit(r)[0] for r in reqs]
    ev = [_drain(h) for h in [h0] + hs]
    sched.shutdown()
    assert sched.merged_jobs == 2 and [h.group_size for h in hs] == [3, 3, 3]
    for r, e in zip(reqs, ev[1:]):
        res = [x["result"] for x in e if x["type"] == "result"]
        assert e[-1]["type"] == "completed"
        assert [x["freq_hz"] for x in res] == [float(np.float32(f)) for f in r["frequencies_hz"]]
        assert [x["solved_count"] for x in e if x["type"] == "result"] == list(range(1, len(res) + 1))
    # the 9 frequencies of the merged jobs ran in two balanced batches (5 + 4), not as 3 + 4 + 2
    assert sorted(len(b) for b in RecordingEngine.batches[-2:]) == [4, 5]


def test_bypass_is_bounded(tmp_path):
    # a large job that does not fit beside running small ones may be passed at most max_bypass times
    small = _mesh(tmp_path, 1, "small.msh")
    big = _mesh(tmp_path, 3, "big.msh")
    from beat_hip import driver as D
    from beat_hip import geometry as G

    def fp(path, nf):
        init, parsed = D.parse_request(_req(path, nf))
        job = D.build_job(_req(path, nf), parsed)
        prep = G.prepare(job.mesh)
        return S.Scheduler._footprint(None, job, prep, min(2, nf))

    budget = fp(big, 4) + fp(small, 4) // 2  # the big job fits alone, not beside a small one
    sched = S.Scheduler(hbm_budget_bytes=budget, slots=3, batch=2, engine_factory=SlowEngine, merge=False,
                        max_bypass=2)
    first = sched.submit(_req(small, nf=12))[0]
    time.sleep(0.15)
    hb = sched.submit(_req(big, nf=4))[0]
    time.sleep(0.15)
    smalls = [sched.submit(_req(small, nf=12))[0] for _ in range(6)]
    for h in [first, hb] + smalls:
        assert _drain(h)[-1]["type"] == "completed"
    sched.shutdown()
    passed = sum(1 for h in smalls if h.t_admit < hb.t_admit)
    assert hb.bypassed <= 2 and passed <= 2


def test_arrival_time_is_the_submission(tmp_path):
    sched = S.Scheduler(hbm_budget_bytes=1 << 40, slots=1, batch=4, engine_factory=FakeEngine)
    t0 = time.perf_counter()
    h, _ = sched.submit(_req(_mesh(tmp_path), nf=2), arrival=t0 - 5.0)
    _drain(h)
    sched.shutdown()
    assert h.latency >= 5.0 and h.t_ready is not None and h.t_admit >= h.t_ready
