"""Worker NDJSON protocol and request semantics, with a test double in place of the GPU engine.

Re-expresses the reference's fake-worker contract tests (tests/test_solver_backends.py:193-267,513-666): event
order, payload schema accepted by the consumer, failure and cancellation paths. The double only fabricates
device results; every host step (parsing, drives, synthesis, SPL, impedance, wire format) is the product's.
"""

import io
import json
import math

import numpy as np
import pytest

from beat_hip import contract as K
from beat_hip import driver as D
from beat_hip import synthetic as syn
from beat_hip import worker as W
from beat_hip.engine import BatchResult


class FakeBatch:
    def __init__(self, prep, qunit, obs, rho, c, nf):
        self.prep, self.qunit, self.obs, self.rho, self.c, self.nf = prep, qunit, obs, rho, c, nf

    def run(self, freqs):
        n = len(freqs)
        V = self.prep.n_vertices
        C = self.qunit.shape[1]
        om = 2 * math.pi * np.asarray(freqs)
        k = om / self.c
        r = np.linalg.norm(self.obs, axis=1)
        field = np.stack([np.exp(1j * kk * r)[:, None] / r[:, None] * np.arange(1, C + 1)[None, :] for kk in k])
        p = np.ones((n, V, C), complex)
        q = (1j * self.rho * om)[:, None, None] * self.qunit[None]
        return BatchResult(freqs=np.asarray(freqs), pressure=p, q_source=q, field=field, residual=np.full(n, 1e-12),
                           growth=np.ones(n), iters=np.ones(n, int), info=np.zeros(n, int), t_assembly=0.1,
                           t_solve=0.2, t_field=0.05)

    def close(self):
        pass


class FakeEngine:
    def __init__(self, prep, qunit):
        self.prep, self.qunit = prep, qunit

    def batch(self, precision, nf, obs, rho, c, **kw):
        return FakeBatch(self.prep, self.qunit, obs, rho, c, nf)

    def close(self):
        pass


@pytest.fixture
def mesh_file(tmp_path):
    V, F, T = syn.icosphere(1, 0.1)
    path = tmp_path / "sphere.msh"
    syn.write_gmsh22(path, V, F, T, scale=0.001)
    return str(path)


def _request(mesh_file, freqs=(500.0, 1000.0), **cfg):
    config = {"mesh_file": mesh_file, "step_size": 30.0, "spherical_sampling_enabled": True,
              "spherical_sampling_points": 20, "rho": 1.205}
    config.update(cfg)
    return {"schema_version": 2, "config": config, "frequencies_hz": list(freqs), "beat_engine_backend": "rocm"}


def _run(req):
    events = []
    D.run_job(req, FakeEngine, events.append, batch=1)
    return events


def test_event_sequence_and_payload(mesh_file):
    ev = _run(_request(mesh_file))
    types = [e["type"] for e in ev]
    assert types[0] == "initialized" and types[-1] == "completed"
    assert types.count("result") == 2 and "status" in types
    init = ev[0]
    assert init["polar_angle_deg"][0] == -180.0 and len(init["polar_angle_deg"]) == 13
    assert init["radiator_names"] == ["Radiator"]
    assert set(init["sphere_metadata"]) == {"r_distance_m", "theta_polar_rad", "phi_azimuth_rad"}
    results = [e for e in ev if e["type"] == "result"]
    assert [r["solved_count"] for r in results] == [1, 2]
    assert all(r["total_count"] == 2 for r in results)
    res = K.frequency_result_from_dict(json.loads(json.dumps(results[0]["result"])))
    assert res.freq_hz == 500.0
    assert res.horizontal_pressure.shape == (1, 13)
    assert res.sphere_pressure.shape == (1, 20)
    assert res.impedance.shape == (1, 2)
    # normalised SPL is 0 dB on axis (solver.jl:568-569)
    on_axis = int(np.argmin(np.abs(np.asarray(init["polar_angle_deg"]))))
    assert abs(res.horizontal_spl_norm_db[on_axis]) < 1e-5
    assert res.diagnostics.backend == "rocm" and res.diagnostics.convergence_info == 0
    assert ev[-1] == {"type": "completed", "solved_count": 2}


def test_channels_and_radiators(mesh_file):
    req = _request(mesh_file, radiators=[
        {"name": "A", "tag": 2, "channel": "hf", "velocity_offset_db": 6.0},
        {"name": "B", "tag": 2, "channel": "lf"}],
        channels=[{"name": "hf", "level_db": -3.0}, {"name": "lf", "polarity": -1,
                                                      "lpf": {"type": "lowpass", "frequency_hz": 800, "order": 2}}])
    ev = _run(req)
    r = [e for e in ev if e["type"] == "result"][0]["result"]
    assert r["channel_names"] == ["hf", "lf"]
    assert len(r["horizontal_pressure"]["real"]) == 2
    assert ev[0]["radiator_names"] == ["A", "B"]


def test_failed_and_validation(mesh_file, tmp_path):
    with pytest.raises(ValueError, match="schema_version"):
        _run({**_request(mesh_file), "schema_version": 7})
    with pytest.raises(ValueError, match="step_size"):
        _run(_request(mesh_file, step_size=0))
    with pytest.raises(ValueError, match="unknown channel"):
        _run(_request(mesh_file, radiators=[{"name": "A", "tag": 2, "channel": "x"}], channels=[{"name": "y"}]))
    with pytest.raises(ValueError, match="No elements found"):
        _run(_request(mesh_file, radiators=[{"name": "A", "tag": 9}]))
    with pytest.raises(ValueError, match="Wavelength"):
        _run(_request(mesh_file, regular_quadrature_mode="wavelength"))


def test_cancel_file(mesh_file, tmp_path):
    cancel = tmp_path / "cancel"
    cancel.write_text("cancel")
    req = _request(mesh_file)
    req["cancel_path"] = str(cancel)
    ev = _run(req)
    assert ev[-1] == {"type": "cancelled", "solved_count": 0}


def test_worker_loop_ready_failed_continue(mesh_file, tmp_path, monkeypatch):
    good = tmp_path / "good.json"
    good.write_text(json.dumps(_request(mesh_file, freqs=(700.0,))))
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps(_request(str(tmp_path / "missing.msh"))))
    monkeypatch.setattr(W, "engine_factory", lambda prep, q: FakeEngine(prep, q))
    out = []
    stdin = io.StringIO("\n".join(json.dumps({"request": str(p), "operation": "solve"}) for p in (bad, good)) + "\n")
    W.worker_loop(stdin=stdin, emit=out.append)
    types = [e["type"] for e in out]
    assert out[0]["type"] == "ready" and out[0]["protocol"] == "boundary_lab_julia_worker"
    assert "failed" in types and types[-1] == "completed"
    assert types.index("failed") < types.index("completed")


def test_one_shot_argument_errors(capsys):
    assert W.main(["--bogus"]) == 1
    assert json.loads(capsys.readouterr().out.strip())["type"] == "failed"
    assert W.main([]) == 1


def test_crossover_and_impedance_formulas():
    # Butterworth: |H(fc)| = 1/sqrt(2) for any order; LR = Butterworth^2 -> -6 dB at fc
    for order in (1, 2, 4, 6):
        h = D.crossover_response({"type": "lowpass", "frequency_hz": 1000.0, "order": order}, 1000.0)
        assert abs(abs(h) - 1 / math.sqrt(2)) < 1e-12
    lr = D.crossover_response({"type": "highpass", "filter": "linkwitz_riley", "frequency_hz": 500, "order": 4}, 500)
    assert abs(abs(lr) - 0.5) < 1e-12
    # first-order high-pass: s / (s + wc)
    h = D.crossover_response({"type": "highpass", "frequency_hz": 100.0, "order": 1}, 300.0)
    s = 2j * math.pi * 300
    assert abs(h - s / (s + 2 * math.pi * 100)) < 1e-12
    from beat_hip import geometry as G

    V, F, T = syn.icosphere(1, 0.1)
    m = G.BoundaryMesh(V, F, T)
    rad = [{"tag": 2, "mesh_id": 1}]
    p = np.full(len(V), 2.0 + 1.0j)
    imp = D.impedance_for_radiators(m, p, rad, [1.0 + 0j])
    force = 10.0 * (2.0 + 1.0j) * m.areas.sum()
    assert np.allclose(imp[0], [np.float32(force.real / 2), np.float32(-force.imag / 2)])


def test_polar_points_follow_reference_conventions():
    angles, h, v, on_axis = D.polar_observation_points({"step_size": 5.0})
    assert len(angles) == 73 and on_axis == 36 and angles.dtype == np.float32
    assert np.allclose(h[on_axis], [0, 0, 2.0]) and np.allclose(v[18], [0, -2.0, 0], atol=1e-6)
    angles, *_ = D.polar_observation_points({"step_size": 7.0, "min_angle": -90, "max_angle": 90})
    assert angles[-1] == 90.0 and angles[0] == -90.0  # max_angle appended when not on the grid
    pts, meta = D.spherical_observation({"spherical_sampling_enabled": True, "spherical_sampling_points": 1000})
    assert pts.shape == (1000, 3) and np.allclose(np.linalg.norm(pts, axis=1), 2.0, atol=1e-5)
    assert abs(pts[0, 2] - 2.0 * (1 - 1 / 1000)) < 1e-5  # z = 1 - (2i+1)/n ordering


def test_symmetry_requests(tmp_path):
    """x / xy symmetry: fundamental-domain meshes are accepted (planes snapped, domain validated), the engine
    receives image tables, impedance carries the reduction factor (solver.jl:607-608); meshes crossing a
    plane fail with the reference's message."""
    V, F, T = syn.uv_sphere(6, 12, 0.1)
    seen = {}

    class Spy(FakeEngine):
        def __init__(self, prep, qunit):
            super().__init__(prep, qunit)
            seen["prep"] = prep

    imps = {}
    for mode in ("off", "x", "xy"):
        v, f, t = (V, F, T) if mode == "off" else syn.symmetric_part(V, F, T, mode)
        path = tmp_path / f"s_{mode}.msh"
        syn.write_gmsh22(path, v, f, t, scale=0.001)
        req = _request(str(path), freqs=(800.0,), symmetry=mode)
        ev = []
        D.run_job(req, Spy, ev.append, batch=1)
        assert ev[-1]["type"] == "completed", ev[-1]
        prep = seen["prep"]
        assert prep.symmetry == mode and len(prep.transforms) == {"off": 1, "x": 2, "xy": 4}[mode]
        if mode != "off":
            assert np.any(prep.pairs.trial >= prep.n_faces) and prep.row_weights.max() > 1
        res = [e for e in ev if e["type"] == "result"][0]["result"]
        imps[mode] = res["impedance"][0][0]
    # the double returns p = 1 on every vertex: force = 10 * area of the radiator * factor, same on every model
    assert imps["x"] == pytest.approx(imps["off"], rel=2e-3) and imps["xy"] == pytest.approx(imps["off"], rel=2e-3)
    bad = tmp_path / "bad.msh"
    syn.write_gmsh22(bad, V, F, T, scale=0.001)
    ev = []
    with pytest.raises(ValueError, match="positive X fundamental domain"):
        D.run_job(_request(str(bad), freqs=(800.0,), symmetry="x"), FakeEngine, ev.append, batch=1)


class FlagEngine(FakeEngine):
    """Test double whose solves report device status bits: `bad[precision]` maps a frequency to its bits."""

    bad = {}
    calls = []

    def batch(self, precision, nf, obs, rho, c, **kw):
        eng = self
        base = FakeBatch(self.prep, self.qunit, obs, rho, c, nf)

        class B:
            def run(self, freqs):
                FlagEngine.calls.append((precision, list(freqs)))
                r = base.run(freqs)
                r.flags = np.array([FlagEngine.bad.get(precision, {}).get(float(f), 0) for f in freqs], np.int32)
                if precision == "fp64":
                    r.pressure = r.pressure * 2.0  # tell the retry's result apart
                return r

            def close(self):
                pass

        return B()


def test_fp64_job_with_lu_breakdown_fails(mesh_file):
    FlagEngine.bad = {"fp64": {1000.0: 1}}
    ev = []
    req = _request(mesh_file, freqs=(500.0, 1000.0, 1500.0), precision="fp64")
    D.run_job(req, FlagEngine, ev.append, batch=3)
    # frequencies before the failing one keep their results, as in the reference's per-frequency loop
    assert [e["type"] for e in ev if e["type"] != "status"] == ["initialized", "result", "failed"]
    assert "LU breakdown" in ev[-1]["error"] and "1000" in ev[-1]["error"]


def test_mixed_failure_retried_in_fp64(mesh_file):
    # a mixed solve that misses its refinement tolerance is solved again on the FP64 path (bounded retry)
    FlagEngine.bad = {"mixed": {1000.0: 4}}
    FlagEngine.calls = []
    ev = []
    D.run_job(_request(mesh_file, freqs=(500.0, 1000.0, 1500.0)), FlagEngine, ev.append, batch=3)
    res = [e for e in ev if e["type"] == "result"]
    assert ev[-1]["type"] == "completed" and [r["solved_count"] for r in res] == [1, 2, 3]
    d = res[1]["result"]["diagnostics"]
    assert d["fp64_retry"] is True and "residual" in d["fp64_retry_reason"] and d["precision"] == "fp64"
    assert "fp64_retry" not in res[0]["result"]["diagnostics"]
    assert ("fp64", [1000.0]) in FlagEngine.calls


def test_mixed_failure_and_failed_retry_is_failed(mesh_file):
    FlagEngine.bad = {"mixed": {1500.0: 2}, "fp64": {1500.0: 2}}
    ev = []
    D.run_job(_request(mesh_file, freqs=(500.0, 1000.0, 1500.0)), FlagEngine, ev.append, batch=3)
    types = [e["type"] for e in ev if e["type"] != "status"]
    assert types == ["initialized", "result", "result", "failed"]
    assert "non-finite" in ev[-1]["error"]


def test_cancellation_is_checked_before_every_frequency(mesh_file, tmp_path):
    # one device batch of 6 frequencies; the cancel file appears after the second result
    cancel = tmp_path / "cancel"
    req = _request(mesh_file, freqs=(300.0, 400.0, 500.0, 600.0, 700.0, 800.0), cancel_path=None)
    req["cancel_path"] = str(cancel)
    ev = []

    def emit(e):
        ev.append(e)
        if e["type"] == "result" and e["solved_count"] == 2:
            cancel.write_text("1")

    D.run_job(req, FakeEngine, emit, batch=6)
    assert ev[-1] == {"type": "cancelled", "solved_count": 2}
    assert sum(e["type"] == "result" for e in ev) == 2


def test_more_than_four_channels(mesh_file, tmp_path):
    # six radiators on six channels: the reference loops over every channel (solver.jl:889-920)
    V, F, T = syn.icosphere(1, 0.1)
    T = (np.arange(len(F)) % 6 + 2).astype(int)
    path = tmp_path / "six.msh"
    syn.write_gmsh22(path, V, F, T, scale=0.001)
    rads = [{"name": f"r{i}", "tag": i + 2, "channel": f"c{i}"} for i in range(6)]
    req = _request(str(path), radiators=rads)
    ev = []
    D.run_job(req, FakeEngine, ev.append, batch=2)
    res = [e["result"] for e in ev if e["type"] == "result"]
    assert ev[-1]["type"] == "completed" and len(res[0]["channel_names"]) == 6
    assert len(res[0]["horizontal_pressure"]["real"]) == 6


def test_fit_batch_counts_the_fp64_companion():
    from types import SimpleNamespace

    from beat_hip.engine import Engine, estimate_footprint

    V, F, T = syn.icosphere(3, 0.1)
    from beat_hip import geometry as G

    prep = G.prepare(G.BoundaryMesh(V, F, T))
    fake = SimpleNamespace(prep=prep, qunit=np.ones((len(F), 1)))
    one = estimate_footprint(prep.n_vertices, prep.n_faces, len(prep.pairs.test), 16, 146, "mixed", 1)
    avail = int(one / 0.9) + 1
    assert Engine.fit_batch(fake, "mixed", 16, 146, avail=avail) == 16
    with_ref = Engine.fit_batch(fake, "mixed", 16, 146, companions=("fp64",), avail=avail)
    assert with_ref < 16
    need = sum(estimate_footprint(prep.n_vertices, prep.n_faces, len(prep.pairs.test), with_ref, 146, p, 1)
               for p in ("mixed", "fp64"))
    assert need <= 0.9 * avail


def test_balanced_batches():
    assert D.balanced_batch(40, 32) == 20 and D.balanced_batch(40, 40) == 40 and D.balanced_batch(41, 40) == 21
    assert D.balanced_batch(7, 3) == 3 and D.balanced_batch(1, 8) == 1 and D.balanced_batch(60, 40) == 30


def test_quadrature_orders_reach_the_mesh(mesh_file):
    """quadrature_order / singular_order of the request select the regular rule (3 points for order 2, as the
    reference's triangle_rule) and the Duffy order; the diagnostics report the order used; a Duffy order outside
    1..4 and the wavelength mode are refused as the reference refuses them on a GPU backend."""
    seen = {}

    class Probe(FakeEngine):
        def __init__(self, prep, qunit, **kw):
            seen["qp"] = prep.qpoints.shape[1]
            seen["orders"] = (prep.regular_order, prep.singular_order)
            super().__init__(prep, qunit, **kw)

    events = []
    D.run_job(_request(mesh_file, quadrature_order=2, singular_order=3), Probe, events.append, batch=1)
    assert seen == {"qp": 3, "orders": (2, 3)}
    res = [e for e in events if e["type"] == "result"][0]["result"]
    assert res["diagnostics"]["regular_quadrature_order"] == 2
    for cfg, msg in (({"singular_order": 5}, "between 1 and 4"), ({"regular_quadrature_mode": "wavelength"},
                                                                  "only for the BEAT CPU backend")):
        with pytest.raises(ValueError, match=msg):
            D.run_job(_request(mesh_file, **cfg), FakeEngine, lambda e: None, batch=1)


def test_reference_interpolation_matches_the_bracket_scan():
    """The flat-target reference value (solver.jl:499-544) takes the first bracket of the periodically extended
    angles that holds the reference angle, scanning in list order; the driver bisects when the angles ascend. Both
    give the same value, bit for bit, for ascending, random-ascending and unordered angle lists."""
    import math

    def scan(values, angles_deg, ref_deg):
        ref = (float(ref_deg) + 180.0) % 360.0 - 180.0
        ang, val = np.asarray(angles_deg, np.float64), np.asarray(values, np.complex128)
        if math.isclose(ang[0], -180.0) and math.isclose(ang[-1], 180.0):
            ang, val = ang[:-1], val[:-1]
        ea, ev = np.concatenate([ang - 360.0, ang, ang + 360.0]), np.concatenate([val, val, val])
        for i in range(len(ea) - 1):
            if ea[i] <= ref <= ea[i + 1]:
                if math.isclose(ea[i], ea[i + 1]):
                    return ev[i]
                frac = (ref - ea[i]) / (ea[i + 1] - ea[i])
                return ev[i] * (1.0 - frac) + ev[i + 1] * frac
        return ev[int(np.argmin(np.abs(ea - ref)))]

    rng = np.random.default_rng(7)
    for trial in range(600):
        n = int(rng.integers(2, 80))
        ang = [np.linspace(-180, 180, n), np.sort(rng.uniform(-180, 180, n)), rng.uniform(-180, 180, n)][trial % 3]
        vals = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        for ref in (0.0, float(rng.uniform(-400, 400)), float(ang[int(rng.integers(0, n))]), 180.0, -180.0):
            assert D.interpolate_complex_reference(vals, ang, ref) == scan(vals, ang, ref)


def test_event_encoding_round_trips_float32_rows():
    """The worker writes lists of floats with nine significant digits: read back as float32 (as the consumer reads
    the result rows) every value is the one json.dumps would give (an integral value may be written without a
    decimal point, the same JSON number); non-finite values, strings, nesting and other types come out as
    json.dumps writes them."""
    import math

    from beat_hip import worker as W

    rng = np.random.default_rng(3)
    rows = (rng.standard_normal(500) * 10.0 ** rng.integers(-30, 30, 500)).astype(np.float32).tolist()
    ev = {"type": "result", "n": 3, "ok": True, "none": None, "name": 'a "q" \\ b',
          "result": {"rows": rows, "nested": {"real": [rows[:7], [-0.0, 1e-45, 3.4e38]]},
                     "bad": [1.0, float("nan"), float("-inf")], "empty": [], "ints": [1, 2, 3]}}
    text = W.encode_event(ev)
    assert "\n" not in text and len(text) < len(json.dumps(ev))
    got, want = json.loads(text), json.loads(json.dumps(ev))

    def same(a, b):
        if isinstance(a, dict):
            return a.keys() == b.keys() and all(same(a[k], b[k]) for k in a)
        if isinstance(a, list):
            return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
        if isinstance(b, float):  # a JSON number (an integral value may read back as int)
            return isinstance(a, (int, float)) and ((math.isnan(a) and math.isnan(b)) or np.float32(a) == np.float32(b))
        return a == b and type(a) is type(b)

    assert same(got, want)
    assert '"bad":[1.0,NaN,-Infinity]' in text


def test_result_rows_formatted_by_the_library_read_back_exactly():
    """Result rows (Float32Row) are written by the library's row formatter: every finite float32 value, from the
    smallest subnormal to the largest normal, at decade edges and rounding carries, reads back as the same
    float32 (parsed as float, then rounded to float32; -0 reads back as 0 as with json.dumps' -0.0 -> "-0.0"
    parsed by float); non-finite rows fall back to the Python writer."""
    from beat_hip import worker as W
    from beat_hip.wire import Float32Row, row_json

    rng = np.random.default_rng(11)
    bits = rng.integers(0, 0x7F800000, 200000, dtype=np.int64).astype(np.uint32)  # every finite magnitude
    vals = bits.view(np.float32) * np.where(rng.random(bits.size) < 0.5, -1, 1).astype(np.float32)
    edges = np.array([1e-45, 1.1754942e-38, 1.17549435e-38, 3.4028235e38, 1.0, 0.1, 1e-4, 9.99999975e-05,
                      999999999.0, 1e9, 123456789.0, 0.99999994, 9.9999999e8, 1e-5, 5e-324 * 0, 0.0], np.float32)
    vals = np.concatenate([vals, edges, -edges])
    text = W.encode_event(Float32Row(vals))
    assert text == row_json(Float32Row(vals))  # the library path was taken
    back = np.array([float(t) for t in text[1:-1].split(",")]).astype(np.float32)
    assert np.array_equal(back, vals)  # -0 == 0 numerically
    assert np.array_equal(np.array(json.loads(text), np.float64).astype(np.float32), vals)  # valid JSON
    # the same digits as the Python writer for lists of floats
    py = ("[" + ",".join(["%.9g"] * len(vals)) + "]") % tuple(vals.astype(np.float64).tolist())
    assert np.array_equal(np.array([float(t) for t in py[1:-1].split(",")]).astype(np.float32), back)
    odd = Float32Row(np.array([1.0, np.nan, -np.inf], np.float32))
    assert row_json(odd) is None and W.encode_event(odd) == "[1.0,NaN,-Infinity]"
