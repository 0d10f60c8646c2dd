"""Serving loop on the device: jobs of uneven size run concurrently on separate slots (and identical jobs merged
into shared frequency batches) give the same results as each job run alone.

Tolerance: the per-frequency solves do not depend on which other matrices share a launch (every matrix sees the
same operations in the same order), so the Float32 wire values are expected to agree exactly; the test allows
1e-6 relative on the wire pressures and 1e-4 dB on the SPL rows.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(tmp_path):
    from beat_hip import synthetic as syn

    # the icosphere's faces all carry the radiator tag 2; the horn's throat disk is tag 2
    specs = [("sphere", syn.icosphere(3, 0.1), [2, 2], [300.0, 900.0, 1700.0, 2600.0, 4100.0, 6000.0]),
             ("horn", syn.horn_in_box(h=0.05), [2], [150.0, 800.0, 2000.0, 5000.0, 7000.0]),
             ("small", syn.icosphere(2, 0.1), [2], [500.0, 2500.0, 3300.0])]
    reqs = []
    for name, (V, F, T), tags, freqs in specs:
        path = tmp_path / f"{name}.msh"
        syn.write_gmsh22(path, V, F, T, scale=1.0)
        rads = [{"name": f"{name}{i}", "tag": int(t), "channel": f"c{i}", "level_db": -3.0 * i}
                for i, t in enumerate(tags)]
        reqs.append({"schema_version": 2, "frequencies_hz": freqs,
                     "config": {"mesh_file": str(path), "scale_factor": 1.0, "step_size": 10.0, "rho": 1.205,
                                "radiators": rads, "frequency_batch": 4}})
    return reqs


def _drain(h):
    out = []
    while True:
        ev = h.events.get(timeout=120)
        out.append(ev)
        if ev["type"] in ("completed", "failed", "cancelled"):
            return out


def _results(events):
    assert events[-1]["type"] == "completed", events[-1]
    return [e["result"] for e in events if e["type"] == "result"]


def _same(a, b):
    assert [x["freq_hz"] for x in a] == [y["freq_hz"] for y in b]
    for x, y in zip(a, b):
        for key in ("horizontal_pressure", "vertical_pressure"):
            for part in ("real", "imag"):
                u, v = np.asarray(x[key][part]), np.asarray(y[key][part])
                assert np.abs(u - v).max() <= 1e-6 * np.abs(v).max(), key
        for key in ("horizontal_spl_db", "vertical_spl_db"):
            assert np.abs(np.asarray(x[key]) - np.asarray(y[key])).max() <= 1e-4


def test_jobs_alone_and_in_the_mix(tmp_path):
    from beat_hip import serving as S

    reqs = _requests(tmp_path)
    sched = S.Scheduler(slots=3, batch=4)
    try:
        alone = []
        for r in reqs:
            print("alone:", r["config"]["mesh_file"], flush=True)
            alone.append(_results(_drain(sched.submit(r)[0])))
        print("mixed: all three at once", flush=True)
        hs = [sched.submit(r)[0] for r in reqs]
        mixed = [_results(_drain(h)) for h in hs]
        assert max(h.group_size for h in hs) == 1
        print("merged: the horn job twice", flush=True)
        twins = [sched.submit(reqs[1])[0], sched.submit(reqs[1])[0]]
        merged = [_results(_drain(h)) for h in twins]
    finally:
        sched.shutdown()
    for a, m in zip(alone, mixed):
        _same(m, a)
    for m in merged:
        _same(m, alone[1])
    assert sched.peak_running >= 2
