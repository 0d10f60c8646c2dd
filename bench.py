"""Benchmark: freq-points/s of the mixed-precision batched solver vs the plain FP64 GPU path at 6,000 triangles.

Workload (synthetic, seeded): exponential horn (throat 25 mm, mouth 300 mm, length 200 mm) flush in a closed
0.4 x 0.6 x 0.3 m box, 6,000 triangles / 3,002 P1 unknowns, +-5 % vertex jitter (seed 1234), unit normal velocity
on the throat disk, 40 log-spaced frequencies 100 Hz - 8 kHz, polar arcs (2 x 73 points) at 2 m.

One step = one request of that job through the worker's job driver (driver.run_job, the path a `beat_rocm`
request takes): from the request dict to the last `result` event, including the Gmsh read, mesh preparation,
upload, device batches (HIP graphs for jobs of several batches), result post-processing and the JSON encoding of
every event, on a warm worker (kernels loaded). Mixed precision: FP32 assembly + FP64 singular pairs, split-bf16
MFMA LU, FP64 refinement to 1e-10. value = 40 * K / (wall time of K requests).
The plain FP64 path (FP64 assembly/LU/field, one frequency per launch) is timed the same way, end to end, and
every mixed result is compared with its FP64 counterpart (far-field dB / phase error). The steady-state rate of
the mixed batch alone (mesh resident, graph replay) is reported beside it.

Also reported: the roofline of the dominant kernel (regular-pair assembly) from HIP events on the library's
stream, the CPU baseline (C/OpenMP restatement of the BEAT CPU algorithm, FP32, LAPACK cgetrf) on a bounded
sample, and the serving loop on the seeded Poisson stream of mixed-size jobs (200 jobs at 2 jobs/s by default):
throughput and p50/p99 job latency from each job's scheduled arrival to its last result.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "boundary-lab_amd"):
    sys.path.insert(0, str(p))

METRIC = "freq-points/s @6k tris, mixed vs FP64 GPU (≤0.05 dB); p50/p99 job latency"
# Algorithmic work of regular assembly per ordered quadrature-point pair (test point, trial point) and frequency
# (fixed from the one-sided kernel's arithmetic, FMA = 2): geometry 9, rsqrt/r 2, phase 2 (+ sin, cos), Green and
# gradient 10, trial-normal projection 8, inner contraction 24, outer contraction 12 (72 flops per test point /
# 6 trial points). `achieved` counts this work over the (F^2 - adjacent) ordered element pairs.
FLOPS_PER_QPAIR = 66
# What k_regular_pair executes per unordered quadrature-point pair and frequency: both blocks of the pair from one
# kernel evaluation. Difference vector 3, |r|^2 5, r and 1/r^2 2, phase 1, Green's function 2, the double-layer
# factor 4 pi G (ik - 1/r) / r 6, its product with the trial point's n_t offset 2 (the n_s offset is constant
# over the test point's trial points and scales the contracted sum, 1), six contracted quantities: inner 19,
# outer 18.5 = 59, i.e. 59 / (2 x 66) = 0.45 of the ordered-pair work.
FLOPS_PER_UNORDERED_QPAIR = 59
# Issue slots of k_regular_pair's compiled point-pair loop: 43 VALU instructions, three of them quarter-rate
# transcendentals (v_sin, v_cos, v_rsq) that occupy four slots each. A SIMD issues one wave64 instruction per
# four clocks; MI355X has 256 CUs x 4 SIMDs at 2.4 GHz.
VALU_SLOTS_PER_QPAIR = 52
VALU_WAVE_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
PIPE_NOTE = ("bound by the FP32 vector ALUs (VALU issue, no MFMA, HBM far from its limit); 157.3 TFLOP/s is MI355X's "
             "FP32 vector peak (packed FP32); achieved counts ordered point pairs, the kernel evaluates each "
             "unordered pair once (executed_tflops); valu_issue_frac is the share of the VALU issue slots the "
             "compiled loop needs")
TRAFFIC_PROFILE = ROOT / "profiles" / "k_regular_pair_horn6k_traffic.md"


def measured_traffic():
    """HBM bytes per k_regular_pair<1> launch from the committed rocprofv3 counter summary (FETCH_SIZE +
    WRITE_SIZE per dispatch, each scaled by the factor a plain copy of known size reads), or None."""
    try:
        vals = {}
        for line in TRAFFIC_PROFILE.read_text().splitlines():
            if ":" in line and line.strip().startswith(("fetch_bytes_per_launch", "write_bytes_per_launch")):
                key, val = line.strip().split(":", 1)
                vals[key] = float(val.strip().split()[0])
        return int(vals["fetch_bytes_per_launch"] + vals["write_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def traffic_probe(workload="horn"):
    """Workload of the rocprofv3 counter passes: a 1 GiB device-to-device copy (calibration: the copy kernel
    fetches and writes exactly 1 GiB), then one eager mixed-precision job on one stream (every kernel of the path:
    assembly, split-bf16 LU, refinement, field) and, for the horn job, one plain FP64 frequency. The first
    k_regular_pair<1> launch of the horn job is the one-launch unit of the committed traffic summary."""
    import torch

    from beat_hip import _lib
    from beat_hip.engine import Engine

    prep, qunit, obs, freqs = horn_job() if workload == "horn" else scale_job()
    eng = Engine(prep, qunit)
    _lib.check(eng.dev.lib.bem_ctx_set_fanout(eng.dev.ctx, 1), "bem_ctx_set_fanout")
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    b = eng.batch("mixed", len(freqs), obs, 1.205, 343.0)
    r = b.run(freqs)
    b.close()
    if workload == "horn":
        b1 = eng.batch("fp64", 1, obs, 1.205, 343.0)
        b1.run(freqs[20:21])
        b1.close()
    eng.dev.sync()
    eng.close()
    log(f"probe: copy 1 GiB, then one eager mixed {workload} job of {len(freqs)} frequencies "
        f"(max residual {r.residual.max():.1e})" + (" and one FP64 frequency" if workload == "horn" else ""))


def field_probe():
    """Workload of the field-kernel counter passes at the field shape of the 20k-triangle loudspeaker job: its
    mesh resident, 24 wavenumbers x 3 sources x the 5,000-point Fibonacci sphere at 2 m, boundary pressure of
    random values (the field kernels' work does not depend on them), one bem_field call after the 1 GiB copy
    calibration. Each step prints before it runs."""
    import ctypes as C

    import torch

    from beat_hip import _lib
    from beat_hip.engine import PRECISIONS, DeviceMesh, Device, _lda, _ptr

    log("probe: building the 20k-triangle loudspeaker mesh")
    prep, qunit, obs, freqs = scale_job()
    dev = Device(0)
    log("probe: uploading the mesh")
    src = np.nonzero(np.any(prep.permute_faces(qunit) != 0, axis=1))[0]
    dm = DeviceMesh(dev, prep, src)
    nf, nch, F, V = len(freqs), qunit.shape[1], prep.n_faces, prep.n_vertices
    ld = _lda(V)
    rng = np.random.default_rng(3)
    with torch.cuda.stream(dev.stream):
        k = torch.from_numpy(2 * np.pi * freqs / 343.0).to(dev.device)
        p = torch.from_numpy(rng.standard_normal((nf, nch, ld)) + 1j * rng.standard_normal((nf, nch, ld))).to(dev.device)
        q = torch.from_numpy(rng.standard_normal((nf, nch, F)) + 1j * rng.standard_normal((nf, nch, F))).to(dev.device)
        ob = torch.from_numpy(np.ascontiguousarray(obs)).to(dev.device)
        out = torch.zeros((nf, nch, len(obs)), dtype=torch.complex128, device=dev.device)
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev.device)
        b = torch.empty_like(a)
        b.copy_(a)
    dev.sync()
    log(f"probe: bem_field, {nf} wavenumbers x {nch} sources x {len(obs)} points, {6 * F} sources")
    _lib.check(dev.lib.bem_field(dev.ctx, dm.handle, PRECISIONS["mixed"], nf, _ptr(k), len(obs), _ptr(ob), _ptr(p),
                                 ld, nch * ld, _ptr(q), nch * F, nch, _ptr(out), len(obs), nch * len(obs)), "bem_field")
    dev.sync()
    log("probe: field done")
    dm.close()
    dev.close()


def lu_probe(workload="horn"):
    """Workload of the LU counter passes at the batch shape of a job (horn: 40 matrices of N = 3,002; scale: 24 of
    N = 10,069), random well-conditioned complex matrices, the tuned launch shapes, one stream: a 1 GiB copy
    (calibration, the run's first copy kernel), then one split-bf16 factorization (after the copy of the matrices
    into place) with the FP64 refinement after it. Each step prints before it runs."""
    import torch

    from beat_hip import _lib
    from beat_hip import autotune as AT
    from beat_hip import tuning as TU
    from beat_hip.engine import Device

    n, batch = (3002, 40) if workload == "horn" else (10069, 24)
    dev = Device(0)
    _lib.check(dev.lib.bem_ctx_set_fanout(dev.ctx, 1), "bem_ctx_set_fanout")
    with torch.cuda.stream(dev.stream):  # the calibration copy is the run's first copy kernel
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev.device)
        b = torch.empty_like(a)
        b.copy_(a)
    dev.sync()
    del a, b
    log(f"probe: LU of {batch} random matrices of N = {n}")
    run = AT.lu_workload(dev, "mixed", n, batch)
    AT._set(dev, TU.table().lu_choice("mixed", n, batch))
    dev.sync()
    log("probe: factorization and refinement")
    run()
    dev.sync()
    info, rel = run.check()
    log(f"probe: done (info {info}, residual {rel:.1e})")
    del run
    dev.close()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def horn_job():
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from beat_hip import driver as D

    V, F, T = syn.horn_in_box()
    mesh = G.BoundaryMesh(V, F, T)
    prep = G.prepare(mesh)
    qunit = (mesh.tags == 2).astype(np.complex128)[:, None]
    angles, hor, ver, _ = D.polar_observation_points({"step_size": 5.0})
    obs = np.concatenate([hor, ver]).astype(np.float64)
    freqs = np.logspace(np.log10(100.0), np.log10(8000.0), 40)
    return prep, qunit, obs, freqs


def scale_job():
    """Two-way loudspeaker-like box (~20,000 triangles): horn throat and two piston cones as three velocity
    sources, 24 log-spaced frequencies 200 Hz-4 kHz, far field on a 5,000-point Fibonacci sphere at 2 m."""
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from beat_hip import driver as D

    V, F, T = syn.loudspeaker_box()
    mesh = G.BoundaryMesh(V, F, T)
    prep = G.prepare(mesh)
    qunit = np.stack([(mesh.tags == tag) for tag in (2, 3, 4)], axis=1).astype(np.complex128)
    obs = D.spherical_observation({"spherical_sampling_enabled": True, "spherical_sampling_points": 5000,
                                   "distance": 2.0})[0].astype(np.float64)
    freqs = np.logspace(np.log10(200.0), np.log10(4000.0), 24)
    return prep, qunit, obs, freqs


def run_scale(args):
    """Mixed batched path on the loudspeaker job (all 24 frequencies in one captured batch), checked against the
    plain FP64 path at the lowest and highest frequency. Prints one JSON line."""
    import torch

    from beat_hip.engine import Engine

    log("bench: building the 20k-triangle two-way loudspeaker job")
    prep, qunit, obs, freqs = scale_job()
    eng = Engine(prep, qunit)
    from beat_hip import _lib

    _lib.check(eng.dev.lib.bem_ctx_set_fanout(eng.dev.ctx, args.fanout), "bem_ctx_set_fanout")
    nb = len(freqs)
    b = eng.batch("mixed", nb, obs, 1.205, 343.0)
    b.capture()
    for _ in range(args.warmup):
        b.run(freqs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = b.run(freqs)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    b.close()
    value = len(freqs) * args.steps / t
    log(f"bench: scale mixed {value:.2f} freq-points/s ({1e3 * t / args.steps:.0f} ms per 24-frequency job)")
    out = {"metric": METRIC, "value": round(value, 4), "unit": "freq-points/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * t / args.steps, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic",
           "config": {"workload": "two-way loudspeaker box ~20k tris, 3 sources, 24 log freqs 200 Hz-4 kHz, mixed "
                                  "precision batched (hipGraph), 5,000-pt Fibonacci sphere at 2 m",
                      "triangles": int(prep.n_faces), "p1_dofs": int(prep.n_vertices), "frequencies": nb,
                      "batch": nb, "channels": int(qunit.shape[1]), "precision": "mixed"},
           "max_refinement_iterations": int(r.iters.max()), "max_relative_residual": float(r.residual.max())}
    be = eng.batch("mixed", nb, obs, 1.205, 343.0)
    be.run(freqs)
    eng.dev.timers(True)
    be.run(freqs)
    out["stages_ms_per_job"] = {name: round(eng.dev.timer(slot)[0], 3) for slot, name in enumerate(_lib.TIMER_NAMES)}
    eng.dev.timers(False)
    be.close()
    if not args.no_fp64:
        b1 = eng.batch("fp64", 1, obs, 1.205, 343.0)
        b1.run(freqs[:1])
        fm, fr = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j in (0, nb - 1):
            fr.append(b1.run(freqs[j : j + 1]).field[0])
            fm.append(r.field[j])
        torch.cuda.synchronize()
        out["fp64_gpu_value"] = round(2 / (time.perf_counter() - t0), 4)
        out["mixed_over_fp64"] = round(value / out["fp64_gpu_value"], 3)
        b1.close()
        out["max_field_error_db"], out["max_field_error_deg"] = field_errors(fm, fr)
        out["fp64_check_freqs_hz"] = [round(float(freqs[0]), 2), round(float(freqs[-1]), 2)]
        log(f"bench: scale field error vs FP64 {out['max_field_error_db']:.4f} dB / "
            f"{out['max_field_error_deg']:.3f} deg")
    eng.close()
    print(json.dumps(out))


def coupled_job(level, layers=3):
    """Coupled FEM-BEM workload: air between spheres of 0.10 and 0.15 m (icosphere level `level`, `layers` prism
    layers of tetrahedra) radiating through the outer sphere; the inner sphere is an electrodynamic rigid piston
    along z (6 ohm, 0.5 mH, Bl 7, 15 g, 0.5 mm/N, 1 kg/s) driven at 2.83 V, and a separate BEM sphere of 0.05 m
    (icosphere level - 2) at (0.32, 0, 0.05) is a unit-velocity source; bulk loss 0.01; order-2 regular and Duffy
    rules (the coupled production defaults); 8 log-spaced frequencies 100 Hz - 2 kHz, 500 Fibonacci points at 2 m."""
    import scipy.sparse as sp

    from beat_hip import fem as FE
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from beat_hip.coupled import CoupledModel, Excitation

    vm, (V, F, T) = syn.spherical_shell(level, 0.1, 0.15, layers)
    V2, F2, _ = syn.icosphere(max(1, level - 2), 0.05)
    V = np.concatenate([V, V2 + np.array([0.32, 0.0, 0.05])])
    F = np.concatenate([F, F2 + (len(V) - len(V2))])
    T = np.concatenate([T, np.full(len(F2), 5)])
    bm = G.BoundaryMesh(V, F, T)
    im = FE.build_conforming_interface_map(vm, bm, syn.TAG_OUTER, syn.TAG_OUTER)
    td = FE.Transducer("driver", [syn.TAG_INNER], [1.0], [], [], np.array([0.0, 0.0, 1.0]), 1.0, 1, 6.0, 5e-4, 7.0,
                       0.015, 5e-4, 1.0)
    sel = np.nonzero(T == 5)[0]
    pv = sp.csr_matrix((np.ones(len(sel)), (sel, np.zeros(len(sel), np.int64))), shape=(len(F), 1))
    model = CoupledModel(vm, bm, im, 1.21, 343.0, [td], pv, np.full(vm.n_vertices, 0.01), [])
    exc = [Excitation("voltage", (), (), 0, 1, 2.83), Excitation("normal_velocity", (), (), 1, 0, 1.0)]
    freqs = np.geomspace(100.0, 2000.0, 8)
    i = np.arange(500) + 0.5
    th, ph = np.arccos(1.0 - 2.0 * i / 500), math.pi * (1.0 + 5.0 ** 0.5) * i
    pts = 2.0 * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    return model, exc, freqs, pts


def field_errors(fields_mixed, fields_ref):
    """Max far-field |dB| and phase error of mixed results against the FP64 path, over all frequencies."""
    db, ph = [], []
    for a, b in zip(fields_mixed, fields_ref):
        with np.errstate(divide="ignore", invalid="ignore"):
            db.append(np.abs(20 * np.log10(np.abs(a) / np.abs(b))).max())
            ph.append(np.degrees(np.abs(np.angle(a / b))).max())
    return float(max(db)), float(max(ph))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prep, qunit, obs, freqs, budget_s=20.0):
    """BEAT CPU algorithm restated (FP32, C/OpenMP + LAPACK) on a bounded sample of the same job."""
    from oracle import beat_cpu as B
    from oracle import bem_oracle as O

    m = prep.original  # qunit is in the original face order
    om = O.Mesh(m.vertices, m.faces, m.tags, m.mesh_ids)
    cm = B.CpuMesh(om, "single")
    done, t0 = 0, time.perf_counter()
    sample = [freqs[len(freqs) // 2], freqs[-1], freqs[0]]
    for f in sample:
        cm.solve_frequency(float(f), qunit, 343.0, 1.205, obs)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "freq-points/s", "cores": B.threads(), "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{done} of the 40 frequencies ({', '.join(f'{f:.0f} Hz' for f in sample[:done])}) of the "
                      f"6k-triangle horn job, BEAT CPU algorithm restated in C/OpenMP FP32 + LAPACK cgetrf, "
                      f"{B.threads()} threads on {cpu_model()}"}


# ---------------------------------------------------------------------------------------------------------------
# end-to-end jobs through the worker's driver
# ---------------------------------------------------------------------------------------------------------------


def horn_request(path, precision, freqs):
    cfg = {"mesh_file": str(path), "scale_factor": 1.0, "rho": 1.205, "sound_speed": 343.0, "step_size": 5.0,
           "precision": precision, "radiators": [{"name": "throat", "tag": 2, "channel": "main"}]}
    if precision == "fp64":
        cfg["frequency_batch"] = 1  # the plain method: one frequency per launch
    return {"schema_version": 2, "config": cfg, "frequencies_hz": [float(f) for f in freqs],
            "beat_engine_backend": "rocm"}


def run_request(req):
    """One request through driver.run_job on the worker's shared device; returns (wall s, result events). Every
    event is JSON-encoded as the worker writes it (worker.encode_event)."""
    from beat_hip import driver as D
    from beat_hip import worker as W

    lines, results = [], []

    def emit(ev):
        lines.append(W.encode_event(ev))
        if ev["type"] == "result":
            results.append(ev["result"])
        elif ev["type"] == "failed":
            raise RuntimeError(ev["error"])

    t0 = time.perf_counter()
    D.run_job(req, W.engine_factory, emit)
    return time.perf_counter() - t0, results


def wire_field(result):
    """Complex far field of a result event (horizontal then vertical arcs, channel 0) from its Float32 rows."""
    rows = []
    for key in ("horizontal_pressure", "vertical_pressure"):
        w = result[key]
        rows.append(np.asarray(w["real"][0], np.float64) + 1j * np.asarray(w["imag"][0], np.float64))
    return np.concatenate(rows)


# ---------------------------------------------------------------------------------------------------------------
# serving stream
# ---------------------------------------------------------------------------------------------------------------


def _serving_mesh(args):
    j, kind, n_tri, seed, path = args
    from beat_hip import synthetic as syn

    V, F, T = syn.mesh_for_triangle_count(kind, n_tri, seed=seed)
    syn.write_gmsh22(path, V, F, T, scale=1.0)
    return j, len(F), sorted(set(T.tolist()))


def serving_stream(n_jobs, rate, seed, tmp):
    """The seeded mixed-size stream: Poisson arrivals at `rate` jobs/s, mesh sizes log-uniform in 1,000-15,000
    triangles over the sphere / horn-in-box / box-with-pistons generators, 8-60 frequencies log-uniform in
    100 Hz - 10 kHz, 1-3 velocity sources. Meshes are generated up front in a process pool (host only)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor

    rng = np.random.default_rng(seed)
    specs, jobs = [], []
    for j in range(n_jobs):
        n_tri = int(np.exp(rng.uniform(np.log(1000), np.log(15000))))
        kind = ["sphere", "horn", "pistons"][j % 3]
        specs.append((j, kind, n_tri, seed + j, str(Path(tmp) / f"job{j}.msh")))
        jobs.append({"nsrc": int(rng.integers(1, 4)), "nf": int(rng.integers(8, 61))})
    for jb in jobs:
        jb["freqs"] = np.sort(np.exp(rng.uniform(np.log(100), np.log(10000), jb["nf"]))).tolist()
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    workers = min(16, os.cpu_count() or 1)
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as pool:
        made = {j: (nf, tags) for j, nf, tags in pool.map(_serving_mesh, specs)}
    stream = []
    for (j, kind, n_tri, _, path), jb, ta in zip(specs, jobs, arrivals):
        n_faces, tags = made[j]
        rtags = [t for t in tags if t != 1] or [2]
        nsrc = min(len(rtags), jb["nsrc"])
        rads = [{"name": f"r{i}", "tag": int(rtags[i]), "channel": f"c{i}"} for i in range(nsrc)]
        req = {"schema_version": 2, "frequencies_hz": jb["freqs"],
               "config": {"mesh_file": path, "scale_factor": 1.0, "radiators": rads, "step_size": 5.0}}
        stream.append({"arrival": float(ta), "request": req, "triangles": n_faces, "kind": kind,
                       "frequencies": jb["nf"], "sources": nsrc})
    return stream


def serving_sample(stream, rate, seed, batch, slots):
    """Run the stream through the serving scheduler: each request is submitted at its scheduled arrival and
    stamped with it, so host preparation and queueing count in its latency."""
    from beat_hip import serving as S

    sched = S.Scheduler(slots=slots, batch=batch)
    handles = []
    t0 = time.perf_counter()
    for job in stream:
        now = time.perf_counter() - t0
        if job["arrival"] > now:
            time.sleep(job["arrival"] - now)
        h, _ = sched.submit(job["request"], arrival=t0 + job["arrival"])
        handles.append(h)
        if len(handles) % 50 == 0:
            log(f"bench: serving, {len(handles)} jobs submitted")
    total_f, failed = 0, []
    for h in handles:
        while True:
            ev = h.events.get()
            if ev["type"] == "result":
                total_f += 1
            if ev["type"] in ("completed", "failed", "cancelled"):
                if ev["type"] == "failed":
                    failed.append(ev["error"])
                break
    wall = time.perf_counter() - t0
    sched.shutdown()
    if failed:
        raise RuntimeError(f"{len(failed)} serving jobs failed, first: {failed[0]}")
    lat = [h.latency for h in handles]
    queue = [h.t_admit - h.arrival for h in handles]
    return {"jobs": len(stream), "rate_jobs_per_s": rate, "seed": seed, "freq_points": total_f, "wall_s": wall,
            "throughput_freq_points_per_s": total_f / wall, "p50_latency_s": S.percentile(lat, 50),
            "p99_latency_s": S.percentile(lat, 99), "mean_latency_s": float(np.mean(lat)),
            "p50_wait_to_admission_s": S.percentile(queue, 50), "p99_wait_to_admission_s": S.percentile(queue, 99),
            "peak_running": sched.peak_running, "merged_jobs": sched.merged_jobs, "slots": slots,
            "frequency_batch": batch, "tri_min": min(j["triangles"] for j in stream),
            "tri_max": max(j["triangles"] for j in stream)}


def run_coupled(args):
    """Coupled FEM-BEM solve per frequency on the device (beat_hip.coupled), float32 production precision (FP32 LU
    + FP64 refinement) with the requested static condensation left to the solver's cost choice, on the default
    workload (icosphere level 3: 2,568 FEM / 684 BEM vertices) and a large one (level 4: 10,248 / 2,724);
    float64 and forced host condensation timed beside; the oracle (FP64, monolithic, dense LAPACK) timed on one
    frequency of the default workload as the CPU baseline and used as the parity check. Prints one JSON line."""
    import torch

    from beat_hip.coupled import CoupledSolver
    from beat_hip.engine import Device

    dev = Device(0)
    out = {"metric": "coupled FEM-BEM freq-points/s (float32 production path vs FP64 oracle)", "n_gpus": 1,
           "unit": "freq-points/s", "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32+fp64", "data": "synthetic", "steps": args.steps, "warmup": args.warmup}

    def timed(model, exc, freqs, pts, **opts):
        s = CoupledSolver(model, dev=dev, **opts)
        try:
            for _ in range(args.warmup):
                s.solve_batch(freqs[: s.batch_capacity()], exc, pts)
            torch.cuda.synchronize()
            stages = {}
            iters = 0
            t0 = time.perf_counter()
            res = None
            nb = s.batch_capacity()
            for _ in range(args.steps):
                for i in range(0, len(freqs), nb):
                    for j, (sols, field, _, tim) in enumerate(s.solve_batch(freqs[i : i + nb], exc, pts)):
                        res = (freqs[i + j], sols, field) if res is None else res
                        iters = max(iters, int(tim.pop("refinement_iterations", 0)))
                        for k, v in tim.items():
                            stages[k] = stages.get(k, 0.0) + v
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            n = args.steps * len(freqs)
            return dict(value=n / el, ms_per_freq=1e3 * el / n, n=s.n, formulation="condensed" if s.condensed
                        else "monolithic", batch=nb, max_refinement_iterations=iters,
                        stages_ms={k: round(1e3 * v / n, 3) for k, v in stages.items()}), res
        finally:
            s.close()

    model, exc, freqs, pts = coupled_job(3)
    log(f"bench: coupled default workload ({model.fem_mesh.n_vertices} FEM / {model.bem_mesh.n_vertices} BEM vertices)")
    r32, (f0, sols, field) = timed(model, exc, freqs, pts, precision="float32", static_condensation=True)
    r64, _ = timed(model, exc, freqs, pts, precision="float64", static_condensation=True)
    rcond, _ = timed(model, exc, freqs[:2], pts, precision="float32", static_condensation=True, formulation="condensed")
    out.update(value=round(r32["value"], 4), ms_per_step=round(r32["ms_per_freq"] * len(freqs), 3))
    out["config"] = {"workload": "coupled FEM-BEM shell (level-3 icosphere, 3 layers), electrodynamic piston + BEM "
                                 "source, 8 log freqs 100 Hz-2 kHz, order-2 rules, 500 field points",
                     "fem_vertices": int(model.fem_mesh.n_vertices), "tetrahedra": int(len(model.fem_mesh.tetrahedra)),
                     "bem_vertices": int(model.bem_mesh.n_vertices), "interface_dofs": int(model.interface_map.count),
                     "frequencies": len(freqs), "excitations": len(exc)}
    out["float32"], out["float64"], out["float32_host_condensed"] = r32, r64, rcond
    # oracle: CPU baseline on one frequency and parity of that frequency
    from oracle import bem_oracle as O
    from oracle import coupled_oracle as CO

    if not args.no_cpu:
        fm, bm, im = model.fem_mesh, model.bem_mesh, model.interface_map
        om = O.Mesh(bm.vertices, bm.faces, bm.tags, np.zeros(bm.n_faces, np.int64))
        imo = dict(fem_vertices=im.fem_vertex_indices, bem_vertices=im.fem_to_bem_vertex_indices,
                   fem_faces=im.fem_face_indices, bem_faces=im.bem_face_indices, sign=im.normal_sign)
        td = model.transducers[0]
        pist = [dict(fem_tags=td.fem_tags, fem_signs=td.fem_signs, bem_tags=[], bem_signs=[], axis=td.motion_axis,
                     completion=1.0, re=td.re_ohm, le=td.le_h, bl=td.bl_n_per_a, mmd=td.mmd_kg, cms=td.cms_m_per_n,
                     rms=td.rms_n_s_per_m)]
        t0 = time.perf_counter()
        ref = CO.solve_coupled(fm.vertices, fm.tetrahedra, fm.boundary_faces, fm.boundary_tags, om, imo, float(f0),
                               343.0, 1.21, [{"kind": "voltage", "transducer": 1, "amplitude": 2.83},
                                             {"kind": "normal_velocity", "bem_source": 1}], pistons=pist,
                               prescribed=model.prescribed_velocity.toarray(), loss_by_vertex=model.bulk_loss_by_vertex,
                               regular_order=2, singular_order=2, points=pts)
        el = time.perf_counter() - t0
        cores = os.cpu_count()
        out["cpu_baseline"] = {"value": round(1.0 / el, 5), "unit": "freq-points/s", "cores": cores, "kind": "port",
                               "cpu_model": cpu_model(),
                               "sample": f"1 of the 8 frequencies ({f0:.0f} Hz), FP64 oracle: per-element FEM loops, "
                                         "dense BEM operators (NumPy), monolithic dense LAPACK solve, field"}
        with np.errstate(divide="ignore", invalid="ignore"):
            db = np.abs(20 * np.log10(np.abs(field) / np.abs(ref["field"])))
            ph = np.degrees(np.abs(np.angle(field / ref["field"])))
        out["parity_vs_oracle"] = {"freq_hz": round(float(f0), 2), "max_field_error_db": float(np.nanmax(db)),
                                   "max_field_error_deg": float(np.nanmax(ph)),
                                   "diaphragm_velocity_rel": float(abs(sols[0].diaphragm_velocity[0]
                                                                       - ref["diaphragm_velocity"][0, 0])
                                                                   / abs(ref["diaphragm_velocity"][0, 0]))}
    if not args.no_fp64:  # the large workload (GPU only)
        model, exc, freqs, pts = coupled_job(4)
        log(f"bench: coupled large workload ({model.fem_mesh.n_vertices} FEM / {model.bem_mesh.n_vertices} BEM)")
        big, _ = timed(model, exc, freqs[::2], pts, precision="float32", static_condensation=True)
        bigc, _ = timed(model, exc, freqs[:1], pts, precision="float32", static_condensation=True,
                        formulation="condensed")
        out["large"] = {"fem_vertices": int(model.fem_mesh.n_vertices), "bem_vertices": int(model.bem_mesh.n_vertices),
                        "interface_dofs": int(model.interface_map.count), "auto": big, "host_condensed": bigc}
    dev.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=40)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=200)
    ap.add_argument("--serving-rate", type=float, default=2.0)
    ap.add_argument("--serving-batch", type=int, default=40, help="frequencies per device batch in the serving loop")
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--fanout", type=int, default=4,
                    help="streams used by the batched LU and refinement (1 = single stream, e.g. for tracing)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run the workload of the counter passes (copy calibration + one eager job of the "
                         "chosen --workload)")
    ap.add_argument("--probe", choices=["job", "field", "lu"], default="job",
                    help="--traffic-probe workload: one eager job of --workload, the field evaluation of the "
                         "20k-triangle job alone, or the LU of --workload's batch shape on random matrices")
    ap.add_argument("--workload", choices=["horn", "scale", "coupled"], default="horn",
                    help="horn: the 6k-triangle horn-in-box job (headline); scale: the 20k-triangle two-way "
                         "loudspeaker job; coupled: the FEM-BEM shell with a piston")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build runs on a single MI355X); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    if args.traffic_probe:
        if args.probe == "field":
            field_probe()
        elif args.probe == "lu":
            lu_probe(args.workload)
        else:
            traffic_probe(args.workload)
        return
    if args.workload == "scale":
        run_scale(args)
        return
    if args.workload == "coupled":
        run_coupled(args)
        return

    import tempfile

    tmp = Path(tempfile.mkdtemp(prefix="beat-bench-"))
    stream = None
    if args.serving_jobs > 0:  # host-only mesh generation, before anything touches the GPU
        log(f"bench: generating the serving stream ({args.serving_jobs} meshes)")
        stream = serving_stream(args.serving_jobs, args.serving_rate, 2024, tmp)

    import torch

    from beat_hip import _lib
    from beat_hip import synthetic as syn
    from beat_hip import worker as W
    from beat_hip.engine import Engine

    log("bench: building 6k-triangle horn-in-box job")
    prep, qunit, obs, freqs = horn_job()
    V, F, T = syn.horn_in_box()
    mesh_path = tmp / "horn6k.msh"
    syn.write_gmsh22(mesh_path, V, F, T, scale=1.0)
    dev = W.shared_device()
    _lib.check(dev.lib.bem_ctx_set_fanout(dev.ctx, args.fanout), "bem_ctx_set_fanout")
    F = prep.n_faces

    # ---- warm worker: one small job of each precision loads the kernels --------------------------------------
    Vs, Fs, Ts = syn.icosphere(2, 0.1)
    warm_path = tmp / "warm.msh"
    syn.write_gmsh22(warm_path, Vs, Fs, Ts, scale=1.0)
    for precision in ("mixed", "fp64"):
        warm = horn_request(warm_path, precision, [500.0, 1500.0, 2500.0])
        warm["config"]["radiators"] = [{"name": "s", "tag": int(Ts[0]), "channel": "main"}]
        run_request(warm)

    # ---- headline: mixed precision, end to end through driver.run_job ------------------------------------------
    req = horn_request(mesh_path, "mixed", freqs)
    log(f"bench: mixed end to end, {args.warmup} warmup + {args.steps} timed requests")
    for _ in range(args.warmup):
        run_request(req)
    torch.cuda.synchronize()
    walls, mixed_results = [], None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w, mixed_results = run_request(req)
        walls.append(w)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    ms_per_step = 1e3 * t_mixed / args.steps
    value = len(freqs) * args.steps / t_mixed
    log(f"bench: mixed end to end {value:.2f} freq-points/s ({ms_per_step:.1f} ms per 40-frequency request)")

    # ---- plain FP64 path, end to end, one frequency per launch ------------------------------------------------
    fp64_value, db_err, ph_err, fp64_walls = None, None, None, []
    if not args.no_fp64:
        fsteps = max(1, min(args.steps, 2))
        log(f"bench: plain FP64 end to end, {fsteps} requests")
        req64 = horn_request(mesh_path, "fp64", freqs)
        for _ in range(fsteps):
            w, fp64_results = run_request(req64)
            fp64_walls.append(w)
        fp64_value = len(freqs) * fsteps / sum(fp64_walls)
        db_err, ph_err = field_errors([wire_field(r) for r in mixed_results], [wire_field(r) for r in fp64_results])
        log(f"bench: fp64 end to end {fp64_value:.2f} freq-points/s; mixed/fp64 = {value / fp64_value:.2f}x; "
            f"max field error {db_err:.4f} dB / {ph_err:.3f} deg")
    max_iters = max(int(r["diagnostics"]["refinement_iterations"]) for r in mixed_results)
    max_res = max(float(r["diagnostics"]["relative_residual"]) for r in mixed_results)

    # ---- steady state: the mixed batch alone, mesh resident, graph replay ---------------------------------------
    eng = Engine(prep, qunit, dev=dev)
    nb = max(1, min(args.batch, len(freqs)))
    b = eng.batch("mixed", nb, obs, 1.205, 343.0)
    b.capture()
    for _ in range(max(1, args.warmup)):
        [b.run(freqs[i : i + nb]) for i in range(0, len(freqs), nb)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        [b.run(freqs[i : i + nb]) for i in range(0, len(freqs), nb)]
    torch.cuda.synchronize()
    t_ss = time.perf_counter() - t0
    steady_value = len(freqs) * args.steps / t_ss
    b.close()
    fp64_steady = None
    if not args.no_fp64:
        b1 = eng.batch("fp64", 1, obs, 1.205, 343.0)
        b1.capture()
        b1.run(freqs[:1])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(len(freqs)):
            b1.run(freqs[i : i + 1])
        torch.cuda.synchronize()
        fp64_steady = len(freqs) / (time.perf_counter() - t0)
        b1.close()
    log(f"bench: steady state (graph replay, mesh resident): mixed {steady_value:.2f}"
        + ("" if fp64_steady is None else f", fp64 {fp64_steady:.2f}") + " freq-points/s")

    # ---- per-stage times of one eager job and the roofline of the regular assembly kernel -----------------------
    be = eng.batch("mixed", nb, obs, 1.205, 343.0)
    be.run(freqs[:nb])
    dev.timers(True)
    for i in range(0, len(freqs), nb):
        be.run(freqs[i : i + nb])
    stages = {name: dev.timer(slot) for slot, name in enumerate(_lib.TIMER_NAMES)}
    dev.timers(False)
    be.close()
    t_ms, units, regions = stages["regular"]
    avg_ms = t_ms / max(1, regions)
    freqs_per_launch = units / max(1, regions)
    qpairs = (F * F - len(prep.pairs.test)) * 36 * freqs_per_launch
    achieved = qpairs * FLOPS_PER_QPAIR / (avg_ms * 1e-3) / 1e12
    executed = 0.5 * qpairs * FLOPS_PER_UNORDERED_QPAIR / (avg_ms * 1e-3) / 1e12
    traffic = measured_traffic()
    log(f"bench: k_regular_pair<1> avg {avg_ms:.3f} ms/launch ({freqs_per_launch:g} freqs) -> "
        f"{achieved:.1f} TFLOP/s of ordered-pair work ({executed:.1f} executed); stages (ms/job): "
        + ", ".join(f"{k} {v[0]:.1f}" for k, v in stages.items()))

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline sample")
        cpu = cpu_baseline(prep, qunit, obs, freqs)
        log(f"bench: cpu {cpu['value']:.3f} freq-points/s on {cpu['cores']} threads")
    eng.close()

    serving = None
    if stream is not None:
        log(f"bench: serving stream, {len(stream)} jobs at {args.serving_rate}/s")
        serving = serving_sample(stream, args.serving_rate, 2024, batch=args.serving_batch, slots=args.serving_slots)
        log(f"bench: serving p50 {serving['p50_latency_s']:.2f} s p99 {serving['p99_latency_s']:.2f} s, "
            f"{serving['throughput_freq_points_per_s']:.1f} freq-points/s")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "freq-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": "horn-in-box 6000 tris / 3002 P1 dofs, 40 log freqs 100 Hz-8 kHz, mixed precision, "
                               "end to end per request through the worker's job driver (Gmsh read, preparation, "
                               "upload, device batches, result events), polar 2x73 pts at 2 m",
                   "triangles": int(F), "p1_dofs": int(prep.n_vertices), "frequencies": int(len(freqs)),
                   "precision": "mixed"},
        "request_latency_s": {"mean": round(float(np.mean(walls)), 4), "max": round(float(np.max(walls)), 4)},
        "fp64_gpu_value": None if fp64_value is None else round(fp64_value, 4),
        "mixed_over_fp64": None if fp64_value is None else round(value / fp64_value, 3),
        "steady_state_value": round(steady_value, 4),
        "steady_state_fp64_value": None if fp64_steady is None else round(fp64_steady, 4),
        "steady_state_mixed_over_fp64": None if fp64_steady is None else round(steady_value / fp64_steady, 3),
        "max_field_error_db": db_err,
        "max_field_error_deg": ph_err,
        "max_refinement_iterations": max_iters,
        "max_relative_residual": max_res,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "kernel": "k_regular_pair<1>", "pipe": PIPE_NOTE, "avg_launch_ms": round(avg_ms, 4),
                     "freqs_per_launch": freqs_per_launch, "flops_per_qpair": FLOPS_PER_QPAIR,
                     "qpairs_per_launch": int(qpairs),
                     "executed_tflops": round(executed, 3),
                     "executed_flops_per_unordered_qpair": FLOPS_PER_UNORDERED_QPAIR,
                     "valu_issue_frac": round(0.5 * qpairs * VALU_SLOTS_PER_QPAIR / 64 / VALU_WAVE_ISSUE_PER_S
                                              / (avg_ms * 1e-3), 4)},
        "stages_ms_per_job": {k: round(v[0], 3) for k, v in stages.items()},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
