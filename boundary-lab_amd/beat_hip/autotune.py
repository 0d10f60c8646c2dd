"""The project's tuner: times candidate launch shapes on the device and writes the table the engine applies
(beat_hip/tuning.json; see beat_hip/tuning.py).

    python -m beat_hip.autotune [--out PATH] [--quick]

LU shapes (per precision, N band and batch band): the split-bf16 update tile, the one-workgroup-per-matrix solve
threshold and the look-ahead update's workgroups, timed as one factorization plus the solve path the engine runs
after it (FP64 refinement for mixed batches, triangular solves and the FP64 residual for FP64 ones) on a batch of
random, well-conditioned complex matrices of the band's representative size. Assembly shapes (per precision and
face-count band): frequencies per strip buffer and the field's slice budget, timed as bem_assemble and bem_field
on a synthetic mesh of the band. Each knob is varied in turn from the library defaults (coordinate descent), each
candidate timed as the median of three runs with HIP events on the library's stream after one untimed run; a
candidate replaces the current choice only when it is at least 2 % faster.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import time
from dataclasses import replace
from pathlib import Path

import numpy as np

from . import _lib
from . import tuning as TU

# representative shapes: (precision, N, batch) and (precision, mesh kind, triangles, frequencies, observation points)
LU_SHAPES = [("mixed", 642, 8), ("mixed", 642, 40), ("mixed", 1500, 8), ("mixed", 1500, 40), ("mixed", 3002, 1),
             ("mixed", 3002, 8), ("mixed", 3002, 40), ("mixed", 6000, 8), ("mixed", 6000, 24),
             ("mixed", 10069, 8), ("mixed", 10069, 24), ("fp64", 642, 1), ("fp64", 3002, 1), ("fp64", 6000, 1)]
ASM_SHAPES = [("mixed", "sphere", 1280, 40, 1146), ("mixed", "horn", 6000, 40, 146), ("mixed", "pistons", 15000, 24, 1146),
              ("fp64", "sphere", 1280, 1, 1146), ("fp64", "horn", 6000, 1, 146)]
LU_CANDIDATES = {"split_tile": (0, 1), "trsv_wg_min_batch": (4, 8, 16, 0), "lookahead_wgs": (0, -1, 128)}
ASM_CANDIDATES = {"strip_nf": (2, 4, 8), "field_slice_budget": (1024, 2048, 4096)}
QUICK_LU = [("mixed", 642, 8), ("mixed", 3002, 40), ("fp64", 642, 1)]
QUICK_ASM = [("mixed", "sphere", 1280, 8, 146)]


def log(msg):
    print(msg, flush=True)


def _time(dev, fn, reps=3):
    torch = dev.torch
    with torch.cuda.stream(dev.stream):
        fn()
    dev.sync()
    out = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(dev.stream):
            e0.record(dev.stream)
            fn()
            e1.record(dev.stream)
        dev.sync()
        out.append(e0.elapsed_time(e1))
    return float(np.median(out))


def _set(dev, lu: TU.LuChoice, budget: int = 0):
    t = _lib.Tuning(lu.split_tile, lu.trsv_wg_min_batch, lu.lookahead_wgs, budget)
    _lib.check(dev.lib.bem_ctx_set_tuning(dev.ctx, C.byref(t)), "bem_ctx_set_tuning")


def lu_workload(dev, precision, n, batch, seed=0):
    """Closure that factors a fresh copy of a random batch and runs the engine's solve path on it."""
    from .engine import PRECISIONS, _lda, _ptr

    torch = dev.torch
    prec = PRECISIONS[precision]
    lda = _lda(n)
    cdt = torch.complex64 if precision == "mixed" else torch.complex128
    g = torch.Generator(device=dev.device)
    g.manual_seed(seed)
    with torch.cuda.stream(dev.stream):
        kw = dict(device=dev.device)
        A0 = torch.zeros((batch, n, lda), dtype=cdt, **kw)
        re = torch.randn((batch, n, n), generator=g, dtype=torch.float32, **kw)
        im = torch.randn((batch, n, n), generator=g, dtype=torch.float32, **kw)
        A0[:, :, :n] = (torch.complex(re, im) / np.sqrt(n)).to(cdt)
        A0[:, torch.arange(n), torch.arange(n)] += 3.0
        del re, im
        A = torch.empty_like(A0)
        b = torch.zeros((batch, 1, lda), dtype=torch.complex128, **kw)
        b[:, 0, :n] = 1.0
        x = torch.zeros_like(b)
        r = torch.zeros_like(b)
        y = torch.zeros((batch, 1, lda), dtype=cdt, **kw)
        ipiv = torch.zeros((batch, n), dtype=torch.int32, **kw)
        perm = torch.zeros_like(ipiv)
        info = torch.zeros(batch, dtype=torch.int32, **kw)
        w = torch.empty(_lib.lu_workspace_bytes(prec, n, batch), dtype=torch.uint8, **kw)
        rel = torch.zeros(batch, dtype=torch.float64, **kw)
        it = torch.zeros(batch, dtype=torch.int32, **kw)
        done = torch.zeros(batch, dtype=torch.int32, **kw)
        scratch = torch.zeros(2 * batch, dtype=torch.int32, **kw)
    dev.sync()
    lib = dev.lib
    _lib.check(lib.bem_ctx_set_lu_products(dev.ctx, _lib.BEM_LU_SPLIT_BF16 if precision == "mixed"
                                           else _lib.BEM_LU_FP32), "products")
    sA = n * lda
    keep = (A0, A, b, x, r, y, ipiv, perm, info, w, rel, it, done, scratch)

    def run():
        with torch.cuda.stream(dev.stream):  # the copy precedes the factorization on the library's stream
            A.copy_(A0)
        _lib.check(lib.bem_lu_factor(dev.ctx, prec, n, batch, _ptr(A), lda, sA, _ptr(ipiv), _ptr(perm), n,
                                     _ptr(info), _ptr(w)), "bem_lu_factor")
        if precision == "mixed":
            _lib.check(lib.bem_refine(dev.ctx, n, batch, _ptr(A), lda, sA, _ptr(perm), n, _ptr(w), _ptr(A0), _ptr(b),
                                      lda, lda, _ptr(x), _ptr(r), _ptr(y), 1, 4, 1e-10, _ptr(rel), _ptr(it),
                                      _ptr(done)), "bem_refine")
        else:
            _lib.check(lib.bem_lu_solve(dev.ctx, prec, n, batch, _ptr(A), lda, sA, _ptr(perm), n, _ptr(w), _ptr(b), lda,
                                        lda, _ptr(x), lda, lda, 1, None), "bem_lu_solve")
            _lib.check(lib.bem_residual(dev.ctx, prec, n, batch, _ptr(A0), lda, sA, _ptr(b), _ptr(x), lda, lda,
                                        _ptr(r), 1, _ptr(rel), _ptr(scratch)), "bem_residual")

    run.keep = keep
    run.check = lambda: (int(info.max().item()), float(rel.max().item()))
    return run


def descend(evaluate, start, candidates, label):
    """Coordinate descent over the fields of a frozen dataclass choice; returns (best, best_ms, start_ms)."""
    best, best_ms = start, evaluate(start)
    start_ms = best_ms
    for name, values in candidates.items():
        for v in values:
            if getattr(best, name) == v:
                continue
            cand = replace(best, **{name: v})
            ms = evaluate(cand)
            log(f"  {label} {name}={v}: {ms:.3f} ms (best {best_ms:.3f})")
            if ms < 0.98 * best_ms:
                best, best_ms = cand, ms
    return best, best_ms, start_ms


def tune_lu(dev, shapes):
    out = []
    for precision, n, batch in shapes:
        log(f"LU {precision} N={n} batch={batch}")
        run = lu_workload(dev, precision, n, batch)

        def evaluate(choice):
            _set(dev, choice)
            return _time(dev, run)

        best, ms, ms0 = descend(evaluate, TU.LuChoice(), LU_CANDIDATES, f"N={n} B={batch}")
        _set(dev, best)
        run()
        dev.sync()
        info, rel = run.check()
        if info != 0 or (precision == "mixed" and rel > 1e-10):
            raise RuntimeError(f"tuned LU shape failed its check: info {info}, residual {rel:.2e}")
        out.append({"precision": precision, "n_max": TU.band(n, TU.N_BANDS), "batch_max": TU.band(batch, TU.BATCH_BANDS),
                    "n": n, "batch": batch, "choice": TU.as_dict(best), "ms": round(ms, 4), "default_ms": round(ms0, 4)})
        log(f"  -> {TU.as_dict(best)}: {ms:.3f} ms (defaults {ms0:.3f} ms)")
        del run
        dev.torch.cuda.empty_cache()
    _set(dev, TU.LuChoice())
    return out


def tune_assembly(dev, shapes):
    from . import geometry as G
    from . import synthetic as syn
    from .engine import Engine

    out = []
    for precision, kind, tris, nf, n_obs in shapes:
        V, F, T = syn.mesh_for_triangle_count(kind, tris, seed=1)
        prep = G.prepare(G.BoundaryMesh(V, F, T))
        log(f"assembly {precision} {kind} {prep.n_faces} triangles, {nf} frequencies, {n_obs} points")
        q = np.ones((prep.n_faces, 1))
        obs = np.stack([np.linspace(-1, 1, n_obs), np.full(n_obs, 0.3), np.full(n_obs, 2.0)], axis=1)
        eng = Engine(prep, q, dev=dev)
        freqs = np.logspace(2, np.log10(8000.0), nf)
        batches = {}

        def evaluate(choice):
            b = batches.get(choice.strip_nf)
            if b is None:
                b = batches[choice.strip_nf] = eng.batch(precision, nf, obs, 1.205, 343.0, strip_nf=choice.strip_nf)
                b.set_frequencies(freqs)
            _set(dev, TU.LuChoice(), choice.field_slice_budget)  # b._assemble re-asserts its strip grouping

            def run():
                b._assemble()
                b._field()

            return _time(dev, run)

        try:
            best, ms, ms0 = descend(evaluate, TU.AssemblyChoice(), ASM_CANDIDATES, f"F={prep.n_faces}")
        finally:
            for b in batches.values():
                b.close()
            eng.close()
            _set(dev, TU.LuChoice())
        out.append({"precision": precision, "faces_max": TU.band(prep.n_faces, TU.FACE_BANDS), "faces": prep.n_faces,
                    "frequencies": nf, "points": n_obs, "choice": TU.as_dict(best), "ms": round(ms, 4),
                    "default_ms": round(ms0, 4)})
        log(f"  -> {TU.as_dict(best)}: {ms:.3f} ms (defaults {ms0:.3f} ms)")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(TU.TABLE_PATH))
    ap.add_argument("--quick", action="store_true", help="a few small shapes (a check of the tuner itself)")
    args = ap.parse_args(argv)
    from .engine import Device

    dev = Device(0)
    t0 = time.perf_counter()
    lu = tune_lu(dev, QUICK_LU if args.quick else LU_SHAPES)
    asm = tune_assembly(dev, QUICK_ASM if args.quick else ASM_SHAPES)
    props = dev.torch.cuda.get_device_properties(dev.device)
    table = {"device": f"{props.name} ({getattr(props, 'gcnArchName', '')})", "seconds": round(time.perf_counter() - t0, 1),
             "method": "coordinate descent from the library defaults, median of 3 timed runs, 2 % threshold",
             "lu": lu, "assembly": asm}
    TU.validate(table)
    Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
    log(f"wrote {args.out} ({len(lu)} LU and {len(asm)} assembly entries, {table['seconds']} s)")


if __name__ == "__main__":
    main()
