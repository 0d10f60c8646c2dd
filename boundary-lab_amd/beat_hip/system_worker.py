"""NDJSON worker for compiled-system solves: the stdin/stdout protocol of the reference's coupled worker
(src/blab/solvers/julia_local/coupled_solver.jl:2515-2575, consumed by src/blab/solvers/coupled_backend.py:158-200),
served by the HIP engine: exterior systems (``beat_hip.system``) and coupled FEM-BEM systems
(``beat_hip.coupled_system``).

* ``--worker``: print ``{"type":"ready"}``, then for every stdin line ``{"request": "<path>", "operation": "solve"}``
  print ``{"type":"result","result":{...}}`` per frequency and one terminal ``completed`` / ``cancelled``
  (``solved_count``) or ``failed`` (``error``). The ``bem_field`` operation is answered with ``failed``.
* no flag: read one request object from stdin and print one bare result object per frequency; exit 1 on error.
"""

from __future__ import annotations

import json
import os
import sys

from .worker import _emit, engine_factory


def coupled_factory(model, **options):
    from .coupled import CoupledSolver
    from .worker import shared_device

    return CoupledSolver(model, dev=shared_device(), **options)


def solve_path(path: str, emit=_emit):
    from .system import solve_system_request

    with open(path, "r", encoding="utf-8") as handle:
        request = json.load(handle)
    cancelled, solved = solve_system_request(request, engine_factory, emit, event_mode=True,
                                             coupled_factory=coupled_factory)
    emit({"type": "cancelled" if cancelled else "completed", "solved_count": solved})


def worker_loop(stdin=None, emit=_emit):
    emit({"type": "ready", "pid": os.getpid()})
    for line in stdin or sys.stdin:
        text = line.strip()
        if not text:
            continue
        try:
            submission = json.loads(text)
            operation = str(submission.get("operation", "solve"))
            if operation == "bem_field":
                raise ValueError("The bem_field operation (retained BEM field caches) is not served by the HIP "
                                 "exterior backend.")
            if operation != "solve":
                raise ValueError(f"Unsupported coupled worker operation: {operation}")
            solve_path(str(submission["request"]), emit)
        except Exception as exc:  # reported as a failed job; the worker keeps serving
            emit({"type": "failed", "error": f"{type(exc).__name__}: {exc}"})


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if "--worker" in argv:
        worker_loop()
        return 0
    try:
        from .system import solve_system_request

        request = json.loads(sys.stdin.read())
        solve_system_request(request, engine_factory, _emit, event_mode=False, coupled_factory=coupled_factory)
        return 0
    except Exception as exc:
        sys.stderr.write(f"{type(exc).__name__}: {exc}\n")
        return 1


if __name__ == "__main__":
    sys.exit(main())
