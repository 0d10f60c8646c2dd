"""NDJSON worker: the stdin/stdout protocol of the reference's Julia worker, served by the HIP engine.

Protocol (src/blab/solvers/julia_local/solver.jl:12-42,1022-1051; consumed by
src/blab/solvers/beat_engine_backend.py:286-462):

* ``--worker``: print ``{"type":"ready","protocol":"boundary_lab_julia_worker","pid":...}``, then for every
  stdin line ``{"request": "<path>", "operation": "solve"}`` run the job and print its events, one JSON object
  per line: ``initialized``, ``status``*, ``result``*, then exactly one of ``completed`` / ``cancelled`` /
  ``failed``. A failed job does not end the worker (the Python adapter retires it).
* ``--request <path>``: one-shot; on failure print ``failed`` and exit 1.
"""

from __future__ import annotations

import json
import os
import sys
import threading

_DEVICE = None
_DEVICE_LOCK = threading.Lock()


def _emit(event: dict, stream=None):
    out = stream or sys.stdout
    out.write(json.dumps(event, separators=(",", ":")) + "\n")
    out.flush()


def shared_device():
    global _DEVICE
    with _DEVICE_LOCK:
        if _DEVICE is None:
            from .engine import Device

            _DEVICE = Device(int(os.environ.get("BEAT_HIP_DEVICE", "0")))
        return _DEVICE


def engine_factory(prep, qunit):
    from .engine import Engine

    return Engine(prep, qunit, dev=shared_device())


def solve_request(request: dict, emit=_emit):
    from .driver import run_job

    run_job(request, engine_factory, emit)


def solve_request_path(path: str, emit=_emit):
    with open(path, "r", encoding="utf-8") as handle:
        request = json.load(handle)
    solve_request(request, emit)


def worker_loop(stdin=None, emit=_emit):
    emit({"type": "ready", "protocol": "boundary_lab_julia_worker", "pid": os.getpid()})
    for line in stdin or sys.stdin:
        text = line.strip()
        if not text:
            continue
        try:
            message = json.loads(text)
            solve_request_path(str(message["request"]), emit)
        except Exception as exc:  # reported as a failed job; the worker keeps serving
            emit({"type": "failed", "error": f"{type(exc).__name__}: {exc}"})


def parse_args(argv):
    request, worker = None, False
    i = 0
    while i < len(argv):
        if argv[i] == "--request" and i + 1 < len(argv):
            request = argv[i + 1]
            i += 2
        elif argv[i] == "--worker":
            worker = True
            i += 1
        else:
            raise ValueError(f"Unknown argument: {argv[i]}")
    return request, worker


def main(argv=None) -> int:
    try:
        request, worker = parse_args(sys.argv[1:] if argv is None else argv)
        if worker:
            worker_loop()
            return 0
        if request is None:
            raise ValueError("Missing --request path.")
        solve_request_path(request)
        return 0
    except Exception as exc:
        _emit({"type": "failed", "error": f"{type(exc).__name__}: {exc}"})
        return 1


if __name__ == "__main__":
    sys.exit(main())
