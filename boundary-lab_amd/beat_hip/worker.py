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

from .wire import Float32Row, row_json

_DEVICE = None
_DEVICE_LOCK = threading.Lock()


_FLOAT_LIST_FORMATS: dict[int, str] = {}
_KEYS: dict[str, str] = {}
_FLOAT_ONLY = {float}
_FINITE_BOUND = 1.7976931348623157e308


def encode_event(event) -> str:
    """One event as compact JSON text. Lists of floats (the Float32 result rows) are written with nine significant
    digits, which return every float32 value exactly (the reference's worker writes its Float32 arrays in their
    float32 form; the consumer reads them back as float32); everything else, and any list holding a non-finite
    value, as json.dumps writes it. Result rows (Float32Row) go through the library's row formatter (the same
    digits); dict keys are encoded once per process, scalars by their repr where that is their JSON text."""
    t = type(event)
    if t is dict:
        parts = []
        for k, v in event.items():
            kt = _KEYS.get(k) if type(k) is str else None
            if kt is None:
                if type(k) is not str:
                    return json.dumps(event, separators=(",", ":"))
                kt = _KEYS.setdefault(k, json.dumps(k))
            parts.append(kt + ":" + encode_event(v))
        return "{" + ",".join(parts) + "}"
    if t is Float32Row:
        text = row_json(event)
        if text is not None:
            return text
        t = list
    if t is list:
        if event and set(map(type, event)) == _FLOAT_ONLY:
            fmt = _FLOAT_LIST_FORMATS.get(len(event))
            if fmt is None:
                fmt = _FLOAT_LIST_FORMATS.setdefault(len(event), "[" + ",".join(["%.9g"] * len(event)) + "]")
            text = fmt % tuple(event)
            if "n" not in text:  # no nan / inf
                return text
        return "[" + ",".join([encode_event(v) for v in event]) + "]"
    if t is float and -_FINITE_BOUND <= event <= _FINITE_BOUND:
        return float.__repr__(event)
    if t is int:
        return int.__repr__(event)
    if event is None:
        return "null"
    return json.dumps(event, separators=(",", ":"))


def _emit(event: dict, stream=None):
    out = stream or sys.stdout
    out.write(encode_event(event) + "\n")
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
