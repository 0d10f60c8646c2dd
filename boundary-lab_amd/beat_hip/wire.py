"""Result rows as they go on the wire.

The reference's worker writes its result rows as Float32 arrays (julia_local/solver.jl:921-996, written by
emit_event, solver.jl:12-17). Here a row is a ``Float32Row``: a list of those float32 values, which is what every
consumer of the events sees, holding its float32 array as well, from which the worker's JSON writer formats the
row in one library call (``bem_host_format_f32_json``)."""

from __future__ import annotations

import ctypes

import numpy as np


class Float32Row(list):
    """A list of the values of a float32 array, keeping the array (``.array``) for the JSON writer."""

    __slots__ = ("array",)

    def __init__(self, values):
        a = np.ascontiguousarray(values, dtype=np.float32)
        super().__init__(a.tolist())
        self.array = a

    def __reduce__(self):  # copies and pickles are plain lists of the same values
        return (list, (list(self),))


_FORMAT = None
_BUF = [ctypes.create_string_buffer(0)]


def row_json(row: Float32Row) -> str | None:
    """JSON text of the row's float32 values, nine significant digits each (every value reads back as the same
    float32); None if a value is not finite or the library is not built (the caller formats the list itself)."""
    global _FORMAT
    if _FORMAT is None:
        try:
            from . import _lib

            _FORMAT = _lib.load().bem_host_format_f32_json
        except Exception:
            _FORMAT = False
    if not _FORMAT:
        return None
    a = row.array
    n = a.shape[0]
    if len(_BUF[0]) < 16 * n + 2:
        _BUF[0] = ctypes.create_string_buffer(32 * n + 64)
    m = _FORMAT(a.ctypes.data, n, _BUF[0], len(_BUF[0]))
    return _BUF[0].raw[:m].decode("ascii") if m > 0 else None
