"""Launch-shape table written by the project's tuner (beat_hip/autotune.py) and read by the engine.

The table (``tuning.json`` next to this file) holds two lists of entries, each measured on the device:

* ``lu``: per (precision, N band, batch band) the bem_tuning fields of the factorization and solves -- split-bf16
  update tile, one-workgroup-per-matrix solve threshold, look-ahead update workgroups.
* ``assembly``: per (precision, face-count band) the frequencies per strip buffer (``strip_nf``) and the field's
  slice budget.

A lookup takes the first entry whose bands contain the shape (bands are upper bounds, entries sorted ascending);
shapes beyond every band, or a missing table, get the library defaults.
"""

from __future__ import annotations

import json
from dataclasses import asdict, dataclass, field
from pathlib import Path

TABLE_PATH = Path(__file__).with_name("tuning.json")

# band upper bounds (inclusive)
N_BANDS = (1024, 2048, 4096, 8192, 24576)
BATCH_BANDS = (1, 7, 23, 64)
FACE_BANDS = (2500, 9000, 50000)


@dataclass(frozen=True)
class LuChoice:
    split_tile: int = 0
    trsv_wg_min_batch: int = 8
    lookahead_wgs: int = 0


@dataclass(frozen=True)
class AssemblyChoice:
    strip_nf: int = 4
    field_slice_budget: int = 2048


@dataclass
class Table:
    lu: list = field(default_factory=list)  # dicts: precision, n_max, batch_max, choice, ms, default_ms
    assembly: list = field(default_factory=list)  # dicts: precision, faces_max, choice, ms, default_ms
    device: str = ""

    def lu_choice(self, precision: str, n: int, batch: int) -> LuChoice:
        nb, bb = band(n, N_BANDS), band(batch, BATCH_BANDS)
        for e in self.lu:
            if e["precision"] == precision and e["n_max"] == nb and e["batch_max"] == bb:
                return LuChoice(**e["choice"])
        return LuChoice()

    def assembly_choice(self, precision: str, faces: int) -> AssemblyChoice:
        fb = band(faces, FACE_BANDS)
        for e in self.assembly:
            if e["precision"] == precision and e["faces_max"] == fb:
                return AssemblyChoice(**e["choice"])
        return AssemblyChoice()


def band(value: int, bounds) -> int | None:
    """The smallest bound >= value, None above every bound."""
    for b in bounds:
        if value <= b:
            return b
    return None


def validate(raw: dict) -> Table:
    """Table from its JSON form; raises ValueError on entries the library would reject."""
    t = Table(lu=list(raw.get("lu", [])), assembly=list(raw.get("assembly", [])), device=str(raw.get("device", "")))
    for e in t.lu:
        c = LuChoice(**e["choice"])
        if e["precision"] not in ("mixed", "fp64") or e["n_max"] not in N_BANDS or e["batch_max"] not in BATCH_BANDS:
            raise ValueError(f"bad LU entry {e}")
        if c.split_tile not in (0, 1) or c.trsv_wg_min_batch < 0 or c.lookahead_wgs < -1:
            raise ValueError(f"bad LU choice {e}")
    for e in t.assembly:
        c = AssemblyChoice(**e["choice"])
        if e["precision"] not in ("mixed", "fp64") or e["faces_max"] not in FACE_BANDS:
            raise ValueError(f"bad assembly entry {e}")
        if c.strip_nf < 1 or c.field_slice_budget < 0:
            raise ValueError(f"bad assembly choice {e}")
    return t


_TABLE = None


def table(path: Path | None = None) -> Table:
    global _TABLE
    if path is not None:
        return validate(json.loads(Path(path).read_text()))
    if _TABLE is None:
        _TABLE = validate(json.loads(TABLE_PATH.read_text())) if TABLE_PATH.exists() else Table()
    return _TABLE


def as_dict(choice) -> dict:
    return asdict(choice)
