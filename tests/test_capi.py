"""The C-ABI library loads and exports every entry point declared in include/bemhip.h (no device calls)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "bemhip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|long long)\s+(bem_[a-z0-9_]+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from beat_hip import _lib

    if not _lib.LIB_PATH.exists():
        import importlib.util

        spec = importlib.util.spec_from_file_location("bem_build", ROOT / "boundary-lab_amd" / "build.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    return ctypes.CDLL(str(_lib.LIB_PATH))


def test_header_declares_expected_surface():
    names = declared_functions()
    for required in ("bem_ctx_create", "bem_mesh_create", "bem_assemble", "bem_lu_factor", "bem_lu_solve",
                     "bem_refine", "bem_field", "bem_last_error", "bem_graph_begin"):
        assert required in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from beat_hip import _lib

    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_error_buffer(lib):
    from beat_hip import _lib

    typed = _lib.load()
    assert typed.bem_version() >= 10000
    buf = ctypes.create_string_buffer(64)
    n = typed.bem_last_error(buf, len(buf))
    assert n >= 0 and buf.value is not None


def test_mesh_desc_layout_matches_header():
    # the ctypes struct mirrors bem_mesh_desc field for field
    from beat_hip import _lib

    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = text[text.index("typedef struct bem_mesh_desc {") : text.index("} bem_mesh_desc;")]
    names = re.findall(r"\b([a-z_0-9]+)(?:\[\d+\])?;", body)
    assert names == [f[0] for f in _lib.MeshDesc._fields_]


def test_lu_workspace_size():
    # one inverted 64 x 64 block of L and one of U per 64-row block, complex in the factor's precision
    from beat_hip import _lib

    assert _lib.lu_workspace_bytes(_lib.BEM_PREC_MIXED, 3000, 5) == 47 * 2 * 4096 * 8 * 5
    assert _lib.lu_workspace_bytes(_lib.BEM_PREC_FP64, 64, 1) == 2 * 4096 * 16
    with pytest.raises(RuntimeError):
        _lib.lu_workspace_bytes(7, 64, 1)


@pytest.fixture(scope="module")
def desc_parts():
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn

    V, F, T = syn.icosphere(2, 0.1)
    prep = G.prepare(G.BoundaryMesh(V, F, T))
    return prep


def _validate(prep, mutate=None, src=(0, 1)):
    from beat_hip import _lib
    from beat_hip.engine import mesh_desc

    import copy

    d, keep = mesh_desc(copy.deepcopy(prep), np.asarray(src))  # mutations must not reach the shared mesh
    if mutate is not None:
        mutate(d, keep)
    lib = _lib.load()
    rc = lib.bem_mesh_validate(ctypes.byref(d))
    return rc, _lib.last_error(lib)


def test_valid_mesh_description_passes(desc_parts):
    from beat_hip import _lib

    rc, msg = _validate(desc_parts)
    assert rc == _lib.BEM_OK, msg


@pytest.mark.parametrize("name,mutate,expect", [
    ("faces", lambda d, k: k["faces"].__setitem__((5, 1), 10 ** 6), "face vertex id"),
    ("faces_neg", lambda d, k: k["faces"].__setitem__((0, 0), -1), "face vertex id"),
    ("vstar_row", lambda d, k: k["vsr"].__setitem__(3, 10 ** 7), "vstar_row"),
    ("vstar_row_swap", lambda d, k: k["vsr"].__setitem__(slice(0, 2), k["vsr"][[1, 0]] + 3), "vstar_row"),
    ("vstar_ptr_neg", lambda d, k: k["vsp"].__setitem__(-1, -5), "vstar_ptr"),
    ("vstar_ptr_order", lambda d, k: k["vsp"].__setitem__(4, k["vsp"][6]), "vstar"),
    ("order_pad", lambda d, k: k["order"].__setitem__(-1, 3), "order"),
    ("sing_rule", lambda d, k: k["sr"].__setitem__(7, 99), "rule index"),
    ("sing_test", lambda d, k: k["st"].__setitem__(7, k["st"][7] + 1), "grouped by test"),
    ("sing_ptr", lambda d, k: k["sp"].__setitem__(-1, k["sp"][-1] + 1), "sing_ptr"),
    ("sing_trial", lambda d, k: k["str"].__setitem__(2, 10 ** 6), "trial key"),
    ("rule_len", lambda d, k: k["rl"].__setitem__(0, 0), "rule offset"),
    ("src", lambda d, k: k["src"].__setitem__(0, 10 ** 6), "radiating face"),
    ("vertex_last", lambda d, k: k["vlast"].__setitem__(0, 999), "vertex_last"),
    ("area", lambda d, k: k["a2"].__setitem__(3, 0.0), "area"),
])
def test_malformed_mesh_description_rejected(desc_parts, name, mutate, expect):
    # every index table the kernels dereference is checked on the host: BEM_E_ARG before any upload or launch
    from beat_hip import _lib

    rc, msg = _validate(desc_parts, mutate)
    assert rc == _lib.BEM_E_ARG
    assert expect in msg, msg


def test_library_binds_to_torchs_hip_runtime():
    # loading the library imports torch first, so that one HIP runtime serves both (a fresh process that only
    # prepares a mesh must not pull in a second runtime before torch)
    import subprocess
    import sys

    code = ("import sys; sys.path[:0] = ['{root}', '{root}/boundary-lab_amd']; "
            "from beat_hip import _lib; assert 'torch' not in sys.modules; _lib.load(); "
            "assert 'torch' in sys.modules; print('ok')").format(root=ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stderr
