"""GPU solve driver: device mesh, batched assemble -> LU -> solve/refine -> field through libbemhip.

PyTorch-ROCm provides device buffers and the stream only; all arithmetic runs in the HIP kernels behind the
C ABI. Two forms, matching the workloads:

* ``precision="fp64"`` -- FP64 assembly, FP64 LU and solve, FP64 field (the plain method).
* ``precision="mixed"`` -- FP32 regular assembly with FP64 singular pairs, FP32-MFMA LU, FP64 residual
  iterative refinement to ``tol``, FP32 field; FP64 pivot-growth check.

``Engine.solve`` processes a list of frequencies in batches (one batch = one set of launches, optionally
captured once as a HIP graph and replayed with new wavenumbers).
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .geometry import PreparedMesh

PRECISIONS = {"fp64": _lib.BEM_PREC_FP64, "mixed": _lib.BEM_PREC_MIXED}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(None)


def _np_ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def estimate_footprint(n_vertices: int, n_faces: int, n_pairs: int, batch: int, n_obs: int,
                       precision: str = "mixed", n_ch: int = 1) -> int:
    """Device bytes a job holds while resident (matrices dominate: 2 x N^2 per frequency in the batch)."""
    cb = 8 if precision == "mixed" else 16
    lda = ((n_vertices + 15) // 16) * 16
    strip_len = ((n_vertices + 63) // 64) * 64  # strip rows hold one column per vertex
    lu_work = batch * ((n_vertices + 63) // 64) * 2 * 4096 * cb  # inverted diagonal blocks
    matrices = batch * 2 * n_vertices * lda * cb + lu_work
    vectors = batch * 8 * n_ch * lda * 16
    # row + column strips; two sets of up to 4 GiB when there is more than one frequency group
    strip = min(batch, 4) * 3 * n_faces * strip_len * cb * (2 if precision == "mixed" else 1)
    strip *= 2 if batch > 4 and strip <= 4 << 30 else 1
    singular = batch * n_pairs * 12 * 16
    mesh = n_faces * (2 * 128 + 6 * 6 * 12 + 200) + n_pairs * 16
    field_ws = batch * n_ch * 6 * n_faces * 2 * cb + batch * n_obs * n_ch * 16 * 64
    return int(matrices + vectors + strip + singular + mesh + field_ws)


class Device:
    """One HIP context on one GPU with its own (default-priority) stream."""

    def __init__(self, index: int = 0):
        import torch

        if not torch.cuda.is_available():
            raise _lib.BemError("no GPU visible: the HIP solver has no CPU fallback")
        self.torch = torch
        self.lib = _lib.load()
        self.index = index
        self.device = torch.device("cuda", index)
        self.stream = torch.cuda.Stream(device=self.device)
        ctx = C.c_void_p()
        _lib.check(self.lib.bem_ctx_create(index, C.c_void_p(self.stream.cuda_stream), C.byref(ctx)), "bem_ctx_create")
        self.ctx = ctx

    def sync(self):
        _lib.check(self.lib.bem_ctx_sync(self.ctx), "bem_ctx_sync")

    def timers(self, enable: bool):
        _lib.check(self.lib.bem_ctx_timers(self.ctx, int(bool(enable))), "bem_ctx_timers")

    def timer(self, slot: int):
        """(accumulated ms, accumulated work units, timed regions) of a kernel timer slot."""
        ms, units, regions = C.c_double(), C.c_longlong(), C.c_longlong()
        _lib.check(self.lib.bem_ctx_timer_read(self.ctx, slot, C.byref(ms), C.byref(units), C.byref(regions)),
                   "bem_ctx_timer_read")
        return float(ms.value), int(units.value), int(regions.value)

    def close(self):
        if self.ctx:
            self.lib.bem_ctx_destroy(self.ctx)
            self.ctx = None


def mesh_desc(prep: PreparedMesh, src_faces) -> tuple[_lib.MeshDesc, dict]:
    """The C description of a prepared mesh (bem_mesh_desc) and the host arrays it points into, which must stay
    alive while the description is used."""
    m = prep.mesh
    F, V = m.n_faces, m.n_vertices
    keep = {}

    def arr(name, a, dtype):
        keep[name] = np.ascontiguousarray(a, dtype=dtype)
        return keep[name]

    d = _lib.MeshDesc()
    d.n_faces = F
    d.n_vertices = V
    d.faces = _np_ptr(arr("faces", m.faces, np.int32), C.c_int)
    d.face_vertices = _np_ptr(arr("fv", m.face_vertices.reshape(F, 9), np.float64), C.c_double)
    d.normals = _np_ptr(arr("n", m.normals, np.float64), C.c_double)
    d.curls = _np_ptr(arr("curl", prep.curls.reshape(F, 9), np.float64), C.c_double)
    d.area2 = _np_ptr(arr("a2", 2.0 * m.areas, np.float64), C.c_double)
    nq = prep.qpoints.shape[1]
    qp18 = np.zeros((F, 18))
    qp18[:, : 3 * nq] = prep.qpoints.reshape(F, 3 * nq)
    d.qpoints = _np_ptr(arr("qp", qp18, np.float64), C.c_double)
    d.n_qp = nq
    s = prep.slots
    d.n_chunks = len(s.slot_count)
    d.order = _np_ptr(arr("order", s.order, np.int32), C.c_int)
    d.slot_count = _np_ptr(arr("sc", s.slot_count, np.int32), C.c_int)
    d.slot_start = _np_ptr(arr("ss", s.slot_start, np.uint16), C.c_uint16)
    d.slot_inc = _np_ptr(arr("si", s.slot_inc, np.uint16), C.c_uint16)
    d.slot_vertex = _np_ptr(arr("sv", s.slot_vertex, np.int32), C.c_int)
    d.slot_pend = _np_ptr(arr("spd", s.slot_pend, np.int32), C.c_int)
    d.slot_prev = _np_ptr(arr("spv", s.slot_prev, np.int32), C.c_int)
    d.elem_pos = _np_ptr(arr("epos", s.elem_pos, np.int32), C.c_int)
    d.chunk_start = _np_ptr(arr("cst", s.chunk_start, np.int32), C.c_int)
    d.vertex_last = _np_ptr(arr("vlast", s.vertex_last, np.int32), C.c_int)
    d.vstar_ptr = _np_ptr(arr("vsp", s.vstar_ptr, np.int32), C.c_int)
    d.vstar_row = _np_ptr(arr("vsr", s.vstar_row, np.int32), C.c_int)
    p = prep.pairs
    d.n_pairs = len(p.test)
    d.sing_test = _np_ptr(arr("st", p.test, np.int32), C.c_int)
    d.sing_trial = _np_ptr(arr("str", p.trial, np.int32), C.c_int)
    d.sing_rule = _np_ptr(arr("sr", p.rule, np.int32), C.c_int)
    d.sing_ptr = _np_ptr(arr("sp", p.ptr, np.int32), C.c_int)
    d.n_rules = len(p.rule_off)
    d.rule_off = _np_ptr(arr("ro", p.rule_off, np.int32), C.c_int)
    d.rule_len = _np_ptr(arr("rl", p.rule_len, np.int32), C.c_int)
    d.rule_pts = _np_ptr(arr("rp", p.rule_pts, np.float64), C.c_double)
    for i in range(9):
        d.m11[i] = float(prep.m11.reshape(-1)[i])
    for i in range(3):
        d.m10[i] = float(prep.m10[i])
    src = arr("src", np.asarray(src_faces, dtype=np.int32).reshape(-1), np.int32)
    d.n_src = len(src)
    d.src_faces = _np_ptr(src, C.c_int) if len(src) else C.POINTER(C.c_int)()
    tf = prep.transforms if prep.transforms is not None else np.ones((1, 3), np.int32)
    d.n_transforms = len(tf)
    for i, v in enumerate(np.asarray(tf, dtype=np.int32).reshape(-1)):
        d.transform_signs[i] = int(v)
    if prep.row_weights is not None and prep.symmetry != "off":
        d.row_weights = _np_ptr(arr("rw", prep.row_weights, np.float64), C.c_double)
        d.plane_mask = _np_ptr(arr("pm", prep.plane_mask, np.int32), C.c_int)
    return d, keep


class DeviceMesh:
    """A prepared mesh resident on the device (geometry, singular pairs, Duffy rules, slot tables)."""

    def __init__(self, dev: Device, prep: PreparedMesh, src_faces: np.ndarray):
        self.dev = dev
        self.prep = prep
        self.F, self.V = prep.mesh.n_faces, prep.mesh.n_vertices
        d, keep = mesh_desc(prep, src_faces)
        h = C.c_void_p()
        _lib.check(dev.lib.bem_mesh_create(dev.ctx, C.byref(d), C.byref(h)), "bem_mesh_create")
        self.handle = h
        self.src_faces = keep["src"]

    def reserve(self, precision: str, nf: int, strip_nf: int, n_obs: int, n_ch: int):
        _lib.check(
            self.dev.lib.bem_mesh_reserve(self.handle, PRECISIONS[precision], nf, strip_nf, n_obs, n_ch),
            "bem_mesh_reserve",
        )

    def device_bytes(self) -> int:
        v = C.c_longlong()
        _lib.check(self.dev.lib.bem_mesh_device_bytes(self.handle, C.byref(v)), "bem_mesh_device_bytes")
        return int(v.value)

    def close(self):
        if self.handle:
            self.dev.lib.bem_mesh_destroy(self.handle)
            self.handle = None


def matrix_bytes(n_vertices: int, precision: str) -> int:
    """Device bytes per frequency held by one batch slot (A + A0 copy + vectors)."""
    lda = _lda(n_vertices)
    cb = 8 if precision == "mixed" else 16
    return 2 * n_vertices * lda * cb + 64 * n_vertices * 16


def _lda(n):
    return ((n + 15) // 16) * 16


@dataclass
class BatchResult:
    freqs: np.ndarray  # (nf,)
    pressure: np.ndarray  # (nf, V, C) complex128
    field: np.ndarray  # (nf, n_obs, C) complex128 or None
    residual: np.ndarray  # (nf,) relative FP64 residual of the solved system (max over channels)
    growth: np.ndarray  # (nf,) pivot growth
    iters: np.ndarray  # (nf,) refinement corrections applied (0 for fp64)
    info: np.ndarray  # (nf,) LU info
    t_assembly: float
    t_solve: float
    t_field: float
    flags: np.ndarray | None = None  # (nf,) solve-status bits (see Batch.status), None if not checked
    # (nf, F, C) complex128 Neumann data i rho omega * unit drives, or a zero-argument callable that returns it
    # (downloaded from the device on first use of .q: the result payloads do not need it)
    q_source: object = None

    @property
    def q(self) -> np.ndarray:
        if callable(self.q_source):
            self.q_source = self.q_source()
        return self.q_source


# solve-status bits (bem_solve_status)
STATUS_BREAKDOWN, STATUS_NONFINITE, STATUS_RESIDUAL, STATUS_GROWTH = 1, 2, 4, 8
# pivot growth above which a factorization is not trusted (partial pivoting bounds growth by 2^(N-1); the
# Burton-Miller matrices here stay near 1-10)
GROWTH_LIMIT = 1e8
# FP64 path: relative residual a backward-stable solve of these matrices reaches with ample margin
FP64_RESIDUAL_LIMIT = 1e-8


class Batch:
    """Device buffers and (optionally) captured HIP graphs for one (mesh, precision, nf) shape.

    Any number of channels: the system is assembled and factored once per frequency, the right-hand sides,
    solves and field run per group of at most four channels on the same factors."""

    def __init__(self, dmesh: DeviceMesh, precision: str, nf: int, qunit: np.ndarray, obs: np.ndarray | None,
                 rho: float, sound_speed: float, tol: float = 1e-10, max_iter: int = 4, strip_nf: int | None = None):
        from . import tuning as TU

        dev = dmesh.dev
        torch = dev.torch
        self.dmesh, self.dev, self.precision, self.nf = dmesh, dev, precision, nf
        self.prec = PRECISIONS[precision]
        self.rho, self.c = float(rho), float(sound_speed)
        self.tol, self.max_iter = float(tol), int(max_iter)
        N, F = dmesh.V, dmesh.F
        # launch shapes of the tuner's table for this shape (library defaults where it has none)
        tab = TU.table()
        self.lu_choice = tab.lu_choice(precision, N, nf)
        self.asm_choice = tab.assembly_choice(precision, F)
        if strip_nf is None:
            strip_nf = self.asm_choice.strip_nf
        qunit = np.asarray(qunit, dtype=np.complex128).reshape(F, -1)
        self.n_ch = qunit.shape[1]
        if self.n_ch < 1:
            raise ValueError("at least one channel is needed")
        G = _lib.BEM_MAX_CHANNELS
        self.groups = [(c0, min(G, self.n_ch - c0)) for c0 in range(0, self.n_ch, G)]
        ng = len(self.groups)
        self.N, self.F = N, F
        self.lda = _lda(N)
        self.ldb = _lda(N)
        cdt = torch.complex64 if precision == "mixed" else torch.complex128
        with torch.cuda.stream(dev.stream):
            kw = dict(device=dev.device)
            self.k = torch.zeros(nf, dtype=torch.float64, **kw)
            self.rhs_scale = torch.zeros(nf, dtype=torch.complex128, **kw)
            self.qunit = torch.from_numpy(np.ascontiguousarray(qunit.T)).to(dev.device)  # (C, F)
            self.A = torch.empty((nf, N, self.lda), dtype=cdt, **kw)
            self.A0 = torch.empty((nf, N, self.lda), dtype=cdt, **kw)
            self.b = torch.zeros((nf, self.n_ch, self.ldb), dtype=torch.complex128, **kw)
            self.x = torch.zeros((nf, self.n_ch, self.ldb), dtype=torch.complex128, **kw)
            self.r = torch.zeros((nf, self.n_ch, self.ldb), dtype=torch.complex128, **kw)
            self.y = torch.zeros((nf, self.n_ch, self.ldb), dtype=cdt, **kw)
            self.rowmax = torch.zeros((nf, N), dtype=torch.float64, **kw)
            self.umax = torch.zeros((nf, N), dtype=torch.float64, **kw)
            self.ipiv = torch.zeros((nf, N), dtype=torch.int32, **kw)
            self.perm = torch.zeros((nf, N), dtype=torch.int32, **kw)
            self.info = torch.zeros(nf, dtype=torch.int32, **kw)
            wbytes = _lib.lu_workspace_bytes(self.prec, N, nf)
            self.luw = torch.empty(wbytes // (8 if precision == "mixed" else 16), dtype=cdt, **kw)
            self.rel = torch.zeros((ng, nf), dtype=torch.float64, **kw)
            self.iters = torch.zeros((ng, nf), dtype=torch.int32, **kw)
            self.done = torch.zeros(nf, dtype=torch.int32, **kw)
            self.scratch = torch.zeros(2 * nf, dtype=torch.int32, **kw)
            self.growth = torch.zeros(nf, dtype=torch.float64, **kw)
            self.q = torch.zeros((nf, self.n_ch, F), dtype=torch.complex128, **kw)
            self.n_obs = 0 if obs is None else int(len(obs))
            if self.n_obs:
                self.obs = torch.from_numpy(np.ascontiguousarray(obs, dtype=np.float64)).to(dev.device)
                self.field = torch.zeros((nf, self.n_ch, self.n_obs), dtype=torch.complex128, **kw)
        # strip buffers (row and column strips of strip_nf frequencies) within the memory still free
        free, _ = torch.cuda.mem_get_info(dev.device)
        strip_len = ((N + 63) // 64) * 64
        per_freq = 2 * 3 * F * strip_len * (8 if precision == "mixed" else 16)
        strip_nf = max(1, min(strip_nf, nf, int(0.9 * free) // per_freq))
        self.strip_nf = strip_nf
        dmesh.reserve(precision, nf, strip_nf, self.n_obs, min(self.n_ch, G))
        self.graphs = None

    # -- launch sequence ------------------------------------------------------------------------
    def _assemble(self):
        lib, dev, m = self.dev.lib, self.dev, self.dmesh
        sB = self.n_ch * self.ldb
        # the mesh's strip grouping is whatever the last reservation set (another batch on the same mesh, e.g. a
        # one-frequency retry, may have changed it): this batch's own reservation, host-only once buffers fit
        m.reserve(self.precision, self.nf, self.strip_nf, self.n_obs, min(self.n_ch, _lib.BEM_MAX_CHANNELS))
        _lib.check(lib.bem_neumann(dev.ctx, self.nf, self.n_ch, self.F, _ptr(self.rhs_scale), _ptr(self.qunit),
                                   _ptr(self.q)), "bem_neumann")
        c0, nc = self.groups[0]
        _lib.check(
            lib.bem_assemble(dev.ctx, m.handle, self.prec, self.nf, _ptr(self.k), _ptr(self.qunit), nc,
                             _ptr(self.rhs_scale), _ptr(self.A), self.lda, self.N * self.lda, _ptr(self.A0),
                             _ptr(self.b), self.ldb, sB, _ptr(self.rowmax)),
            "bem_assemble",
        )
        for c0, nc in self.groups[1:]:
            _lib.check(lib.bem_assemble_rhs(dev.ctx, m.handle, self.prec, self.nf, _ptr(self.k),
                                            C.c_void_p(self.qunit[c0].data_ptr()), nc, _ptr(self.rhs_scale),
                                            C.c_void_p(self.b[:, c0].data_ptr()), self.ldb, sB), "bem_assemble_rhs")

    def _factor_solve(self):
        lib, dev = self.dev.lib, self.dev
        sA, sB, sP = self.N * self.lda, self.n_ch * self.ldb, self.N
        # mixed: the large LU updates as split-bf16 products; the factors only precondition the FP64 refinement
        _lib.check(lib.bem_ctx_set_lu_products(dev.ctx, _lib.BEM_LU_SPLIT_BF16 if self.precision == "mixed"
                                               else _lib.BEM_LU_FP32), "bem_ctx_set_lu_products")
        _lib.check(lib.bem_lu_factor(dev.ctx, self.prec, self.N, self.nf, _ptr(self.A), self.lda, sA,
                                     _ptr(self.ipiv), _ptr(self.perm), sP, _ptr(self.info), _ptr(self.luw)),
                   "bem_lu_factor")
        for g, (c0, nc) in enumerate(self.groups):
            bp, xp = C.c_void_p(self.b[:, c0].data_ptr()), C.c_void_p(self.x[:, c0].data_ptr())
            rp, yp = C.c_void_p(self.r[:, c0].data_ptr()), C.c_void_p(self.y[:, c0].data_ptr())
            relp = C.c_void_p(self.rel[g].data_ptr())
            if self.precision == "mixed":
                _lib.check(lib.bem_refine(dev.ctx, self.N, self.nf, _ptr(self.A), self.lda, sA, _ptr(self.perm), sP,
                                          _ptr(self.luw), _ptr(self.A0), bp, self.ldb, sB, xp, rp, yp, nc,
                                          self.max_iter, self.tol, relp, C.c_void_p(self.iters[g].data_ptr()),
                                          _ptr(self.done)), "bem_refine")
            else:
                _lib.check(lib.bem_lu_solve(dev.ctx, self.prec, self.N, self.nf, _ptr(self.A), self.lda, sA,
                                            _ptr(self.perm), sP, _ptr(self.luw), bp, self.ldb, sB, xp, self.ldb, sB,
                                            nc, None), "bem_lu_solve")
                _lib.check(lib.bem_residual(dev.ctx, self.prec, self.N, self.nf, _ptr(self.A0), self.lda, sA, bp, xp,
                                            self.ldb, sB, rp, nc, relp, _ptr(self.scratch)), "bem_residual")
        _lib.check(lib.bem_growth(dev.ctx, self.prec, self.N, self.nf, _ptr(self.A), self.lda, sA,
                                  _ptr(self.rowmax), _ptr(self.umax), _ptr(self.growth)), "bem_growth")

    def _field(self):
        if not self.n_obs:
            return
        lib, dev = self.dev.lib, self.dev
        for c0, nc in self.groups:
            _lib.check(lib.bem_field(dev.ctx, self.dmesh.handle, self.prec, self.nf, _ptr(self.k), self.n_obs,
                                     _ptr(self.obs), C.c_void_p(self.x[:, c0].data_ptr()), self.ldb,
                                     self.n_ch * self.ldb, C.c_void_p(self.q[:, c0].data_ptr()), self.n_ch * self.F,
                                     nc, C.c_void_p(self.field[:, c0].data_ptr()), self.n_obs,
                                     self.n_ch * self.n_obs), "bem_field")

    def _stages(self):
        return (self._assemble, self._factor_solve, self._field)

    def _apply_tuning(self):
        c = self.lu_choice
        t = _lib.Tuning(c.split_tile, c.trsv_wg_min_batch, c.lookahead_wgs, self.asm_choice.field_slice_budget)
        _lib.check(self.dev.lib.bem_ctx_set_tuning(self.dev.ctx, C.byref(t)), "bem_ctx_set_tuning")

    def capture(self, warm: bool = True):
        """Capture the launch sequence once, one graph per stage (assembly, factor + solve, field) so that
        replays keep per-stage timings; later batches replay them with new wavenumbers. `warm` runs the sequence
        eagerly first (kernel code objects loaded before the capture)."""
        torch = self.dev.torch
        graphs = []
        self._apply_tuning()
        try:
            with torch.cuda.stream(self.dev.stream):
                if warm:
                    self.set_frequencies(np.full(self.nf, 1000.0))
                    for st in self._stages():
                        st()
                    self.dev.sync()
                for st in self._stages():
                    _lib.check(self.dev.lib.bem_graph_begin(self.dev.ctx), "bem_graph_begin")
                    try:
                        st()
                    finally:
                        g = C.c_void_p()
                        rc = self.dev.lib.bem_graph_end(self.dev.ctx, C.byref(g))
                    _lib.check(rc, "bem_graph_end")
                    graphs.append(g)
        except Exception:
            for g in graphs:
                self.dev.lib.bem_graph_destroy(g)
            raise
        self.graphs = graphs

    @property
    def graph(self):
        return self.graphs

    def set_frequencies(self, freqs):
        torch = self.dev.torch
        f = np.asarray(freqs, dtype=np.float64)
        if len(f) != self.nf:
            f = np.concatenate([f, np.full(self.nf - len(f), f[-1])])
        omega = 2.0 * math.pi * f
        scale = (1j * self.rho * omega).astype(np.complex128)
        with torch.cuda.stream(self.dev.stream):
            self.k.copy_(torch.from_numpy(omega / self.c))
            self.rhs_scale.copy_(torch.from_numpy(scale))

    def launch(self, freqs):
        """Enqueue the whole sequence for up to nf frequencies (no wait); returns the stage events."""
        torch = self.dev.torch
        self.set_frequencies(freqs)
        if self.graphs is None:
            self._apply_tuning()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        with torch.cuda.stream(self.dev.stream):
            ev[0].record(self.dev.stream)
            for i, st in enumerate(self._stages()):
                if self.graphs is not None:
                    _lib.check(self.dev.lib.bem_graph_launch(self.dev.ctx, self.graphs[i]), "bem_graph_launch")
                else:
                    st()
                ev[i + 1].record(self.dev.stream)
        return ev

    def status(self, n: int | None = None, growth_limit: float = GROWTH_LIMIT) -> np.ndarray:
        """Solve-status bits of the first n matrices after their solve (synchronizes): STATUS_BREAKDOWN (LU zero
        pivot), STATUS_NONFINITE (non-finite solution), STATUS_RESIDUAL (relative residual above the refinement
        tolerance, or for FP64 above FP64_RESIDUAL_LIMIT), STATUS_GROWTH (pivot growth above growth_limit)."""
        n = self.nf if n is None else int(n)
        lib, dev = self.dev.lib, self.dev
        tol = self.tol if self.precision == "mixed" else FP64_RESIDUAL_LIMIT
        out = np.zeros(n, dtype=np.int32)
        for g, (c0, nc) in enumerate(self.groups):
            fl = np.zeros(n, dtype=np.int32)
            rc = lib.bem_solve_status(dev.ctx, self.N, n, _ptr(self.info), C.c_void_p(self.x[:, c0].data_ptr()),
                                      self.ldb, self.n_ch * self.ldb, nc, C.c_void_p(self.rel[g].data_ptr()), tol,
                                      _ptr(self.growth) if g == 0 else C.c_void_p(None), float(growth_limit),
                                      _np_ptr(fl, C.c_int))
            if rc not in (_lib.BEM_OK, _lib.BEM_E_SINGULAR, _lib.BEM_E_ACCURACY):
                _lib.check(rc, "bem_solve_status")
            out |= fl
        return out

    def run(self, freqs, check: bool = True) -> BatchResult:
        """Solve up to nf frequencies (padded internally); returns host arrays."""
        freqs = np.asarray(freqs, dtype=np.float64)
        n = len(freqs)
        ev = self.launch(freqs)
        self.dev.stream.synchronize()
        t_as = ev[0].elapsed_time(ev[1]) / 1e3
        t_so = ev[1].elapsed_time(ev[2]) / 1e3
        t_fi = ev[2].elapsed_time(ev[3]) / 1e3
        flags = self.status(n) if check else None
        N = self.N
        prep = self.dmesh.prep
        # back to the caller's vertex / face numbering
        p = prep.unpermute_vertices(self.x[:n, :, :N].cpu().numpy().transpose(0, 2, 1), axis=1)
        with self.dev.torch.cuda.stream(self.dev.stream):  # ordered before the batch's next run reuses its buffer
            qdev = self.q[:n].clone()

        def neumann():
            with self.dev.torch.cuda.stream(self.dev.stream):
                qi = qdev.cpu().numpy().transpose(0, 2, 1)
            q = np.empty_like(qi)
            q[:, prep.fperm, :] = qi
            return q

        field = self.field[:n].cpu().numpy().transpose(0, 2, 1) if self.n_obs else None
        return BatchResult(
            freqs=freqs, pressure=p, q_source=neumann, field=field,
            residual=self.rel[:, :n].cpu().numpy().max(axis=0), growth=self.growth[:n].cpu().numpy(),
            iters=self.iters[:, :n].cpu().numpy().max(axis=0) if self.precision == "mixed" else np.zeros(n, np.int32),
            info=self.info[:n].cpu().numpy(), t_assembly=t_as, t_solve=t_so, t_field=t_fi, flags=flags,
        )

    def close(self):
        if self.graphs is not None:
            for g in self.graphs:
                self.dev.lib.bem_graph_destroy(g)
            self.graphs = None


class Engine:
    """Convenience wrapper: mesh on device + batches per precision."""

    def __init__(self, prep: PreparedMesh, qunit: np.ndarray, dev: Device | None = None):
        self.dev = dev or Device(0)
        self.prep = prep
        # unit drives arrive in the caller's (original) face order; the device works in locality order
        self.qunit = prep.permute_faces(np.asarray(qunit, dtype=np.complex128).reshape(prep.n_faces, -1))
        src = np.nonzero(np.any(self.qunit != 0, axis=1))[0]
        self.dmesh = DeviceMesh(self.dev, prep, src)

    def batch(self, precision, nf, obs=None, rho=1.21, sound_speed=343.0, **kw) -> Batch:
        return Batch(self.dmesh, precision, nf, self.qunit, obs, rho, sound_speed, **kw)

    def available_bytes(self) -> int:
        """Device memory still available to this process: free device memory plus what torch's allocator holds
        unused."""
        torch = self.dev.torch
        free, _ = torch.cuda.mem_get_info(self.dev.device)
        return int(free + torch.cuda.memory_reserved(self.dev.device) - torch.cuda.memory_allocated(self.dev.device))

    def fit_batch(self, precision: str, nf: int, n_obs: int, share: float = 0.9, companions=(),
                  avail: int | None = None) -> int:
        """Largest frequency batch <= nf whose estimated device footprint fits in `share` of the available
        memory. `companions` lists the precisions of further batches of the same size held at the same time
        (e.g. the FP64 check batch beside a mixed one); their footprints count too."""
        avail = self.available_bytes() if avail is None else int(avail)
        p = self.prep
        n_ch = self.qunit.shape[1]
        nf = max(1, int(nf))

        def need(n):
            return sum(estimate_footprint(p.n_vertices, p.n_faces, len(p.pairs.test), n, n_obs, pr, n_ch)
                       for pr in (precision, *companions))

        while nf > 1 and need(nf) > share * avail:
            nf -= 1
        return nf

    def solve(self, freqs, precision="mixed", batch=None, obs=None, rho=1.21, sound_speed=343.0, graph=False,
              **kw):
        freqs = np.asarray(freqs, dtype=np.float64)
        nb = len(freqs) if batch is None else int(batch)
        b = self.batch(precision, nb, obs, rho, sound_speed, **kw)
        try:
            if graph:
                b.capture()
            out = []
            t0 = time.perf_counter()
            for i in range(0, len(freqs), nb):
                out.append(b.run(freqs[i : i + nb]))
            wall = time.perf_counter() - t0
        finally:
            b.close()
        return out, wall

    def close(self):
        self.dmesh.close()
