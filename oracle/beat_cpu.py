"""CPU restatement of the BEAT CPU engine (C/OpenMP assembly + LAPACK), used as oracle and CPU baseline.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.

Per frequency, as src/blab/solvers/julia_local/src/BeatEngineCpuSolve.jl:55-71 and solver.jl:851-920:
assemble S, D, D*, H (colour-threaded, BeatEngineCpuAssembly.jl:636-886) -> form the Burton-Miller LHS and RHS
operator -> LAPACK getrf once -> getrs per channel -> field at the observation points (threaded).
FP32 ("single") is the reference's production precision; FP64 ("double") serves as a fast oracle.
"""

from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import time
from pathlib import Path

import numpy as np

from . import bem_oracle as O

HERE = Path(__file__).resolve().parent
LIB = HERE / "libbeatcpu.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
    return _lib


def color_groups(faces, n_vertices):
    """Greedy vertex-disjoint colouring of test elements in index order (BeatEngineCpuAssembly.jl:102-130)."""
    vertex_colors = [set() for _ in range(n_vertices)]
    groups = []
    for e, f in enumerate(faces):
        used = vertex_colors[f[0]] | vertex_colors[f[1]] | vertex_colors[f[2]]
        c = 0
        while c in used:
            c += 1
        while len(groups) <= c:
            groups.append([])
        groups[c].append(e)
        for v in f:
            vertex_colors[v].add(c)
    return groups


class CpuMesh:
    """Flat arrays of one mesh for the C kernels (singular cache and colouring computed once)."""

    def __init__(self, mesh: O.Mesh, precision: str = "single"):
        self.mesh = mesh
        self.precision = precision
        self.dt = np.float32 if precision == "single" else np.float64
        self.ct = np.complex64 if precision == "single" else np.complex128
        F, V = mesh.n_faces, mesh.n_vertices
        self.F, self.V = F, V
        dt = self.dt
        self.fv = np.ascontiguousarray(mesh.face_vertices.reshape(F, 9), dtype=dt)
        self.nrm = np.ascontiguousarray(mesh.normals, dtype=dt)
        self.area = np.ascontiguousarray(mesh.areas, dtype=dt)
        self.faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
        self.curls = np.ascontiguousarray(O.surface_curls(mesh.face_vertices, mesh.normals).reshape(F, 9), dtype=dt)
        groups = color_groups(mesh.faces.tolist(), V)
        self.n_colors = len(groups)
        self.color_ptr = np.cumsum([0] + [len(g) for g in groups]).astype(np.int32)
        self.color_elem = np.asarray([e for g in groups for e in g], dtype=np.int32)
        cache = O.build_singular_cache(mesh, 4)
        order = np.argsort(cache.test, kind="stable")
        test = cache.test[order]
        self.sp_trial = np.ascontiguousarray(cache.trial[order], dtype=np.int32)
        self.sp_rule = np.ascontiguousarray(cache.rule[order], dtype=np.int32)
        self.sp_ptr = np.searchsorted(test, np.arange(F + 1)).astype(np.int32)
        offs, lens, pts = [], [], []
        off = 0
        for tp, sp, w in cache.rules:
            offs.append(off)
            lens.append(len(w))
            pts.append(np.column_stack([tp, sp, w]))
            off += len(w)
        self.rule_off = np.asarray(offs, dtype=np.int32)
        self.rule_len = np.asarray(lens, dtype=np.int32)
        self.rule_pts = np.ascontiguousarray(np.concatenate(pts), dtype=dt)
        self.M11, self.M10 = O.identity_matrices(mesh)

    def _p(self, a, ct=None):
        return a.ctypes.data_as(C.c_void_p)

    def assemble(self, k: float, regular=True, singular=True):
        suf = "_f32" if self.precision == "single" else "_f64"
        fn = getattr(lib(), "beat_cpu_assemble" + suf)
        real = C.c_float if self.precision == "single" else C.c_double
        fn.restype = C.c_int
        F, V = self.F, self.V
        S = np.zeros((F, V), self.ct)  # column-major (V x F) == row-major (F x V)
        Ds = np.zeros((F, V), self.ct)
        D = np.zeros((V, V), self.ct)
        H = np.zeros((V, V), self.ct)
        rc = fn(C.c_int(F), C.c_int(V), self._p(self.fv), self._p(self.nrm), self._p(self.area), self._p(self.faces),
                self._p(self.curls), real(k), C.c_int(self.n_colors), self._p(self.color_ptr), self._p(self.color_elem),
                self._p(self.sp_ptr), self._p(self.sp_trial), self._p(self.sp_rule), self._p(self.rule_off),
                self._p(self.rule_len), self._p(self.rule_pts), C.c_int(int(regular)), C.c_int(int(singular)),
                self._p(S), self._p(D), self._p(Ds), self._p(H))
        if rc != 0:
            raise RuntimeError("beat_cpu_assemble failed")
        # column-major buffers: transpose views give [row, col]
        return O.Operators(S.T, D.T, Ds.T, H.T)

    def field(self, k, p, q, obs):
        suf = "_f32" if self.precision == "single" else "_f64"
        fn = getattr(lib(), "beat_cpu_field" + suf)
        real = C.c_float if self.precision == "single" else C.c_double
        obs = np.ascontiguousarray(obs, dtype=self.dt)
        out = np.zeros(len(obs), self.ct)
        pp = np.ascontiguousarray(p, dtype=self.ct)
        qq = np.ascontiguousarray(q, dtype=self.ct)
        fn(C.c_int(self.F), self._p(self.fv), self._p(self.nrm), self._p(self.area), self._p(self.faces), real(k),
           self._p(pp), self._p(qq), C.c_int(len(obs)), self._p(obs), self._p(out))
        return out

    def solve_frequency(self, freq, q_unit, sound_speed, rho, obs=None):
        """One frequency of the BEAT CPU algorithm: assemble, factor once, solve each channel, field."""
        import scipy.linalg.lapack as lapack

        omega = 2.0 * math.pi * freq
        k = omega / sound_speed
        t0 = time.perf_counter()
        ops = self.assemble(k)
        t1 = time.perf_counter()
        c = 1j / k
        lhs = (0.5 * self.M11 - ops.double_layer + c * ops.hypersingular).astype(self.ct)
        rhs_op = (-ops.single_layer - c * (ops.adjoint_double_layer + 0.5 * self.M10)).astype(self.ct)
        getrf, getrs = (lapack.cgetrf, lapack.cgetrs) if self.precision == "single" else (lapack.zgetrf, lapack.zgetrs)
        lu, piv, info = getrf(lhs, overwrite_a=True)
        if info != 0:
            raise RuntimeError(f"getrf info={info}")
        q = (1j * rho * omega * np.asarray(q_unit, complex).reshape(self.F, -1)).astype(self.ct)
        p = np.zeros((self.V, q.shape[1]), self.ct)
        for ch in range(q.shape[1]):
            rhs = rhs_op @ q[:, ch]
            x, info = getrs(lu, piv, rhs)
            p[:, ch] = x
        t2 = time.perf_counter()
        field = None
        if obs is not None:
            field = np.stack([self.field(k, p[:, ch], q[:, ch], obs) for ch in range(q.shape[1])], axis=1)
        t3 = time.perf_counter()
        return {"pressure": p, "field": field, "q": q, "timings": (t1 - t0, t2 - t1, t3 - t2)}


def threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
