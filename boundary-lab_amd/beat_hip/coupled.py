"""Coupled FEM-BEM solve: FEM interior condensed on the host, the dense coupled system assembled, factored and
solved on the GPU.

Replaces the reference's ROCm coupled path (src/blab/solvers/julia_local/src/BeatEngineCoupled.jl):
``prepare_coupled_cache`` :916-1159, ``build_coupled_system`` :1670-2163 with its ROCm hybrid condensation
``_build_rocm_hybrid_fem_condensation`` :1585-1661 (UMFPACK Schur complement on the CPU, dense LU with rocSOLVER)
and ``solve_coupled_excitations`` :2360-2554 with ``_solve_rocm_hybrid_condensed_excitations`` :2285-2358.

Unknowns, in the reference's block order (:1857-1880):

    [ FEM (all vertices; the retained set Gamma when condensed) | BEM P1 | interface flux | piston velocity |
      voice-coil current ]

    FEM rows    : (K - k^2 M - i k^2 M_loss - i rho w sum Y_w M_w) p_F - fem_load q_I - i rho w fem_surface u = f_F
                  (condensed: S = A_GG - A_GI A_II^-1 A_IG on Gamma, right-hand side f_G - A_GI A_II^-1 f_I)
    BEM rows    : (1/2 M11 - D + (i/k) H) p_B + [-(B bem_flux)] q_I + [-(B i rho w bem_nv)] u = B g_prescribed
                  with B = -S - (i/k)(D* + 1/2 M10) the Burton-Miller RHS operator
    flux rows   : fem_trace p_F - bem_trace p_B = 0
    piston rows : -fem_force^T p_F + bem_force^T p_B + Z_m u - Bl i = 0,  Z_m = R_ms + i (1/(w C_ms) - w M_md)
    coil rows   : Bl u + (R_e - i w L_e) i = e

What runs where: FEM matrices, the interface and piston operators, the Schur complement (SciPy's SuperLU on the
FEM interior, in complex128 whatever the request's precision, as the reference factors A_II in ComplexF64) and
the reduction / recovery of right-hand sides stay on the host; the BEM block (bem_assemble), the BEM interface
and piston blocks (bem_rhs_columns), the placement of every sparse and dense block (bem_sparse_scatter,
bem_dense_block), an exact power-of-two row equilibration (bem_row_equilibrate), the LU (bem_lu_factor) and the
solve -- FP32 LU with FP64 iterative refinement against the assembled matrix for float32 requests, FP64 LU and
solve for float64 -- and the far field run on the device.
The device matrix is row-major with BEM rows and columns in the mesh's locality order; solutions are returned
in the caller's numbering.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import _lib
from . import fem as FE
from . import geometry as G
from . import tuning as TU
from .engine import PRECISIONS, Device, DeviceMesh, _lda, _np_ptr, _ptr

REFINE_TOL = 1e-10
# bem_refine launches max_iter correction rounds whatever the convergence (its kernels exit on the device once a
# matrix has converged): row-equilibrated coupled systems converge in 2-3, so a small cap keeps the idle rounds'
# launch overhead low; a system that does not converge within it is reported by the solve status
REFINE_MAX_ITER = 10
FP64_RESIDUAL_LIMIT = 1e-8


@dataclass
class CoupledModel:
    """Frequency-invariant inputs of a coupled solve (the arguments of prepare_coupled_cache and
    build_coupled_system that do not change with frequency)."""

    fem_mesh: FE.VolumeMesh
    bem_mesh: G.BoundaryMesh
    interface_map: FE.InterfaceMap
    density: float
    sound_speed: float
    transducers: list = field(default_factory=list)
    prescribed_velocity: sp.csr_matrix | None = None  # (F, n_prescribed) BEM normal velocity columns
    bulk_loss_by_vertex: np.ndarray | None = None
    wall_impedances: list = field(default_factory=list)  # dicts: tag, thickness_m, flow_resistivity_pa_s_per_m2


@dataclass
class Excitation:
    """One right-hand side (solve_coupled_excitations, BeatEngineCoupled.jl:2360-2420)."""

    kind: str  # "normal_velocity" or "voltage"
    fem_tags: tuple = ()
    fem_weights: tuple = ()
    bem_source_index: int = 0  # 1-based column of the prescribed BEM velocity, 0 for none
    transducer_index: int = 0  # 1-based, voltage excitations
    amplitude: complex = 1.0


@dataclass
class CoupledSolution:
    fem_pressure: np.ndarray
    bem_pressure: np.ndarray
    interface_flux: np.ndarray
    bem_neumann: np.ndarray
    diaphragm_velocity: np.ndarray
    voice_coil_current: np.ndarray
    pressure_continuity_error: float
    flux_conservation_error: float
    fem_interior_residual: float | None = None
    relative_residual: float | None = None
    all_bem_replay_error: float | None = None
    fem_rhs_condensation_s: float | None = None
    fem_reconstruction_s: float | None = None


def schur_complement(A: sp.spmatrix, retained: np.ndarray, block: int = 256):
    """S = A_GG - A_GI A_II^-1 A_IG (dense complex128) with A_II factored by SuperLU; returns (S, parts) where
    parts holds the factorization and blocks for right-hand-side reduction and interior recovery
    (_build_rocm_hybrid_fem_condensation / _blocked_umfpack_schur_complement, BeatEngineCoupled.jl:1530-1661)."""
    A = sp.csr_matrix(A, dtype=np.complex128)
    n = A.shape[0]
    keep = np.zeros(n, bool)
    keep[retained] = True
    interior = np.nonzero(~keep)[0]
    t0 = time.perf_counter()
    A_r = A[retained]
    A_i = A[interior]
    S = A_r[:, retained].toarray()
    A_GI = A_r[:, interior].tocsr()
    A_IG = A_i[:, retained].tocsc()
    A_II = A_i[:, interior].tocsc()
    t_part = time.perf_counter() - t0
    lu = None
    t0 = time.perf_counter()
    if len(interior):
        lu = spla.splu(A_II)
    t_fact = time.perf_counter() - t0
    t0 = time.perf_counter()
    if lu is not None:
        for c0 in range(0, len(retained), block):
            c1 = min(len(retained), c0 + block)
            X = lu.solve(A_IG[:, c0:c1].toarray())
            S[:, c0:c1] -= A_GI @ X
    t_schur = time.perf_counter() - t0
    parts = dict(lu=lu, interior=interior, retained=np.asarray(retained), A_GI=A_GI, A_IG=A_IG.tocsr(),
                 A_II=A_II.tocsr(), timings=dict(partition_s=t_part, factorization_s=t_fact,
                                                  schur_extraction_s=t_schur))
    return S, parts


def monolithic_is_cheaper(n_fem: int, n_retained: int, n_mono: int) -> bool:
    """Choose the dense monolithic system on the device over host condensation when it fits the dense LU and its
    estimated time is lower. Host condensation is one SuperLU factorization of the interior plus one sparse
    triangular solve pair per retained vertex: measured about 2.6e6 nonzeros in L + U for a 7,700-vertex
    tetrahedral interior (fill growing like n^(4/3)) at 6 GFLOP/s of complex solves. The device LU of the
    monolithic system (one matrix, panel-chain bound) is estimated at 0.35 s x (n / 15,500)^3, and is never
    below 10 ms."""
    if n_mono > _lib.BEM_LU_MAX_N:
        return False
    n_int = max(0, n_fem - n_retained)
    nnz_lu = 2.6e6 * (n_int / 7700.0) ** (4.0 / 3.0)
    host = 8.0 * nnz_lu * n_retained / 6e9
    device = max(0.01, 0.35 * (n_mono / 15500.0) ** 3)
    return device < host


def _colset_arrays(mat: sp.spmatrix, finv: np.ndarray):
    """Column-set tables (internal face numbering) of an (F x n_cols) sparse matrix of Neumann columns."""
    m = sp.csc_matrix(mat, dtype=np.complex128)
    m.sum_duplicates()
    m.sort_indices()
    faces = finv[m.indices].astype(np.int64)
    src = np.unique(faces)
    idx = np.searchsorted(src, faces)
    return (np.ascontiguousarray(src, np.int32), np.ascontiguousarray(m.indptr, np.int32),
            np.ascontiguousarray(idx, np.int32), np.ascontiguousarray(m.data, np.complex128))


class CoupledSolver:
    """The coupled system of one model on one device, rebuilt and solved per frequency."""

    def __init__(self, model: CoupledModel, precision: str = "float32", static_condensation: bool = True,
                 validation_diagnostics: bool = False, regular_order: int = 2, singular_order: int = 2,
                 symmetry: str = "off", dev: Device | None = None, formulation: str = "auto"):
        """formulation: "auto" follows static_condensation unless condensation was requested and the
        monolithic system fits the dense LU and is estimated cheaper (monolithic_is_cheaper); "condensed" /
        "monolithic" force one."""
        if precision not in ("float32", "float64"):
            raise ValueError("Unsupported coupled precision. Expected float32 or float64.")
        if static_condensation and validation_diagnostics:
            raise ValueError("FEM static condensation cannot be combined with full-matrix validation diagnostics.")
        if formulation not in ("auto", "condensed", "monolithic"):
            raise ValueError("formulation must be auto, condensed or monolithic")
        self.model = m = model
        self.precision = precision
        self.engine_precision = "mixed" if precision == "float32" else "fp64"
        self.prec = PRECISIONS[self.engine_precision]
        self.condensation_requested = bool(static_condensation)
        self.condensed = bool(static_condensation) if formulation == "auto" else formulation == "condensed"
        self.validation = bool(validation_diagnostics)
        self.symmetry = G.normalized_symmetry_mode(symmetry)
        self.regular_order, self.singular_order = int(regular_order), int(singular_order)
        self.dev = dev if dev is not None else Device(0)
        self._own_dev = dev is None
        self.timings = {}
        fm, bm = m.fem_mesh, m.bem_mesh
        if fm.is_quadratic:
            raise ValueError("Quadratic tetrahedra are currently supported only by the pure interior FEM solve.")
        eps = float(np.finfo(np.float32 if precision == "float32" else np.float64).eps)
        t0 = time.perf_counter()
        self.K, self.M = FE.p1_fem_matrices(fm, eps)
        loss = np.zeros(fm.n_vertices) if m.bulk_loss_by_vertex is None else np.asarray(m.bulk_loss_by_vertex, float)
        if len(loss) != fm.n_vertices:
            raise ValueError("Per-region FEM bulk-loss vector must contain one value per FEM vertex.")
        if not np.all(np.isfinite(loss) & (loss >= 0.0) & (loss <= 1.0)):
            raise ValueError("Per-region FEM bulk-loss factors must be finite and between 0 and 1.")
        self.Mloss = (sp.diags(loss) @ self.M).tocsr()
        self.walls = []
        for w in m.wall_impedances:
            faces = np.nonzero(fm.boundary_tags == int(w["tag"]))[0]
            if not len(faces):
                raise ValueError(f"Wall-impedance boundary tag {w['tag']} has no FEM faces.")
            self.walls.append((FE.boundary_mass_matrix(fm, faces, np.arange(fm.n_vertices)), float(w["thickness_m"]),
                               float(w["flow_resistivity_pa_s_per_m2"])))
        self.timings["fem_matrix_cache_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.iops = FE.interface_operators(fm, bm, m.interface_map)
        self.timings["interface_operator_cache_s"] = time.perf_counter() - t0
        self.tops = FE.transducer_operators(fm, bm, m.transducers)
        self.nt = len(m.transducers)
        tfem = np.unique(self.tops["fem_surface"].nonzero()[0])
        self.retained = np.unique(np.concatenate([m.interface_map.fem_vertex_indices, tfem])).astype(np.int64)
        n_mono = fm.n_vertices + bm.n_vertices + m.interface_map.count + 2 * self.nt
        if formulation == "auto" and self.condensed and monolithic_is_cheaper(fm.n_vertices, len(self.retained),
                                                                            n_mono):
            self.condensed = False
        if self.condensed:
            outside = np.ones(fm.n_vertices, bool)
            outside[self.retained] = False
            if self.iops.fem_load[np.nonzero(outside)[0]].nnz:
                raise ValueError("FEM static condensation requires interface loads to have support only on "
                                 "retained nodes.")
        pv = m.prescribed_velocity
        self.pvel = sp.csr_matrix((bm.n_faces, 0)) if pv is None else sp.csr_matrix(pv, dtype=np.float64)
        if self.pvel.shape[0] != bm.n_faces:
            raise ValueError("Prescribed BEM normal velocity must contain one row per BEM face.")
        # BEM mesh on the device (locality order inside)
        t0 = time.perf_counter()
        self.prep = G.prepare(bm, self.singular_order, symmetry=self.symmetry, regular_order=self.regular_order)
        self.dmesh = DeviceMesh(self.dev, self.prep, np.zeros(0, np.int32))
        self.timings["bem_device_cache_s"] = time.perf_counter() - t0
        V, F = bm.n_vertices, bm.n_faces
        self.vinv = np.empty(V, np.int64)
        self.vinv[self.prep.vperm] = np.arange(V)
        self.finv = np.empty(F, np.int64)
        self.finv[self.prep.fperm] = np.arange(F)
        # block layout
        self.n_fem = fm.n_vertices
        self.nF = len(self.retained) if self.condensed else self.n_fem
        self.n_if = m.interface_map.count
        self.V, self.F = V, F
        self.bem_off = self.nF
        self.flux_off = self.bem_off + V
        self.mech_off = self.flux_off + self.n_if
        self.elec_off = self.mech_off + self.nt
        self.n = self.elec_off + self.nt
        if self.n > _lib.BEM_LU_MAX_N:
            raise ValueError(f"Coupled system order {self.n} exceeds the dense LU limit {_lib.BEM_LU_MAX_N}.")
        self.ldc = _lda(self.n)
        self._handles = []
        self._build_device_tables()
        self._alloc()

    # -- frequency-invariant device tables ------------------------------------------------------------
    def _sparse(self, rows, cols, vals):
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        vals = np.ascontiguousarray(vals, np.complex128)
        h = C.c_void_p()
        _lib.check(self.dev.lib.bem_sparse_create(self.dev.ctx, self.n, self.n, len(vals), _np_ptr(rows, C.c_int),
                                                  _np_ptr(cols, C.c_int), _np_ptr(vals.view(np.float64), C.c_double),
                                                  C.byref(h)), "bem_sparse_create")
        self._handles.append(("sparse", h))
        return h

    def _block(self, mat, row_idx, col_idx, scale=1.0):
        """Sparse object of a block: entry (i, j) of mat goes to (row_idx[i], col_idx[j]) of the coupled matrix."""
        c = sp.coo_matrix(mat)
        c.sum_duplicates()
        keep = c.data != 0
        return self._sparse(np.asarray(row_idx)[c.row[keep]], np.asarray(col_idx)[c.col[keep]], scale * c.data[keep])

    def _colset(self, mat):
        src, ptr, idx, w = _colset_arrays(mat, self.finv)
        h = C.c_void_p()
        _lib.check(self.dev.lib.bem_colset_create(self.dev.ctx, self.dmesh.handle, len(src), _np_ptr(src, C.c_int),
                                                  len(ptr) - 1, _np_ptr(ptr, C.c_int), _np_ptr(idx, C.c_int),
                                                  _np_ptr(w.view(np.float64), C.c_double), C.byref(h)),
                   "bem_colset_create")
        self._handles.append(("colset", h))
        return h, len(ptr) - 1

    def _build_device_tables(self):
        io, to = self.iops, self.tops
        fem_pos = self.retained if self.condensed else np.arange(self.n_fem)  # FEM vertex of each FEM unknown
        fem_row = np.full(self.n_fem, -1, np.int64)
        fem_row[fem_pos] = np.arange(len(fem_pos))
        flux = self.flux_off + np.arange(self.n_if)
        bem = self.bem_off + self.vinv  # coupled index of each BEM vertex (caller numbering)
        self.sp_fem = []
        if self.condensed:
            self.sp_fem_load = self._block(io.fem_load[self.retained], np.arange(self.nF), flux, -1.0)
            self.sp_fem_trace = self._block(io.fem_trace[:, self.retained], flux, np.arange(self.nF))
        else:
            idx = np.arange(self.n_fem)
            self.sp_fem = [self._block(self.K, idx, idx), self._block(self.M, idx, idx),
                           self._block(self.Mloss, idx, idx)]
            self.sp_walls = [self._block(w[0], idx, idx) for w in self.walls]
            self.sp_fem_load = self._block(io.fem_load, idx, flux, -1.0)
            self.sp_fem_trace = self._block(io.fem_trace, flux, idx)
        self.sp_bem_trace = self._block(io.bem_trace, flux, bem, -1.0)
        self.cs_if, _ = self._colset(io.bem_flux)
        self.sp_motion = None
        if self.nt:
            mech = self.mech_off + np.arange(self.nt)
            self.cs_motion, _ = self._colset(to["bem_normal_velocity"])
            self.sp_fem_motion = self._block(to["fem_surface"][fem_pos], np.arange(self.nF), mech)
            self.sp_fem_force = self._block(to["fem_force"][fem_pos].T, mech, np.arange(self.nF), -1.0)
            self.sp_bem_force = self._block(to["bem_force"].T, mech, bem)
        self.n_pres = self.pvel.shape[1]
        if self.n_pres:
            self.cs_pres, _ = self._colset(self.pvel)

    def _alloc(self):
        self._cap_nb = 0
        self._rhs_cap = (0, 0)

    def _ensure(self, nb: int):
        """Device buffers for batches of nb frequencies (grown on demand)."""
        if nb <= self._cap_nb:
            return
        torch = self.dev.torch
        cdt = torch.complex64 if self.engine_precision == "mixed" else torch.complex128
        n, ldc, V, F = self.n, self.ldc, self.V, self.F
        kw = dict(device=self.dev.device)
        with torch.cuda.stream(self.dev.stream):
            self.Cm = self.C0 = None
            self.Cm = torch.empty((nb, n, ldc), dtype=cdt, **kw)
            self.C0 = torch.empty((nb, n, ldc), dtype=cdt, **kw)  # the assembled matrices, for residuals
            self.k = torch.zeros(nb, dtype=torch.float64, **kw)
            self.scal = torch.zeros((4, nb), dtype=torch.complex128, **kw)  # i rho w, -1, -i rho w, 1 per frequency
            self.qzero = torch.zeros((1, F), dtype=torch.complex128, **kw)
            self.bdummy = torch.zeros((nb, 1, _lda(V)), dtype=torch.complex128, **kw)
            self.rowmax = torch.zeros((nb, V), dtype=torch.float64, **kw)
            self.P = torch.zeros((nb, V, max(1, self.n_pres)), dtype=torch.complex128, **kw)
            self.ipiv = torch.zeros((nb, n), dtype=torch.int32, **kw)
            self.perm = torch.zeros((nb, n), dtype=torch.int32, **kw)
            self.info = torch.zeros(nb, dtype=torch.int32, **kw)
            wbytes = _lib.lu_workspace_bytes(self.prec, n, nb)
            self.luw = torch.empty(wbytes // (8 if self.engine_precision == "mixed" else 16), dtype=cdt, **kw)
            self.done = torch.zeros(nb, dtype=torch.int32, **kw)
            self.scratch = torch.zeros(2 * nb, dtype=torch.int32, **kw)
            self.rscale = torch.ones((nb, n), dtype=torch.float64, **kw)
        self._cap_nb = nb
        self._rhs_cap = (0, 0)

    def _rhs_buffers(self, nb: int, n_exc: int):
        ng = (n_exc + 3) // 4
        if self._rhs_cap == (nb, ng):
            return
        torch = self.dev.torch
        cdt = torch.complex64 if self.engine_precision == "mixed" else torch.complex128
        kw = dict(device=self.dev.device)
        with torch.cuda.stream(self.dev.stream):
            self.b = torch.zeros((nb, ng * 4, self.ldc), dtype=torch.complex128, **kw)
            self.x = torch.zeros((nb, ng * 4, self.ldc), dtype=torch.complex128, **kw)
            self.r = torch.zeros((nb, ng * 4, self.ldc), dtype=torch.complex128, **kw)
            self.y = torch.zeros((nb, ng * 4, self.ldc), dtype=cdt, **kw)
            self.rel = torch.zeros((ng, nb), dtype=torch.float64, **kw)
            self.iters = torch.zeros((ng, nb), dtype=torch.int32, **kw)
        self._rhs_cap = (nb, ng)

    def batch_capacity(self, share: float = 0.5) -> int:
        """Frequencies per device batch that fit in `share` of the free device memory (at most 8)."""
        torch = self.dev.torch
        free, _ = torch.cuda.mem_get_info(self.dev.device)
        cb = 8 if self.engine_precision == "mixed" else 16
        per = 2 * self.n * self.ldc * cb + self.n * 64 * 16
        return int(max(1, min(8, share * free // per)))

    def fem_system(self, freq: float):
        omega = 2.0 * math.pi * float(freq)
        k = omega / self.model.sound_speed
        A = FE.dynamic_stiffness(self.K, self.M, k, bulk_loss_mass=self.Mloss)
        ys = [FE.miki_admittance(float(freq), self.model.sound_speed, self.model.density, th, sig)
              for _, th, sig in self.walls]
        for (mat, _, _), y in zip(self.walls, ys):
            A = A - (1j * self.model.density * omega * y) * mat
        return A.tocsr(), ys

    def _ptr_at(self, t, f, row, col):
        return C.c_void_p(t.data_ptr() + ((f * self.n + row) * self.ldc + col) * t.element_size())

    def _scatter(self, h, f, alpha=1.0):
        a = complex(alpha)
        _lib.check(self.dev.lib.bem_sparse_scatter(self.dev.ctx, h, self.prec, a.real, a.imag, 1,
                                                   self._ptr_at(self.Cm, f, 0, 0), self.ldc), "bem_sparse_scatter")

    def solve(self, freq: float, excitations, points=None):
        """One frequency: (solutions, exterior pressures at `points` or None, wall admittances, timings)."""
        return self.solve_batch([freq], excitations, points)[0]

    def solve_batch(self, freqs, excitations, points=None):
        """Assemble, factor and solve the coupled systems of several frequencies as one device batch, for every
        excitation; returns per frequency (solutions, exterior pressures at `points` or None, wall admittances,
        timings; the batch's device times divided evenly)."""
        torch, lib, dev = self.dev.torch, self.dev.lib, self.dev
        exc = list(excitations)
        if not exc:
            raise ValueError("At least one coupled excitation is required.")
        freqs = [float(f) for f in freqs]
        nb = len(freqs)
        self._ensure(nb)
        m = self.model
        rho = m.density
        omegas = [2.0 * math.pi * f for f in freqs]
        ks = [w / m.sound_speed for w in omegas]
        iros = [1j * rho * w for w in omegas]
        tim = {"fem_system_s": 0.0, "fem_condensation_s": 0.0}
        host = []
        for f in freqs:
            t0 = time.perf_counter()
            A_fem, ys = self.fem_system(f)
            tim["fem_system_s"] += time.perf_counter() - t0
            t0 = time.perf_counter()
            S = parts = None
            if self.condensed:
                S, parts = schur_complement(A_fem, self.retained)
                for k2, v in parts["timings"].items():
                    tim["fem_condensation_" + k2] = tim.get("fem_condensation_" + k2, 0.0) + v
            tim["fem_condensation_s"] += time.perf_counter() - t0
            host.append((A_fem, ys, S, parts))
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        sc_host = np.array([iros, [-1.0] * nb, [-x for x in iros], [1.0] * nb], dtype=np.complex128)
        with torch.cuda.stream(dev.stream):
            ev[0].record(dev.stream)
            self.k[:nb].copy_(torch.tensor(ks, dtype=torch.float64))
            self.scal[:, :nb].copy_(torch.from_numpy(sc_host))
            _lib.check(lib.bem_dense_block(dev.ctx, self.prec, nb * self.n, self.n, None, 0, 0.0, 0.0, 0,
                                           _ptr(self.Cm), self.ldc), "bem_dense_block")
            sA = self.n * self.ldc
            ldb = self.bdummy.shape[-1]
            _lib.check(lib.bem_assemble(dev.ctx, self.dmesh.handle, self.prec, nb, _ptr(self.k), _ptr(self.qzero), 1,
                                        _ptr(self.scal), self._ptr_at(self.Cm, 0, self.bem_off, self.bem_off), self.ldc,
                                        sA, None, _ptr(self.bdummy), ldb, ldb, _ptr(self.rowmax)), "bem_assemble")
            ev[1].record(dev.stream)
            row = lambda i: C.c_void_p(self.scal.data_ptr() + 16 * i * self.scal.shape[1])  # noqa: E731
            _lib.check(lib.bem_rhs_columns(dev.ctx, self.dmesh.handle, self.cs_if, self.prec, nb, _ptr(self.k), row(1),
                                           self._ptr_at(self.Cm, 0, self.bem_off, self.flux_off), self.ldc, sA),
                       "bem_rhs_columns")
            if self.nt:
                _lib.check(lib.bem_rhs_columns(dev.ctx, self.dmesh.handle, self.cs_motion, self.prec, nb, _ptr(self.k),
                                               row(2), self._ptr_at(self.Cm, 0, self.bem_off, self.mech_off), self.ldc,
                                               sA), "bem_rhs_columns")
            if self.n_pres:
                _lib.check(lib.bem_rhs_columns(dev.ctx, self.dmesh.handle, self.cs_pres, _lib.BEM_PREC_FP64, nb,
                                               _ptr(self.k), row(0), _ptr(self.P), self.n_pres, self.V * self.n_pres),
                           "bem_rhs_columns")
            ev[2].record(dev.stream)
            S_dev = []
            diags = []
            for j in range(nb):
                A_fem, ys, S, parts = host[j]
                k, iro = ks[j], iros[j]
                if self.condensed:
                    S_dev.append(torch.from_numpy(np.ascontiguousarray(S)).to(dev.device))
                    _lib.check(lib.bem_dense_block(dev.ctx, self.prec, self.nF, self.nF, _ptr(S_dev[-1]), self.nF, 1.0,
                                                   0.0, 1, self._ptr_at(self.Cm, j, 0, 0), self.ldc), "bem_dense_block")
                else:
                    self._scatter(self.sp_fem[0], j, 1.0)
                    self._scatter(self.sp_fem[1], j, -k * k)
                    self._scatter(self.sp_fem[2], j, -1j * k * k)
                    for h, y in zip(self.sp_walls, ys):
                        self._scatter(h, j, -iro * y)
                self._scatter(self.sp_fem_load, j)
                self._scatter(self.sp_fem_trace, j)
                self._scatter(self.sp_bem_trace, j)
                if self.nt:
                    self._scatter(self.sp_fem_motion, j, -iro)
                    self._scatter(self.sp_fem_force, j)
                    self._scatter(self.sp_bem_force, j)
                    w = omegas[j]
                    tds = m.transducers
                    mech = self.mech_off + np.arange(self.nt)
                    elec = self.elec_off + np.arange(self.nt)
                    zm = np.array([complex(t.rms_n_s_per_m, -w * t.mmd_kg + 1.0 / (w * t.cms_m_per_n)) for t in tds])
                    ze = np.array([complex(t.re_ohm, -w * t.le_h) for t in tds])
                    bl = np.array([t.bl_n_per_a for t in tds], dtype=np.complex128)
                    diags.append(self._sparse(np.concatenate([mech, mech, elec, elec]),
                                              np.concatenate([mech, elec, mech, elec]),
                                              np.concatenate([zm, -bl, bl, ze])))
                    self._scatter(diags[-1], j)
            # exact power-of-two row scaling: FEM, BEM, trace and lumped rows differ in scale by orders of magnitude
            _lib.check(lib.bem_row_equilibrate(dev.ctx, self.prec, nb * self.n, self.n, _ptr(self.Cm), self.ldc,
                                               _ptr(self.rscale)), "bem_row_equilibrate")
            self.C0[:nb].copy_(self.Cm[:nb])
            ev[3].record(dev.stream)
        # right-hand sides (host), as the reference builds them (BeatEngineCoupled.jl:2360-2420, 2285-2310)
        n_exc = len(exc)
        ng = (n_exc + 3) // 4
        dev.stream.synchronize()
        P_all = self.P[:nb].cpu().numpy() if self.n_pres else None
        rs_all = self.rscale[:nb].cpu().numpy()
        rhs = np.zeros((nb, ng * 4, self.ldc), np.complex128)
        fem_rhs_all, pres_q_all, t_rhs = [], [], []
        for j in range(nb):
            t0 = time.perf_counter()
            fem_rhs, bem_rhs, pres_q, elec_rhs = self._excitation_rhs(exc, omegas[j], iros[j],
                                                                      None if P_all is None else P_all[j])
            parts = host[j][3]
            if self.condensed:
                g = fem_rhs[self.retained]
                if parts["lu"] is not None:
                    g = g - parts["A_GI"] @ parts["lu"].solve(np.ascontiguousarray(fem_rhs[parts["interior"]]))
                rhs[j, :n_exc, : self.nF] = g.T
            else:
                rhs[j, :n_exc, : self.nF] = fem_rhs.T
            rhs[j, :n_exc, self.bem_off : self.bem_off + self.V] = bem_rhs.T
            rhs[j, :n_exc, self.elec_off : self.elec_off + self.nt] = elec_rhs.T
            rhs[j, :, : self.n] *= rs_all[j][None, :]  # the rows' equilibration factors
            fem_rhs_all.append(fem_rhs)
            pres_q_all.append(pres_q)
            t_rhs.append(time.perf_counter() - t0)
        self._rhs_buffers(nb, n_exc)
        sA, sB = self.n * self.ldc, ng * 4 * self.ldc
        with torch.cuda.stream(dev.stream):
            self.b.copy_(torch.from_numpy(rhs))
            self.x.zero_()
            _lib.check(lib.bem_ctx_set_lu_products(dev.ctx, _lib.BEM_LU_SPLIT_BF16 if self.prec == _lib.BEM_PREC_MIXED
                                                   else _lib.BEM_LU_FP32), "bem_ctx_set_lu_products")
            c = TU.table().lu_choice(self.engine_precision, self.n, nb)
            t = _lib.Tuning(c.split_tile, c.trsv_wg_min_batch, c.lookahead_wgs, 2048)
            _lib.check(lib.bem_ctx_set_tuning(dev.ctx, C.byref(t)), "bem_ctx_set_tuning")
            _lib.check(lib.bem_lu_factor(dev.ctx, self.prec, self.n, nb, _ptr(self.Cm), self.ldc, sA, _ptr(self.ipiv),
                                         _ptr(self.perm), self.n, _ptr(self.info), _ptr(self.luw)), "bem_lu_factor")
            for gi in range(ng):
                nc = min(4, n_exc - 4 * gi)
                grp = lambda t_: C.c_void_p(t_.data_ptr() + 4 * gi * self.ldc * t_.element_size())  # noqa: E731
                relp = C.c_void_p(self.rel.data_ptr() + 8 * gi * nb)
                if self.engine_precision == "mixed":
                    _lib.check(lib.bem_refine(dev.ctx, self.n, nb, _ptr(self.Cm), self.ldc, sA, _ptr(self.perm),
                                              self.n, _ptr(self.luw), _ptr(self.C0), grp(self.b), self.ldc, sB,
                                              grp(self.x), grp(self.r), grp(self.y), nc, REFINE_MAX_ITER, REFINE_TOL,
                                              relp, C.c_void_p(self.iters.data_ptr() + 4 * gi * nb),
                                              _ptr(self.done)), "bem_refine")
                else:
                    _lib.check(lib.bem_lu_solve(dev.ctx, self.prec, self.n, nb, _ptr(self.Cm), self.ldc, sA,
                                                _ptr(self.perm), self.n, _ptr(self.luw), grp(self.b), self.ldc, sB,
                                                grp(self.x), self.ldc, sB, nc, None), "bem_lu_solve")
                    _lib.check(lib.bem_residual(dev.ctx, self.prec, self.n, nb, _ptr(self.C0), self.ldc, sA,
                                                grp(self.b), grp(self.x), self.ldc, sB, grp(self.r), nc, relp,
                                                _ptr(self.scratch)), "bem_residual")
            ev[4].record(dev.stream)
        dev.stream.synchronize()
        dt = {"bem_operator_s": ev[0].elapsed_time(ev[1]) / 1e3 / nb,
              "bem_matrix_s": ev[1].elapsed_time(ev[2]) / 1e3 / nb,
              "block_assembly_s": ev[2].elapsed_time(ev[3]) / 1e3 / nb,
              "coupled_factorization_s": ev[3].elapsed_time(ev[4]) / 1e3 / nb}
        tol = REFINE_TOL if self.engine_precision == "mixed" else FP64_RESIDUAL_LIMIT
        for gi in range(ng):
            nc = min(4, n_exc - 4 * gi)
            flags = np.zeros(nb, np.int32)
            rc = lib.bem_solve_status(dev.ctx, self.n, nb, _ptr(self.info),
                                      C.c_void_p(self.x.data_ptr() + 4 * gi * self.ldc * 16), self.ldc, sB, nc,
                                      C.c_void_p(self.rel.data_ptr() + 8 * gi * nb), tol, None, 0.0,
                                      _np_ptr(flags, C.c_int))
            if rc not in (_lib.BEM_OK, _lib.BEM_E_SINGULAR, _lib.BEM_E_ACCURACY):
                _lib.check(rc, "bem_solve_status")
            bad = np.nonzero(flags)[0]
            if len(bad):
                from .driver import status_text

                j = int(bad[0])
                raise RuntimeError(f"Coupled dense solve failed at {freqs[j]:g} Hz: {status_text(int(flags[j]))}.")
        Xall = self.x[:nb, :n_exc, : self.n].cpu().numpy()  # (nb, n_exc, n)
        rel = self.rel.cpu().numpy()  # (ng, nb): relative residual of the row-equilibrated systems
        iters = self.iters.cpu().numpy() if self.engine_precision == "mixed" else None
        qs = []
        out = []
        for j in range(nb):
            X = Xall[j].T
            parts = host[j][3]
            fem_rhs = fem_rhs_all[j]
            t0 = time.perf_counter()
            fem_p = np.zeros((self.n_fem, n_exc), np.complex128)
            interior_res = None
            if self.condensed:
                pG = X[: self.nF]
                fem_p[self.retained] = pG
                interior_res = np.zeros(n_exc)
                if parts["lu"] is not None:
                    I = parts["interior"]
                    f_I = fem_rhs[I]
                    u_I = parts["lu"].solve(np.ascontiguousarray(f_I - parts["A_IG"] @ pG))
                    fem_p[I] = u_I
                    t_int = parts["A_II"] @ u_I
                    t_ret = parts["A_IG"] @ pG
                    res = t_int + t_ret - f_I
                    interior_res = np.array([
                        np.linalg.norm(res[:, e]) / max(np.linalg.norm(t_int[:, e]), np.linalg.norm(t_ret[:, e]),
                                                       np.linalg.norm(f_I[:, e]), np.finfo(float).eps)
                        for e in range(n_exc)])
            else:
                fem_p[:] = X[: self.nF]
            t_rec = time.perf_counter() - t0
            bem_p = X[self.bem_off + self.vinv]  # caller numbering (V, n_exc)
            flux = X[self.flux_off : self.flux_off + self.n_if]
            u = X[self.mech_off : self.mech_off + self.nt]
            cur = X[self.elec_off : self.elec_off + self.nt]
            q = self.iops.bem_flux @ flux + iros[j] * (self.tops["bem_normal_velocity"] @ u) + pres_q_all[j]
            qs.append(q)
            sols = []
            for e in range(n_exc):
                pc, fc = self._interface_errors(fem_p[:, e], bem_p[:, e], flux[:, e], q[:, e])
                sols.append(CoupledSolution(
                    fem_pressure=fem_p[:, e], bem_pressure=bem_p[:, e], interface_flux=flux[:, e],
                    bem_neumann=q[:, e], diaphragm_velocity=u[:, e], voice_coil_current=cur[:, e],
                    pressure_continuity_error=pc, flux_conservation_error=fc,
                    fem_interior_residual=None if interior_res is None else float(interior_res[e]),
                    relative_residual=float(rel[e // 4, j]),
                    fem_rhs_condensation_s=t_rhs[j] if self.condensed else None,
                    fem_reconstruction_s=t_rec if self.condensed else None))
            tj = dict(dt)
            tj["fem_system_s"] = tim["fem_system_s"] / nb
            tj["fem_condensation_s"] = tim["fem_condensation_s"] / nb
            for k2 in ("partition_s", "factorization_s", "schur_extraction_s"):
                if "fem_condensation_" + k2 in tim:
                    tj["fem_condensation_" + k2] = tim["fem_condensation_" + k2] / nb
            if iters is not None:
                tj["refinement_iterations"] = int(iters[:, j].max())
            out.append([sols, None, host[j][1], tj])
        if points is not None and len(points):
            t0 = time.perf_counter()
            fields = self._field(np.asarray(points, np.float64), nb, n_exc, qs)
            tf = (time.perf_counter() - t0) / nb
            for j in range(nb):
                out[j][1] = fields[j]
                out[j][3]["field_s"] = tf
        if self.validation:
            for j in range(nb):
                self._replay(freqs[j], out[j][0], qs[j])
        for h in diags:
            lib.bem_sparse_destroy(h)
        ids = {id(h) for h in diags}
        self._handles = [h for h in self._handles if id(h[1]) not in ids]
        return [tuple(o) for o in out]

    def _excitation_rhs(self, exc, omega, iro, P):
        """FEM, BEM (internal vertex order), prescribed Neumann and coil right-hand sides of every excitation
        (solve_coupled_excitations, BeatEngineCoupled.jl:2360-2420)."""
        m = self.model
        n_exc = len(exc)
        fem_rhs = np.zeros((self.n_fem, n_exc), np.complex128)
        bem_rhs = np.zeros((self.V, n_exc), np.complex128)
        pres_q = np.zeros((self.F, n_exc), np.complex128)
        elec_rhs = np.zeros((self.nt, n_exc), np.complex128)
        for e, x in enumerate(exc):
            amp = complex(x.amplitude)
            if x.kind == "normal_velocity":
                tags, wts = list(x.fem_tags), list(x.fem_weights) or [1.0] * len(x.fem_tags)
                if len(tags) != len(wts):
                    raise ValueError("Prescribed FEM boundary tags and weights must have the same length.")
                if not all(np.isfinite(w) and w > 0.0 for w in wts):
                    raise ValueError("Prescribed FEM boundary weights must be finite and greater than zero.")
                for tag, w in zip(tags, wts):
                    fem_rhs[:, e] += FE.prescribed_velocity_load(m.fem_mesh, tag, m.density, omega, amp * w)
                if x.bem_source_index > 0:
                    if x.bem_source_index > self.n_pres:
                        raise ValueError(f"Prescribed BEM source index {x.bem_source_index} is unavailable.")
                    bem_rhs[:, e] = amp * P[:, x.bem_source_index - 1]
                    pres_q[:, e] = amp * iro * self.pvel[:, x.bem_source_index - 1].toarray().reshape(-1)
                if not tags and x.bem_source_index == 0:
                    raise ValueError("A prescribed-velocity excitation must own at least one FEM or BEM moving "
                                     "boundary.")
            elif x.kind == "voltage":
                if not 1 <= x.transducer_index <= self.nt:
                    raise ValueError(f"Voltage excitation references invalid transducer index {x.transducer_index}.")
                elec_rhs[x.transducer_index - 1, e] = amp
            else:
                raise ValueError(f"Unsupported coupled excitation kind: {x.kind}.")
        return fem_rhs, bem_rhs, pres_q, elec_rhs

    def _interface_errors(self, p_fem, p_bem, flux, q):
        """Pressure continuity and integrated-flux conservation on the interface
        (_coupled_solution_from_parts, BeatEngineCoupled.jl:2181-2261)."""
        io, im = self.iops, self.model.interface_map
        tf, tb = io.fem_trace @ p_fem, io.bem_trace @ p_bem
        scale = max(np.linalg.norm(tf), np.linalg.norm(tb), np.finfo(float).eps)
        pc = float(np.linalg.norm(tf - tb) / scale)
        fm, bm = self.model.fem_mesh, self.model.bem_mesh
        dof = np.full(fm.n_vertices, -1, np.int64)
        dof[im.fem_vertex_indices] = np.arange(im.count)
        faces = fm.boundary_faces[im.fem_face_indices]
        fem_int = np.sum(FE.triangle_areas(fm.vertices, faces) * flux[dof[faces]].sum(axis=1) / 3.0)
        bem_int = np.sum(im.normal_sign * bm.areas[im.bem_face_indices] * q[im.bem_face_indices])
        fscale = max(abs(fem_int), abs(bem_int), np.finfo(float).eps)
        return pc, float(abs(fem_int - bem_int) / fscale)

    def _field(self, points, nb, n_exc, qs):
        """Exterior pressure of every frequency of the batch from the BEM trace and the coupled Neumann data
        (bem_field over the batch, 4 excitations per call). Returns (n_obs, n_exc) per frequency."""
        torch, lib, dev = self.dev.torch, self.dev.lib, self.dev
        n_obs = len(points)
        ng = (n_exc + 3) // 4
        q_int = np.stack([np.ascontiguousarray(q[self.prep.fperm].T) for q in qs])  # (nb, n_exc, F) internal order
        with torch.cuda.stream(dev.stream):
            obs = torch.from_numpy(np.ascontiguousarray(points)).to(dev.device)
            qd = torch.from_numpy(q_int).to(dev.device)
            out = torch.zeros((nb, n_exc, n_obs), dtype=torch.complex128, device=dev.device)
            for c0 in range(0, n_exc, 4):
                nc = min(4, n_exc - c0)
                p = C.c_void_p(self.x.data_ptr() + (c0 * self.ldc + self.bem_off) * 16)
                _lib.check(lib.bem_field(dev.ctx, self.dmesh.handle, self.prec, nb, _ptr(self.k), n_obs, _ptr(obs), p,
                                         self.ldc, ng * 4 * self.ldc, C.c_void_p(qd.data_ptr() + c0 * self.F * 16),
                                         n_exc * self.F, nc, C.c_void_p(out.data_ptr() + c0 * n_obs * 16), n_obs,
                                         n_exc * n_obs), "bem_field")
        dev.stream.synchronize()
        return [o.T for o in out.cpu().numpy()]

    def _replay(self, freq, sols, q):
        """BEM-only replay error of the validation diagnostics (BeatEngineCoupled.jl:2236-2244): the BEM block
        solved alone with the coupled Neumann data must reproduce the coupled BEM pressure."""
        from .engine import Engine

        iro = 1j * self.model.density * 2.0 * math.pi * float(freq)
        eng = Engine(self.prep, q / iro, dev=self.dev)  # q in the caller's face order, as Engine takes it
        try:
            b = eng.batch(self.engine_precision, 1, None, self.model.density, self.model.sound_speed)
            r = b.run([float(freq)])
            b.close()
        finally:
            eng.close()
        for e, s in enumerate(sols):
            rp = r.pressure[0][:, e]
            scale = max(np.linalg.norm(s.bem_pressure), np.linalg.norm(rp), np.finfo(float).eps)
            s.all_bem_replay_error = float(np.linalg.norm(s.bem_pressure - rp) / scale)

    def close(self):
        lib = self.dev.lib
        for kind, h in self._handles:
            (lib.bem_colset_destroy if kind == "colset" else lib.bem_sparse_destroy)(h)
        self._handles = []
        if getattr(self, "dmesh", None) is not None:
            self.dmesh.close()
            self.dmesh = None
        if self._own_dev:
            self.dev.close()
