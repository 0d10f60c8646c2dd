"""GPU parity of the HIP path against the FP64 oracle (pulsating sphere, 1,280 triangles)."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A_SPHERE = 0.1
FREQ = 1000.0
RHO = 1.205
C0 = 343.0


@pytest.fixture(scope="module")
def sphere():
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn

    V, F, T = syn.icosphere(3, A_SPHERE)
    return G.prepare(G.BoundaryMesh(V, F, T))


@pytest.fixture(scope="module")
def oracle_sphere(sphere):
    from oracle import bem_oracle as O

    m = sphere.original  # oracle in the original numbering; the engine reports in it too
    om = O.Mesh(m.vertices, m.faces, m.tags, m.mesh_ids)
    pts = O.fibonacci_observation({"spherical_sampling_enabled": True, "spherical_sampling_points": 1000,
                                   "distance": 2.0})[0].astype(np.float64)
    out = O.solve_frequency(om, FREQ, np.ones((m.n_faces, 1)), C0, RHO, points=pts)
    out["points"] = pts
    return out


def _engine(dev, prep):
    from beat_hip.engine import Engine

    return Engine(prep, np.ones((prep.n_faces, 1)), dev=dev)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_assembly_matches_oracle(gpu_device, sphere, oracle_sphere, precision, tol):
    eng = _engine(gpu_device, sphere)
    b = eng.batch(precision, 1, None, RHO, C0)
    b.set_frequencies([FREQ])
    with gpu_device.torch.cuda.stream(gpu_device.stream):
        b._assemble()
    gpu_device.sync()
    N = sphere.n_vertices
    A = b.A0[0, :, :N].cpu().numpy().astype(np.complex128)
    rhs = b.b[0, 0, :N].cpu().numpy()
    vp = sphere.vperm  # device rows/columns are in locality order
    lhs_ref = oracle_sphere["lhs"][np.ix_(vp, vp)]
    rhs_ref = oracle_sphere["rhs"][vp, 0]
    print(precision, "A rel", _rel(A, lhs_ref), "b rel", _rel(rhs, rhs_ref))
    assert _rel(A, lhs_ref) < tol
    assert _rel(rhs, rhs_ref) < tol
    eng.close()


def test_singular_precision_of_mixed_assembly(gpu_device, sphere, oracle_sphere):
    """Mixed assemblies sum the Duffy rules in FP32 by default (as the reference's ROCm singular kernels in the
    working precision) or in FP64 on request: both within the mixed tolerance of the oracle, and within 2e-6 of each
    other (FP32 rounding of the singular blocks only); unknown modes are refused."""
    from beat_hip import _lib

    lib, ctx = gpu_device.lib, gpu_device.ctx
    eng = _engine(gpu_device, sphere)
    vp = sphere.vperm
    lhs_ref = oracle_sphere["lhs"][np.ix_(vp, vp)]
    N = sphere.n_vertices
    A = {}
    try:
        for mode in (_lib.BEM_SING_FP64, _lib.BEM_SING_WORKING):
            _lib.check(lib.bem_ctx_set_singular_precision(ctx, mode), "bem_ctx_set_singular_precision")
            b = eng.batch("mixed", 2, None, RHO, C0)
            b.set_frequencies([FREQ, 3 * FREQ])
            with gpu_device.torch.cuda.stream(gpu_device.stream):
                b._assemble()
            gpu_device.sync()
            A[mode] = b.A0[:, :, :N].cpu().numpy().astype(np.complex128)
            b.close()
            print("singular mode", mode, "A rel", _rel(A[mode][0], lhs_ref))
            assert _rel(A[mode][0], lhs_ref) < 1e-5
        for f in range(2):
            assert _rel(A[_lib.BEM_SING_WORKING][f], A[_lib.BEM_SING_FP64][f]) < 2e-6
        assert lib.bem_ctx_set_singular_precision(ctx, 2) == _lib.BEM_E_ARG
    finally:
        lib.bem_ctx_set_singular_precision(ctx, _lib.BEM_SING_WORKING)
        eng.close()


@pytest.fixture(scope="module")
def fan_sphere():
    """Latitude/longitude sphere whose two poles are shared by 80 triangles each (1,120 triangles): star rows
    beyond the 64 the row gather keeps in LDS, and vertices that span two 64-element chunks."""
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from oracle import bem_oracle as O

    V, F, T = syn.uv_sphere(8, 80, A_SPHERE)
    prep = G.prepare(G.BoundaryMesh(V, F, T))
    m = prep.original
    ref = O.solve_frequency(O.Mesh(m.vertices, m.faces, m.tags, m.mesh_ids), FREQ, np.ones((m.n_faces, 1)), C0, RHO)
    return prep, ref


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_high_valence_vertices_match_oracle(gpu_device, fan_sphere, precision, tol):
    prep, ref = fan_sphere
    assert np.bincount(prep.original.faces.ravel()).max() == 80
    eng = _engine(gpu_device, prep)
    b = eng.batch(precision, 1, None, RHO, C0)
    b.set_frequencies([FREQ])
    with gpu_device.torch.cuda.stream(gpu_device.stream):
        b._assemble()
    gpu_device.sync()
    N = prep.n_vertices
    A = b.A0[0, :, :N].cpu().numpy().astype(np.complex128)
    rhs = b.b[0, 0, :N].cpu().numpy()
    vp = prep.vperm
    print(precision, "A rel", _rel(A, ref["lhs"][np.ix_(vp, vp)]), "b rel", _rel(rhs, ref["rhs"][vp, 0]))
    assert _rel(A, ref["lhs"][np.ix_(vp, vp)]) < tol
    assert _rel(rhs, ref["rhs"][vp, 0]) < tol
    res, _ = eng.solve([FREQ], precision=precision, rho=RHO, sound_speed=C0)
    p = res[0].pressure[0, :, 0]
    assert res[0].info[0] == 0
    assert _rel(p, ref["pressure"][:, 0]) < (1e-10 if precision == "fp64" else 1e-4)
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_and_field_match_oracle(gpu_device, sphere, oracle_sphere, precision):
    eng = _engine(gpu_device, sphere)
    res, _ = eng.solve([FREQ], precision=precision, obs=oracle_sphere["points"], rho=RHO, sound_speed=C0)
    r = res[0]
    p = r.pressure[0, :, 0]
    u = r.field[0, :, 0]
    po, uo = oracle_sphere["pressure"][:, 0], oracle_sphere["field"][:, 0]
    db = np.abs(20 * np.log10(np.abs(u) / np.abs(uo)))
    ph = np.degrees(np.abs(np.angle(u / uo)))
    print(precision, "p rel", _rel(p, po), "field max dB", db.max(), "phase", ph.max(), "res", r.residual,
          "growth", r.growth, "iters", r.iters)
    assert r.info[0] == 0
    if precision == "fp64":
        assert _rel(p, po) < 1e-10
        assert _rel(u, uo) < 1e-10
        assert r.residual[0] < 1e-12
    else:
        assert _rel(p, po) < 1e-4
        assert db.max() < 0.05 and ph.max() < 0.5
        assert r.residual[0] <= 1e-10
    # analytic pulsating sphere (discretisation error of 1,280 flat triangles)
    from oracle.bem_oracle import pulsating_sphere_pressure

    k = 2 * math.pi * FREQ / C0
    pa = pulsating_sphere_pressure(np.linalg.norm(oracle_sphere["points"], axis=1), A_SPHERE, k, RHO,
                                   2 * math.pi * FREQ)
    assert np.abs(20 * np.log10(np.abs(u) / np.abs(pa))).max() < 0.05
    eng.close()


# sizes: one partial half block (40), one trailing block with a partial second half (100), partial trailing
# blocks (190, 200) and the register-panel thresholds (1500, 2100), and the panels with row sets in LDS (3,500 to
# 6,500 rows: three to seven row sets per thread); batches of 8 at N <= 4096 take the one-workgroup-per-matrix
# solve, the others the blocked multi-workgroup sweeps; above 12,288 rows the panels hold 24 row sets per thread
# (checked by the backward error only: a CPU solve of that size takes minutes). Leading dimensions: padded to 16
# (the sub-panels move whole 16-byte granules) except where given: an odd one (entry by entry), and odd orders whose
# last sub-panel ends inside a granule; 3,073 and 6,145 rows change the sub-panel width inside the first half
# (8 to 16 and 4 to 8 columns), so the U rows and inverse blocks of one sub-panel feed a sub-panel of another width.
@pytest.mark.parametrize("precision,n,batch,lda", [
    ("mixed", 40, 2, None), ("fp64", 100, 1, None), ("mixed", 100, 2, None), ("mixed", 190, 2, None),
    ("fp64", 200, 2, None), ("mixed", 200, 3, None), ("fp64", 1500, 1, None), ("mixed", 2100, 2, None),
    ("fp64", 3500, 1, None), ("mixed", 4500, 1, None), ("mixed", 6500, 1, None), ("fp64", 6500, 1, None),
    ("fp64", 300, 8, None), ("mixed", 1100, 8, None), ("fp64", 12400, 1, None), ("mixed", 13000, 1, None),
    ("mixed", 1001, 2, 1001), ("fp64", 517, 2, 517), ("mixed", 3073, 1, 3074), ("mixed", 6145, 1, 6146)])
def test_lu_random(gpu_device, precision, n, batch, lda):
    import ctypes as C

    from beat_hip import _lib
    from beat_hip.engine import PRECISIONS, _ptr

    torch = gpu_device.torch
    rng = np.random.default_rng(n + batch)
    lda = ((n + 15) // 16) * 16 if lda is None else lda
    M = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))
    M += 4 * np.sqrt(n) * np.eye(n)[None] * (rng.standard_normal(batch)[:, None, None] > 5)  # usually no boost
    B = rng.standard_normal((batch, 2, n)) + 1j * rng.standard_normal((batch, 2, n))
    cdt = torch.complex64 if precision == "mixed" else torch.complex128
    dev = gpu_device.device
    with torch.cuda.stream(gpu_device.stream):
        A = torch.zeros((batch, n, lda), dtype=cdt, device=dev)
        A[:, :, :n] = torch.from_numpy(M).to(dev).to(cdt)
        b = torch.zeros((batch, 2, lda), dtype=torch.complex128, device=dev)
        b[:, :, :n] = torch.from_numpy(B).to(dev)
        y = torch.zeros((batch, 2, lda), dtype=cdt, device=dev)
        ipiv = torch.zeros((batch, n), dtype=torch.int32, device=dev)
        perm = torch.zeros((batch, n), dtype=torch.int32, device=dev)
        info = torch.zeros(batch, dtype=torch.int32, device=dev)
        prec = PRECISIONS[precision]
        work = torch.empty(_lib.lu_workspace_bytes(prec, n, batch), dtype=torch.uint8, device=dev)
    gpu_device.sync()
    lib = gpu_device.lib
    _lib.check(lib.bem_ctx_set_lu_products(gpu_device.ctx, _lib.BEM_LU_FP32), "products")
    _lib.check(lib.bem_lu_factor(gpu_device.ctx, prec, n, batch, _ptr(A), lda, n * lda, _ptr(ipiv), _ptr(perm), n,
                                 _ptr(info), _ptr(work)), "factor")
    _lib.check(lib.bem_lu_solve(gpu_device.ctx, prec, n, batch, _ptr(A), lda, n * lda, _ptr(perm), n, _ptr(work),
                                _ptr(b), lda, 2 * lda, _ptr(y), lda, 2 * lda, 2, None), "solve")
    gpu_device.sync()
    X = y[:, :, :n].cpu().numpy().astype(np.complex128)
    # forward error vs LAPACK, and the backward error |B - M X| / (|M|_2 |X|) that partial-pivoting LU bounds.
    # LAPACK's own cgetrf reaches 6.6e-4 forward / 1.6e-5 backward on a matrix of this family at N = 2100, and
    # 1.1e-2 / 4.7e-5 at N = 6,500 (the forward error grows with the condition number).
    tol_fwd, tol_bwd = (1e-10, 1e-13) if precision == "fp64" else ((3e-3 if n <= 2100 else 5e-2), 1e-4)
    for m in range(batch):
        err = 0.0
        if n <= 8000:
            ref = np.linalg.solve(M[m], B[m].T).T
            err = np.linalg.norm(X[m] - ref) / np.linalg.norm(ref)
        # for large n, |M|_2 of an iid complex Gaussian matrix is close to 2 |M|_F / sqrt(n) (no CPU SVD)
        mnorm = np.linalg.norm(M[m], 2) if n <= 8000 else np.linalg.norm(M[m]) / np.sqrt(n) * 2
        bwd = np.linalg.norm(B[m].T - M[m] @ X[m].T) / (mnorm * np.linalg.norm(X[m]))
        print(precision, n, m, "rel err", err, "backward", bwd)
        assert err < tol_fwd and bwd < tol_bwd
    assert int(info.cpu().numpy().max()) == 0


# sizes: partial 128-row and 64-column tiles of the split update (700), the mr = 2 and mr = 3 panels (2100)
@pytest.mark.parametrize("n,batch", [(700, 3), (2100, 2)])
def test_lu_split_bf16_products_precondition_refinement(gpu_device, n, batch):
    """Split-bf16 trailing updates (three bf16 MFMA products per FP32 product): the factors of a matrix with a
    condition number in the tens have a backward error near 2^-16 sqrt(N), and FP64 refinement with them reaches
    a relative residual of 1e-10 in as many steps as with FP32 factors (at most three here). One bf16 product
    per FP32 product (BEM_LU_BF16): a backward error near 2^-8, refinement to 1e-10 in at most six steps."""
    from beat_hip import _lib
    from beat_hip.engine import PRECISIONS, _ptr

    torch = gpu_device.torch
    rng = np.random.default_rng(5 + n)
    lda = ((n + 15) // 16) * 16
    M = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))) / np.sqrt(2 * n)
    M += (1.5 + 0.5j) * np.eye(n)[None]
    B = rng.standard_normal((batch, 1, n)) + 1j * rng.standard_normal((batch, 1, n))
    prec = PRECISIONS["mixed"]
    lib, dev = gpu_device.lib, gpu_device.device
    results = {}
    try:
        for mode in (_lib.BEM_LU_FP32, _lib.BEM_LU_SPLIT_BF16, _lib.BEM_LU_BF16):
            _lib.check(lib.bem_ctx_set_lu_products(gpu_device.ctx, mode), "products")
            with torch.cuda.stream(gpu_device.stream):
                A = torch.zeros((batch, n, lda), dtype=torch.complex64, device=dev)
                A[:, :, :n] = torch.from_numpy(M).to(dev).to(torch.complex64)
                ipiv = torch.zeros((batch, n), dtype=torch.int32, device=dev)
                perm = torch.zeros((batch, n), dtype=torch.int32, device=dev)
                info = torch.zeros(batch, dtype=torch.int32, device=dev)
                work = torch.empty(_lib.lu_workspace_bytes(prec, n, batch), dtype=torch.uint8, device=dev)
                b = torch.zeros((batch, 1, lda), dtype=torch.complex128, device=dev)
                y = torch.zeros((batch, 1, lda), dtype=torch.complex64, device=dev)
            gpu_device.sync()
            _lib.check(lib.bem_lu_factor(gpu_device.ctx, prec, n, batch, _ptr(A), lda, n * lda, _ptr(ipiv),
                                         _ptr(perm), n, _ptr(info), _ptr(work)), "factor")
            gpu_device.sync()
            assert int(info.cpu().numpy().max()) == 0

            def solve(rhs):
                b[:, :, :n] = torch.from_numpy(rhs).to(dev)
                gpu_device.sync()
                _lib.check(lib.bem_lu_solve(gpu_device.ctx, prec, n, batch, _ptr(A), lda, n * lda, _ptr(perm), n,
                                            _ptr(work), _ptr(b), lda, lda, _ptr(y), lda, lda, 1, None), "solve")
                gpu_device.sync()
                return y[:, :, :n].cpu().numpy().astype(np.complex128)

            X = solve(B)
            bwd = [np.linalg.norm(B[m, 0] - M[m] @ X[m, 0]) / (np.linalg.norm(M[m], 2) * np.linalg.norm(X[m, 0]))
                   for m in range(batch)]
            hist = []
            for it in range(6):
                R = B - np.einsum("mij,mj->mi", M, X[:, 0])[:, None]
                rel = max(np.linalg.norm(R[m]) / np.linalg.norm(B[m]) for m in range(batch))
                hist.append(rel)
                if rel < 1e-10:
                    break
                X = X + solve(R)
            results[mode] = (max(bwd), hist, A.cpu().numpy())
            print("products", mode, "n", n, "backward", max(bwd), "residuals", hist)
    finally:
        _lib.check(lib.bem_ctx_set_lu_products(gpu_device.ctx, _lib.BEM_LU_FP32), "products")
    bwd32, hist32, f32 = results[_lib.BEM_LU_FP32]
    bwd16, hist16, f16 = results[_lib.BEM_LU_SPLIT_BF16]
    bwd8, hist8, f8 = results[_lib.BEM_LU_BF16]
    assert not np.array_equal(f32, f16) and not np.array_equal(f16, f8)  # the bf16 products are in use
    assert bwd32 < 1e-5 and bwd16 < 1e-4 and bwd8 < 2e-2
    assert hist16[-1] < 1e-10 and len(hist16) <= 4 and len(hist16) <= len(hist32) + 1
    assert hist8[-1] < 1e-10 and len(hist8) <= 6


@pytest.mark.parametrize("precision,products", [("mixed", "fp32"), ("mixed", "split_bf16"), ("mixed", "bf16"),
                                                ("fp64", "fp32")])
def test_lu_streams_do_not_change_the_factors(gpu_device, precision, products):
    """Look-ahead (fan-out >= 2: a second stream for the trailing updates beyond the next block) computes every
    entry with the same operations in the same order as one stream: bitwise equal factors, pivots and
    workspace. With split-bf16 products the look-ahead splits an update between the 256 x 128 and the 128 x 64
    tile kernels, which form every entry with the same products in the same order."""
    from beat_hip import _lib
    from beat_hip.engine import PRECISIONS, _ptr

    torch = gpu_device.torch
    n, batch = 1100, 8  # the first pair's update spans 844 rows x 716 columns beyond the next block
    rng = np.random.default_rng(11)
    lda = ((n + 15) // 16) * 16
    M = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))
    cdt = torch.complex64 if precision == "mixed" else torch.complex128
    prec = PRECISIONS[precision]
    lib = gpu_device.lib
    outs = []
    mode = {"split_bf16": _lib.BEM_LU_SPLIT_BF16, "bf16": _lib.BEM_LU_BF16}.get(products, _lib.BEM_LU_FP32)
    try:
        _lib.check(lib.bem_ctx_set_lu_products(gpu_device.ctx, mode), "products")
        for fanout in (1, 4):
            _lib.check(lib.bem_ctx_set_fanout(gpu_device.ctx, fanout), "fanout")
            with torch.cuda.stream(gpu_device.stream):
                A = torch.zeros((batch, n, lda), dtype=cdt, device=gpu_device.device)
                A[:, :, :n] = torch.from_numpy(M).to(gpu_device.device).to(cdt)
                ipiv = torch.zeros((batch, n), dtype=torch.int32, device=gpu_device.device)
                perm = torch.zeros((batch, n), dtype=torch.int32, device=gpu_device.device)
                info = torch.zeros(batch, dtype=torch.int32, device=gpu_device.device)
                work = torch.zeros(_lib.lu_workspace_bytes(prec, n, batch), dtype=torch.uint8,
                                   device=gpu_device.device)
            gpu_device.sync()
            _lib.check(lib.bem_lu_factor(gpu_device.ctx, prec, n, batch, _ptr(A), lda, n * lda, _ptr(ipiv),
                                         _ptr(perm), n, _ptr(info), _ptr(work)), "factor")
            gpu_device.sync()
            outs.append((A.cpu().numpy(), ipiv.cpu().numpy(), perm.cpu().numpy(), work.cpu().numpy()))
    finally:
        _lib.check(lib.bem_ctx_set_fanout(gpu_device.ctx, 4), "fanout")
        _lib.check(lib.bem_ctx_set_lu_products(gpu_device.ctx, _lib.BEM_LU_FP32), "products")
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def _corrupt_slots(s, what):
    s.slot_start = s.slot_start.copy()
    s.slot_inc = s.slot_inc.copy()
    s.slot_pend = s.slot_pend.copy()
    if what == "start":  # an incidence range running backwards
        s.slot_start[0, 1], s.slot_start[0, 2] = s.slot_start[0, 2], s.slot_start[0, 1]
    elif what == "lane":  # an incidence code naming lane 64
        s.slot_inc[0, 0] = 64 << 2
    elif what == "corner":  # corner 3 of a triangle
        s.slot_inc[0, 0] = int(s.slot_inc[0, 0]) | 3
    elif what == "pend":  # an LDS partial-sum slot past the 96 the kernels hold
        s.slot_pend[0, 0] = 96


@pytest.mark.parametrize("what,message", [("start", "incidence range"), ("lane", "incidence code"),
                                          ("corner", "incidence code"), ("pend", "partial-sum slot")])
def test_mesh_upload_rejects_bad_slot_tables(gpu_device, sphere, what, message):
    """The slot tables index LDS arrays inside the assembly kernels; mesh upload refuses tables that would send
    an index past them (the error names the table) instead of launching on them."""
    import copy
    import dataclasses

    from beat_hip import _lib
    from beat_hip.engine import DeviceMesh

    slots = copy.copy(sphere.slots)
    _corrupt_slots(slots, what)
    prep = dataclasses.replace(sphere, slots=slots)
    with pytest.raises(_lib.BemError, match=message):
        DeviceMesh(gpu_device, prep, np.arange(prep.n_faces, dtype=np.int32))


def test_golden_fixtures(gpu_device):
    """Committed oracle fixtures: sphere (1,280 triangles) pressure and field; two disjoint tetrahedra (ragged
    chunk of 8 elements, two unwelded parts) operators recombined into the Burton-Miller LHS."""
    from pathlib import Path

    from beat_hip import geometry as G
    from beat_hip.engine import Engine

    gold = Path(__file__).resolve().parent / "golden"
    g = np.load(gold / "sphere1280_1khz.npz")
    prep = G.prepare(G.BoundaryMesh(g["vertices"], g["faces"], np.full(len(g["faces"]), 2)))
    eng = Engine(prep, np.ones((prep.n_faces, 1)), dev=gpu_device)
    res, _ = eng.solve([FREQ], precision="fp64", obs=g["points"], rho=RHO, sound_speed=C0)
    assert _rel(res[0].pressure[0, :, 0], g["pressure"]) < 1e-10
    assert _rel(res[0].field[0, :, 0], g["field"]) < 1e-10
    eng.close()

    t = np.load(gold / "tetra_pair_ops.npz")
    prep = G.prepare(G.BoundaryMesh(t["vertices"], t["faces"], np.ones(len(t["faces"]))))
    from oracle import bem_oracle as O

    om = O.Mesh(t["vertices"], t["faces"], np.ones(8, int), np.ones(8, int))
    M11, M10 = O.identity_matrices(om)
    k = float(t["k"])
    lhs = 0.5 * M11 - t["D"] + (1j / k) * t["H"]
    q = np.zeros((8, 1))
    q[[0, 5], 0] = 1.0
    rhs = (-t["S"] - (1j / k) * (t["Ds"] + 0.5 * M10)) @ q[:, 0]
    eng = Engine(prep, q, dev=gpu_device)
    for precision, tol in (("fp64", 1e-12), ("mixed", 2e-6)):
        b = eng.batch(precision, 1, None, 1.0, k)  # sound speed = k -> omega = 2 pi f, k(f = 1/(2 pi)) = 1
        b.set_frequencies([k * k / (2 * np.pi)])  # k_device = 2 pi f / c = k
        with gpu_device.torch.cuda.stream(gpu_device.stream):
            b._assemble()
        gpu_device.sync()
        A = b.A0[0, :, :8].cpu().numpy().astype(np.complex128)
        bb = b.b[0, 0, :8].cpu().numpy() / (1j * 1.0 * 2 * np.pi * k * k / (2 * np.pi))
        vp = prep.vperm
        assert _rel(A, lhs[np.ix_(vp, vp)]) < tol, (precision, _rel(A, lhs[np.ix_(vp, vp)]))
        assert _rel(bb, rhs[vp]) < tol, (precision, _rel(bb, rhs[vp]))
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_no_radiating_faces_gives_zero_field(gpu_device, sphere, precision):
    """Edge case: no face carries a drive -> b = 0, x = 0, finite diagnostics, no NaN."""
    from beat_hip.engine import Engine

    eng = Engine(sphere, np.zeros((sphere.n_faces, 1)), dev=gpu_device)
    pts = np.array([[0.0, 0.0, 2.0], [1.0, 1.0, 1.0]])
    res, _ = eng.solve([500.0, 900.0], precision=precision, obs=pts, rho=RHO, sound_speed=C0)
    r = res[0]
    assert np.all(r.pressure == 0) and np.all(r.field == 0)
    assert np.all(np.isfinite(r.residual)) and np.all(r.info == 0)
    eng.close()


def test_ragged_multi_batch_and_graph_replay(gpu_device):
    """Horn-in-box with 1,528 triangles (ragged last chunk), 5 frequencies in batches of 2 (padded last
    batch), eager vs captured-graph replays: identical results on every replay."""
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from beat_hip.engine import Engine

    V, F, T = syn.horn_in_box(h=0.05)
    prep = G.prepare(G.BoundaryMesh(V, F, T))
    assert prep.n_faces % 64 != 0
    q = (T == 2).astype(float)[:, None]
    eng = Engine(prep, q, dev=gpu_device)
    pts = np.array([[0.0, 0.0, 2.0], [0.5, 0.0, 1.8]])
    freqs = [150.0, 700.0, 1500.0, 3000.0, 4500.0]
    eager, _ = eng.solve(freqs, precision="mixed", batch=2, obs=pts, rho=RHO, sound_speed=C0)
    b = eng.batch("mixed", 2, pts, RHO, C0)
    b.capture()
    for rep in range(3):
        got = [b.run(freqs[i : i + 2]) for i in range(0, 5, 2)]
        for e, g in zip(eager, got):
            assert np.array_equal(e.field, g.field), rep
            assert np.array_equal(e.pressure, g.pressure), rep
        gpu_device.sync()
    b.close()
    # every frequency also agrees with the FP64 path to 0.05 dB
    ref, _ = eng.solve(freqs, precision="fp64", batch=5, obs=pts, rho=RHO, sound_speed=C0)
    fm = np.concatenate([e.field for e in eager])
    fr = ref[0].field
    assert np.abs(20 * np.log10(np.abs(fm) / np.abs(fr))).max() < 0.05
    eng.close()


@pytest.fixture(scope="module")
def sym_models():
    """Half (x) and quarter (xy) of a mirror-symmetric latitude/longitude sphere, with oracle solutions."""
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from oracle import bem_oracle as O

    V, F, T = syn.uv_sphere(10, 20, 0.1)
    pts = np.array([[0.3, 0.2, 1.5], [1.0, 0.5, -0.3], [0.2, 1.1, 0.4], [-0.7, -0.4, 0.9]])
    out = {}
    for mode in ("x", "xy"):
        v, f, t = syn.symmetric_part(V, F, T, mode)
        prep = G.prepare(G.BoundaryMesh(v, f, t), symmetry=mode)
        m = prep.original
        om = O.Mesh(m.vertices, m.faces, m.tags, m.mesh_ids)
        ref = O.solve_frequency_symmetric(om, FREQ, np.ones((len(f), 1)), mode, C0, RHO, points=pts)
        out[mode] = (prep, ref, pts)
    return out


@pytest.mark.parametrize("mode", ["x", "xy"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_symmetric_assembly_matches_oracle(gpu_device, sym_models, mode, precision, tol):
    """Images of every trial element, image-singular Duffy pairs and orbit row weights: A and b of the
    fundamental domain against the oracle's symmetric operators."""
    prep, ref, _ = sym_models[mode]
    eng = _engine(gpu_device, prep)
    b = eng.batch(precision, 1, None, RHO, C0)
    b.set_frequencies([FREQ])
    with gpu_device.torch.cuda.stream(gpu_device.stream):
        b._assemble()
    gpu_device.sync()
    N = prep.n_vertices
    A = b.A0[0, :, :N].cpu().numpy().astype(np.complex128)
    rhs = b.b[0, 0, :N].cpu().numpy()
    vp = prep.vperm
    print(mode, precision, "A rel", _rel(A, ref["lhs"][np.ix_(vp, vp)]), "b rel", _rel(rhs, ref["rhs"][vp, 0]))
    assert _rel(A, ref["lhs"][np.ix_(vp, vp)]) < tol
    assert _rel(rhs, ref["rhs"][vp, 0]) < tol
    eng.close()


@pytest.mark.parametrize("mode", ["x", "xy"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_symmetric_solve_and_field_match_oracle(gpu_device, sym_models, mode, precision):
    prep, ref, pts = sym_models[mode]
    eng = _engine(gpu_device, prep)
    res, _ = eng.solve([FREQ], precision=precision, obs=pts, rho=RHO, sound_speed=C0)
    p, u = res[0].pressure[0, :, 0], res[0].field[0, :, 0]
    po, uo = ref["pressure"][:, 0], ref["field"][:, 0]
    db = np.abs(20 * np.log10(np.abs(u) / np.abs(uo)))
    print(mode, precision, "p rel", _rel(p, po), "field rel", _rel(u, uo), "max dB", db.max())
    if precision == "fp64":
        assert _rel(p, po) < 1e-10 and _rel(u, uo) < 1e-10
    else:
        assert _rel(p, po) < 1e-4 and db.max() < 0.05
    # the reduced model also meets the pulsating-sphere closed form
    from oracle.bem_oracle import pulsating_sphere_pressure

    k = 2 * math.pi * FREQ / C0
    pa = pulsating_sphere_pressure(np.linalg.norm(pts, axis=1), A_SPHERE, k, RHO, 2 * math.pi * FREQ)
    assert np.abs(20 * np.log10(np.abs(u) / np.abs(pa))).max() < 0.2
    eng.close()


@pytest.mark.parametrize("precision,tol", [("float64", 1e-9), ("float32", 1e-4)])
def test_system_exterior_request_matches_oracle(gpu_device, tmp_path, precision, tol):
    """Compiled-system exterior request (one BEM surface, two excitation ports on disjoint moving boundaries)
    through the coupled-worker wire: boundary pressure, field and impedance against the FP64 oracle."""
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from beat_hip import system as SY
    from beat_hip.engine import Engine
    from oracle import bem_oracle as O

    V, F, _ = syn.icosphere(3, A_SPHERE)
    T = np.where(V[F].mean(axis=1)[:, 2] > 0, 4, 9)
    path = tmp_path / "s.msh"
    syn.write_gmsh22(path, V, F, T, scale=1.0)
    pts = [[0.0, 0.0, 2.0], [1.5, 0.0, 1.0], [0.0, -2.0, 0.3]]
    req = {
        "schema_version": 1,
        "compiled_system": {
            "meshes": [{"id": "m", "file": str(path), "purpose": "bem_surface", "scale_to_m": 1.0,
                        "translation_m": [0.0, 0.0, 0.0]}],
            "regions": [{"id": "air", "kind": "unbounded_air", "mesh_ids": ["m"], "sound_speed_m_per_s": C0,
                         "density_kg_per_m3": RHO}],
            "boundaries": [{"id": "top", "region_id": "air", "kind": "moving", "group": {"mesh_id": "m", "tag": 4}},
                           {"id": "bot", "region_id": "air", "kind": "moving",
                            "group": {"mesh_id": "m", "tag": 9}}],
            "components": [{"id": "c1", "kind": "ideal_velocity_source", "boundary_ids": ["top"],
                            "parameters": {"boundary_motion_weights": {"top": 0.5}}},
                           {"id": "c2", "kind": "ideal_velocity_source", "boundary_ids": ["bot"]}],
            "excitation_ports": [{"id": "e1", "component_id": "c1", "kind": "normal_velocity"},
                                 {"id": "e2", "component_id": "c2", "kind": "normal_velocity"}],
        },
        "frequencies_hz": [FREQ],
        "excitation_port_ids": ["e1", "e2"],
        "outputs": [{"id": "far", "quantity": "exterior_pressure", "options": {"points_m": pts}},
                    {"id": "p", "quantity": "bem_boundary_pressure"},
                    {"id": "z", "quantity": "radiation_impedance"}],
        "solver_options": {"precision": precision, "bem_backend": "rocm"},
    }
    events = []
    SY.solve_system_request(req, lambda prep, q: Engine(prep, q, dev=gpu_device), events.append)
    assert len(events) == 1
    qs = {q["id"]: q for q in events[0]["result"]["quantities"]}

    def arr(w):
        return (np.asarray(w["real"]) + 1j * np.asarray(w["imag"])).reshape(w["shape"])

    job = SY.parse_exterior_request(req)
    m = job["mesh"]
    drives = SY.unit_drives(m, job["excitations"])
    freq = float(np.float32(FREQ)) if precision == "float32" else FREQ
    c, rho = job["sound_speed"], job["density"]
    ref = O.solve_frequency(O.Mesh(m.vertices, m.faces, m.tags, m.mesh_ids), freq, drives, c, rho,
                            points=np.asarray(pts))
    p = arr(qs["p"]["values"])
    assert p.shape == (2, m.n_vertices)
    assert _rel(p.T, ref["pressure"]) < tol
    assert _rel(arr(qs["far"]["values"]).T, ref["field"]) < tol
    z = arr(qs["z"]["values"])
    for e, ex in enumerate(job["excitations"]):
        assert abs(z[e] - SY.component_impedance(m, ref["pressure"][:, e], ex, 1)) < tol * abs(z[e]) * 10
