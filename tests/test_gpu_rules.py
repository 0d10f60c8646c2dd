"""GPU parity of the lower-order regular rules (three-point order 2, centroid order 1) with lower-order Duffy
rules against the FP64 oracle under the same rules: assembled A and b, boundary pressure and far field on the
1,280-triangle sphere. The coupled FEM-BEM path integrates with order 2 / 2 by default, so these are its BEM
operators. Tolerances as for the order-4 path: A, b within 1e-11 (FP64) / 1e-5 (mixed) relative."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FREQ = 1700.0
RHO = 1.205
C0 = 343.0


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def models():
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from oracle import bem_oracle as O

    V, F, T = syn.icosphere(3, 0.1)
    out = {}
    pts = O.fibonacci_observation({"spherical_sampling_enabled": True, "spherical_sampling_points": 200,
                                   "distance": 2.0})[0].astype(np.float64)
    rng = np.random.default_rng(5)
    qu = np.stack([np.ones(len(F)), rng.uniform(-1.0, 1.0, len(F))], axis=1)
    for ro, so in ((2, 2), (1, 1), (2, 4)):
        prep = G.prepare(G.BoundaryMesh(V, F, T), regular_order=ro, singular_order=so)
        m = prep.original
        ref = O.solve_frequency(O.Mesh(m.vertices, m.faces, m.tags, m.mesh_ids), FREQ, qu, C0, RHO, points=pts,
                                regular_order=ro, singular_order=so)
        out[(ro, so)] = (prep, ref, qu, pts)
    return out


@pytest.mark.parametrize("orders", [(2, 2), (1, 1), (2, 4)], ids=lambda o: f"regular{o[0]}-singular{o[1]}")
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_lower_order_rules_match_oracle(gpu_device, models, orders, precision, tol):
    from beat_hip.engine import Engine

    prep, ref, qu, pts = models[orders]
    print(f"rules {orders} {precision}", flush=True)
    eng = Engine(prep, qu, dev=gpu_device)  # drives in the caller's face order; the engine permutes them
    try:
        b = eng.batch(precision, 1, pts, RHO, C0)
        b.set_frequencies([FREQ])
        with gpu_device.torch.cuda.stream(gpu_device.stream):
            b._assemble()
        gpu_device.sync()
        N = prep.n_vertices
        A = b.A0[0, :, :N].cpu().numpy().astype(np.complex128)
        rhs = b.b[0, :2, :N].cpu().numpy().T
        vp = prep.vperm
        ea, eb = _rel(A, ref["lhs"][np.ix_(vp, vp)]), _rel(rhs, ref["rhs"][vp])
        r = b.run([FREQ])
        ep = _rel(r.pressure[0], ref["pressure"])
        ef = _rel(r.field[0], ref["field"])
        b.close()
    finally:
        eng.close()
    print(f"  A {ea:.2e} b {eb:.2e} pressure {ep:.2e} field {ef:.2e}")
    assert ea < tol and eb < tol
    assert ep < (1e-9 if precision == "fp64" else 1e-5)
    assert ef < (1e-9 if precision == "fp64" else 1e-5)
