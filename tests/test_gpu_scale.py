"""GPU checks at the 20k-triangle scale (two-way loudspeaker box, three velocity sources, 5,000 field points).

The oracle does not finish this size in seconds, so the mixed path is held to size-independent properties:
the plain FP64 GPU path at the same frequencies (far field within 0.05 dB and 0.5 degrees), the refined FP64
residual, and linearity in the drives (the summed drive gives the sum of the per-source fields).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RHO, C0 = 1.205, 343.0
FREQS = np.array([200.0, 4000.0])


@pytest.fixture(scope="module")
def loudspeaker():
    from beat_hip import driver as D
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn

    V, F, T = syn.loudspeaker_box()
    mesh = G.BoundaryMesh(V, F, T)
    prep = G.prepare(mesh)
    q3 = np.stack([(mesh.tags == tag) for tag in (2, 3, 4)], axis=1).astype(np.complex128)
    obs = D.spherical_observation({"spherical_sampling_enabled": True, "spherical_sampling_points": 5000,
                                   "distance": 2.0})[0].astype(np.float64)
    return prep, q3, obs


def test_scale_mixed_matches_fp64_and_is_linear(gpu_device, loudspeaker):
    from beat_hip.engine import Engine

    prep, q3, obs = loudspeaker
    assert prep.n_faces > 19000 and q3.sum(axis=0).min() > 0
    print("faces", prep.n_faces, "dofs", prep.n_vertices, flush=True)
    eng = Engine(prep, q3, dev=gpu_device)
    mixed, _ = eng.solve(FREQS, precision="mixed", obs=obs, rho=RHO, sound_speed=C0)
    ref, _ = eng.solve(FREQS, precision="fp64", obs=obs, rho=RHO, sound_speed=C0, batch=1)
    eng.close()
    um = mixed[0].field  # (nf, n_obs, 3)
    ur = np.concatenate([r.field for r in ref], axis=0)
    db = np.abs(20 * np.log10(np.abs(um) / np.abs(ur)))
    ph = np.degrees(np.abs(np.angle(um / ur)))
    print("max dB", db.max(), "max deg", ph.max(), "residual", mixed[0].residual, "iters", mixed[0].iters,
          flush=True)
    assert mixed[0].info.max() == 0
    assert mixed[0].residual.max() <= 1e-10
    assert db.max() < 0.05 and ph.max() < 0.5

    # the job driver's batch sizing: a batch that cannot fit the device shrinks; a small one is kept
    eng2 = Engine(prep, q3, dev=gpu_device)
    big = eng2.fit_batch("mixed", 100000, len(obs))
    assert 1 <= big < 100000 and eng2.fit_batch("mixed", 2, len(obs)) == 2
    print("largest fitting mixed batch", big, flush=True)
    eng2.close()

    eng1 = Engine(prep, q3.sum(axis=1, keepdims=True), dev=gpu_device)
    summed, _ = eng1.solve(FREQS, precision="mixed", obs=obs, rho=RHO, sound_speed=C0)
    eng1.close()
    u1 = summed[0].field[..., 0]
    us = um.sum(axis=2)
    rel = np.linalg.norm(u1 - us) / np.linalg.norm(us)
    print("linearity rel", rel, flush=True)
    assert rel < 1e-5
