"""Every launch shape of the tuner's committed table works at its own size: LU entries factor and solve a batch of
random well-conditioned matrices at the entry's N and batch (no breakdown; FP64 relative residual after refinement
<= 1e-10 for mixed, after the solve <= 1e-12 for FP64), assembly entries reproduce the default shapes' results
(FP64 field within 1e-12 relative, mixed within 1e-6: only the summation order of field slices changes)."""

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _entries(kind):
    from beat_hip import tuning as TU

    if not TU.TABLE_PATH.exists():
        return []
    return json.loads(TU.TABLE_PATH.read_text())[kind]


@pytest.mark.parametrize("entry", _entries("lu"), ids=lambda e: f"{e['precision']}-N{e['n']}-B{e['batch']}")
def test_tuned_lu_shape(gpu_device, entry):
    from beat_hip import autotune as AT
    from beat_hip import tuning as TU

    print(f"tuned LU {entry['precision']} N={entry['n']} batch={entry['batch']}: {entry['choice']}", flush=True)
    run = AT.lu_workload(gpu_device, entry["precision"], entry["n"], entry["batch"], seed=11)
    try:
        AT._set(gpu_device, TU.LuChoice(**entry["choice"]))
        run()
        gpu_device.sync()
        info, rel = run.check()
    finally:
        AT._set(gpu_device, TU.LuChoice())
        del run
        gpu_device.torch.cuda.empty_cache()
    print(f"  info {info} residual {rel:.2e}")
    assert info == 0
    assert rel <= (1e-10 if entry["precision"] == "mixed" else 1e-12)


@pytest.mark.parametrize("entry", _entries("assembly"), ids=lambda e: f"{e['precision']}-F{e['faces']}")
def test_tuned_assembly_shape(gpu_device, entry):
    from beat_hip import autotune as AT
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn
    from beat_hip import tuning as TU
    from beat_hip.engine import Engine

    V, F, T = syn.icosphere(3, 0.1)
    prep = G.prepare(G.BoundaryMesh(V, F, T))
    obs = np.array([[0.0, 0.0, 2.0], [1.2, 0.4, -0.9], [-0.5, 1.5, 0.6]])
    eng = Engine(prep, np.ones((len(F), 1)), dev=gpu_device)
    freqs = [300.0, 1200.0, 2500.0, 5000.0]
    try:
        out = []
        for choice in (TU.AssemblyChoice(), TU.AssemblyChoice(**entry["choice"])):
            b = eng.batch(entry["precision"], len(freqs), obs, 1.205, 343.0, strip_nf=choice.strip_nf)
            b.asm_choice = choice
            out.append(b.run(freqs))
            b.close()
    finally:
        AT._set(gpu_device, TU.LuChoice())
        eng.close()
    tol = 1e-12 if entry["precision"] == "fp64" else 1e-6
    d = np.linalg.norm(out[0].field - out[1].field) / np.linalg.norm(out[0].field)
    print(f"tuned assembly {entry['choice']}: field relative difference {d:.2e}")
    assert d <= tol
