"""bem_refine_noalloc (static per-device scratch, no allocation) against bem_refine (caller workspace): identical
solutions, residuals and iteration counts on two calls in a row, the second smaller; oversized calls rejected."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(torch, dev, n, batch, nrhs, seed):
    rng = np.random.default_rng(seed)
    lda = (n + 15) // 16 * 16
    A = np.zeros((batch, n, lda), np.complex64)
    A[:, :, :n] = (rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n))) / np.sqrt(n)
    A[:, np.arange(n), np.arange(n)] += 3.0
    b = np.zeros((batch, nrhs, lda), np.complex128)
    b[:, :, :n] = rng.standard_normal((batch, nrhs, n)) + 1j * rng.standard_normal((batch, nrhs, n))
    with torch.cuda.stream(dev.stream):
        out = lda, *(torch.from_numpy(a).to(dev.device) for a in (A, A.copy(), b))
    dev.sync()
    return out


def _factor(dev, lib, _lib, n, batch, lda, A):
    # every buffer is allocated on the library's stream and outlives the factorization that writes it
    torch = dev.torch
    with torch.cuda.stream(dev.stream):
        ipiv = torch.zeros((batch, n), dtype=torch.int32, device=dev.device)
        perm = torch.zeros((batch, n), dtype=torch.int32, device=dev.device)
        info = torch.zeros(batch, dtype=torch.int32, device=dev.device)
        w = torch.empty(_lib.lu_workspace_bytes(_lib.BEM_PREC_MIXED, n, batch) // 8, dtype=torch.complex64,
                        device=dev.device)
    p = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
    _lib.check(lib.bem_lu_factor(dev.ctx, _lib.BEM_PREC_MIXED, n, batch, p(A), lda, n * lda, p(ipiv), p(perm), n,
                                 p(info), p(w)), "bem_lu_factor")
    dev.sync()
    return ipiv, perm, info, w


def test_noalloc_refine_matches_refine(gpu_device):
    from beat_hip import _lib

    dev, torch = gpu_device, gpu_device.torch
    lib = dev.lib
    p = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
    for n, batch, nrhs, seed in ((700, 6, 3, 1), (257, 3, 2, 2)):  # second call smaller
        print(f"no-alloc refine: n={n} batch={batch} nrhs={nrhs}", flush=True)
        lda, A, A0, b = _system(torch, dev, n, batch, nrhs, seed)
        ipiv, perm, info, w = _factor(dev, lib, _lib, n, batch, lda, A)
        sB = nrhs * lda
        outs = []
        keep = []
        for variant in ("workspace", "noalloc"):
            with torch.cuda.stream(dev.stream):
                x = torch.zeros_like(b)
                rel = torch.zeros(batch, dtype=torch.float64, device=dev.device)
                it = torch.zeros(batch, dtype=torch.int32, device=dev.device)
                r = torch.zeros_like(b)
                y = torch.zeros(b.shape, dtype=torch.complex64, device=dev.device)
                done = torch.zeros(batch, dtype=torch.int32, device=dev.device)
            dev.sync()
            keep.append((x, rel, it, r, y, done))
            if variant == "workspace":
                rc = lib.bem_refine(dev.ctx, n, batch, p(A), lda, n * lda, p(perm), n, p(w), p(A0), p(b), lda, sB,
                                    p(x), p(r), p(y), nrhs, 4, 1e-10, p(rel), p(it), p(done))
            else:
                rc = lib.bem_refine_noalloc(dev.ctx, n, batch, p(A), lda, n * lda, p(perm), n, p(w), p(A0), p(b), lda,
                                            sB, p(x), nrhs, 4, 1e-10, p(rel), p(it))
            _lib.check(rc, variant)
            dev.sync()
            outs.append((x.cpu().numpy(), rel.cpu().numpy(), it.cpu().numpy()))
        (xa, ra, ia), (xb, rb, ib) = outs
        assert (info.cpu().numpy() == 0).all()
        assert np.array_equal(xa, xb) and np.array_equal(ra, rb) and np.array_equal(ia, ib)
        assert ra.max() <= 1e-10
        # and the solution solves the system
        An = A0.cpu().numpy()[:, :, :n].astype(np.complex128)
        for m in range(batch):
            for c in range(nrhs):
                res = An[m] @ xb[m, c, :n] - b.cpu().numpy()[m, c, :n]
                assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(b.cpu().numpy()[m, c, :n])


def test_noalloc_refine_rejects_oversized_calls(gpu_device):
    # batch * strideB * 24 bytes above the 64 MiB scratch: refused before anything is enqueued
    from beat_hip import _lib

    dev = gpu_device
    big_batch = _lib.BEM_NOALLOC_SCRATCH_BYTES // (24 * 4 * 4096) + 1
    rc = dev.lib.bem_refine_noalloc(dev.ctx, 4096, big_batch, C.c_void_p(16), 4096, 4096 * 4096, C.c_void_p(16),
                                    4096, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 4096, 4 * 4096,
                                    C.c_void_p(16), 4, 4, 1e-10, C.c_void_p(16), C.c_void_p(16))
    assert rc == _lib.BEM_E_ARG
    assert "scratch" in _lib.last_error(dev.lib)
