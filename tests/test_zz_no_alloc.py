"""pwk_dev_dense_eigh_noalloc (the eigensolver with its static per-GPU workspace, for callers
that cannot allocate) against pwk_dev_dense_eigh (workspace allocated by the context), through
the C ABI on device buffers: two calls in a row, the second with a smaller batch, each equal to
the allocating entry point on the same matrices (same kernels: bit for bit), and the eigenvalues
within 1e-10 of LAPACK. A batch above the static workspace is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hermitian_batch(rng, n, nk):
    a = rng.standard_normal((nk, n, n)) + 1j * rng.standard_normal((nk, n, n))
    return 0.5 * (a + np.conj(np.transpose(a, (0, 2, 1))))


def _run(torch, lib, h, n, nb, nk, fn):
    # column-major n×n per matrix: the transpose of the row-major numpy array
    hc = np.ascontiguousarray(np.transpose(h, (0, 2, 1)))
    dh = torch.from_numpy(hc.reshape(-1).view(np.float64).copy()).to("cuda:0")
    dw = torch.zeros(nk * nb, dtype=torch.float64, device="cuda:0")
    dv = torch.zeros(nk * n * nb * 2, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    st = getattr(lib, fn)(_ctx_handle[0], n, nb, nk, C.c_void_p(dh.data_ptr()),
                          C.c_void_p(dw.data_ptr()), C.c_void_p(dv.data_ptr()))
    return st, dw, dv


_ctx_handle = [None]


def test_noalloc_eigh_equals_allocating_entry_point(gpu, pwk_mod):
    import torch

    lib = pwk_mod.lib()
    _ctx_handle[0] = gpu.h
    rng = np.random.default_rng(20261203)
    for n, nb, nk in ((200, 12, 6), (60, 5, 3)):  # second call smaller (and the one-wave solver)
        print(f"noalloc eigh n={n} nb={nb} nk={nk}", flush=True)
        h = _hermitian_batch(rng, n, nk)
        st_a, w_a, v_a = _run(torch, lib, h, n, nb, nk, "pwk_dev_dense_eigh")
        assert st_a == 0, lib.pwk_last_error()
        gpu.sync()
        st_s, w_s, v_s = _run(torch, lib, h, n, nb, nk, "pwk_dev_dense_eigh_noalloc")
        assert st_s == 0, lib.pwk_last_error()
        gpu.sync()
        torch.cuda.synchronize()
        assert torch.equal(w_a, w_s)
        assert torch.equal(v_a, v_s)
        ref = np.stack([np.linalg.eigvalsh(m)[:nb] for m in h])
        assert np.abs(w_s.cpu().numpy().reshape(nk, nb) - ref).max() < 1e-10
        # eigenvectors: H v = λ v for every returned pair
        v = v_s.cpu().numpy().view(np.complex128).reshape(nk, nb, n)
        for k in range(nk):
            r = h[k] @ v[k].T - v[k].T * ref[k][None, :]
            assert np.abs(r).max() < 1e-9


def test_noalloc_eigh_refuses_batch_above_static_workspace(gpu, pwk_mod):
    lib = pwk_mod.lib()
    # 216 matrices of 1189 rows, 24 pairs: 216·(4·1189 + 6·1189·24) doubles, far above 2^21;
    # refused on the host before anything is launched (the pointers are never touched)
    dummy = C.c_void_p(16)
    st = lib.pwk_dev_dense_eigh_noalloc(gpu.h, 1189, 24, 216, dummy, dummy, dummy)
    assert st == 1  # PWK_E_INVALID
    assert b"static workspace" in lib.pwk_last_error()
