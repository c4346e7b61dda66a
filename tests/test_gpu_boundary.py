"""The k × band batch entry points and the warm-start Rayleigh-Ritz, through the C ABI.

  * pwk_hpsi_batch: ragged batches (a different band count per k-point entry) equal the
    per-k-point pwk_hpsi on every band (same kernels, same arithmetic: bit for bit on the fused
    small-grid path and the stick path) and the oracle's dense H·ψ to 1e-12 (FP64).
  * pwk_density_accumulate_batch: equals pwk_density_accumulate over the same packed bands and
    the oracle's density to 1e-12.
  * pwk_subspace_rr (diagonalize_subspace, src/eigensolver/dense.rs:183-260): with the exact
    eigenvectors as the start block the gate passes and the Ritz values equal LAPACK's to
    1e-10 and the rotated block spans the same space (projector check); with a perturbed start
    block the residual gate (1e-6, dense.rs:37) sends that k-point to the full solve, whose
    result is bit for bit pwk_dense_eigh's (dense.rs:184-186: "bit-identical to what
    diagonalize_lowest would have returned"); without a start block every k-point does.
"""
import numpy as np
import pytest

from conftest import SI_UPF

pytestmark = pytest.mark.gpu

A_SI = 5.431
SI_FRAC = [[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]
KS = [np.array([0.11, -0.23, 0.05]), np.zeros(3), np.array([-0.3, 0.2, 0.1])]


def si_lat():
    h = A_SI / 2.0
    return np.array([[0.0, h, h], [h, 0.0, h], [h, h, 0.0]])


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _setup(gpu, oracle, ecut, dims=None):
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], ecut, dims=dims)
    gpu.prepare_buffers(s.dims, s.g2())
    gpu.set_basis(s.miller())
    gpu.set_nk(len(KS))
    rng = np.random.default_rng(17)
    veff = np.fft.fftn(rng.uniform(-1, 1, s.dims) * 13.6).reshape(-1) / s.n_grid
    gpu.set_veff(veff)
    for ik, k in enumerate(KS):
        gpu.set_kinetic(ik, s.kinetic(k))
        B, DBH, _ = s.nonlocal_(k)
        gpu.set_nonlocal(ik, B, DBH)
    return s, veff


@pytest.mark.parametrize("fused", [True, False])
def test_hpsi_batch_ragged_equals_per_k(gpu, oracle, fused):
    # 150 eV: 18³ grid (fused small-grid kernel with a band table); fused off: stick path per entry
    s, veff = _setup(gpu, oracle, 150.0)
    iks, nbs = [2, 0, 1, 0], [3, 5, 1, 2]
    rng = np.random.default_rng(3)
    tot = sum(nbs)
    psi = rng.standard_normal((tot, s.npw)) + 1j * rng.standard_normal((tot, s.npw))
    print("hpsi batch", fused, s.dims, s.npw, flush=True)
    try:
        gpu.set_fused_hpsi(1 if fused else 0)
        for prec, tol in (("fp64", 1e-12), ("fp32", 2e-5)):
            out = gpu.hpsi_batch(iks, nbs, psi, prec)
            off = 0
            for ik, nb in zip(iks, nbs):
                ref_k = gpu.hpsi(ik, psi[off:off + nb], prec)
                assert np.array_equal(out[off:off + nb], ref_k), (prec, ik)
                H = s.dense_h(KS[ik], veff, with_nl=True)
                assert rel(out[off:off + nb], (H @ psi[off:off + nb].T).T) < tol, (prec, ik)
                off += nb
    finally:
        gpu.set_fused_hpsi(1)


def test_hpsi_batch_rejects_bad_entries(gpu, oracle, pwk_mod):
    s, _ = _setup(gpu, oracle, 150.0)
    psi = np.ones((2, s.npw), complex)
    with pytest.raises(pwk_mod.PwkError) as e:
        gpu.hpsi_batch([0, 5], [1, 1], psi)  # k-point slot out of range
    assert e.value.code == 1
    with pytest.raises(pwk_mod.PwkError) as e:
        gpu.hpsi_batch([0, 1], [2, 0], psi)  # empty entry
    assert e.value.code == 1


def test_density_batch_equals_packed_density_and_oracle(gpu, oracle):
    s, _ = _setup(gpu, oracle, 150.0)
    iks, nbs = [1, 0, 2], [4, 2, 3]
    rng = np.random.default_rng(8)
    tot = sum(nbs)
    psi = rng.standard_normal((tot, s.npw)) + 1j * rng.standard_normal((tot, s.npw))
    fw = rng.uniform(0.0, 0.3, tot)
    fw[4] = 0.0  # skipped band (f·w < 1e-15, src/scf/density.rs:14,72)
    rho_b = gpu.density_batch(iks, nbs, psi, fw)
    rho_p = gpu.density(psi, fw)
    assert np.array_equal(rho_b, rho_p)
    ref = np.zeros(s.n_grid)
    for b in range(tot):
        if fw[b] < 1e-15:
            continue
        g = np.zeros(s.n_grid, complex)
        g[s.g_to_fft()] = psi[b]
        r = np.fft.ifftn(g.reshape(s.dims)).reshape(-1) * s.n_grid
        ref += fw[b] * np.abs(r) ** 2
    assert rel(rho_b, ref) < 1e-12


def test_subspace_rr_gate_and_fallback(gpu, oracle):
    s, veff = _setup(gpu, oracle, 150.0)
    nb = 8
    hs = np.stack([s.dense_h(k, veff, True) for k in KS[:2]])
    exact = []
    for hm in hs:
        wl, vl = np.linalg.eigh(hm)
        exact.append(vl[:, :nb])
    exact = np.stack(exact)
    # k0: the exact lowest eigenvectors (gate passes); k1: perturbed and re-orthonormalised
    rng = np.random.default_rng(4)
    pert = exact[1] + 1e-2 * (rng.standard_normal(exact[1].shape) + 1j * rng.standard_normal(exact[1].shape))
    pert, _ = np.linalg.qr(pert)
    vprev = np.stack([exact[0], pert])
    w, v, used, res = gpu.subspace_rr(hs, vprev, nb)
    print("subspace rr used_full", used, "max residual", res, flush=True)
    assert list(used) == [0, 1]
    assert res[0] < 1e-6 < res[1]
    wf, vf = gpu.dense_eigh(hs, nb)
    # the gated k-point is the full solve, bit for bit
    assert np.array_equal(w[1], wf[1]) and np.array_equal(v[1], vf[1])
    # the warm-start k-point: Ritz values of span(V_prev) (numpy restatement of the projection)
    hsub = exact[0].conj().T @ hs[0] @ exact[0]
    lam = np.linalg.eigvalsh(0.5 * (hsub + hsub.conj().T))
    assert np.abs(w[0] - lam).max() < 1e-10
    assert np.abs(w[0] - np.linalg.eigvalsh(hs[0])[:nb]).max() < 1e-9
    # same subspace: the projectors onto the rotated and the start block agree
    P0 = v[0] @ v[0].conj().T
    Pe = exact[0] @ exact[0].conj().T
    assert np.abs(P0 - Pe).max() < 1e-10
    assert np.abs(hs[0] @ v[0] - v[0] * w[0]).max() < 1e-6
    # no start block: every k-point takes the full solve
    w2, v2, used2, res2 = gpu.subspace_rr(hs, None, nb)
    assert list(used2) == [1, 1] and np.all(res2 == 0.0)
    assert np.array_equal(w2, wf) and np.array_equal(v2, vf)
