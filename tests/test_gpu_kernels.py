"""GPU parity of libpwk against the FP64 oracle, through the C ABI.

Tolerances: FP64 paths agree with the oracle to ~1e-12 relative (different
summation order only); FP32 paths to the single-precision bound stated per
test (the reference's own GPU-vs-CPU tolerances are 0.01/0.015 eV on XC,
tests/gpu_consistency.rs:144-193).
"""
import numpy as np
import pytest

from conftest import SI_UPF

pytestmark = pytest.mark.gpu

A_SI = 5.431


def si_lat():
    h = A_SI / 2.0
    return np.array([[0.0, h, h], [h, 0.0, h], [h, h, 0.0]])


SI_FRAC = [[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def si_sys(oracle):
    # Si FCC at 150 eV: reference grid rule gives an 18³ grid
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 150.0)
    return s


@pytest.fixture()
def acc(gpu, si_sys):
    gpu.prepare_buffers(si_sys.dims, si_sys.g2())
    return gpu


def test_grid_dims_of_deck(si_sys):
    print("deck dims", si_sys.dims, "npw", si_sys.npw, flush=True)
    assert si_sys.dims[0] == si_sys.dims[1]


# ---- grid kernels ---------------------------------------------------------------
def test_hartree_veff(acc, si_sys, oracle):
    rng = np.random.default_rng(11)
    n = si_sys.n_grid
    rho_g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    vl, vx = si_sys.vloc(), rng.standard_normal(n) + 1j * rng.standard_normal(n)
    g2 = si_sys.g2()
    fe = 4 * np.pi * 14.399645351950548
    vh_ref = oracle.hartree(rho_g, g2)
    veff_ref = oracle.veff(vl, vh_ref, vx)
    vh = acc.hartree_potential(rho_g, fe, "fp64")
    assert np.array_equal(vh, vh_ref)  # same operations, same order
    assert np.array_equal(acc.v_eff_assembly(vl, vh_ref, vx, "fp64"), veff_ref)
    assert np.array_equal(acc.hartree_veff(rho_g, fe, vl, vx, "fp64"), veff_ref)
    assert vh[0] == 0  # G = 0 removed
    assert rel(acc.hartree_potential(rho_g, fe, "fp32"), vh_ref) < 2e-6
    assert rel(acc.hartree_veff(rho_g, fe, vl, vx, "fp32"), veff_ref) < 2e-6


def test_lda_xc(acc, si_sys, oracle):
    n = si_sys.n_grid
    rho = np.logspace(-4, 1, n)  # the reference's GPU sweep range (gpu_consistency.rs:144-193)
    rho[:5] = [0.0, 1e-31, 1e-20, 0.25, 3.0]
    e_ref, v_ref = oracle.lda_xc(rho)
    e, v = acc.lda_xc(rho, "fp64")
    assert rel(e, e_ref) < 1e-14 and rel(v, v_ref) < 1e-14
    assert e[0] == 0 and v[0] == 0 and e[1] == 0
    e32, v32 = acc.lda_xc(rho, "fp32")
    assert np.abs(e32 - e_ref).max() < 0.01 and np.abs(v32 - v_ref).max() < 0.015
    assert rel(e32, e_ref) < 1e-5


def test_lsda_xc(acc, si_sys, oracle):
    n = si_sys.n_grid
    rng = np.random.default_rng(2)
    rt = np.logspace(-4, 1, n)
    zeta = rng.uniform(-1, 1, n)
    zeta[:3] = [-1.0, 0.0, 1.0]
    ru, rd = rt * (1 + zeta) / 2, rt * (1 - zeta) / 2
    ref = oracle.lsda_xc(ru, rd)
    out = acc.lsda_xc(ru, rd, "fp64")
    for a, b in zip(out, ref):
        assert rel(a, b) < 1e-13
    out32 = acc.lsda_xc(ru, rd, "fp32")
    for a, b in zip(out32, ref):
        assert np.abs(a - b).max() < 0.015


# ---- MFMA complex GEMM (the projector and subspace products) -----------------------
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 2e-5), ("split_bf16", 5e-5)])
def test_zgemm_mfma_shapes(gpu, prec, tol):
    # edge tiles (M, N, K not multiples of 64 / 16 / 16), split-K (long K, few tiles), 144-wide
    # column blocks, accumulation and a strided batch, against numpy complex128
    import torch

    rng = np.random.default_rng(21)
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    for (m, n, k, acc, batch) in [(1, 1, 1, False, 1), (17, 5, 3, True, 1), (64, 144, 16, False, 1),
                                  (65, 145, 300, True, 1), (40, 24, 5000, False, 1), (700, 144, 512, True, 1),
                                  (33, 20, 169, False, 8)]:
        A = rng.standard_normal((batch, m, k)) + 1j * rng.standard_normal((batch, m, k))
        B = rng.standard_normal((batch, k, n)) + 1j * rng.standard_normal((batch, k, n))
        C0 = rng.standard_normal((batch, m, n)) + 1j * rng.standard_normal((batch, m, n))
        ref = A @ B + (C0 if acc else 0)
        # column-major storage of each product = row-major transposes, batches back to back
        dA = torch.from_numpy(np.ascontiguousarray(A.transpose(0, 2, 1))).to(dt).cuda()
        dB = torch.from_numpy(np.ascontiguousarray(B.transpose(0, 2, 1))).to(dt).cuda()
        dC = torch.from_numpy(np.ascontiguousarray(C0.transpose(0, 2, 1))).to(dt).cuda()
        torch.cuda.synchronize()
        gpu.zgemm_dev(m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, dC.data_ptr(), m, precision=prec,
                      accumulate=acc, batch=batch, strides=(m * k, k * n, m * n))
        gpu.sync()
        out = dC.cpu().numpy().astype(np.complex128).transpose(0, 2, 1)
        err = np.abs(out - ref).max() / np.abs(ref).max()
        print("zgemm", prec, (m, n, k, acc, batch), f"{err:.2e}", flush=True)
        assert err < tol, (m, n, k, acc, batch)


# ---- FFT ------------------------------------------------------------------------
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 3e-6)])
def test_fft3d_matches_dft(gpu, prec, tol):
    rng = np.random.default_rng(4)
    for dims in [(15, 15, 15), (16, 18, 20), (24, 25, 27), (60, 60, 60), (72, 75, 80), (81, 81, 64), (90, 96, 100),
                 (125, 128, 135), (256, 18, 20), (20, 243, 16), (1024, 16, 18), (18, 20, 1000), (16, 729, 15)]:
        gpu.prepare_buffers(dims, np.zeros(int(np.prod(dims))))
        x = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
        f = gpu.fft3d(x.reshape(-1), -1, prec).reshape(dims)
        assert rel(f, np.fft.fftn(x)) < tol, dims
        b = gpu.fft3d(x.reshape(-1), +1, prec).reshape(dims)
        assert rel(b, np.fft.ifftn(x) * x.size) < tol, dims


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 3e-6)])
def test_fft_plans_agree_and_autotune_picks_one(gpu, prec, tol):
    # every full-grid FFT plan (radix order x lanes per workgroup x LDS-staged contiguous pass)
    # computes the DFT; plans of one radix order give bit-identical results (they only move
    # data); the autotuned plan is one of them, no slower than the default plan in the same
    # timing run, and the transform it runs equals that plan forced
    rng = np.random.default_rng(21)
    try:
        for dims in [(24, 18, 18), (25, 25, 25), (45, 45, 45), (60, 60, 60)]:
            gpu.prepare_buffers(dims, np.zeros(int(np.prod(dims))))
            x = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
            ref = np.fft.fftn(x)
            outs = {}
            for o in (0, 1):
                for tpb in (64, 128, 256):
                    for z in (0, 1):
                        if z and tpb * (dims[2] + 1) * (16 if prec == "fp64" else 8) > 160 * 1024:
                            continue
                        gpu.set_fft_plan(o, tpb, z)
                        f = gpu.fft3d(x.reshape(-1), -1, prec).reshape(dims)
                        assert rel(f, ref) < tol, (dims, o, tpb, z)
                        outs[(o, tpb, z)] = f
            for (o, tpb, z), f in outs.items():
                assert np.array_equal(f, outs[(o, 64, 0)]), (dims, o, tpb, z)
            gpu.set_fft_plan(-1)
            pl = gpu.fft_plan(1, prec)
            print("fft plan", dims, prec, pl, flush=True)
            assert pl["candidates"] >= 3 and pl["us_tuned"] <= pl["us_default"]
            f = gpu.fft3d(x.reshape(-1), -1, prec).reshape(dims)
            assert np.array_equal(f, outs[(pl["ord"], pl["tpb"], pl["zlds"])]), dims
    finally:
        gpu.set_fft_plan(-1)


def test_unsupported_grid_is_an_error(gpu, pwk_mod):
    gpu.prepare_buffers((7, 7, 7), np.zeros(343))
    with pytest.raises(pwk_mod.PwkError) as e:
        gpu.fft3d(np.zeros(343, complex), -1)
    assert e.value.code == 5


# ---- Hamiltonian -----------------------------------------------------------------
@pytest.fixture()
def ham(acc, si_sys, oracle):
    rng = np.random.default_rng(7)
    vr = rng.uniform(-1, 1, si_sys.dims) * 13.6
    veff = np.fft.fftn(vr).reshape(-1) / si_sys.n_grid
    k = np.array([0.11, -0.23, 0.05])
    acc.set_basis(si_sys.miller())
    acc.set_nk(2)
    acc.set_kinetic(0, si_sys.kinetic(k))
    acc.set_kinetic(1, si_sys.kinetic([0, 0, 0]))
    acc.set_veff(veff)
    B, DBH, _ = si_sys.nonlocal_(k)
    return dict(veff=veff, k=k, B=B, DBH=DBH)


@pytest.mark.parametrize("dims", [(72, 72, 75), (80, 80, 81), (81, 81, 72), (90, 96, 100), (128, 128, 120),
                                  (256, 250, 18), (512, 24, 20), (20, 18, 1024)])
def test_hpsi_on_large_grids(gpu, oracle, dims):
    # the largest register line lengths through the z-stick and plane kernels, and beyond them
    # (an axis of 90-1024 points) the full-grid path with LDS-staged lines: Si FCC at 150 eV on a
    # forced grid (no aliasing: the product is the same as on the rule's grid)
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 150.0, dims=list(dims))
    print("hpsi grid", dims, "npw", s.npw, flush=True)
    gpu.prepare_buffers(s.dims, s.g2())
    gpu.set_basis(s.miller())
    gpu.set_nk(1)
    k = np.array([0.07, 0.02, -0.11])
    gpu.set_kinetic(0, s.kinetic(k))
    rng = np.random.default_rng(11)
    veff = np.fft.fftn(rng.uniform(-1, 1, s.dims) * 13.6).reshape(-1) / s.n_grid
    gpu.set_veff(veff)
    B, DBH, _ = s.nonlocal_(k)
    gpu.set_nonlocal(0, B, DBH)
    psi = rng.standard_normal((3, s.npw)) + 1j * rng.standard_normal((3, s.npw))
    ref = s.hpsi_with_nl(k, veff, B, DBH, psi)
    assert rel(gpu.hpsi(0, psi, "fp64"), ref) < 1e-12
    assert rel(gpu.hpsi(0, psi, "fp32"), ref) < 2e-5


@pytest.mark.parametrize("dims", [(15, 18, 20), (24, 18, 18), (20, 16, 25)])
def test_noncubic_grid_hpsi_and_density(gpu, oracle, dims):
    # nx != ny: the full-grid path (batched 3D FFTs) for H·psi and the density
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 120.0, dims=list(dims))
    print("noncubic grid", dims, "npw", s.npw, flush=True)
    gpu.prepare_buffers(s.dims, s.g2())
    gpu.set_basis(s.miller())
    gpu.set_nk(1)
    k = np.array([0.05, -0.08, 0.13])
    gpu.set_kinetic(0, s.kinetic(k))
    rng = np.random.default_rng(15)
    veff = np.fft.fftn(rng.uniform(-1, 1, s.dims) * 13.6).reshape(-1) / s.n_grid
    gpu.set_veff(veff)
    B, DBH, _ = s.nonlocal_(k)
    gpu.set_nonlocal(0, B, DBH)
    psi = rng.standard_normal((5, s.npw)) + 1j * rng.standard_normal((5, s.npw))
    ref = s.hpsi_with_nl(k, veff, B, DBH, psi)
    assert rel(gpu.hpsi(0, psi, "fp64"), ref) < 1e-12
    assert rel(gpu.hpsi(0, psi, "fp32"), ref) < 2e-5
    fw = np.array([2.0, 0.5, 1e-16, 1.0, 0.25])
    rho0 = rng.uniform(0, 1, s.n_grid)
    dref = s.density(psi, fw, rho0)
    assert rel(gpu.density(psi, fw, rho0, "fp64"), dref) < 1e-12
    assert rel(gpu.density(psi, fw, rho0, "fp32"), dref) < 1e-5


def test_state_errors(gpu, si_sys, pwk_mod):
    gpu.prepare_buffers(si_sys.dims, si_sys.g2())
    gpu.set_basis(si_sys.miller())
    gpu.set_nk(1)
    gpu.set_kinetic(0, si_sys.kinetic([0, 0, 0]))
    psi = np.ones((1, si_sys.npw), complex)
    with pytest.raises(pwk_mod.PwkError) as e:
        gpu.hpsi(0, psi)  # V_eff not set
    assert e.value.code == 6
    with pytest.raises(pwk_mod.PwkError) as e:
        gpu.hpsi(3, psi)  # k index out of range
    assert e.value.code == 1


def test_dense_h_local_bitwise(acc, ham, si_sys):
    h = acc.dense_h(0, with_nl=False)
    ref = si_sys.dense_h(ham["k"], ham["veff"], with_nl=False)
    assert np.array_equal(h, ref)


def test_dense_h_with_nonlocal(acc, ham, si_sys):
    acc.set_nonlocal(0, ham["B"], ham["DBH"])
    h = acc.dense_h(0, with_nl=True)
    ref = si_sys.dense_h(ham["k"], ham["veff"], with_nl=True)
    assert rel(h, ref) < 1e-13
    assert rel(h, h.conj().T) < 1e-13


def test_device_nonlocal_build(acc, ham, si_sys, oracle):
    from pwk import decks

    u = decks.read_upf(SI_UPF)
    nch = acc.build_nonlocal(0, ham["k"], si_sys.gvec(), np.array(SI_FRAC) @ si_lat(), [0, 0],
                             [dict(l=u["l"], dij=u["dij"], r=u["r"], rab=u["rab"], chi=u["chi"])], si_sys.omega)
    assert nch == ham["B"].shape[1] == 36
    h = acc.dense_h(0, with_nl=True)
    ref = si_sys.dense_h(ham["k"], ham["veff"], with_nl=True)
    assert rel(h, ref) < 1e-12


@pytest.mark.parametrize("with_nl", [False, True])
def test_hpsi_fp64_matches_dense(acc, ham, si_sys, with_nl):
    rng = np.random.default_rng(9)
    nb = 7
    psi = rng.standard_normal((nb, si_sys.npw)) + 1j * rng.standard_normal((nb, si_sys.npw))
    if with_nl:
        acc.set_nonlocal(0, ham["B"], ham["DBH"])
    else:
        acc.set_nonlocal(0, None, None)
    H = si_sys.dense_h(ham["k"], ham["veff"], with_nl=with_nl)
    ref = (H @ psi.T).T
    out = acc.hpsi(0, psi, "fp64")
    assert rel(out, ref) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-5), ("split_bf16", 2e-5)])
def test_hpsi_single_precision(acc, ham, si_sys, prec, tol):
    rng = np.random.default_rng(10)
    nb = 5
    psi = rng.standard_normal((nb, si_sys.npw)) + 1j * rng.standard_normal((nb, si_sys.npw))
    acc.set_nonlocal(0, ham["B"], ham["DBH"])
    ref = si_sys.hpsi_with_nl(ham["k"], ham["veff"], ham["B"], ham["DBH"], psi)
    out = acc.hpsi(0, psi, prec)
    assert rel(out, ref) < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_hpsi_fused_and_stick_paths_agree(acc, ham, si_sys, prec, tol):
    # the fused one-launch kernel (small cubic grids) and the z-stick kernels + projector GEMMs
    # compute the same product; both against the oracle, with and without projectors
    rng = np.random.default_rng(14)
    psi = rng.standard_normal((7, si_sys.npw)) + 1j * rng.standard_normal((7, si_sys.npw))
    for with_nl in (True, False):
        if with_nl:
            acc.set_nonlocal(0, ham["B"], ham["DBH"])
            ref = si_sys.hpsi_with_nl(ham["k"], ham["veff"], ham["B"], ham["DBH"], psi)
        else:
            acc.set_nonlocal(0, None, None)
            ref = (si_sys.dense_h(ham["k"], ham["veff"], with_nl=False) @ psi.T).T
        try:
            acc.set_fused_hpsi(True)
            fused = acc.hpsi(0, psi, prec)
            acc.set_fused_hpsi(False)
            stick = acc.hpsi(0, psi, prec)
        finally:
            acc.set_fused_hpsi(True)
        assert rel(fused, ref) < tol, with_nl
        assert rel(stick, ref) < tol, with_nl
    # the density: fused kernel (band groups, FP64 partial sums) and the plane kernels
    fw = np.array([2.0, 1e-16, 0.5, 1.0, 0.0, 0.3, 1.2])
    rho0 = rng.uniform(0, 1, si_sys.n_grid)
    dref = si_sys.density(psi, fw, rho0)
    try:
        acc.set_fused_hpsi(True)
        dfused = acc.density(psi, fw, rho0, prec)
        acc.set_fused_hpsi(False)
        dstick = acc.density(psi, fw, rho0, prec)
    finally:
        acc.set_fused_hpsi(True)
    dtol = 1e-12 if prec == "fp64" else 1e-5
    assert rel(dfused, dref) < dtol
    assert rel(dstick, dref) < dtol


def test_hpsi_second_kpoint_and_single_band(acc, ham, si_sys):
    rng = np.random.default_rng(12)
    psi = rng.standard_normal((1, si_sys.npw)) + 1j * rng.standard_normal((1, si_sys.npw))
    acc.set_nonlocal(1, None, None)
    H = si_sys.dense_h([0, 0, 0], ham["veff"], with_nl=False)
    assert rel(acc.hpsi(1, psi, "fp64"), (H @ psi.T).T) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_density_accumulate(acc, ham, si_sys, prec, tol):
    rng = np.random.default_rng(13)
    nb = 6
    psi = rng.standard_normal((nb, si_sys.npw)) + 1j * rng.standard_normal((nb, si_sys.npw))
    fw = np.array([2.0, 1.5, 1e-16, 0.25, 0.0, 0.75])  # bands below 1e-15 are skipped
    rho0 = rng.uniform(0, 1, si_sys.n_grid)
    ref = si_sys.density(psi, fw, rho0)
    out = acc.density(psi, fw, rho0, prec)
    assert rel(out, ref) < tol


# ---- dense eigensolver (diagonalize_lowest, src/eigensolver/dense.rs:87-98) -----
def _hermitian_with_spectrum(rng, lam):
    n = len(lam)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return (q * lam) @ q.conj().T


def _check_pairs(h, w, v, w_ref, tol_w, tol_res):
    n = h.shape[0]
    nb = w.shape[0]
    assert np.all(np.diff(w) >= -1e-12 * max(1.0, np.abs(w).max()))  # ascending
    assert np.abs(w - w_ref[:nb]).max() <= tol_w * max(1.0, np.abs(w_ref).max())
    res = h @ v - v * w
    assert np.abs(res).max() <= tol_res * max(1.0, np.abs(w_ref).max())
    assert np.abs(v.conj().T @ v - np.eye(nb)).max() < 1e-10
    assert v.shape == (n, nb)


def test_dense_eigh_random_and_degenerate(gpu):
    rng = np.random.default_rng(5)
    mats, refs = [], []
    # generic spectrum, an exactly degenerate one (3-fold and 2-fold clusters), a
    # near-degenerate one (splitting 1e-9) and a diagonal matrix (zero reflectors)
    lam_sets = [
        np.sort(rng.standard_normal(40)) * 10,
        np.array([-5.0] * 3 + [-1.0] * 2 + list(np.linspace(0.5, 9.0, 35))),
        np.array([-2.0, -2.0 + 1e-9, -2.0 + 2e-9] + list(np.linspace(-1.0, 4.0, 37))),
    ]
    for lam in lam_sets:
        mats.append(_hermitian_with_spectrum(rng, lam))
        refs.append(np.sort(lam))
    mats.append(np.diag(np.arange(40, dtype=float)[::-1]).astype(np.complex128))
    refs.append(np.arange(40, dtype=float))
    h = np.stack(mats)
    for nb in (1, 8, 40):
        w, v = gpu.dense_eigh(h, nb)
        for i in range(len(mats)):
            _check_pairs(mats[i], w[i], v[i], refs[i], 1e-12, 1e-10)


def test_dense_eigh_tiny_sizes(gpu):
    rng = np.random.default_rng(9)
    for n in (1, 2, 3):
        hm = _hermitian_with_spectrum(rng, np.sort(rng.standard_normal(n)))
        w, v = gpu.dense_eigh(hm, n)
        _check_pairs(hm, w[0], v[0], np.linalg.eigvalsh(hm), 1e-12, 1e-12)


def test_dense_eigh_jacobi_sizes(gpu):
    # the LDS Jacobi solver (n <= 32) in both of its forms: one barrier a round (a thread for every
    # rotation block, n <= 8 at one wave) and rotations shared through LDS (larger n); generic,
    # exactly degenerate and diagonal spectra, odd n (padded pairing)
    rng = np.random.default_rng(21)
    for n in (5, 8, 12, 17, 32):
        spectra = [np.sort(rng.standard_normal(n)) * 3,
                   np.array([-1.0] * 2 + [0.5] * 2 + list(np.linspace(1.0, 4.0, n - 4)))]
        mats = [_hermitian_with_spectrum(rng, lam) for lam in spectra]
        mats.append(np.diag(np.arange(n, dtype=float)[::-1]).astype(np.complex128))
        w, v = gpu.dense_eigh(np.stack(mats), n)
        for i, hm in enumerate(mats):
            _check_pairs(hm, w[i], v[i], np.linalg.eigvalsh(hm), 1e-12, 1e-10)


def test_dense_eigh_kohn_sham(gpu, oracle, si_sys):
    # dense H of Si at two k-points: eigenvalues equal the oracle's FP64 EVD
    rng = np.random.default_rng(3)
    n = si_sys.n_grid
    veff = np.fft.fftn(rng.uniform(-1, 1, si_sys.dims)).ravel() / n
    ks = [np.zeros(3), np.array([0.1, -0.2, 0.05])]
    hs = [si_sys.dense_h(k, veff, True) for k in ks]
    nb = 12
    w, v = gpu.dense_eigh(np.stack(hs), nb)
    for i, hm in enumerate(hs):
        w_ref, _ = oracle.eigh(hm)
        _check_pairs(hm, w[i], v[i], w_ref, 1e-12, 1e-9)
