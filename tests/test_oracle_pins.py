"""Pins the FP64 oracle against the reference's own known answers.

Every check here is a value the reference's tests or data files hold (file:line
cited, paths relative to pwdft/pwdft-core/), or a closed form / textbook
identity chosen so that a dropped term or a swapped index fails it.
"""
import csv
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, SI_UPF, C_UPF, FE_UPF

A_SI = 5.431


def si_lat(a=A_SI):
    h = a / 2.0
    return np.array([[0.0, h, h], [h, 0.0, h], [h, h, 0.0]])


SI_FRAC = [[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]]


def rows(name):
    with open(os.path.join(GOLDEN, "csv", name)) as f:
        return list(csv.DictReader(f))


# ---- tables -----------------------------------------------------------------
def test_basis_count_si_200ev(oracle):
    # src/basis.rs:121-130: Si 5.431 Å at 200 eV has 259 plane waves
    m, _ = oracle.basis(si_lat(), 200.0)
    assert len(m) == 259
    # enumeration order n1 -> n2 -> n3 ascending (src/basis.rs:47-59)
    key = m[:, 0] * 10000 + m[:, 1] * 100 + m[:, 2]
    assert np.all(np.diff(key) > 0)


def test_basis_brute_force_membership(oracle):
    # |G|² <= ecut/(ħ²/2m) decided over a generous Miller box by an independent loop
    lat = si_lat()
    m, g = oracle.basis(lat, 150.0)
    rec = 2 * np.pi * np.linalg.inv(lat).T
    hb = oracle.const("HBAR2_OVER_2M")
    r = np.arange(-9, 10)
    n = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    gg = n @ rec
    inside = (gg ** 2).sum(1) * hb <= 150.0
    # only the exactly-on-the-sphere ties could differ; there are none at this cutoff
    assert inside.sum() == len(m)
    assert set(map(tuple, n[inside])) == set(map(tuple, m))


def test_fft_grid_size_kats(oracle):
    # src/fft.rs:323-328
    assert [oracle.fft_grid_size(n) for n in (3, 4, 5)] == [8, 9, 12]


def test_fd3m_has_48_ops(oracle):
    # src/symmetry/detect.rs:246
    rot, tau = oracle.symmetry(si_lat(), [14, 14], SI_FRAC)
    assert len(rot) == 48
    dets = np.round(np.linalg.det(rot.astype(float)))
    assert set(dets) == {1.0, -1.0}


@pytest.mark.parametrize("shift,expected", [((0, 0, 0), 8), ((1, 1, 1), 10)])
def test_ibz_counts_si_444(oracle, shift, expected):
    # src/symmetry/kpoints.rs:24-27: Γ-centred -> 8, MP1976 -> 10, weights sum to 1
    k, w = oracle.kpoints(si_lat(), [14, 14], SI_FRAC, [4, 4, 4], shift, mode=2)
    assert len(k) == expected
    assert abs(w.sum() - 1.0) < 1e-12


def test_identity_reduction_keeps_full_grid(oracle):
    # src/symmetry/mod.rs (reduce_kpoints with identity only preserves the grid)
    k, w = oracle.kpoints(si_lat(), [14, 14], SI_FRAC, [3, 3, 3], (0, 0, 0), mode=1)
    kf, wf = oracle.kpoints(si_lat(), [14, 14], SI_FRAC, [3, 3, 3], (0, 0, 0), mode=0)
    assert len(k) == 27 and np.array_equal(k, kf) and np.allclose(w, 1 / 27)


# ---- radial transforms and projectors --------------------------------------------
def test_spherical_bessel_kats(oracle):
    # src/potential/nonlocal.rs:667-702
    assert abs(oracle.bessel(4, 3.0) - 0.056149714328844) < 1e-10
    assert abs(oracle.bessel(5, 5.0) - 0.106811161456505) < 1e-10
    assert abs(oracle.bessel(4, 5.0) - 0.187017655344889) < 1e-10
    x = 2.5
    j2 = (3 / x ** 2 - 1) * math.sin(x) / x - 3 * math.cos(x) / x ** 2
    assert abs(oracle.bessel(2, x) - j2) < 1e-12


def test_ylm_addition_theorem(oracle):
    # src/potential/nonlocal.rs:973-1002: Σ_m Y_lm(a)Y_lm(b) = (2l+1)/(4π) P_l(â·b̂)
    from numpy.polynomial import legendre

    rng = np.random.default_rng(1)
    lmax = 3
    for _ in range(20):
        a, b = rng.standard_normal(3), rng.standard_normal(3)
        ya, yb = oracle.ylm(a, lmax), oracle.ylm(b, lmax)
        c = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
        for l in range(lmax + 1):
            s = slice(l * l, (l + 1) ** 2)
            lhs = ya[s] @ yb[s]
            rhs = (2 * l + 1) / (4 * np.pi) * legendre.legval(c, [0] * l + [1])
            assert abs(lhs - rhs) < 1e-12


def test_beta_form_factors_csv(oracle):
    # data/csv/beta_q_si_reference.csv, tol 1e-4 Bohr^3/2 (tests/vgcmp_beta_q_cross_check.rs:16)
    pp = oracle.Upf(SI_UPF)
    b = oracle.const("BOHR_TO_ANG")
    worst = 0.0
    for r in rows("beta_q_si_reference.csv"):
        assert pp.proj_l(int(r["projector_index"])) == int(r["l"])
        f = pp.beta_ff(int(r["projector_index"]), float(r["q_bohr_inv"]) / b) / b ** 1.5
        worst = max(worst, abs(f - float(r["F_l_q_bohr_3halves"])))
    assert worst < 1e-4


def test_vloc_shells_csv(oracle):
    # data/csv/vloc_g_si_reference.csv, tol 1e-4 Ry (tests/vgcmp_vloc_cross_check.rs:21)
    pp = oracle.Upf(SI_UPF)
    b, ry = oracle.const("BOHR_TO_ANG"), oracle.const("RY_TO_EV")
    omega = abs(np.linalg.det(si_lat()))
    data = rows("vloc_g_si_reference.csv")
    assert len(data) == 20
    for r in data:
        v = pp.vloc_g(float(r["g_bohr_inv"]) / b, omega) / ry
        assert abs(v - float(r["v_local_g_ry"])) < 1e-4


def test_dij_csv(oracle):
    # data/csv/dij_si_reference.csv, tol 1e-12 eV (tests/vgcmp_dij_cross_check.rs:16)
    pp = oracle.Upf(SI_UPF)
    d = pp.dij()
    ry = oracle.const("RY_TO_EV")
    for r in rows("dij_si_reference.csv"):
        assert abs(d[int(r["i"]), int(r["j"])] - float(r["dij_ry"]) * ry) < 1e-12


def test_rho_core_csv(oracle):
    # data/csv/rho_core_g_reference.csv rows (si, shell 0/1); src/pseudopotential/upf/parse.rs:286-335
    pp = oracle.Upf(SI_UPF)
    assert pp.has_nlcc
    r, rab, core = pp.array(0), pp.array(1), pp.array(4)
    omega = A_SI ** 3 / 4
    want = {int(x["shell_index"]): float(x["rho_core_g_e_per_ang3"]) for x in rows("rho_core_g_reference.csv")
            if x["system"] == "si"}
    for shell, g in ((0, 0.0), (1, 2 * np.pi / A_SI * np.sqrt(3.0))):
        gr = g * r
        j0 = np.where(gr < 1e-10, 1 - gr * gr / 6, np.sin(gr) / np.where(gr < 1e-10, 1, gr))
        val = 4 * np.pi * oracle.lib().orc_simpson(len(r), oracle._d(oracle.f64(core * r * r * j0)),
                                                    oracle._d(oracle.f64(rab))) / omega
        assert abs(val - want[shell]) < 1e-5


def test_assembled_h_diagonal_csv(oracle):
    # data/csv/vgcmp_phase4_reference.csv: kinetic and H diagonal at Γ, tol 1e-4 Ry
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 200.0)
    B, DBH, _ = s.nonlocal_([0, 0, 0])
    hnl_diag = np.einsum("ic,ci->i", B, DBH).real
    t = s.kinetic([0, 0, 0])
    m = s.miller()
    ry = oracle.const("RY_TO_EV")
    data = rows("vgcmp_phase4_reference.csv")
    assert len(data) == 5
    for r in data:
        mm = [int(r["miller_n1"]), int(r["miller_n2"]), int(r["miller_n3"])]
        i = int(np.where((m == mm).all(1))[0][0])
        assert abs(t[i] / ry - float(r["kinetic_ry"])) < 1e-4
        assert abs(hnl_diag[i] / ry - float(r["v_nl_diag_ry"])) < 1e-4
        assert abs((t[i] + hnl_diag[i]) / ry - float(r["h_diag_ry"])) < 1e-4


def test_nonlocal_l2_closed_form(oracle):
    # src/potential/nonlocal.rs:871-965: single l=2 channel, H_NL(G,G') =
    # (D/Ω)·F(|G|)·F(|G'|)·(5/4π)·P_2(cosθ); check the addition-theorem form on a real cell
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 120.0)
    B, DBH, meta = s.nonlocal_([0.1, -0.2, 0.05])
    H = B @ DBH
    assert np.abs(H - H.conj().T).max() < 1e-12 * np.abs(H).max()  # Hermitian
    # projector-level reconstruction: H = Σ_a Σ_ij D_ij Σ_m B_(a,i,m) B*_(a,j,m)
    pp = oracle.Upf(SI_UPF)
    D = pp.dij()
    H2 = np.zeros_like(H)
    for ci in range(B.shape[1]):
        for cj in range(B.shape[1]):
            a1, i1, m1, l1 = meta[ci]
            a2, i2, m2, l2 = meta[cj]
            if a1 == a2 and m1 == m2 and l1 == l2:
                H2 += D[i1, i2] * np.outer(B[:, ci], B[:, cj].conj())
    assert np.abs(H - H2).max() < 1e-12 * np.abs(H).max()


# ---- FFT, dense H, eigensolver -------------------------------------------------------
@pytest.mark.parametrize("shape", [(15, 15, 15), (16, 12, 18), (25, 24, 20)])
def test_fft_matches_dft_definition(oracle, shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.abs(oracle.fft3d(x, -1) - np.fft.fftn(x)).max() < 1e-10
    assert np.abs(oracle.fft3d(x, +1) - np.fft.ifftn(x) * x.size).max() < 1e-10


def test_fft_hpsi_equals_dense_h(oracle):
    # circular convolution = wrapped Miller lookup of fill_hamiltonian_with_v_eff
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 150.0)
    rng = np.random.default_rng(3)
    vr = rng.uniform(-1, 1, s.dims)
    veff = np.fft.fftn(vr).reshape(-1) / s.n_grid
    k = np.array([0.1, 0.2, -0.05])
    H = s.dense_h(k, veff)
    psi = rng.standard_normal((3, s.npw)) + 1j * rng.standard_normal((3, s.npw))
    ref = (H @ psi.T).T
    assert np.abs(s.hpsi(k, veff, psi) - ref).max() < 1e-12 * np.abs(ref).max()


def test_eigh_against_lapack(oracle):
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 40):
        lam = np.round(rng.standard_normal(n) * 2) / 2  # repeated eigenvalues
        q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        a = (q * lam) @ q.conj().T
        a = (a + a.conj().T) / 2
        w, v = oracle.eigh(a)
        assert np.abs(w - np.linalg.eigvalsh(a)).max() < 1e-12
        assert np.abs(a @ v - v * w).max() < 1e-12
        assert np.abs(v.conj().T @ v - np.eye(n)).max() < 1e-12


def test_eigh_2x2_kat(oracle):
    # src/eigensolver/dense.rs:272-358 style: [[2, 1],[1, 2]] -> 1, 3
    w, v = oracle.eigh(np.array([[2, 1], [1, 2]], dtype=complex))
    assert np.allclose(w, [1, 3], atol=1e-14)


# ---- XC ------------------------------------------------------------------------------
def test_lsda_unpolarised_limit_equals_lda(oracle):
    # src/potential/xc.rs:1865-1891
    rho = np.logspace(-4, 1, 200)
    e, v = oracle.lda_xc(rho)
    es, vu, vd = oracle.lsda_xc(rho / 2, rho / 2)
    assert np.abs(es - e).max() < 1e-10 * np.abs(e).max()
    assert np.abs(vu - v).max() < 1e-10 * np.abs(v).max()
    assert np.abs(vd - v).max() < 1e-10 * np.abs(v).max()


def _pz_c(rs, polarised):
    """PZ81 correlation ε_c(r_s) in Ha and dε_c/dr_s, from the published parametrisation
    (Perdew & Zunger, PRB 23, 5048 (1981), Table XII): r_s ≥ 1 γ/(1 + β1√r_s + β2 r_s),
    r_s < 1 A ln r_s + B + C r_s ln r_s + D r_s."""
    rs = np.asarray(rs, dtype=float)
    if polarised:
        g, b1, b2, A, B, C, D = -0.0843, 1.3981, 0.2611, 0.01555, -0.0269, 0.0007, -0.0048
    else:
        g, b1, b2, A, B, C, D = -0.1423, 1.0529, 0.3334, 0.0311, -0.048, 0.0020, -0.0116
    hi = rs >= 1.0
    sr = np.sqrt(rs)
    den = 1 + b1 * sr + b2 * rs
    lr = np.log(rs)
    ec = np.where(hi, g / den, A * lr + B + C * rs * lr + D * rs)
    dec = np.where(hi, -g * (b1 / (2 * sr) + b2) / den ** 2, A / rs + C * (lr + 1) + D)
    return ec, dec


def _rs_of(oracle, rho_ang):
    b = oracle.const("BOHR_TO_ANG")
    return (3 / (4 * np.pi * rho_ang * b ** 3)) ** (1 / 3)


@pytest.mark.parametrize("rs_target", [0.3, 0.7, 1.5, 4.0, 9.0])
def test_lsda_fully_polarised_closed_form(oracle, rs_target):
    # ζ = 1 (all density spin-up): ε_x = 2^{1/3}·ε_x^unpol, ε_c = ε_c^pol of the polarised PZ81
    # parameters, v↑ = (4/3)ε_x + ε_c^pol − (r_s/3)dε_c^pol/dr_s; v↓ keeps the ζ-derivative of
    # f(ζ) = [(1+ζ)^{4/3} + (1−ζ)^{4/3} − 2]/(2^{4/3} − 2): v↓ = v↑_c − 2 f'(1)(ε_c^pol − ε_c^unpol),
    # with f'(1) = (4/3)·2^{1/3}/(2^{4/3} − 2) and no exchange (ρ↓ = 0). Both r_s branches.
    b = oracle.const("BOHR_TO_ANG")
    ha = oracle.const("HA_TO_EV")
    rho = 3 / (4 * np.pi * rs_target ** 3) / b ** 3  # e/Å³
    rb = rho * b ** 3
    ex_unpol = -0.75 * (3 * rb / np.pi) ** (1 / 3)
    ex = 2 ** (1 / 3) * ex_unpol
    ecp, decp = _pz_c(rs_target, True)
    ecu, _ = _pz_c(rs_target, False)
    vcp = ecp - rs_target / 3 * decp
    fprime1 = (4 / 3) * 2 ** (1 / 3) / (2 ** (4 / 3) - 2)
    e, vu, vd = oracle.lsda_xc(np.array([rho]), np.array([0.0]))
    assert abs(e[0] - (ex + ecp) * ha) < 1e-12 * abs(e[0])
    assert abs(vu[0] - ((4 / 3) * ex + vcp) * ha) < 1e-12 * abs(vu[0])
    assert abs(vd[0] - (vcp - 2 * fprime1 * (ecp - ecu)) * ha) < 1e-12 * abs(vu[0])
    # ζ = −1 is the mirror image
    e2, vu2, vd2 = oracle.lsda_xc(np.array([0.0]), np.array([rho]))
    assert e2[0] == e[0] and vu2[0] == vd[0] and vd2[0] == vu[0]


def test_lsda_reference_polarised_and_swap_kats(oracle):
    # src/potential/xc.rs:1893-1908: ε_xc(0.1, 0) < 0 and below the unpolarised ε_xc(0.1)
    e, _, _ = oracle.lsda_xc(np.array([0.1]), np.array([0.0]))
    eu, _ = oracle.lda_xc(np.array([0.1]))
    assert e[0] < 0 and e[0] < eu[0]
    # src/potential/xc.rs:1910-1935: swapping ρ↑ and ρ↓ keeps ε_xc and swaps the potentials (1e-12)
    e1, u1, d1 = oracle.lsda_xc(np.array([0.15]), np.array([0.05]))
    e2, u2, d2 = oracle.lsda_xc(np.array([0.05]), np.array([0.15]))
    assert abs(e1[0] - e2[0]) < 1e-12 and abs(u1[0] - d2[0]) < 1e-12 and abs(d1[0] - u2[0]) < 1e-12


@pytest.mark.parametrize("zeta", [-0.8, -0.3, 0.0, 0.25, 0.6, 0.95])
@pytest.mark.parametrize("rs_target", [0.5, 2.0, 6.0])
def test_lsda_potentials_are_energy_derivatives(oracle, zeta, rs_target):
    # v_σ = ∂(ρ ε_xc)/∂ρ_σ by central differences: a wrong chain-rule sign, a dropped (±1 − ζ)
    # factor or a wrong f'(ζ) fails this at 1e-7 relative
    b = oracle.const("BOHR_TO_ANG")
    rho = 3 / (4 * np.pi * rs_target ** 3) / b ** 3
    ru, rd = rho * (1 + zeta) / 2, rho * (1 - zeta) / 2
    h = 1e-5 * rho

    def energy(u, d):
        e, _, _ = oracle.lsda_xc(np.array([u]), np.array([d]))
        return (u + d) * e[0]

    _, vu, vd = oracle.lsda_xc(np.array([ru]), np.array([rd]))
    fu = (energy(ru + h, rd) - energy(ru - h, rd)) / (2 * h)
    fd = (energy(ru, rd + h) - energy(ru, rd - h)) / (2 * h)
    assert abs(fu - vu[0]) < 1e-7 * abs(vu[0])
    assert abs(fd - vd[0]) < 1e-7 * abs(vd[0])


def test_lsda_zeta_interpolation_between_limits(oracle):
    # at fixed r_s, ε_c(ζ) − ε_c(0) = f(ζ)(ε_c^pol − ε_c^unpol) and ε_x(ζ) = ε_x^unpol·[(1+ζ)^{4/3} +
    # (1−ζ)^{4/3}]/2; checked at r_s = 2 over ζ ∈ [−1, 1]
    b = oracle.const("BOHR_TO_ANG")
    ha = oracle.const("HA_TO_EV")
    rs = 2.0
    rho = 3 / (4 * np.pi * rs ** 3) / b ** 3
    rb = rho * b ** 3
    z = np.linspace(-1, 1, 41)
    e, _, _ = oracle.lsda_xc(rho * (1 + z) / 2, rho * (1 - z) / 2)
    f = ((1 + z) ** (4 / 3) + (1 - z) ** (4 / 3) - 2) / (2 ** (4 / 3) - 2)
    ecp, _ = _pz_c(rs, True)
    ecu, _ = _pz_c(rs, False)
    exu = -0.75 * (3 * rb / np.pi) ** (1 / 3)
    ref = (exu * ((1 + z) ** (4 / 3) + (1 - z) ** (4 / 3)) / 2 + ecu + f * (ecp - ecu)) * ha
    assert np.abs(e - ref).max() < 1e-12 * np.abs(ref).max()


def test_lda_closed_form_exchange_and_regimes(oracle):
    # ε_x = -(3/4)(3ρ/π)^{1/3} Ha (ρ in Bohr⁻³); the two PZ81 branches meet at r_s = 1 up to the
    # parametrisation's own jump: γ/(1+β1+β2) - (B + D) = -3.1e-5 Ha
    b = oracle.const("BOHR_TO_ANG")
    ha = oracle.const("HA_TO_EV")
    rs = np.array([0.999999, 1.000001])
    rho = 3 / (4 * np.pi * rs ** 3) / b ** 3
    e, v = oracle.lda_xc(rho)
    jump = (-0.1423 / (1 + 1.0529 + 0.3334) - (-0.048 - 0.0116)) * ha
    assert abs((e[1] - e[0]) - jump) < 1e-4
    r = 0.05
    rb = r * b ** 3
    ex = -0.75 * (3 * rb / np.pi) ** (1 / 3) * ha
    rs0 = (3 / (4 * np.pi * rb)) ** (1 / 3)
    sr = np.sqrt(rs0)
    ec = -0.1423 / (1 + 1.0529 * sr + 0.3334 * rs0) * ha
    e1, _ = oracle.lda_xc(np.array([r]))
    assert abs(e1[0] - (ex + ec)) < 1e-12
    e0, v0 = oracle.lda_xc(np.array([0.0, 1e-31]))
    assert np.all(e0 == 0) and np.all(v0 == 0)


def test_hartree_g0_zero_and_scaling(oracle):
    rho = np.array([1 + 0j, 2 + 1j, -1j])
    g2 = np.array([0.0, 2.0, 4.0])
    v = oracle.hartree(rho, g2)
    fe = 4 * np.pi * oracle.const("E2_COULOMB")
    assert v[0] == 0 and np.allclose(v[1:], rho[1:] * fe / g2[1:], rtol=1e-15)


def test_nlcc_double_counting_probe():
    # src/scf/energy.rs:866-918: E_xc[ρ_val+ρ_core] − ∫ρ_val v_xc = −1.7 eV on the 2-point probe
    rho_val = np.array([0.5, 1.5])
    rho_xc = rho_val + np.array([0.1, 0.3])
    exc, vxc = np.array([-1.0, -2.0]), np.array([-0.5, -1.5])
    dv = 2.0 / 2
    assert abs((np.sum(rho_xc * exc * dv) - np.sum(rho_val * vxc * dv)) - (-1.7)) < 1e-12


# ---- end to end -----------------------------------------------------------------------
def test_scf_cpu_baseline_pin(oracle):
    # tests/cpu_scf_baseline.rs:43-74: Si 5.431 Å, 100 eV, fixed 16³, Γ, 4 bands, σ 0.05 eV,
    # β 0.3, ndim 4, identity symmetry -> E = -213.028339 eV, E_F = 8.052064 eV (tol 1e-3)
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 100.0, dims=[16, 16, 16])
    r = s.scf([[0, 0, 0]], [1.0], n_bands=4, max_iter=40, conv=1e-6, ethr=1e-5, beta=0.3, ndim=4, sigma=0.05)
    assert r["converged"] == 1.0 and r["n_iter"] < 40
    assert abs(r["e_total"] - (-213.028339)) < 1e-3
    assert abs(r["fermi"] - 8.052064) < 1e-3
    # the per-component identity closes to O(δρ) (tests/vgc5_per_component_si.rs:326-333 asserts 1e-5 eV
    # at conv 1e-8; this deck stops at 1e-6)
    comp = r["e_kinetic"] + r["e_local"] + r["e_nonlocal"] + r["e_hartree"] + r["e_xc"] + r["e_ewald"] + \
        r["e_smearing"]
    assert abs(comp - r["e_total"]) < 1e-4


def test_scf_qe_gamma_centred_pin(oracle):
    # tests/qe_validation.rs:476-513: Si 4×4×4 Γ-centred (full grid, 48-op density symmetrisation),
    # 15 Ry, 8 bands, FD 0.01 Ry, conv 1e-8 -> E = -231.6544 eV within 1 meV
    ry = oracle.const("RY_TO_EV")
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 15 * ry)
    k, w = oracle.kpoints(si_lat(), [14, 14], SI_FRAC, [4, 4, 4], (0, 0, 0), mode=0)
    r = s.scf(k, w, n_bands=8, max_iter=80, conv=1e-8, ethr=1e-6, beta=0.3, ndim=8, sigma=0.01 * ry, use_sym=True)
    assert r["converged"] == 1.0
    assert abs(r["e_total"] - (-231.6544)) < 1e-3
    # QE reference (data/qe/reference_data.toml) within the reference's 60 meV
    assert abs(r["e_total"] - (-17.02299344 * ry)) < 0.060


VGC5_PINS = {  # tests/vgc5_per_component_si.rs:310-320, eV, tolerance 0.05 eV each
    "e_band": 7.7748, "e_kinetic": 83.4120, "e_local": -52.9772, "e_nonlocal": 35.7641, "e_hartree": 14.8249,
    "e_xc": -84.3474, "e_ewald": -228.5192, "e_total": -231.8429,
}


def vgc5_deck(oracle):
    # tests/vgc5_per_component_si.rs:240-270: Si 5.431 Å, 15 Ry, 4×4×4 MP-1976 (shifted), 8 bands,
    # FD 0.01 Ry, β 0.3, ndim 8, conv 1e-8, energy threshold 1e-6 eV, the crystal's 48 operations.
    # The k-list is the irreducible wedge of the shifted grid (10 points, time reversal included).
    ry = oracle.const("RY_TO_EV")
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 15 * ry)
    k, w = oracle.kpoints(si_lat(), [14, 14], SI_FRAC, [4, 4, 4], (1, 1, 1), mode=2)
    return s, k, w, dict(n_bands=8, max_iter=80, conv=1e-8, ethr=1e-6, beta=0.3, ndim=8, sigma=0.01 * ry,
                         use_sym=True)


def test_scf_vgc5_per_component_pins(oracle):
    # every component within the reference's 50 meV; the split closes to 1e-5 eV
    # (tests/vgc5_per_component_si.rs:321-333). The recorded E_total predates the −TS term
    # (≈ −3.5 meV here), which is why the total sits 3.4 meV from its pin while every component
    # is within 1e-4 eV of its own.
    s, k, w, kw = vgc5_deck(oracle)
    assert len(w) == 10
    r = s.scf(k, w, **kw)
    assert r["converged"] == 1.0 and r["n_iter"] < 80
    for name, pin in VGC5_PINS.items():
        assert abs(r[name] - pin) < 0.05, name
    for name in ("e_band", "e_kinetic", "e_local", "e_nonlocal", "e_hartree", "e_xc", "e_ewald"):
        assert abs(r[name] - VGC5_PINS[name]) < 1e-4, name
    comp = sum(r[t] for t in ("e_kinetic", "e_local", "e_nonlocal", "e_hartree", "e_xc", "e_ewald", "e_smearing"))
    assert abs(comp - r["e_total"]) < 1e-5
    assert abs((r["e_total"] - r["e_smearing"]) - VGC5_PINS["e_total"]) < 1e-3


def test_scf_aloc_pin(oracle):
    # tests/aloc_f5_hamiltonian_cache.rs:69-119: Si 5.431 Å, 100 eV, 2×2×2 MP-1976 (full list, the
    # density symmetrised by the 48 operations), 8 bands, FD σ 0.1 eV, β 0.3, ndim 8, conv 1e-5,
    # energy threshold 1e-4 eV -> E = -229.0566 eV within 50 meV
    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 100.0)
    k, w = oracle.kpoints(si_lat(), [14, 14], SI_FRAC, [2, 2, 2], (1, 1, 1), mode=0)
    assert len(w) == 8
    r = s.scf(k, w, n_bands=8, max_iter=30, conv=1e-5, ethr=1e-4, beta=0.3, ndim=8, sigma=0.1, use_sym=True)
    assert r["converged"] == 1.0
    assert abs(r["e_total"] - (-229.0566)) < 0.05
    assert abs(r["e_total"] - (-229.0566)) < 1e-4  # observed 1.5e-5 eV


# ---- nspin = 2 (src/scf/driver_spin.rs) ---------------------------------------------------
def _si_gamma(oracle):
    return oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 100.0, dims=[16, 16, 16])


def test_scf_spin_unpolarised_limit_pin(oracle):
    # tests/spin_polarization.rs:54-139: Si nspin = 2 with no starting moment matches nspin = 1
    # within 1e-5 eV and relaxes to |M| < 1e-4
    s = _si_gamma(oracle)
    kw = dict(n_bands=4, max_iter=40, conv=1e-6, ethr=1e-5, beta=0.3, ndim=4, sigma=0.05)
    r1 = s.scf([[0, 0, 0]], [1.0], **kw)
    r2 = s.scf([[0, 0, 0]], [1.0], nspin=2, **kw)
    assert r1["converged"] == 1.0 and r2["converged"] == 1.0 and r2["n_iter"] < 40
    assert abs(r1["e_total"] - r2["e_total"]) < 1e-5
    assert abs(r2["magnetization"]) < 1e-4
    # both channels carry the unpolarised eigenvalues
    assert np.allclose(r2["eigenvalues"][0], r1["eigenvalues"][0], atol=1e-6)
    assert np.allclose(r2["eigenvalues"][1], r1["eigenvalues"][0], atol=1e-6)


def test_scf_spin_xc_consistency_pin(oracle):
    # tests/spin_polarization.rs:143-225: Si nspin = 2 from starting magnetisation 0.2, conv and
    # energy thresholds 1e-6: |E_HF - E_KS| < 1e-5 eV (O(Δρ²)), M < 0.05
    s = _si_gamma(oracle)
    r = s.scf([[0, 0, 0]], [1.0], n_bands=4, max_iter=60, conv=1e-6, ethr=1e-6, beta=0.3, ndim=4, sigma=0.05,
              nspin=2, start_mag=0.2)
    assert r["converged"] == 1.0
    assert abs(r["e_harris"] - r["e_total"]) < 1e-5
    assert r["magnetization"] < 0.05
    # a wrong channel weighting (spin factor 2, or ρ_core added in full per channel) moves E by eV
    assert abs(r["e_total"] - (-213.028339)) < 1e-3


def _fe_bcc():
    a = 2.87
    return a / 2 * np.array([[-1.0, 1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 1.0, -1.0]]), np.zeros((1, 3))


def test_scf_fe_ccmx_kerker_pin(oracle):
    # tests/spin_polarization.rs:229-336: Fe BCC 15 Ry, 4×4×4 MP-1976 (full list, 48-op density
    # symmetrisation), 8 bands, FD 0.02 Ry, Kerker, starting magnetisation 0.5, conv 1e-3:
    # converges (recorded: ~14 iterations, |HF-KS| ~ 1e-4 eV, M ~ 0), |HF-KS| < 1e-3, M < 0.05,
    # final Δρ < 1e-2 (the pre-coupled-channel mixer locked at 0.254)
    ry = oracle.const("RY_TO_EV")
    lat, frac = _fe_bcc()
    s = oracle.System(lat, [0], [26], frac, [FE_UPF], 15 * ry)
    k, w = oracle.kpoints(lat, [26], frac, [4, 4, 4], (1, 1, 1), mode=0)
    assert len(w) == 64
    r = s.scf(k, w, n_bands=8, max_iter=80, conv=1e-3, ethr=1e-3, beta=0.3, ndim=8, sigma=0.02 * ry, use_sym=True,
              nspin=2, start_mag=0.5, kerker=True)
    assert r["converged"] == 1.0 and r["n_iter"] < 80
    assert 10 <= r["n_iter"] <= 18
    assert abs(r["e_harris"] - r["e_total"]) < 1e-3
    assert r["magnetization"] < 0.05
    assert r["final_delta"] < 1e-2


def test_scf_fe_fixed_moment_does_not_converge(oracle):
    # tests/spin_polarization.rs:340-430 (inverted detector): Fe BCC fixed total moment 2 on this
    # pseudopotential has no SCF fixed point; 100 iterations end unconverged
    ry = oracle.const("RY_TO_EV")
    lat, frac = _fe_bcc()
    s = oracle.System(lat, [0], [26], frac, [FE_UPF], 15 * ry)
    k, w = oracle.kpoints(lat, [26], frac, [4, 4, 4], (1, 1, 1), mode=0)
    r = s.scf(k, w, n_bands=8, max_iter=100, conv=1e-6, ethr=1e-5, beta=0.2, ndim=8, sigma=0.02 * ry, use_sym=True,
              nspin=2, tot_mag=2.0, kerker=True)
    assert r["converged"] == 0.0
    assert r["final_delta"] > 1e-6


def test_scf_fe_qe_config_pin(oracle):
    # tests/qe_validation.rs:710-765: Fe BCC nspin = 2, 8×8×8 Γ-centred full grid, 15 Ry, 12 bands,
    # Kerker, FD 0.02 Ry, conv 1e-8; the reference records E = -3049.016 eV (its 11.14 eV gap
    # to QE is the reference's own, tracked there) and a non-magnetic collapse
    ry = oracle.const("RY_TO_EV")
    lat, frac = _fe_bcc()
    s = oracle.System(lat, [0], [26], frac, [FE_UPF], 15 * ry)
    k, w = oracle.kpoints(lat, [26], frac, [8, 8, 8], (0, 0, 0), mode=0)
    r = s.scf(k, w, n_bands=12, max_iter=80, conv=1e-8, ethr=1e-6, beta=0.3, ndim=8, sigma=0.02 * ry, use_sym=True,
              nspin=2, start_mag=0.5, kerker=True)
    assert r["converged"] == 1.0
    assert abs(r["e_total"] - (-3049.016)) < 1e-3
    assert r["magnetization"] < 1e-3
    assert abs(r["e_total"] - (-224.91744934 * ry)) < 12.0  # the reference's own QE tolerance


def test_ewald_closed_form_independent(oracle):
    # Ewald energy does not depend on the splitting parameter: compare with a direct
    # real+reciprocal sum evaluated here at a different η
    from scipy.special import erfc

    s = oracle.System(si_lat(), [0, 0], [14, 14], SI_FRAC, [SI_UPF], 100.0, dims=[16, 16, 16])
    e_or = s.ewald()
    lat = si_lat()
    pos = np.array(SI_FRAC) @ lat
    q = np.array([4.0, 4.0])
    om = abs(np.linalg.det(lat))
    e2 = 14.399645351950548
    eta = 0.8
    rec = 2 * np.pi * np.linalg.inv(lat).T
    r = np.arange(-8, 9)
    n = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    G = n @ rec
    G2 = (G ** 2).sum(1)
    m = G2 > 1e-12
    S = (q[None, :] * np.exp(1j * G[m] @ pos.T)).sum(1)
    erec = 2 * np.pi * e2 / om * np.sum(np.abs(S) ** 2 * np.exp(-G2[m] / (4 * eta ** 2)) / G2[m])
    T = n @ lat
    ereal = 0.0
    for i in range(2):
        for j in range(2):
            d = np.linalg.norm(pos[i] - pos[j] + T, axis=1)
            d = d[d > 1e-10]
            ereal += q[i] * q[j] * np.sum(erfc(eta * d) / d)
    ereal *= 0.5 * e2
    eself = -eta / np.sqrt(np.pi) * np.sum(q ** 2) * e2
    ebg = -np.pi * q.sum() ** 2 * e2 / (2 * om * eta ** 2)
    assert abs(e_or - (ereal + erec + eself + ebg)) < 1e-8
    # VGC5 pin for the 5.431 Å cell (tests/vgc5_per_component_si.rs:318): -228.5192 eV
    assert abs(e_or - (-228.5192)) < 1e-3
