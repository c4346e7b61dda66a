"""Host integer tables of libpwk (basis, FFT grid, symmetry, k-points) against
the oracle and the reference's own known answers. No GPU: these entry points
run on the host side of the C ABI."""
import numpy as np
import pytest

from conftest import C_UPF, SI_UPF

A_SI = 5.431


def fcc(a):
    h = a / 2.0
    return np.array([[0.0, h, h], [h, 0.0, h], [h, h, 0.0]])


def cubic_diamond(a, n=(1, 1, 1)):
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.25, .25, .25], [.25, .75, .75],
                     [.75, .25, .75], [.75, .75, .25]])
    frac = []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                frac += [[(f[0] + i) / n[0], (f[1] + j) / n[1], (f[2] + k) / n[2]] for f in base]
    lat = np.diag([a * n[0], a * n[1], a * n[2]]).astype(float)
    return lat, np.array(frac)


@pytest.fixture(scope="module")
def scf_mod(pwk_mod):
    from pwk import scf

    return scf


# (lattice, ecut eV): Si FCC KAT deck, the 12 Ry small deck, Si8 16 Ry, C64 30 Ry, a strained triclinic-ish cell
DECKS = [
    (fcc(A_SI), 200.0),
    (fcc(10.26 * 0.529177210903), 12 * 13.605693122994),
    (cubic_diamond(10.26 * 0.529177210903)[0], 16 * 13.605693122994),
    (cubic_diamond(13.48 * 0.529177210903, (2, 2, 2))[0], 30 * 13.605693122994),
    (np.array([[5.0, 0.1, 0.0], [0.3, 4.7, 0.2], [0.0, -0.4, 6.1]]), 180.0),
]


@pytest.mark.parametrize("deck", range(len(DECKS)))
def test_basis_and_grid_bit_exact(scf_mod, oracle, deck):
    lat, ecut = DECKS[deck]
    m = scf_mod.basis(lat, ecut)
    m_ref, _ = oracle.basis(lat, ecut)
    assert np.array_equal(m, m_ref)
    dims = scf_mod.fft_dims(m)
    # grid rule: smallest 2,3,5-smooth >= 2*2*n_max+1 (ratio 4)
    for d in range(3):
        assert dims[d] == oracle.fft_grid_size(2 * int(np.abs(m[:, d]).max()))


def test_basis_kat(scf_mod):
    # src/basis.rs:121-130: Si 5.431 Å at 200 eV has 259 plane waves
    assert scf_mod.basis(fcc(A_SI), 200.0).shape == (259, 3)


def test_grid_sizes_of_configs(scf_mod):
    m = scf_mod.basis(fcc(10.26 * 0.529177210903), 12 * 13.605693122994)
    assert m.shape[0] == 169 and scf_mod.fft_dims(m) == (15, 15, 15)
    lat, _ = cubic_diamond(10.26 * 0.529177210903, (2, 2, 2))
    m = scf_mod.basis(lat, 20 * 13.605693122994)
    assert m.shape[0] == 13133 and scf_mod.fft_dims(m) == (60, 60, 60)


def _op_set(rot, tau):
    return sorted((tuple(r.reshape(9)), tuple(np.round(t, 9))) for r, t in zip(rot, tau))


def test_symmetry_matches_oracle(scf_mod, oracle):
    # Si FCC: 48 operations (src/symmetry/detect.rs:246)
    rot, tau = scf_mod.symmetry(fcc(A_SI), [14, 14], [[0, 0, 0], [.25, .25, .25]])
    assert len(rot) == 48
    r2, t2 = oracle.symmetry(fcc(A_SI), [14, 14], [[0, 0, 0], [.25, .25, .25]])
    assert _op_set(rot, tau) == _op_set(r2, t2)
    # same order as the oracle, too
    assert np.array_equal(rot, r2) and np.array_equal(tau, t2)
    # displaced cubic diamond: fewer operations, still identical to the oracle
    lat, frac = cubic_diamond(5.43)
    frac = frac.copy()
    frac[3] += [0.01, 0.0, 0.0]
    z = [14] * 8
    rot, tau = scf_mod.symmetry(lat, z, frac)
    r2, t2 = oracle.symmetry(lat, z, frac)
    assert np.array_equal(rot, r2) and np.array_equal(tau, t2)
    assert 1 <= len(rot) < 48


@pytest.mark.parametrize("shift,n_ibz", [((0, 0, 0), 8), ((1, 1, 1), 10)])
def test_kpoints_kat(scf_mod, oracle, shift, n_ibz):
    # src/symmetry/kpoints.rs:24-27: Si 4×4×4 → 8 (Γ-centred) / 10 (shifted) IBZ points
    args = (fcc(A_SI), [14, 14], [[0, 0, 0], [.25, .25, .25]], (4, 4, 4), shift)
    k, w = scf_mod.kpoints(*args, use_symmetry=True)
    assert len(k) == n_ibz
    k2, w2 = oracle.kpoints(*args, mode=2)
    assert np.array_equal(k, k2) and np.array_equal(w, w2)
    assert abs(w.sum() - 1.0) < 1e-14


def test_kpoints_identity_only(scf_mod, oracle):
    lat, frac = cubic_diamond(10.26 * 0.529177210903)
    args = (lat, [14] * 8, frac, (6, 6, 6), (0, 0, 0))
    k, w = scf_mod.kpoints(*args, use_symmetry=False)
    assert len(k) == 216
    k2, w2 = oracle.kpoints(*args, mode=1)
    assert np.array_equal(k, k2) and np.array_equal(w, w2)


def test_scf_tables_of_small_deck(scf_mod):
    a = 10.26 * 0.529177210903
    deck = scf_mod.CrystalDeck(fcc(a), [[0, 0, 0], [.25, .25, .25]], [0, 0], [SI_UPF])
    p = scf_mod.ScfParams(ecut_ev=12 * 13.605693122994, kgrid=(4, 4, 4))
    t = scf_mod.scf_tables(deck, p)
    assert t == {"npw": 169, "dims": (15, 15, 15), "nk": 8, "nops": 48}
    p2 = scf_mod.ScfParams(ecut_ev=12 * 13.605693122994, kgrid=(2, 2, 2), use_symmetry=False)
    assert scf_mod.scf_tables(deck, p2)["nk"] == 8 and scf_mod.scf_tables(deck, p2)["nops"] == 1
    # full list with a symmetrised density (the reference's run_scf fed an unreduced MP grid)
    p3 = scf_mod.ScfParams(ecut_ev=12 * 13.605693122994, kgrid=(4, 4, 4), use_symmetry=2)
    assert scf_mod.scf_tables(deck, p3)["nk"] == 64 and scf_mod.scf_tables(deck, p3)["nops"] == 48


def test_upf_reader_fields(pwk_mod):
    from pwk.decks import read_upf

    for path, zz in ((SI_UPF, 14), (C_UPF, 6)):
        u = read_upf(path)
        assert u["atomic_number"] == zz
        assert u["v_local"].shape == u["r"].shape
        assert u["core"] is not None  # both ship core_correction="T"
