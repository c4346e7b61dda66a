"""CPU checks of the boundary and of the host-side setup (no GPU calls)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import C_UPF, GOLDEN, SI_UPF


def test_library_exports_every_header_symbol(pwk_mod):
    syms = pwk_mod.header_symbols()
    assert len(syms) >= 30
    L = ctypes.CDLL(pwk_mod.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []
    assert pwk_mod.lib().pwk_version() == 101


def test_struct_layout_matches_header(pwk_mod, tmp_path):
    """The ctypes mirrors of the ABI structs have the C compiler's offsets and sizes."""
    import subprocess
    from pwk import scf

    structs = {"pwk_species": scf.Species, "pwk_crystal": scf.Crystal, "pwk_scf_params": scf.Params,
               "pwk_scf_result": scf.Result}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pwk.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(pwk_mod.HEADER), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_status_codes_match_header(pwk_mod):
    text = open(pwk_mod.HEADER).read()
    for code, name in pwk_mod.STATUS.items():
        assert f"{name} = {code}" in text


def test_null_context_is_rejected(pwk_mod):
    L = pwk_mod.lib()
    rc = L.pwk_ctx_sync(None)
    assert rc == 1  # PWK_E_INVALID
    assert b"null context" in L.pwk_last_error()


def test_no_device_means_none(pwk_mod):
    # GpuAccelerator::try_new returns None without a device (src/gpu/mod.rs:148-183)
    h = ctypes.c_void_p()
    rc = pwk_mod.lib().pwk_ctx_create(0, ctypes.byref(h))
    if rc == 0:
        pwk_mod.lib().pwk_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == 2
    assert pwk_mod.GpuAccelerator.try_new() is None


@pytest.mark.parametrize("deck", ["si_fcc_200", "si_fcc_12ry_1026", "si64_20ry", "c64_30ry", "si8_16ry"])
def test_host_basis_and_grid_bit_identical_to_oracle(oracle, deck):
    from pwk import decks

    if deck == "si_fcc_200":
        lat, frac = decks.fcc_primitive(5.431)
        ecut = 200.0
    elif deck == "si_fcc_12ry_1026":
        lat, frac = decks.fcc_primitive(10.26 * decks.BOHR_TO_ANG)
        ecut = 12 * decks.RY_TO_EV
    elif deck == "si64_20ry":
        lat, frac = decks.diamond_supercell(10.26 * decks.BOHR_TO_ANG, (2, 2, 2))
        ecut = 20 * decks.RY_TO_EV
    elif deck == "c64_30ry":
        lat, frac = decks.diamond_supercell(6.74 * decks.BOHR_TO_ANG, (2, 2, 2))
        ecut = 30 * decks.RY_TO_EV
    else:
        lat, frac = decks.diamond_supercell(10.26 * decks.BOHR_TO_ANG)
        ecut = 16 * decks.RY_TO_EV
    m, g = decks.basis(lat, ecut)
    mo, go = oracle.basis(lat, ecut)
    assert np.array_equal(m, mo)
    assert np.array_equal(g, go)  # bitwise: same operation order
    dims = decks.fft_dims(m)
    if deck == "si_fcc_200":
        s = oracle.System(lat, [0] * len(frac), [14] * len(frac), frac, [SI_UPF], ecut)
        assert tuple(s.dims) == dims
        assert np.array_equal(decks.grid_g2(lat, dims), s.g2())
    # the workload sizes quoted for the configs
    expected = {"si_fcc_12ry_1026": (169, (15, 15, 15)), "si64_20ry": (13133, (60, 60, 60)),
                "c64_30ry": (6859, (45, 45, 45)), "si8_16ry": (1189, (25, 25, 25))}
    if deck in expected:
        assert (len(m), dims) == expected[deck]


def test_host_upf_reader_matches_oracle(oracle):
    from pwk import decks

    for path in (SI_UPF, C_UPF):
        u = decks.read_upf(path)
        o = oracle.Upf(path)
        assert u["z_valence"] == o.z_valence
        assert u["l"] == [o.proj_l(i) for i in range(o.n_proj)]
        assert np.array_equal(u["r"], o.array(0)) and np.array_equal(u["rab"], o.array(1))
        assert np.array_equal(u["dij"], o.dij())
        for i in range(o.n_proj):
            assert np.array_equal(u["chi"][i], o.array(10 + i))


def test_seeded_inputs_are_reproducible():
    from pwk import decks

    g2 = np.linspace(0, 10, 50)
    a = decks.seeded_psi(g2, 3, 20261202)
    b = decks.seeded_psi(g2, 3, 20261202)
    assert np.array_equal(a, b)
    assert np.allclose(np.linalg.norm(a, axis=1), 1.0)
    v = decks.seeded_veff_g((6, 6, 6), 7)
    vr = np.fft.ifftn(v.reshape(6, 6, 6)) * 216
    assert np.abs(vr.imag).max() < 1e-12
    assert np.abs(vr.real).max() <= decks.RY_TO_EV


def _si_deck_and_params():
    from pwk import decks, scf

    lat, frac = decks.fcc_primitive(5.431)
    deck = scf.CrystalDeck(lat, frac, [0, 0], [SI_UPF])
    return deck, scf.ScfParams(ecut_ev=100.0, kgrid=(2, 2, 2))


def _tables_rc(pwk_mod, deck, prm):
    from pwk import scf

    npw, nk, nops = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    dims = (ctypes.c_int * 3)()
    L = scf._lib()
    return L.pwk_scf_tables(ctypes.byref(deck.c), ctypes.byref(prm), ctypes.byref(npw), dims, ctypes.byref(nk),
                            ctypes.byref(nops))


def test_malformed_requests_are_rejected_before_any_table(pwk_mod):
    # the checks every SCF entry point (pwk_scf_tables, pwk_scf_run, pwk_server_submit) runs first:
    # a bad species index, missing arrays or out-of-range parameters give PWK_E_INVALID instead of
    # host reads out of bounds
    from pwk import scf

    deck, p = _si_deck_and_params()
    assert _tables_rc(pwk_mod, deck, scf._params(p)) == 0
    bad_species = np.array([0, 3], dtype=np.int32)
    deck.c.species = bad_species.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    assert _tables_rc(pwk_mod, deck, scf._params(p)) == 1
    assert b"species out of range" in pwk_mod.lib().pwk_last_error()
    deck, p = _si_deck_and_params()
    deck.c.frac = ctypes.POINTER(ctypes.c_double)()
    assert _tables_rc(pwk_mod, deck, scf._params(p)) == 1
    deck, p = _si_deck_and_params()
    deck.c.nspecies = 9
    assert _tables_rc(pwk_mod, deck, scf._params(p)) == 1
    deck, p = _si_deck_and_params()
    deck._sp[0].r = ctypes.POINTER(ctypes.c_double)()
    assert _tables_rc(pwk_mod, deck, scf._params(p)) == 1
    assert b"radial mesh" in pwk_mod.lib().pwk_last_error()
    for field, value in (("ecut_ev", -1.0), ("n_bands", 0), ("mixing_beta", 0.0), ("mixing_beta", 1.5),
                         ("nspin", 3), ("mixing_ndim", 33), ("conv_threshold", 0.0), ("eigensolver", 2),
                         ("smearing_sigma", -0.1), ("max_iter", 0)):
        deck, p = _si_deck_and_params()
        prm = scf._params(p)
        setattr(prm, field, value)
        assert _tables_rc(pwk_mod, deck, prm) == 1, field
    prm = scf._params(p)
    prm.precision = 7
    assert _tables_rc(pwk_mod, deck, prm) == 1
    prm = scf._params(p)
    prm.mixing_ndim = 32  # the largest history the mixer holds
    assert _tables_rc(pwk_mod, deck, prm) == 0


# ---- the library's UPF reader (pwk_upf_load): the reference parser's own tests
# (src/pseudopotential/upf/parse.rs:153-240) and its error cases (xml.rs) --------------------
def test_upf_reader_reference_checks():
    from pwk import decks

    u = decks.read_upf(SI_UPF)
    # test_parse_header
    assert u["element"] == "Si" and u["atomic_number"] == 14
    assert abs(u["z_valence"] - 4.0) < 1e-10 and u["l_max"] >= 1 and len(u["l"]) > 0
    # test_radial_grid_monotonic, test_radial_grid_units
    assert np.all(np.diff(u["r"]) > 0)
    assert u["r"][0] < 0.01 and u["r"][-1] > 1.0
    # test_v_local_coulomb_tail: V_local → -Z e²/r within 1% ten points from the end
    e2 = 14.399645351950548
    r_far, v_far = u["r"][-10], u["v_local"][-10]
    v_c = -u["z_valence"] * e2 / r_far
    assert abs((v_far - v_c) / v_c) < 0.01
    # test_dij_matrix_size
    assert u["dij"].shape == (len(u["l"]), len(u["l"]))
    # test_si_core_charge_integrates_to_partial_core: ∫4πr²ρ_core dr ≈ 0.74 e
    q = np.sum(4.0 * np.pi * u["r"] ** 2 * u["core"] * u["rab"])
    assert abs(q - 0.74) < 0.05


def test_upf_reader_dij_equals_reference_csv():
    from pwk import decks

    u = decks.read_upf(SI_UPF)
    rows = np.loadtxt(os.path.join(GOLDEN, "csv", "dij_si_reference.csv"), delimiter=",", skiprows=1)
    d = np.zeros_like(u["dij"])
    for i, j, v in rows:
        d[int(i), int(j)] = v * 13.605693122994
    assert np.array_equal(u["dij"], d)


def _parse(pwk_mod, text):
    import ctypes as C

    L = pwk_mod.lib()
    h = C.c_void_p()
    b = text.encode()
    st = L.pwk_upf_parse(b, len(b), C.byref(h))
    if st == 0:
        L.pwk_upf_free(h)
    return st, L.pwk_last_error().decode()


def test_upf_reader_errors(pwk_mod):
    good = open(SI_UPF).read()
    assert _parse(pwk_mod, good)[0] == 0
    # not UPF at all
    st, msg = _parse(pwk_mod, "<qe_pp:pseudo>\n</qe_pp:pseudo>")
    assert st == 1 and "unrecognized" in msg
    # a block with one number fewer than mesh_size
    i = good.index("</PP_LOCAL>")
    j = good.rindex(" ", 0, good.rindex("E", 0, i))
    st, msg = _parse(pwk_mod, good[:j] + good[i:])
    assert st == 1 and "PP_LOCAL" in msg and "expected" in msg
    # a Fortran D exponent is not a float (the reference parses with f64::from_str)
    a = good.index("E+00", good.index("<PP_LOCAL"))
    st, msg = _parse(pwk_mod, good[:a] + "D+00" + good[a + 4:])
    assert st == 1 and "parse" in msg
    # missing projector block
    st, msg = _parse(pwk_mod, good.replace("<PP_BETA.2", "<PP_XETA.2").replace("</PP_BETA.2>", "</PP_XETA.2>"))
    assert st == 1 and "PP_BETA.2" in msg
    # unknown element symbol
    k = good.index('element="') + len('element="')
    st, msg = _parse(pwk_mod, good[:k] + "Xx" + good[k + 2:])
    assert st == 1 and "unknown element" in msg
