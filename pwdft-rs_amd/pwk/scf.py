"""SCF and integer-table entry points of libpwk, mirroring the reference's
``run_scf`` (src/scf/mod.rs:326-400, with src/scf/driver.rs:353-735 for nspin = 1 and
src/scf/driver_spin.rs:37-839 for nspin = 2), ``BasisSet::new`` (src/basis.rs:36-63),
``find_symmetry_operations`` (src/symmetry/detect.rs:22-178) and
``reduce_kpoints`` (src/symmetry/kpoints.rs:31-143).

Everything here goes through the C ABI of ``include/pwk.h``; the table
functions run on the host, ``run_scf`` on the context's device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import PwkError, _check, _prec, lib
from .decks import read_upf

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
MAX_SPECIES = 8  # PWK_MAX_SPECIES


class Species(C.Structure):
    _fields_ = [("atomic_number", C.c_int), ("z_valence", C.c_double), ("mesh", C.c_int), ("r", _dp), ("rab", _dp),
                ("v_local", _dp), ("nproj", C.c_int), ("proj_l", _ip), ("beta", _dp), ("dij", _dp),
                ("rho_atom", _dp), ("rho_core", _dp)]


class Crystal(C.Structure):
    _fields_ = [("lattice", C.c_double * 9), ("natoms", C.c_int), ("frac", _dp), ("species", _ip),
                ("nspecies", C.c_int), ("species_data", C.POINTER(Species))]


class Params(C.Structure):
    _fields_ = [("ecut_ev", C.c_double), ("ecutrho_ratio", C.c_int), ("fft_dims", C.c_int * 3),
                ("kgrid", C.c_int * 3), ("kshift", C.c_int * 3), ("use_symmetry", C.c_int), ("symprec", C.c_double),
                ("n_bands", C.c_int), ("max_iter", C.c_int), ("conv_threshold", C.c_double),
                ("energy_threshold", C.c_double), ("mixing_beta", C.c_double), ("mixing_ndim", C.c_int),
                ("smearing_sigma", C.c_double), ("sad_sigma", C.c_double), ("precision", C.c_int),
                ("eigensolver", C.c_int), ("cheb_degree", C.c_int), ("verbose", C.c_int), ("nspin", C.c_int),
                ("starting_magnetization", C.c_double * MAX_SPECIES), ("fixed_magnetization", C.c_int),
                ("tot_magnetization", C.c_double), ("mixing_mode", C.c_int), ("kerker_q_tf", C.c_double)]


class Result(C.Structure):
    _fields_ = [("converged", C.c_int), ("n_iter", C.c_int), ("nk", C.c_int), ("n_bands", C.c_int), ("npw", C.c_int),
                ("n_ops", C.c_int), ("dims", C.c_int * 3)] + [
        (n, C.c_double) for n in ("e_total", "e_harris", "fermi", "ts", "final_delta", "integrated_charge", "e_band",
                                  "e_kinetic", "e_local", "e_nonlocal", "e_hartree", "e_xc", "e_vxc", "e_ewald",
                                  "e_smearing", "seconds")] + [
        ("nspin", C.c_int), ("magnetization", C.c_double), ("n_up", C.c_double), ("n_down", C.c_double),
        ("cheb_degree", C.c_int), ("filter_passes", C.c_int)]


_SIGS = {
    "pwk_basis": (C.c_int, [_dp, C.c_double, C.c_int, _ip, C.POINTER(C.c_int)]),
    "pwk_fft_dims": (C.c_int, [_ip, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "pwk_symmetry": (C.c_int, [_dp, C.c_int, _dp, _ip, C.c_double, C.c_int, _ip, _dp, C.POINTER(C.c_int)]),
    "pwk_kpoints": (C.c_int, [_dp, C.c_int, _dp, _ip, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_double,
                              C.c_int, _dp, _dp, C.POINTER(C.c_int)]),
    "pwk_scf_run": (C.c_int, [C.c_void_p, C.POINTER(Crystal), C.POINTER(Params), C.POINTER(Result), _dp, _dp]),
    "pwk_scf_tables": (C.c_int, [C.POINTER(Crystal), C.POINTER(Params), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]),
}
_bound = set()


def _lib():
    L = lib()
    for name, (res, args) in _SIGS.items():
        if name not in _bound:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
            _bound.add(name)
    return L


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(_ip)


# ---- integer tables -------------------------------------------------------------------
def basis(lattice, ecut_ev):
    """Miller indices of the G-sphere, (npw, 3) int32 in the reference's loop order."""
    lat = _f(lattice).reshape(9)
    n = C.c_int()
    _check(_lib().pwk_basis(_pd(lat), float(ecut_ev), 0, None, C.byref(n)))
    m = np.empty((n.value, 3), dtype=np.int32)
    _check(_lib().pwk_basis(_pd(lat), float(ecut_ev), n.value, _pi(m), C.byref(n)))
    return m


def fft_dims(miller, ratio=4):
    m = _i(miller)
    d = (C.c_int * 3)()
    _check(_lib().pwk_fft_dims(_pi(m), m.shape[0], int(ratio), d))
    return tuple(d)


def symmetry(lattice, z, frac, symprec=1e-5):
    """(rot (nops, 3, 3) int32, tau (nops, 3))."""
    lat, fr, zz = _f(lattice).reshape(9), _f(frac).reshape(-1), _i(z)
    n = C.c_int()
    _check(_lib().pwk_symmetry(_pd(lat), len(zz), _pd(fr), _pi(zz), symprec, 0, None, None, C.byref(n)))
    rot = np.empty((n.value, 3, 3), dtype=np.int32)
    tau = np.empty((n.value, 3))
    _check(_lib().pwk_symmetry(_pd(lat), len(zz), _pd(fr), _pi(zz), symprec, n.value, _pi(rot), _pd(tau),
                               C.byref(n)))
    return rot, tau


def kpoints(lattice, z, frac, grid, shift=(0, 0, 0), use_symmetry=True, symprec=1e-5):
    """Irreducible Monkhorst-Pack k-points: (k (nk, 3) Cartesian 1/Å, weights (nk,))."""
    lat, fr, zz = _f(lattice).reshape(9), _f(frac).reshape(-1), _i(z)
    g = (C.c_int * 3)(*grid)
    s = (C.c_int * 3)(*shift)
    n = C.c_int()
    _check(_lib().pwk_kpoints(_pd(lat), len(zz), _pd(fr), _pi(zz), g, s, int(use_symmetry), symprec, 0, None, None,
                              C.byref(n)))
    k = np.empty((n.value, 3))
    w = np.empty(n.value)
    _check(_lib().pwk_kpoints(_pd(lat), len(zz), _pd(fr), _pi(zz), g, s, int(use_symmetry), symprec, n.value,
                              _pd(k), _pd(w), C.byref(n)))
    return k, w


# ---- SCF ------------------------------------------------------------------------------
@dataclass
class ScfParams:
    """The ScfParams fields the drivers read (src/scf/mod.rs:72-402).

    nspin = 2 runs LSDA; ``starting_magnetization`` maps a species index to its
    starting moment fraction (the reference keys it by element), and
    ``tot_magnetization`` (not None) fixes the total moment. ``mixing_mode``
    is "plain" or "kerker" (``kerker_q_tf`` ≤ 0: Thomas-Fermi estimate)."""
    ecut_ev: float
    kgrid: tuple = (1, 1, 1)
    kshift: tuple = (0, 0, 0)
    use_symmetry: bool = True
    symprec: float = 1e-5
    n_bands: int = 8
    max_iter: int = 100
    conv_threshold: float = 1e-6
    energy_threshold: float = 1e-5
    mixing_beta: float = 0.3
    mixing_ndim: int = 8
    smearing_sigma: float = 0.01
    sad_sigma: float = 1.0
    ecutrho_ratio: int = 4
    fft_dims: tuple = (0, 0, 0)
    precision: str = "fp64"
    eigensolver: int = -1  # 0 dense, 1 Chebyshev-filtered subspace iteration, -1 by size
    cheb_degree: int = 0
    verbose: bool = False
    nspin: int = 1
    starting_magnetization: dict = field(default_factory=dict)
    tot_magnetization: float | None = None
    mixing_mode: str = "plain"
    kerker_q_tf: float = 0.0


class CrystalDeck:
    """A crystal plus its pseudopotentials, held in the C layout pwk_scf_run reads."""

    def __init__(self, lattice, frac, species, upf_paths):
        self.lattice = _f(lattice).reshape(3, 3)
        self.frac = _f(frac).reshape(-1, 3)
        self.species = _i(species)
        self._keep = []
        self.upfs = [read_upf(p) for p in upf_paths]
        sp = (Species * len(self.upfs))()
        for i, u in enumerate(self.upfs):
            arrs = {k: (_f(u[k]) if u.get(k) is not None else None)
                    for k in ("r", "rab", "v_local", "rho_atom", "core")}
            beta = _f(u["chi"]).reshape(-1)
            dij = _f(u["dij"]).reshape(-1)
            pl = _i(u["l"])
            self._keep += [beta, dij, pl] + [a for a in arrs.values() if a is not None]
            sp[i] = Species(u["atomic_number"], u["z_valence"], len(arrs["r"]), _pd(arrs["r"]), _pd(arrs["rab"]),
                            _pd(arrs["v_local"]), len(u["l"]), _pi(pl), _pd(beta), _pd(dij), _pd(arrs["rho_atom"]),
                            _pd(arrs["core"]))
        self._sp = sp
        self.c = Crystal((C.c_double * 9)(*self.lattice.reshape(9)), len(self.species), _pd(self.frac),
                         _pi(self.species), len(self.upfs), sp)

    @property
    def z(self):
        return np.array([self.upfs[s]["atomic_number"] for s in self.species], dtype=np.int32)


_MIXING = {"plain": 0, "kerker": 1}


def _params(p: ScfParams):
    if p.mixing_mode not in _MIXING:
        raise PwkError(1, f"mixing_mode must be one of {sorted(_MIXING)}")
    mag = [0.0] * MAX_SPECIES
    for sp, m in p.starting_magnetization.items():
        if not 0 <= int(sp) < MAX_SPECIES:
            raise PwkError(1, f"species index {sp} out of range")
        mag[int(sp)] = float(m)
    return Params(p.ecut_ev, p.ecutrho_ratio, (C.c_int * 3)(*p.fft_dims), (C.c_int * 3)(*p.kgrid),
                  (C.c_int * 3)(*p.kshift), int(p.use_symmetry), p.symprec, p.n_bands, p.max_iter, p.conv_threshold,
                  p.energy_threshold, p.mixing_beta, p.mixing_ndim, p.smearing_sigma, p.sad_sigma, _prec(p.precision),
                  int(p.eigensolver), int(p.cheb_degree), int(p.verbose), int(p.nspin),
                  (C.c_double * MAX_SPECIES)(*mag), int(p.tot_magnetization is not None),
                  float(p.tot_magnetization or 0.0), _MIXING[p.mixing_mode], float(p.kerker_q_tf))


def scf_tables(deck: CrystalDeck, p: ScfParams):
    prm = _params(p)
    npw, nk, nops = C.c_int(), C.c_int(), C.c_int()
    dims = (C.c_int * 3)()
    _check(_lib().pwk_scf_tables(C.byref(deck.c), C.byref(prm), C.byref(npw), dims, C.byref(nk), C.byref(nops)))
    return {"npw": npw.value, "dims": tuple(dims), "nk": nk.value, "nops": nops.value}


def run_scf(acc, deck: CrystalDeck, p: ScfParams):
    """One SCF solve on ``acc``'s device. Returns a dict of the ScfResult fields
    plus ``eigenvalues`` (nspin·nk, n_bands; spin-up k-points first) and ``rho``
    (the converged total density, grid order)."""
    t = scf_tables(deck, p)
    nb = min(p.n_bands, t["npw"])
    ev = np.zeros(((2 if p.nspin == 2 else 1) * t["nk"], nb))
    rho = np.zeros(int(np.prod(t["dims"])))
    prm = _params(p)
    res = Result()
    _check(_lib().pwk_scf_run(acc.h, C.byref(deck.c), C.byref(prm), C.byref(res), _pd(ev), _pd(rho)))
    out = {name: getattr(res, name) for name, _ in Result._fields_ if name != "dims"}
    out["dims"] = tuple(res.dims)
    out["eigenvalues"] = ev
    out["rho"] = rho
    return out




# ---- serving --------------------------------------------------------------------------
_SIGS.update({
    "pwk_server_create": (C.c_int, [C.c_int, C.c_int, C.c_double, C.POINTER(C.c_void_p)]),
    "pwk_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Crystal), C.POINTER(Params), C.POINTER(C.c_int64)]),
    "pwk_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Result), _dp, _dp, C.POINTER(C.c_double)]),
    "pwk_server_destroy": (C.c_int, [C.c_void_p]),
    "pwk_server_set_pass_batching": (C.c_int, [C.c_void_p, C.c_int]),
    "pwk_server_pass_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
})


class Server:
    """Request loop of SCF solves on one device (``pwk_server_*``): ``submit``
    returns a ticket at once; ``wait`` blocks until that solve is done and
    returns the same dict as :func:`run_scf` plus ``latency_s``."""

    def __init__(self, device=0, n_slots=8, hbm_fraction=0.8):
        h = C.c_void_p()
        _check(_lib().pwk_server_create(device, n_slots, hbm_fraction, C.byref(h)))
        self.h = h
        self._pending = {}

    def submit(self, deck: CrystalDeck, p: ScfParams):
        prm = _params(p)
        t = C.c_int64()
        _check(_lib().pwk_server_submit(self.h, C.byref(deck.c), C.byref(prm), C.byref(t)))
        tab = scf_tables(deck, p)
        self._pending[t.value] = (deck, prm, tab, min(p.n_bands, tab["npw"]), 2 if p.nspin == 2 else 1)
        return t.value

    def wait(self, ticket):
        deck, prm, t, nb, ns = self._pending.pop(ticket)
        ev = np.zeros((ns * t["nk"], nb))
        rho = np.zeros(int(np.prod(t["dims"])))
        res = Result()
        lat = C.c_double()
        _check(_lib().pwk_server_wait(self.h, ticket, C.byref(res), _pd(ev), _pd(rho), C.byref(lat)))
        out = {name: getattr(res, name) for name, _ in Result._fields_ if name != "dims"}
        out["dims"] = tuple(res.dims)
        out["eigenvalues"] = ev
        out["rho"] = rho
        out["latency_s"] = lat.value
        return out

    def set_pass_batching(self, window_us: int):
        """Gathering window of cross-request filter-pass batches in microseconds (0: off)."""
        _check(_lib().pwk_server_set_pass_batching(self.h, int(window_us)))

    def pass_stats(self):
        """(batched pass launches, solve passes they carried) since creation."""
        a, b = C.c_int64(), C.c_int64()
        _check(_lib().pwk_server_pass_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            _check(_lib().pwk_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
