"""ctypes binding of the FP64 CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, always as the checker. The product library
(libpwk) never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


def build() -> None:
    """Compile the oracle with its Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_const": (C.c_double, [C.c_char_p]),
            "orc_fft_grid_size": (C.c_int, [C.c_int]),
            "orc_bessel": (C.c_double, [C.c_int, C.c_double]),
            "orc_ylm": (None, [_dp, C.c_int, _dp]),
            "orc_basis": (C.c_int, [_dp, C.c_double, C.c_int, _ip, _dp]),
            "orc_lda_xc": (None, [C.c_long, _dp, _dp, _dp]),
            "orc_lsda_xc": (None, [C.c_long, _dp, _dp, _dp, _dp, _dp]),
            "orc_hartree": (None, [C.c_long, _dp, _dp, _dp]),
            "orc_veff": (None, [C.c_long, _dp, _dp, _dp, _dp]),
            "orc_fft3d": (None, [_ip, _dp, C.c_int]),
            "orc_eigh": (C.c_int, [C.c_int, _dp, _dp, _dp]),
            "orc_upf_load": (C.c_void_p, [C.c_char_p]),
            "orc_upf_free": (None, [C.c_void_p]),
            "orc_upf_info": (None, [C.c_void_p, _dp]),
            "orc_upf_proj_l": (C.c_int, [C.c_void_p, C.c_int]),
            "orc_upf_dij": (None, [C.c_void_p, _dp]),
            "orc_upf_array": (None, [C.c_void_p, C.c_int, _dp]),
            "orc_upf_vloc_g": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
            "orc_upf_beta_ff": (C.c_double, [C.c_void_p, C.c_int, C.c_double]),
            "orc_simpson": (C.c_double, [C.c_long, _dp, _dp]),
            "orc_sys_new": (C.c_void_p, [_dp, C.c_int, _ip, _ip, _dp, C.c_int, C.POINTER(C.c_char_p),
                                          C.c_double, C.c_int, _ip]),
            "orc_sys_free": (None, [C.c_void_p]),
            "orc_sys_info": (None, [C.c_void_p, _dp]),
            "orc_sys_miller": (None, [C.c_void_p, _ip]),
            "orc_sys_gvec": (None, [C.c_void_p, _dp]),
            "orc_sys_g_to_fft": (None, [C.c_void_p, _lp]),
            "orc_sys_g2": (None, [C.c_void_p, _dp]),
            "orc_sys_vloc": (None, [C.c_void_p, _dp]),
            "orc_sys_core": (C.c_int, [C.c_void_p, _dp]),
            "orc_sys_ewald": (C.c_double, [C.c_void_p]),
            "orc_sys_sad": (None, [C.c_void_p, C.c_double, _dp]),
            "orc_sys_kinetic": (None, [C.c_void_p, _dp, _dp]),
            "orc_sys_nonlocal": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _ip]),
            "orc_sys_dense_h": (None, [C.c_void_p, _dp, _dp, C.c_int, _dp]),
            "orc_sys_hpsi": (None, [C.c_void_p, _dp, _dp, C.c_int, C.c_int, _dp, _dp]),
            "orc_hpsi_with_nl": (None, [C.c_void_p, _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp]),
            "orc_sys_density": (None, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
            "orc_symmetry": (C.c_int, [_dp, C.c_int, _ip, _dp, C.c_double, C.c_int, _ip, _dp]),
            "orc_kpoints": (C.c_int, [_dp, C.c_int, _ip, _dp, _ip, _ip, C.c_int, C.c_double, _dp, _dp]),
            "orc_scf": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_set_threads": (C.c_int, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def set_threads(n: int) -> int:
    """Sets the OpenMP thread count of the oracle; returns the count in effect."""
    return lib().orc_set_threads(n)


def const(name: str) -> float:
    return lib().orc_const(name.encode())


def fft_grid_size(n_max: int) -> int:
    return lib().orc_fft_grid_size(n_max)


def bessel(l: int, x: float) -> float:
    return lib().orc_bessel(l, x)


def ylm(q, lmax):
    out = np.zeros((lmax + 1) ** 2)
    lib().orc_ylm(_d(f64(q)), lmax, _d(out))
    return out


def basis(lattice, ecut_ev):
    lat = f64(lattice).reshape(9)
    n = lib().orc_basis(_d(lat), ecut_ev, 0, None, None)
    m = np.zeros((n, 3), dtype=np.int32)
    g = np.zeros((n, 3))
    lib().orc_basis(_d(lat), ecut_ev, n, _i(m), _d(g))
    return m, g


def lda_xc(rho):
    rho = f64(rho)
    exc = np.empty_like(rho)
    vxc = np.empty_like(rho)
    lib().orc_lda_xc(rho.size, _d(rho), _d(exc), _d(vxc))
    return exc, vxc


def lsda_xc(ru, rd):
    ru, rd = f64(ru), f64(rd)
    e, u, d = np.empty_like(ru), np.empty_like(ru), np.empty_like(ru)
    lib().orc_lsda_xc(ru.size, _d(ru), _d(rd), _d(e), _d(u), _d(d))
    return e, u, d


def hartree(rho_g, g2):
    rho_g = c128(rho_g)
    out = np.empty_like(rho_g)
    lib().orc_hartree(rho_g.size, _d(rho_g.view(np.float64)), _d(f64(g2)), _d(out.view(np.float64)))
    return out


def veff(vl, vh, vxc):
    vl, vh, vxc = c128(vl), c128(vh), c128(vxc)
    out = np.empty_like(vl)
    lib().orc_veff(vl.size, _d(vl.view(np.float64)), _d(vh.view(np.float64)), _d(vxc.view(np.float64)),
                   _d(out.view(np.float64)))
    return out


def fft3d(data, sign):
    a = c128(data).copy()
    dims = np.array(a.shape, dtype=np.int32)
    lib().orc_fft3d(_i(dims), _d(a.view(np.float64)), sign)
    return a


def eigh(h):
    h = np.asfortranarray(h, dtype=np.complex128)
    n = h.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n), dtype=np.complex128, order="F")
    rc = lib().orc_eigh(n, _d(h.reshape(-1, order="F").view(np.float64)), _d(w),
                        _d(v.reshape(-1, order="F").view(np.float64)))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return w, v


class Upf:
    def __init__(self, path):
        self.h = lib().orc_upf_load(path.encode())
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        info = np.zeros(5)
        lib().orc_upf_info(self.h, _d(info))
        self.z_valence, self.l_max = info[0], int(info[1])
        self.mesh, self.n_proj, self.has_nlcc = int(info[2]), int(info[3]), bool(info[4])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_upf_free(self.h)

    def proj_l(self, i):
        return lib().orc_upf_proj_l(self.h, i)

    def dij(self):
        d = np.zeros(self.n_proj ** 2)
        lib().orc_upf_dij(self.h, _d(d))
        return d.reshape(self.n_proj, self.n_proj)

    def array(self, which):
        out = np.zeros(self.mesh)
        lib().orc_upf_array(self.h, which, _d(out))
        return out

    def vloc_g(self, g, omega):
        return lib().orc_upf_vloc_g(self.h, g, omega)

    def beta_ff(self, i, q):
        return lib().orc_upf_beta_ff(self.h, i, q)


class System:
    """Oracle counterpart of ScfContext: basis, grid maps, V_loc, ρ_core."""

    def __init__(self, lattice, species, z, frac, upf_paths, ecut_ev, ratio=4, dims=None):
        lat = f64(lattice).reshape(9)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        zz = np.ascontiguousarray(z, dtype=np.int32)
        fr = f64(frac).reshape(-1)
        arr = (C.c_char_p * len(upf_paths))(*[p.encode() for p in upf_paths])
        d = np.ascontiguousarray(dims, dtype=np.int32) if dims is not None else None
        self.h = lib().orc_sys_new(_d(lat), len(sp), _i(sp), _i(zz), _d(fr), len(upf_paths), arr, ecut_ev, ratio,
                                   _i(d) if d is not None else None)
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        info = np.zeros(6)
        lib().orc_sys_info(self.h, _d(info))
        self.npw = int(info[0])
        self.dims = (int(info[1]), int(info[2]), int(info[3]))
        self.n_grid = self.dims[0] * self.dims[1] * self.dims[2]
        self.omega, self.n_electrons = info[4], info[5]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sys_free(self.h)

    def miller(self):
        m = np.zeros((self.npw, 3), dtype=np.int32)
        lib().orc_sys_miller(self.h, _i(m))
        return m

    def gvec(self):
        g = np.zeros((self.npw, 3))
        lib().orc_sys_gvec(self.h, _d(g))
        return g

    def g_to_fft(self):
        out = np.zeros(self.npw, dtype=np.int64)
        lib().orc_sys_g_to_fft(self.h, out.ctypes.data_as(_lp))
        return out

    def g2(self):
        out = np.zeros(self.n_grid)
        lib().orc_sys_g2(self.h, _d(out))
        return out

    def vloc(self):
        out = np.zeros(self.n_grid, dtype=np.complex128)
        lib().orc_sys_vloc(self.h, _d(out.view(np.float64)))
        return out

    def core(self):
        out = np.zeros(self.n_grid)
        return out if lib().orc_sys_core(self.h, _d(out)) else None

    def ewald(self):
        return lib().orc_sys_ewald(self.h)

    def sad(self, sigma=1.0):
        out = np.zeros(self.n_grid)
        lib().orc_sys_sad(self.h, sigma, _d(out))
        return out

    def kinetic(self, k):
        out = np.zeros(self.npw)
        lib().orc_sys_kinetic(self.h, _d(f64(k)), _d(out))
        return out

    def nonlocal_(self, k):
        """(B npw×nch, DBH nch×npw, meta nch×4)."""
        kk = f64(k)
        nch = lib().orc_sys_nonlocal(self.h, _d(kk), None, None, None)
        B = np.zeros((nch, self.npw), dtype=np.complex128)  # column-major B == row-major B^T
        DBH = np.zeros((self.npw, nch), dtype=np.complex128)
        meta = np.zeros((nch, 4), dtype=np.int32)
        lib().orc_sys_nonlocal(self.h, _d(kk), _d(B.view(np.float64)), _d(DBH.view(np.float64)), _i(meta))
        return B.T.copy(), DBH.T.copy(), meta

    def dense_h(self, k, veff_g, with_nl=True):
        h = np.zeros((self.npw, self.npw), dtype=np.complex128)  # filled column-major
        lib().orc_sys_dense_h(self.h, _d(f64(k)), _d(c128(veff_g).view(np.float64)), int(with_nl),
                              _d(h.view(np.float64)))
        return h.T.copy()

    def hpsi(self, k, veff_g, psi, with_nl=True):
        """psi: (nb, npw) complex128 (band-major). Returns (nb, npw)."""
        psi = c128(psi)
        out = np.zeros_like(psi)
        lib().orc_sys_hpsi(self.h, _d(f64(k)), _d(c128(veff_g).view(np.float64)), int(with_nl), psi.shape[0],
                           _d(psi.view(np.float64)), _d(out.view(np.float64)))
        return out

    def hpsi_with_nl(self, k, veff_g, B, DBH, psi):
        """B: (npw, nch), DBH: (nch, npw); psi (nb, npw)."""
        psi = c128(psi)
        out = np.zeros_like(psi)
        nch = B.shape[1] if B is not None else 0
        Bc = c128(B.T) if nch else np.zeros(1, np.complex128)
        Dc = c128(DBH.T) if nch else np.zeros(1, np.complex128)
        lib().orc_hpsi_with_nl(self.h, _d(f64(k)), _d(c128(veff_g).view(np.float64)), nch,
                               _d(Bc.view(np.float64)), _d(Dc.view(np.float64)), psi.shape[0],
                               _d(psi.view(np.float64)), _d(out.view(np.float64)))
        return out

    def density(self, psi, fw, rho_acc=None):
        psi = c128(psi)
        acc = np.zeros(self.n_grid) if rho_acc is None else f64(rho_acc).copy()
        lib().orc_sys_density(self.h, psi.shape[0], _d(psi.view(np.float64)), _d(f64(fw)), _d(acc))
        return acc

    def scf(self, kpts, weights, n_bands=8, max_iter=100, conv=1e-6, ethr=1e-5, beta=0.3, ndim=8, sigma=0.01,
            gaussian_sigma=1.0, use_sym=False, sym_tol=1e-5, verbose=False, nspin=1, start_mag=0.0,
            tot_mag=None, kerker=False, q_tf=0.0):
        """nspin=2: LSDA; start_mag is the mean per-atom starting magnetisation, tot_mag a fixed
        total moment (one Fermi level per channel). Eigenvalues are (nspin·nk, nb), up channel first.
        kerker: Kerker-preconditioned Anderson (q_tf ≤ 0: Thomas-Fermi estimate)."""
        k = f64(kpts).reshape(-1)
        w = f64(weights)
        nk = len(w)
        params = f64([n_bands, max_iter, conv, ethr, beta, ndim, sigma, gaussian_sigma, float(use_sym), sym_tol,
                      float(verbose), nspin, start_mag, float(tot_mag is not None), tot_mag or 0.0,
                      float(kerker), q_tf])
        out = np.zeros(19)
        nb = min(n_bands, self.npw)
        evals = np.zeros((nspin * nk, nb))
        rho = np.zeros(self.n_grid)
        rc = lib().orc_scf(self.h, nk, _d(k), _d(w), _d(params), _d(out), _d(evals), _d(rho))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        keys = ["e_total", "e_harris", "fermi", "ts", "n_iter", "converged", "final_delta", "e_band", "e_kinetic",
                "e_local", "e_nonlocal", "e_hartree", "e_xc", "e_vxc", "e_ewald", "e_smearing", "magnetization",
                "n_up", "n_dn"]
        res = dict(zip(keys, out.tolist()))
        res["eigenvalues"] = evals
        res["rho"] = rho
        return res


def symmetry(lattice, z, frac, tol=1e-5):
    lat = f64(lattice).reshape(9)
    zz = np.ascontiguousarray(z, dtype=np.int32)
    fr = f64(frac).reshape(-1)
    n = lib().orc_symmetry(_d(lat), len(zz), _i(zz), _d(fr), tol, 0, None, None)
    rot = np.zeros((n, 3, 3), dtype=np.int32)
    tau = np.zeros((n, 3))
    lib().orc_symmetry(_d(lat), len(zz), _i(zz), _d(fr), tol, n, _i(rot), _d(tau))
    return rot, tau


def kpoints(lattice, z, frac, grid, shift=(0, 0, 0), mode=2, tol=1e-5):
    """mode 0: full MP grid; 1: identity-only reduction; 2: symmetry (+TR) reduction."""
    lat = f64(lattice).reshape(9)
    zz = np.ascontiguousarray(z, dtype=np.int32)
    fr = f64(frac).reshape(-1)
    g = np.ascontiguousarray(grid, dtype=np.int32)
    s = np.ascontiguousarray(shift, dtype=np.int32)
    n = lib().orc_kpoints(_d(lat), len(zz), _i(zz), _d(fr), _i(g), _i(s), mode, tol, None, None)
    k = np.zeros((n, 3))
    w = np.zeros(n)
    lib().orc_kpoints(_d(lat), len(zz), _i(zz), _d(fr), _i(g), _i(s), mode, tol, _d(k), _d(w))
    return k, w
