"""Python binding of libpwk (include/pwk.h) over ctypes.

Mirrors the reference's GPU backend (pwdft/pwdft-core/src/gpu/mod.rs,
``GpuAccelerator``) and its Hamiltonian-apply calls
(src/scf/potentials.rs:163-190, src/potential/nonlocal.rs:432-450,
src/scf/density.rs:41-111). Host-array methods take and return numpy arrays;
``*_dev`` methods take device pointers (e.g. ``torch.Tensor.data_ptr()``).

There is no CPU fallback: if libpwk.so is missing or no GPU is present,
construction raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(HERE, "libpwk.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pwk.h")

PWK_FP64, PWK_FP32, PWK_SPLIT_BF16 = 0, 1, 2
PRECISIONS = {"fp64": PWK_FP64, "fp32": PWK_FP32, "split_bf16": PWK_SPLIT_BF16}
STATUS = {0: "PWK_OK", 1: "PWK_E_INVALID", 2: "PWK_E_NO_DEVICE", 3: "PWK_E_DEVICE", 4: "PWK_E_OOM",
          5: "PWK_E_UNSUPPORTED", 6: "PWK_E_STATE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p
_lib = None


class PwkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def header_symbols():
    """Entry points declared in include/pwk.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pwk_[a-z0-9_]+)\s*\(", text)))


_SIGS = {
    "pwk_last_error": (C.c_char_p, []),
    "pwk_version": (C.c_int, []),
    "pwk_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pwk_ctx_destroy": (C.c_int, [_vp]),
    "pwk_ctx_sync": (C.c_int, [_vp]),
    "pwk_ctx_stream": (_vp, [_vp]),
    "pwk_ctx_set_fused_hpsi": (C.c_int, [_vp, C.c_int]),
    "pwk_ctx_set_graphs": (C.c_int, [_vp, C.c_int]),
    "pwk_ctx_set_resident": (C.c_int, [_vp, C.c_int]),
    "pwk_ctx_set_fft_autotune": (C.c_int, [_vp, C.c_int]),
    "pwk_ctx_set_fft_plan": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "pwk_fft_plan_query": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), _dp, _dp, C.POINTER(C.c_int)]),
    "pwk_dev_zgemm": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_longlong, _vp, C.c_int,
                                C.c_longlong, _vp, C.c_int, C.c_longlong, C.c_int, C.c_int, C.c_int]),
    "pwk_profile_enable": (C.c_int, [_vp, C.c_int]),
    "pwk_profile_reset": (C.c_int, [_vp]),
    "pwk_profile_read": (C.c_int, [_vp, C.c_int, C.c_char_p, _dp, C.POINTER(C.c_int64)]),
    "pwk_prepare": (C.c_int, [_vp, _ip, _dp]),
    "pwk_hartree": (C.c_int, [_vp, _dp, C.c_double, _dp, C.c_int]),
    "pwk_veff": (C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_int]),
    "pwk_hartree_veff": (C.c_int, [_vp, _dp, C.c_double, _dp, _dp, _dp, C.c_int]),
    "pwk_lda_xc": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int]),
    "pwk_lsda_xc": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_int]),
    "pwk_dev_hartree_veff": (C.c_int, [_vp, _vp, C.c_double, _vp, _vp, _vp, C.c_int]),
    "pwk_dev_lda_xc": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int]),
    "pwk_dev_lsda_xc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
    "pwk_fft3d": (C.c_int, [_vp, _dp, C.c_int, C.c_int]),
    "pwk_dev_fft3d": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int]),
    "pwk_set_basis": (C.c_int, [_vp, C.c_int, _ip]),
    "pwk_set_nk": (C.c_int, [_vp, C.c_int]),
    "pwk_set_kinetic": (C.c_int, [_vp, C.c_int, _dp]),
    "pwk_set_veff": (C.c_int, [_vp, _dp]),
    "pwk_dev_set_veff": (C.c_int, [_vp, _vp]),
    "pwk_set_nonlocal": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "pwk_build_nonlocal": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int, _dp, _ip, C.c_int, _ip, _ip, _dp, _ip, _dp,
                                     _dp, _dp, C.c_double]),
    "pwk_nonlocal_channels": (C.c_int, [_vp, C.c_int]),
    "pwk_hpsi": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, C.c_int]),
    "pwk_dev_hpsi": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int]),
    "pwk_dense_h": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "pwk_dev_dense_h": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
    "pwk_dense_eigh": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]),
    "pwk_dev_dense_eigh": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "pwk_dev_dense_eigh_noalloc": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "pwk_density_accumulate": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, C.c_int]),
    "pwk_dev_density_accumulate": (C.c_int, [_vp, C.c_int, _vp, _dp, _vp, C.c_int]),
    "pwk_upf_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "pwk_upf_parse": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(_vp)]),
    "pwk_upf_species": (C.c_int, [_vp, _vp]),
    "pwk_upf_info": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_int)]),
    "pwk_upf_free": (C.c_int, [_vp]),
    "pwk_hpsi_batch": (C.c_int, [_vp, _vp, _dp, _dp, C.c_int]),
    "pwk_dev_hpsi_batch": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int]),
    "pwk_density_accumulate_batch": (C.c_int, [_vp, _vp, _dp, _dp, _dp, C.c_int]),
    "pwk_dev_density_accumulate_batch": (C.c_int, [_vp, _vp, _vp, _dp, _vp, C.c_int]),
    "pwk_subspace_rr": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, _dp, _dp, _ip, _dp]),
    "pwk_dev_subspace_rr": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_double, _vp, _vp, _ip, _dp]),
}


class KBatch(C.Structure):
    """pwk_kbatch: per entry a k-point slot and a band count; the entries' bands are packed
    back to back."""
    _fields_ = [("nentries", C.c_int), ("ik", _ip), ("nbands", _ip)]


def kbatch(iks, nbands):
    """A pwk_kbatch over (k-point slot, band count) pairs; keeps its arrays alive."""
    ik = np.ascontiguousarray(iks, dtype=np.int32)
    nb = np.ascontiguousarray(nbands, dtype=np.int32)
    if ik.shape != nb.shape or ik.ndim != 1:
        raise ValueError("one band count per k-point entry")
    b = KBatch(len(ik), ik.ctypes.data_as(_ip), nb.ctypes.data_as(_ip))
    b._keep = (ik, nb)
    return b


def lib():
    """Loads libpwk.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing; build it with pwdft-rs_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise PwkError(code, lib().pwk_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _prec(p):
    return PRECISIONS[p] if isinstance(p, str) else int(p)


class GpuAccelerator:
    """Counterpart of the reference's ``GpuAccelerator`` (src/gpu/mod.rs:36-44).

    ``try_new`` returns None when no device is present, like the reference;
    the constructor raises instead.
    """

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().pwk_ctx_create(device, C.byref(h)))
        self.h = h
        self.dims = None
        self.n_grid = 0
        self.npw = 0

    @classmethod
    def try_new(cls, device: int = 0):
        try:
            return cls(device)
        except PwkError as e:
            if e.code == 2:
                return None
            raise

    def close(self):
        if getattr(self, "h", None):
            lib().pwk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().pwk_ctx_stream(self.h)

    def sync(self):
        _check(lib().pwk_ctx_sync(self.h))

    def zgemm_dev(self, m, n, k, a_ptr, lda, b_ptr, ldb, c_ptr, ldc, precision="fp64", accumulate=False, batch=1,
                  strides=(0, 0, 0)):
        """C (+)= A·B on device pointers (column-major), pwk_dev_zgemm."""
        _check(lib().pwk_dev_zgemm(self.h, m, n, k, a_ptr, lda, strides[0], b_ptr, ldb, strides[1], c_ptr, ldc,
                                   strides[2], batch, int(accumulate), _prec(precision)))

    def set_graphs(self, on=True):
        """Replay captured filter passes in SCF solves (opt-in) or launch kernel by kernel (default)."""
        _check(lib().pwk_ctx_set_graphs(self.h, int(on)))

    def set_resident(self, on=True):
        """Device-resident SCF iterations (default on) or the host-synchronised loop."""
        _check(lib().pwk_ctx_set_resident(self.h, int(on)))

    def set_fft_autotune(self, on=True):
        """Tune the full-grid FFT plan on first use of a grid (default on)."""
        _check(lib().pwk_ctx_set_fft_autotune(self.h, int(on)))

    def set_fft_plan(self, ord=-1, tpb=64, zlds=0):
        """Force the full-grid FFT plan (radix order, lanes per workgroup, LDS-staged
        contiguous pass); ord < 0 returns to the tuned plan."""
        _check(lib().pwk_ctx_set_fft_plan(self.h, int(ord), int(tpb), int(zlds)))

    def fft_plan(self, batch=1, precision="fp64"):
        """The plan this context uses for its grid: dict(ord, tpb, zlds, us_tuned, us_default,
        candidates); tunes it now if it has not been tuned in this process."""
        plan = (C.c_int * 3)()
        tb, td, nc = C.c_double(), C.c_double(), C.c_int()
        _check(lib().pwk_fft_plan_query(self.h, int(batch), _prec(precision), plan, C.byref(tb), C.byref(td),
                                        C.byref(nc)))
        return {"ord": plan[0], "tpb": plan[1], "zlds": plan[2], "us_tuned": tb.value, "us_default": td.value,
                "candidates": nc.value}

    def set_fused_hpsi(self, on=True):
        """Fused one-launch H·psi on small cubic grids: True / 1 (default) also runs each
        Chebyshev filter pass as one launch, 2 keeps one launch per filter step, False / 0
        selects the z-stick path."""
        _check(lib().pwk_ctx_set_fused_hpsi(self.h, int(on)))

    # ---- kernel timing -----------------------------------------------------------
    def profile(self, on=True):
        _check(lib().pwk_profile_enable(self.h, int(on)))

    def profile_reset(self):
        _check(lib().pwk_profile_reset(self.h))

    def profile_read(self):
        """{kernel name: (total ms, launches)} since the last reset."""
        n = lib().pwk_profile_read(self.h, 0, None, None, None)
        if n < 0:
            _check(-n)
        names = C.create_string_buffer(64 * max(n, 1))
        ms = np.zeros(max(n, 1))
        cnt = np.zeros(max(n, 1), dtype=np.int64)
        lib().pwk_profile_read(self.h, n, names, _d(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64)))
        raw = names.raw
        out = {}
        for i in range(n):
            nm = raw[64 * i:64 * i + 64].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(cnt[i]))
        return out

    # ---- src/gpu/mod.rs:197-364 -------------------------------------------------
    def prepare_buffers(self, dims, g_squared):
        d = np.ascontiguousarray(dims, dtype=np.int32)
        g2 = _f64(g_squared)
        assert g2.size == int(np.prod(d))
        _check(lib().pwk_prepare(self.h, d.ctypes.data_as(_ip), _d(g2)))
        self.dims = tuple(int(x) for x in d)
        self.n_grid = int(np.prod(d))

    # ---- src/gpu/mod.rs:398-494 -------------------------------------------------
    def hartree_potential(self, rho_g, fourpi_e2, precision="fp64"):
        r = _c128(rho_g)
        out = np.empty_like(r)
        _check(lib().pwk_hartree(self.h, _d(r.view(np.float64)), fourpi_e2, _d(out.view(np.float64)),
                                 _prec(precision)))
        return out

    # ---- src/gpu/mod.rs:508-610 -------------------------------------------------
    def v_eff_assembly(self, v_local, v_h, v_xc, precision="fp64"):
        a, b, c = _c128(v_local), _c128(v_h), _c128(v_xc)
        out = np.empty_like(a)
        _check(lib().pwk_veff(self.h, _d(a.view(np.float64)), _d(b.view(np.float64)), _d(c.view(np.float64)),
                              _d(out.view(np.float64)), _prec(precision)))
        return out

    def hartree_veff(self, rho_g, fourpi_e2, v_local, v_xc, precision="fp64"):
        r, a, c = _c128(rho_g), _c128(v_local), _c128(v_xc)
        out = np.empty_like(a)
        _check(lib().pwk_hartree_veff(self.h, _d(r.view(np.float64)), fourpi_e2, _d(a.view(np.float64)),
                                      _d(c.view(np.float64)), _d(out.view(np.float64)), _prec(precision)))
        return out

    # ---- src/gpu/mod.rs:627-725 -------------------------------------------------
    def lda_xc(self, rho_r, precision="fp64"):
        r = _f64(rho_r)
        e, v = np.empty_like(r), np.empty_like(r)
        _check(lib().pwk_lda_xc(self.h, _d(r), _d(e), _d(v), _prec(precision)))
        return e, v

    def lsda_xc(self, rho_up, rho_dn, precision="fp64"):
        u, d = _f64(rho_up), _f64(rho_dn)
        e, vu, vd = np.empty_like(u), np.empty_like(u), np.empty_like(u)
        _check(lib().pwk_lsda_xc(self.h, _d(u), _d(d), _d(e), _d(vu), _d(vd), _prec(precision)))
        return e, vu, vd

    def fft3d(self, data, sign, precision="fp64"):
        a = _c128(data).copy()
        _check(lib().pwk_fft3d(self.h, _d(a.view(np.float64)), int(sign), _prec(precision)))
        return a

    # ---- Hamiltonian apply ----------------------------------------------------------
    def set_basis(self, miller):
        m = np.ascontiguousarray(miller, dtype=np.int32)
        _check(lib().pwk_set_basis(self.h, m.shape[0], m.ctypes.data_as(_ip)))
        self.npw = int(m.shape[0])

    def set_nk(self, nk):
        _check(lib().pwk_set_nk(self.h, int(nk)))

    def set_kinetic(self, ik, t):
        _check(lib().pwk_set_kinetic(self.h, ik, _d(_f64(t))))

    def set_veff(self, v_eff_g):
        v = _c128(v_eff_g)
        _check(lib().pwk_set_veff(self.h, _d(v.view(np.float64))))

    def set_veff_dev(self, ptr):
        _check(lib().pwk_dev_set_veff(self.h, ptr))

    def set_nonlocal(self, ik, B, DBH):
        """B: (npw, nch) complex; DBH: (nch, npw) complex (the reference's faer layouts)."""
        if B is None or B.shape[1] == 0:
            _check(lib().pwk_set_nonlocal(self.h, ik, 0, None, None))
            return
        Bc = _c128(np.asarray(B).T)   # column-major npw×nch
        Dc = _c128(np.asarray(DBH).T)  # column-major nch×npw
        _check(lib().pwk_set_nonlocal(self.h, ik, B.shape[1], _d(Bc.view(np.float64)), _d(Dc.view(np.float64))))

    def build_nonlocal(self, ik, k, gvec, tau, atom_species, species, omega):
        """Builds B and D·B^H of k-point ik on the device (src/potential/nonlocal.rs:190-382).

        species: list of dicts with keys l (projector l's), dij (nproj×nproj, eV),
        r, rab (radial mesh, Å) and chi (nproj×mesh, χ = r·β in Å^-1/2).
        Returns the number of channels.
        """
        nproj = [len(s["l"]) for s in species]
        proj_l = np.concatenate([np.asarray(s["l"], dtype=np.int32) for s in species])
        dij = np.concatenate([_f64(s["dij"]).reshape(-1) for s in species])
        mesh = [len(s["r"]) for s in species]
        r = np.concatenate([_f64(s["r"]) for s in species])
        rab = np.concatenate([_f64(s["rab"]) for s in species])
        chi = np.concatenate([_f64(s["chi"]).reshape(-1) for s in species])
        return self.build_nonlocal_raw(ik, k, gvec, tau, atom_species, nproj, proj_l, dij, mesh, r, rab, chi, omega)

    def build_nonlocal_raw(self, ik, k, gvec, tau, atom_species, nproj, proj_l, dij, mesh, r, rab, chi, omega):
        asp = np.ascontiguousarray(atom_species, dtype=np.int32)
        npj = np.ascontiguousarray(nproj, dtype=np.int32)
        pl = np.ascontiguousarray(proj_l, dtype=np.int32)
        ms = np.ascontiguousarray(mesh, dtype=np.int32)
        _check(lib().pwk_build_nonlocal(self.h, ik, _d(_f64(k)), _d(_f64(gvec).reshape(-1)), len(asp),
                                        _d(_f64(tau).reshape(-1)), asp.ctypes.data_as(_ip), len(npj),
                                        npj.ctypes.data_as(_ip), pl.ctypes.data_as(_ip), _d(_f64(dij)),
                                        ms.ctypes.data_as(_ip), _d(_f64(r)), _d(_f64(rab)), _d(_f64(chi)),
                                        float(omega)))
        return lib().pwk_nonlocal_channels(self.h, ik)

    def hpsi(self, ik, psi, precision="fp64"):
        """psi: (nb, npw) complex128, band-major. Returns (nb, npw)."""
        p = _c128(psi)
        out = np.empty_like(p)
        _check(lib().pwk_hpsi(self.h, ik, p.shape[0], _d(p.view(np.float64)), _d(out.view(np.float64)),
                              _prec(precision)))
        return out

    def hpsi_dev(self, ik, nb, psi_ptr, hpsi_ptr, precision="fp32"):
        _check(lib().pwk_dev_hpsi(self.h, ik, nb, psi_ptr, hpsi_ptr, _prec(precision)))

    def hpsi_batch(self, iks, nbands, psi, precision="fp64"):
        """H·psi of a k × band batch (pwk_hpsi_batch): entry e holds nbands[e] bands of k-point
        slot iks[e]; psi (sum(nbands), npw) with the entries' bands back to back."""
        p = _c128(psi)
        b = kbatch(iks, nbands)
        if p.shape[0] != int(np.sum(nbands)):
            raise ValueError("psi rows must equal the batch's band count")
        out = np.empty_like(p)
        _check(lib().pwk_hpsi_batch(self.h, C.byref(b), _d(p.view(np.float64)), _d(out.view(np.float64)),
                                    _prec(precision)))
        return out

    def hpsi_batch_dev(self, iks, nbands, psi_ptr, hpsi_ptr, precision="fp32"):
        b = kbatch(iks, nbands)
        _check(lib().pwk_dev_hpsi_batch(self.h, C.byref(b), psi_ptr, hpsi_ptr, _prec(precision)))

    def density_batch(self, iks, nbands, psi, fw, rho=None, precision="fp64"):
        """rho += Σ_b fw_b |ψ_b(r)|² over a k × band batch (pwk_density_accumulate_batch)."""
        p = _c128(psi)
        b = kbatch(iks, nbands)
        if p.shape[0] != int(np.sum(nbands)) or len(fw) != p.shape[0]:
            raise ValueError("psi rows and fw must match the batch's band count")
        acc = np.zeros(self.n_grid) if rho is None else _f64(rho).copy()
        _check(lib().pwk_density_accumulate_batch(self.h, C.byref(b), _d(p.view(np.float64)), _d(_f64(fw)), _d(acc),
                                                  _prec(precision)))
        return acc

    def subspace_rr(self, h, v_prev, nb, tol=0.0):
        """diagonalize_subspace (src/eigensolver/dense.rs:183-260) batched over k: h (nk, n, n),
        v_prev (nk, n, nb) or None. Returns (w (nk, nb), v (nk, n, nb), used_full (nk,),
        max_residual (nk,))."""
        h = np.asarray(h, dtype=np.complex128)
        if h.ndim == 2:
            h = h[None]
        nk, n, _ = h.shape
        hc = np.ascontiguousarray(np.transpose(h, (0, 2, 1)))
        vp = None
        if v_prev is not None:
            v_prev = np.asarray(v_prev, dtype=np.complex128).reshape(nk, n, -1)[:, :, :nb]
            vp = np.ascontiguousarray(np.transpose(v_prev, (0, 2, 1)))
        w = np.empty((nk, nb))
        v = np.empty((nk, nb, n), dtype=np.complex128)
        used = np.zeros(nk, dtype=np.int32)
        res = np.zeros(nk)
        _check(lib().pwk_subspace_rr(self.h, n, nb, nk, _d(hc.view(np.float64)),
                                     _d(vp.view(np.float64)) if vp is not None else None, float(tol), _d(w),
                                     _d(v.view(np.float64)), used.ctypes.data_as(_ip), _d(res)))
        return w, np.transpose(v, (0, 2, 1)), used, res

    def dense_h(self, ik, with_nl=True):
        h = np.empty((self.npw, self.npw), dtype=np.complex128)  # filled column-major
        _check(lib().pwk_dense_h(self.h, ik, int(with_nl), _d(h.view(np.float64))))
        return h.T.copy()

    def dense_eigh(self, h, nb):
        """Lowest ``nb`` eigenpairs of each Hermitian matrix in ``h`` (nk, n, n):
        diagonalize_lowest (src/eigensolver/dense.rs:87-98) batched over k.
        Returns (w (nk, nb) ascending, v (nk, n, nb))."""
        h = np.asarray(h, dtype=np.complex128)
        if h.ndim == 2:
            h = h[None]
        nk, n, _ = h.shape
        hc = np.ascontiguousarray(np.transpose(h, (0, 2, 1)))  # column-major blocks
        w = np.empty((nk, nb))
        v = np.empty((nk, nb, n), dtype=np.complex128)
        _check(lib().pwk_dense_eigh(self.h, n, nb, nk, _d(hc.view(np.float64)), _d(w), _d(v.view(np.float64))))
        return w, np.transpose(v, (0, 2, 1))

    def density(self, psi, fw, rho=None, precision="fp64"):
        p = _c128(psi)
        acc = np.zeros(self.n_grid) if rho is None else _f64(rho).copy()
        _check(lib().pwk_density_accumulate(self.h, p.shape[0], _d(p.view(np.float64)), _d(_f64(fw)), _d(acc),
                                            _prec(precision)))
        return acc

    def density_dev(self, nb, psi_ptr, fw, rho_ptr, precision="fp32"):
        _check(lib().pwk_dev_density_accumulate(self.h, nb, psi_ptr, _d(_f64(fw)), rho_ptr, _prec(precision)))
