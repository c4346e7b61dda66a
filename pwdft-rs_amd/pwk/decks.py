"""Crystal decks and host-side setup arithmetic for the benchmark workloads.

Geometry follows the reference's conventions: lattice vectors as rows in Å,
fractional atom positions, FCC primitive cell (a/2)·[(0,1,1),(1,0,1),(1,1,0)]
(pwdft/pwdft-core/tests/cpu_scf_baseline.rs:21-31). The basis enumeration and
G-vector arithmetic repeat the reference's floating-point operation order
(src/basis.rs:36-63, src/crystal.rs:108-121) so the tables are bit-identical.
"""
from __future__ import annotations

import math
import os

import numpy as np

# src/consts.rs:15-49
HA_TO_EV = 27.211386245988
RY_TO_EV = 13.605693122994
BOHR_TO_ANG = 0.529177210903
E2_COULOMB = 14.399645351950548
_H_SI = 6.62607015e-34
_HBAR_SI = _H_SI / (2.0 * math.pi)
_M_E = 9.1093837015e-31
HBAR2_OVER_2M = _HBAR_SI * _HBAR_SI / (2.0 * _M_E) * 6.241509074e18 * 1e10 * 1e10

PSEUDO_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                          "tests", "golden", "pseudo")

DIAMOND_BASIS = [(0.0, 0.0, 0.0), (0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0),
                 (0.25, 0.25, 0.25), (0.25, 0.75, 0.75), (0.75, 0.25, 0.75), (0.75, 0.75, 0.25)]


def fcc_primitive(a_ang):
    h = a_ang / 2.0
    lat = np.array([[h * 0.0, h * 1.0, h * 1.0], [h * 1.0, h * 0.0, h * 1.0], [h * 1.0, h * 1.0, h * 0.0]])
    frac = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    return lat, frac


def diamond_supercell(a_ang, n=(1, 1, 1)):
    """Conventional 8-atom diamond cell repeated n1×n2×n3 (simple cubic lattice)."""
    n1, n2, n3 = n
    lat = np.array([[a_ang * n1, 0.0, 0.0], [0.0, a_ang * n2, 0.0], [0.0, 0.0, a_ang * n3]])
    frac = []
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                for f in DIAMOND_BASIS:
                    frac.append(((i + f[0]) / n1, (j + f[1]) / n2, (k + f[2]) / n3))
    return lat, np.array(frac)


def cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def volume(lat):
    return abs(dot(cross(lat[0], lat[1]), lat[2]))


def reciprocal(lat):
    a, b, c = lat
    triple = dot(cross(a, b), c)
    f = 2.0 * math.pi / triple
    return np.array([cross(b, c) * f, cross(c, a) * f, cross(a, b) * f])


def basis(lat, ecut_ev):
    """Miller indices (npw, 3) and Cartesian G (npw, 3) in the reference's order."""
    rec = reciprocal(lat)
    g_max_sq = ecut_ev / HBAR2_OVER_2M
    g_max = math.sqrt(g_max_sq)
    nmax = [int(math.ceil(g_max / math.sqrt(dot(b, b)))) for b in rec]
    r = [np.arange(-n, n + 1, dtype=np.int64) for n in nmax]
    n1, n2, n3 = np.meshgrid(r[0], r[1], r[2], indexing="ij")
    n1, n2, n3 = n1.ravel(), n2.ravel(), n3.ravel()
    f1, f2, f3 = n1.astype(np.float64), n2.astype(np.float64), n3.astype(np.float64)
    g = np.empty((n1.size, 3))
    for d in range(3):
        g[:, d] = (f1 * rec[0][d] + f2 * rec[1][d]) + f3 * rec[2][d]
    g2 = (g[:, 0] * g[:, 0] + g[:, 1] * g[:, 1]) + g[:, 2] * g[:, 2]
    keep = g2 <= g_max_sq
    miller = np.stack([n1[keep], n2[keep], n3[keep]], axis=1).astype(np.int32)
    return miller, g[keep]


def _smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def fft_grid_size(n_max):
    """src/fft.rs:257-266"""
    n = 2 * n_max + 1
    while not _smooth(n):
        n += 1
    return n


def fft_dims(miller, ratio=4):
    """src/scf/grid.rs:44-86 (no explicit override)"""
    nmax = np.abs(miller).max(axis=0)
    s = math.isqrt(ratio)
    scale = s if s * s == ratio else s + 1
    return tuple(fft_grid_size(int(scale * m)) for m in nmax)


def grid_g2(lat, dims):
    """|G|² per grid point in grid order (src/scf/grid.rs:120-129, context.rs:146-156)."""
    rec = reciprocal(lat)
    nx, ny, nz = dims
    idx = [np.arange(n) for n in dims]
    m = [np.where(i > n // 2, i - n, i).astype(np.float64) for i, n in zip(idx, dims)]
    m1, m2, m3 = np.meshgrid(m[0], m[1], m[2], indexing="ij")
    g = np.empty(m1.shape + (3,))
    for d in range(3):
        g[..., d] = (m1 * rec[0][d] + m2 * rec[1][d]) + m3 * rec[2][d]
    return ((g[..., 0] * g[..., 0] + g[..., 1] * g[..., 1]) + g[..., 2] * g[..., 2]).reshape(-1)


def kinetic(gvec, k):
    q = gvec + np.asarray(k)[None, :]
    return HBAR2_OVER_2M * ((q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2])


def read_upf(path):
    """UPF v2 pseudopotential through libpwk's reader (pwk_upf_load, the reference's
    src/pseudopotential/upf/parse.rs:20-140), as numpy arrays in eV / Å units: r, rab,
    v_local, chi (nproj × mesh, χ = r·β), dij (nproj × nproj), l, rho_atom (None when absent
    or all zero), core (None without the core-correction flag)."""
    import ctypes as C

    from . import PwkError, lib

    L = lib()
    h = C.c_void_p()
    st = L.pwk_upf_load(os.fsencode(path), C.byref(h))
    if st != 0:
        raise PwkError(st, L.pwk_last_error().decode())
    try:
        from .scf import Species

        sp = Species()
        st = L.pwk_upf_species(h, C.byref(sp))
        if st != 0:
            raise PwkError(st, L.pwk_last_error().decode())
        sym = C.create_string_buffer(4)
        lmax = C.c_int()
        L.pwk_upf_info(h, sym, C.byref(lmax))
        m, npj = sp.mesh, sp.nproj

        def arr(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if ptr and n else None

        out = {
            "element": sym.value.decode(),
            "atomic_number": sp.atomic_number,
            "z_valence": sp.z_valence,
            "l_max": lmax.value,
            "r": arr(sp.r, m),
            "rab": arr(sp.rab, m),
            "v_local": arr(sp.v_local, m),
            "l": [int(x) for x in (arr(sp.proj_l, npj) if npj else [])],
            "chi": arr(sp.beta, npj * m).reshape(npj, m) if npj else np.zeros((0, m)),
            "dij": arr(sp.dij, npj * npj).reshape(npj, npj) if npj else np.zeros((0, 0)),
            "rho_atom": arr(sp.rho_atom, m),
            "core": arr(sp.rho_core, m),
        }
    finally:
        L.pwk_upf_free(h)
    if out["rho_atom"] is not None and not np.any(np.abs(out["rho_atom"]) > 1e-20):
        out["rho_atom"] = None
    return out


def upf_path(element):
    return os.path.join(PSEUDO_DIR, f"{element}.upf")


def seeded_psi(gvec_bohr2, nb, seed):
    """ψ_b(G) = (N(0,1) + i N(0,1))·exp(-|G|²_bohr/8), each band ℓ2-normalised (band-major, complex128)."""
    rng = np.random.Generator(np.random.Philox(seed))
    npw = gvec_bohr2.shape[0]
    env = np.exp(-gvec_bohr2 / 8.0)
    psi = (rng.standard_normal((nb, npw)) + 1j * rng.standard_normal((nb, npw))) * env[None, :]
    psi /= np.linalg.norm(psi, axis=1)[:, None]
    return psi


def seeded_veff_g(dims, seed, amplitude_ev=RY_TO_EV):
    """V_eff(G) = FFT(V(r))/N for V(r) uniform in [-1, 1] Ry (in eV)."""
    rng = np.random.Generator(np.random.Philox(seed))
    vr = rng.uniform(-1.0, 1.0, size=dims) * amplitude_ev
    return (np.fft.fftn(vr) / vr.size).reshape(-1)
