"""Crystal decks and host-side setup arithmetic for the benchmark workloads.

Geometry follows the reference's conventions: lattice vectors as rows in Å,
fractional atom positions, FCC primitive cell (a/2)·[(0,1,1),(1,0,1),(1,1,0)]
(pwdft/pwdft-core/tests/cpu_scf_baseline.rs:21-31). The basis enumeration and
G-vector arithmetic repeat the reference's floating-point operation order
(src/basis.rs:36-63, src/crystal.rs:108-121) so the tables are bit-identical.
"""
from __future__ import annotations

import math
import re
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


ATOMIC_NUMBER = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10, "Na": 11,
                 "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18, "K": 19, "Ca": 20, "Ti": 22,
                 "Fe": 26, "Co": 27, "Ni": 28, "Cu": 29, "Zn": 30, "Ga": 31, "Ge": 32, "As": 33, "Mn": 25}


def read_upf(path):
    """Minimal UPF v2 reader for the radial tables (src/pseudopotential/upf/parse.rs:20-140)."""
    text = open(path).read()

    def attr(s, name):
        i = s.index(name + '="') + len(name) + 2
        return s[i:s.index('"', i)].strip()

    def block(tag, n):
        i = text.index("<" + tag)
        i = text.index(">", i) + 1
        j = text.index("</" + tag + ">", i)
        v = np.array(text[i:j].split(), dtype=np.float64)
        assert v.size == n, tag
        return v

    mesh = int(attr(text, "mesh_size"))
    nproj = int(attr(text, "number_of_proj"))
    out = {
        "element": attr(text, "element"),
        "z_valence": float(attr(text, "z_valence")),
        "r": block("PP_R", mesh) * BOHR_TO_ANG,
        "rab": block("PP_RAB", mesh) * BOHR_TO_ANG,
        "dij": block("PP_DIJ", nproj * nproj).reshape(nproj, nproj) * RY_TO_EV,
    }
    ls, chi = [], []
    for i in range(1, nproj + 1):
        tag = f"PP_BETA.{i}"
        s = text.index("<" + tag)
        head = text[s:text.index(">", s)]
        ls.append(int(attr(head, "angular_momentum")))
        chi.append(block(tag, mesh) / math.sqrt(BOHR_TO_ANG))
    out["l"] = ls
    out["chi"] = np.array(chi)
    out["v_local"] = block("PP_LOCAL", mesh) * RY_TO_EV
    out["atomic_number"] = ATOMIC_NUMBER[out["element"]]
    out["rho_atom"] = None
    if "<PP_RHOATOM" in text:
        rho = block("PP_RHOATOM", mesh) / BOHR_TO_ANG
        if np.any(np.abs(rho) > 1e-20):
            out["rho_atom"] = rho
    out["core"] = None
    if re.search(r'core_correction="[Tt]"', text) or "nlcc=.true." in text:
        out["core"] = block("PP_NLCC", mesh) / (BOHR_TO_ANG * BOHR_TO_ANG * BOHR_TO_ANG) if "<PP_NLCC" in text else np.zeros(mesh)
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
