"""Writes oracle_decks.json: converged energy terms of FP64 oracle solves whose dense
eigensolve takes minutes on the host, so the device tests compare against these stored
values instead of re-running the oracle. Run from the repository root after building the
oracle (make -C oracle):

    python tests/golden/oracle/make_oracle_decks.py

Decks:
  si8_cubic_16ry_k222: 8-atom cubic Si cell a = 10.26 bohr, ecutwfc 16 Ry (1189 plane waves,
  25³ grid) — the batched uneven-k cell — on a 2×2×2 Γ-centred MP grid without symmetry
  (8 k-points, weight 1/8), 24 bands, FD 0.01 Ry, Anderson β 0.3 ndim 8, conv 1e-8 e/Å³ and
  1e-8 Ry.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "pwdft-rs_amd"))

import numpy as np  # noqa: E402
import pyoracle  # noqa: E402
from pwk import decks  # noqa: E402

RY = 13.605693122994
BOHR = 0.529177210903
SI_UPF = os.path.join(ROOT, "tests", "golden", "pseudo", "Si.upf")
KEYS = ["e_total", "e_harris", "fermi", "ts", "n_iter", "converged", "final_delta", "e_band", "e_kinetic", "e_local",
        "e_nonlocal", "e_hartree", "e_xc", "e_vxc", "e_ewald", "e_smearing"]


def si8_cubic_16ry_k222():
    lat, frac = decks.diamond_supercell(10.26 * BOHR)
    s = pyoracle.System(lat, [0] * 8, [14] * 8, frac, [SI_UPF], 16 * RY)
    k, w = pyoracle.kpoints(lat, [14] * 8, frac, [2, 2, 2], (0, 0, 0), mode=1)
    r = s.scf(k, w, n_bands=24, max_iter=60, conv=1e-8, ethr=1e-8 * RY, beta=0.3, ndim=8, sigma=0.01 * RY)
    out = {key: float(r[key]) for key in KEYS}
    out.update(npw=int(s.npw), dims=list(s.dims), nk=len(w), eigenvalues_k0=r["eigenvalues"][0].tolist())
    return out


def main():
    threads = pyoracle.set_threads(os.cpu_count() or 1)
    res = {}
    for fn in (si8_cubic_16ry_k222,):
        t0 = time.perf_counter()
        res[fn.__name__] = fn()
        res[fn.__name__]["oracle_seconds"] = round(time.perf_counter() - t0, 1)
        res[fn.__name__]["oracle_threads"] = threads
        print(fn.__name__, res[fn.__name__]["e_total"], res[fn.__name__]["n_iter"], flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_decks.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
