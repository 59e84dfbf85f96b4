"""Writes the golden fixtures under tests/golden/ from the oracle (run from
the repository root: python tests/golden/make_golden.py). Each fixture holds
inputs and expected outputs only:

* park_miller.json    -- Park-Miller uniform and polar-Gaussian stream prefixes
                         from seed 1 and the published check value after
                         10,000 draws (1043618065)
* gauss_hermite.json  -- physicists' Gauss-Hermite nodes and weights, n = 1..8
* eam_cu32.npz        -- 32-atom jittered fcc Cu cell (Johnson table, Nr =
                         Nrho = 2000): forces, energy, virial, rho, half list
* dmd_cuh16.npz       -- 16-site Cu-H DMD cell: f_x, f_alpha, mu, free energy
* gcmc_nih.npz        -- a 3x3x3 Ni-H cell under SGCMC exchange at mu = -1.9
                         eV: 300 decisions and exchange energies

The oracle reproduces them bit for bit (test_golden.py); a change of the
oracle that moves any of them is a change of the restated algorithm.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
from mapp4py_amd import potentials as pot  # noqa: E402


def eam_case():
    p = pot.johnson_cu(nr=2000, nrho=2000)
    x, H = pot.fcc(3.615, (2, 2, 2))
    x = x + np.random.default_rng(42).normal(0, 0.08, x.shape)
    elem = np.zeros(len(x), np.uint8)
    md = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=0.5)
    f, e, w, rho = md.force()
    i, j, img = md.pairs()
    return dict(x=x, H=H, elem=elem, f=f, energy=np.array(e), virial=w, rho=rho, pi=i, pj=j, pimg=img)


def dmd_case():
    p = pot.johnson_cuh(nr=2000, nrho=2000)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((1, 2, 2), seed=12)
    x = x + np.random.default_rng(13).normal(0, 0.05, x.shape)
    d = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    f, fa, mu, fe, FH, S = d.force()
    return dict(x=x, H=H, elem=elem, alpha=alpha, c=c, f=f, f_alpha=fa, mu=mu, fe=np.array(fe), FH=FH,
                S=np.array(S))


def gcmc_case():
    rng = np.random.default_rng(21)
    x, H = pot.fcc(3.52, (3, 3, 3))
    x = x + rng.normal(0, 0.03, x.shape)
    xh = x[:10] + np.array([1.76, 0, 0])
    xa = np.concatenate([x, xh])
    elem = np.concatenate([np.zeros(len(x), np.uint8), np.full(10, 2, np.uint8)])
    masses = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
    g = O.OGCMC(O.OEam.read(pot.NIALH_JEA), H, xa, elem, 2, -1.9, 300.0, masses, seed=777)
    d, du = g.attempts(300)
    return dict(x=xa, H=H, elem=elem, decisions=d, dU=du, final_elem=g.state()["elem"], final_seed=np.array(g.rng_seed))


def main():
    pm = {"seed": 1, "uniform": O.rng_stream(1, 12).tolist(), "gaussian": O.rng_stream(1, 12, True).tolist(),
          "seed_after_10000": int(O.rng_seed_after(1, 10000))}
    with open(os.path.join(HERE, "park_miller.json"), "w") as fh:
        json.dump(pm, fh, indent=1)
    gh = {str(n): dict(zip(("x", "w"), (v.tolist() for v in O.gauss_hermite(n)))) for n in range(1, 9)}
    with open(os.path.join(HERE, "gauss_hermite.json"), "w") as fh:
        json.dump(gh, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "eam_cu32.npz"), **eam_case())
    np.savez_compressed(os.path.join(HERE, "dmd_cuh16.npz"), **dmd_case())
    np.savez_compressed(os.path.join(HERE, "gcmc_nih.npz"), **gcmc_case())


if __name__ == "__main__":
    main()
