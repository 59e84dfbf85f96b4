"""The oracle against the committed golden fixtures (tests/golden/, written by
make_golden.py): bit-for-bit reproduction, plus the fixture values that have
independent published or closed-form values."""
import json
import math
import os

import numpy as np

import oracle_lib as O
from mapp4py_amd import potentials as pot

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_park_miller_fixture():
    d = json.load(open(os.path.join(G, "park_miller.json")))
    assert d["seed_after_10000"] == 1043618065  # Park & Miller's published check value
    assert O.rng_stream(1, 12).tolist() == d["uniform"]
    assert O.rng_stream(1, 12, True).tolist() == d["gaussian"]
    assert d["uniform"][0] == 16807 / 2147483647


def test_gauss_hermite_fixture():
    d = json.load(open(os.path.join(G, "gauss_hermite.json")))
    x3, w3 = d["3"]["x"], d["3"]["w"]
    sp = math.sqrt(math.pi)
    assert np.allclose(x3, [-math.sqrt(1.5), 0.0, math.sqrt(1.5)], atol=1e-15)
    assert np.allclose(w3, [sp / 6, 2 * sp / 3, sp / 6], atol=1e-15)
    for n in range(1, 9):
        x, w = O.gauss_hermite(n)
        assert x.tolist() == d[str(n)]["x"] and w.tolist() == d[str(n)]["w"]


def test_eam_fixture():
    g = load("eam_cu32.npz")
    p = pot.johnson_cu(nr=2000, nrho=2000)
    md = O.OMD(O.OEam.from_knots(p), g["H"], g["x"], g["elem"], skin=0.5)
    f, e, w, rho = md.force()
    assert np.array_equal(f, g["f"]) and e == g["energy"] and np.array_equal(w, g["virial"])
    assert np.array_equal(rho, g["rho"])
    i, j, img = md.pairs()
    assert np.array_equal(i, g["pi"]) and np.array_equal(j, g["pj"]) and np.array_equal(img, g["pimg"])
    assert abs(g["f"].sum(0)).max() < 1e-10  # Newton's third law


def test_dmd_fixture():
    g = load("dmd_cuh16.npz")
    p = pot.johnson_cuh(nr=2000, nrho=2000)
    d = O.ODMD(O.OEam.from_knots(p, dmd=True), g["H"], g["x"], g["elem"], g["alpha"], g["c"], nquad=3,
               masses=p.masses)
    f, fa, mu, fe, FH, S = d.force()
    assert np.array_equal(f, g["f"]) and np.array_equal(fa, g["f_alpha"]) and np.array_equal(mu, g["mu"])
    assert fe == g["fe"] and np.array_equal(FH, g["FH"]) and S == g["S"]


def test_gcmc_fixture():
    g = load("gcmc_nih.npz")
    masses = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
    o = O.OGCMC(O.OEam.read(pot.NIALH_JEA), g["H"], g["x"], g["elem"], 2, -1.9, 300.0, masses, seed=777)
    d, du = o.attempts(300)
    assert np.array_equal(d, g["decisions"]) and np.array_equal(du, g["dU"])
    assert np.array_equal(o.state()["elem"], g["final_elem"]) and o.rng_seed == int(g["final_seed"])
    assert {1, -1, 2, -2} <= set(d.tolist())
