"""The GPU path against the committed golden fixtures (tests/golden/): EAM
forces within 1e-5 eV/A and the half list bit-exact, DMD within 1e-9
relative, GCMC decisions identical."""
import os

import numpy as np
import pytest

from mapp4py_amd import potentials as pot

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_eam_golden(M):
    g = load("eam_cu32.npz")
    md = M.AtomsMD(M.ForceFieldEAM.from_knots(pot.johnson_cu(nr=2000, nrho=2000)), g["H"], g["x"], g["elem"],
                   skin=0.5)
    f, e, w = md.force_calc()
    assert np.abs(f - g["f"]).max() < 1e-5
    assert abs(e - float(g["energy"])) < 5e-8 * len(f) + 1e-6
    i, j, img = md.half_pairs()
    got = {(int(a), int(b), *map(int, m)) for a, b, m in zip(i, j, img)}
    want = {(int(a), int(b), *map(int, m)) for a, b, m in zip(g["pi"], g["pj"], g["pimg"])}
    assert got == want


def test_dmd_golden(M):
    g = load("dmd_cuh16.npz")
    p = pot.johnson_cuh(nr=2000, nrho=2000)
    d = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), g["H"], g["x"], g["elem"], g["alpha"], g["c"], nquad=3,
                   masses=p.masses)
    f, fa, mu, fe, FH, S = d.derivative()
    for a, b in ((f, g["f"]), (fa, g["f_alpha"]), (mu, g["mu"]), ([fe], [g["fe"]])):
        a, b = np.asarray(a), np.asarray(b)
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def test_gcmc_golden(M):
    g = load("gcmc_nih.npz")
    masses = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
    gb = M.GCMCBatch(M.ForceFieldEAM.setfl(pot.NIALH_JEA), [(g["x"], g["H"], g["elem"])], 2, -1.9, 300.0, masses,
                     [777])
    d, du = gb.attempts(300)
    assert np.array_equal(d[0], g["decisions"])
    assert np.abs(du[0] - g["dU"]).max() < 1e-9
    c = gb.cell(0)
    assert np.array_equal(c["elem"], g["final_elem"]) and c["seed"] == int(g["final_seed"])
