"""Spline coefficients of the library's tables (mapp_eam_coefs) against the
oracle's (o_eam_coefs): the Hermite slopes and cubic/quartic coefficients of
ImportEAM (src/import_eam.cpp:62-115), the r*rho rescale of the DMD tables
(src/ff_eam_dmd.cpp:69-124) and the three readers (setfl, fs, funcfl;
src/import_eam.h:68-344,582-851) give the same FP64 numbers in both. The two
sides evaluate the same formulas in the same order, so the bar is a few ulp
(relative 1e-13 of the row's largest coefficient), not a physics tolerance.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot
from test_cfg import unpack
from test_funcfl import funcs, write

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


def compare(g, o, label, dmd):
    ne, nr, nrho, _, _ = o.info()
    worst = 0.0
    for which in (0, 1, 2):
        for a in range(ne):
            for b in range(ne if which < 2 else 1):
                # F rows hold the quartic's five coefficients in the DMD tables
                ncols = 5 if (which == 2 and dmd) else 4
                co = o.coefs(which, a, b, ncols=ncols)
                cg = g.coefs(which, a, b, nrows=co.shape[0], ncols=ncols)
                scale = np.maximum(np.abs(co).max(axis=1, keepdims=True), 1e-300)
                err = float((np.abs(cg - co) / scale).max())
                worst = max(worst, err)
                assert err < 1e-13, (label, which, a, b, err)
    say(f"{label}: {ne} elements, nr {nr}, nrho {nrho}: max relative coefficient difference {worst:.1e}")


@pytest.mark.parametrize("dmd", [False, True])
def test_setfl_coefficients(M, tmp_path, dmd):
    path = unpack("Cu_mishin.eam.alloy.gz", tmp_path)
    compare(M.ForceFieldEAM.setfl(path, dmd=dmd), O.OEam.read(path, 0, dmd=dmd), f"Cu_mishin setfl dmd={dmd}", dmd)


def test_three_element_setfl_coefficients(M):
    path = os.path.join(ROOT, "mapp4py_amd", "data", "NiAlH_jea.eam.alloy")
    compare(M.ForceFieldEAM.setfl(path), O.OEam.read(path, 0), "NiAlH_jea setfl", False)


@pytest.mark.parametrize("dmd", [False, True])
def test_fs_coefficients(M, tmp_path, dmd):
    p = pot.johnson_feni_fs(nr=3000, nrho=3000)
    path = tmp_path / "feni.eam.fs"
    pot.write_fs(path, p)
    compare(M.ForceFieldEAM.fs(str(path), dmd=dmd), O.OEam.read(path, 1, dmd=dmd), f"Fe-Ni fs file dmd={dmd}", dmd)
    compare(M.ForceFieldEAM.from_knots(p, dmd=dmd), O.OEam.from_knots(p, dmd=dmd), f"Fe-Ni fs knots dmd={dmd}", dmd)


def test_funcfl_coefficients(M, tmp_path):
    Za, rhoa, Fa, _ = funcs()
    Zb, rhob, Fb, _ = funcs(scale=1.3)
    paths = [tmp_path / "a.funcfl", tmp_path / "b.funcfl"]
    write(paths[0], Za, rhoa, Fa, 2800, 0.002, 3000, 0.02)
    write(paths[1], Zb, rhob, Fb, 2800, 0.002, 3000, 0.02)
    compare(M.ForceFieldEAM.funcfl([str(p) for p in paths]), O.OEam.funcfl(paths), "funcfl two files", False)
