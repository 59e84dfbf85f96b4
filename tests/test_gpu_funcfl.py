"""funcfl import in the library (mapp_eam_create_funcfl) against the oracle's
ImportEAM::funcfl: forces and energy of a jittered lattice through the GPU
passes, on a spacing where the reference's interval-k reading of Z changes
knots (dr = 0.002) and with two files."""
import sys

import numpy as np
import pytest

import mapp4py_amd as M
import oracle_lib as O
from mapp4py_amd import potentials as pot
from test_funcfl import funcs, write

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("two", [False, True])
def test_funcfl_tables_match_the_oracle(tmp_path, two):
    Za, rhoa, Fa, _ = funcs()
    Zb, rhob, Fb, _ = funcs(scale=1.3)
    paths = [tmp_path / "a.funcfl"]
    write(paths[0], Za, rhoa, Fa, 2800, 0.002, 3000, 0.02)
    if two:
        paths.append(tmp_path / "b.funcfl")
        write(paths[1], Zb, rhob, Fb, 2800, 0.002, 3000, 0.02)
    x, H = pot.fcc(3.615, (6, 6, 6))
    x = x + np.random.default_rng(7).normal(0, 0.05, x.shape)
    elem = ((np.random.default_rng(8).uniform(size=len(x)) < 0.4) if two else np.zeros(len(x))).astype(np.uint8)
    g = M.AtomsMD(M.ForceFieldEAM.funcfl([str(p) for p in paths]), H, x, elem)
    f, e, FH = g.force_calc()
    fo, eo, Wo, _ = O.OMD(O.OEam.funcfl(paths), H, x, elem).force()
    df = np.abs(f - fo).max()
    say(f"funcfl {'two files' if two else 'one file'}: max|df| {df:.2e} eV/A, dE/N {abs(e - eo) / len(x):.2e} eV")
    assert df < 1e-5
    assert abs(e - eo) < 5e-8 * len(x) + 1e-6  # FP32 pair splines summed in FP64


def test_funcfl_rejects_a_short_file(tmp_path):
    Z, rho, F, _ = funcs()
    pa, pb = tmp_path / "a.funcfl", tmp_path / "b.funcfl"
    write(pa, Z, rho, F, 2800, 0.002, 3000, 0.02)
    write(pb, Z, rho, F, 1000, 0.004, 3000, 0.02)
    with pytest.raises(M.MappError, match="fewer r knots"):
        M.ForceFieldEAM.funcfl([str(pa), str(pb)])
