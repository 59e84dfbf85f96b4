"""The GPU BDF driver (mapp_dmd_bdf_run) against the oracle's DAEBDF on the
same system: the same steps, orders, rejections and tally times, and the
same concentrations at the end."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
from o_bdf import OBdf  # noqa: E402

pytestmark = pytest.mark.gpu
R_CRD, ZETA = [3.5, 3.5], [1.7, 1.7]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("cells,T", [((2, 2, 2), 3000.0), ((3, 3, 3), 1e5)])
def test_bdf_matches_the_oracle(cells, T):
    import mapp4py_amd as M

    p = pot.johnson_cuh(nr=3000, nrho=3000)
    x, H, elem, alpha, c = pot.dmd_cuh_sites(cells, seed=3)
    x = x + np.random.default_rng(1).normal(0, 0.05, x.shape)
    s = x @ np.linalg.inv(H)
    s -= np.floor(s)
    x = s @ H
    g = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    rep, tally = g.bdf_run(R_CRD, ZETA, T, a_tol=1e-8, max_nsteps=400, ntally=10)
    cg = g.state()[2]
    o = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    b = OBdf(o, R_CRD, ZETA, p.masses, a_tol=1e-8, max_nsteps=400, ntally=10, x=x, alpha=alpha)
    n = b.run(T)
    co = o.state_c()
    say(f"BDF {cells} to t={T:g}: GPU {rep['nsteps']} steps (max order {rep['max_q']}, {rep['integ_rej']} rejected), "
        f"oracle {n} steps ({b.counts['integ_rej']} rejected); max|dc| {np.abs(cg - co).max():.2e}")
    assert rep["nsteps"] == n and rep["t"] == pytest.approx(b.t_cur, rel=1e-12)
    for k in ("integ_acc", "integ_rej", "intpol_acc", "intpol_rej", "nonlin_acc", "nonlin_rej"):
        assert rep[k] == b.counts[k], k
    to = np.array(b.tally)
    assert tally.shape == to.shape
    np.testing.assert_array_equal(tally[:, 0], to[:, 0])
    # the step control works on |c - y_0| at the a_tol level, so the Newton
    # solutions' last digits reach the step sizes: equal step and rejection
    # counts above, times to 1e-5
    np.testing.assert_allclose(tally[:, 1], to[:, 1], rtol=1e-5)
    np.testing.assert_allclose(tally[:, 2], to[:, 2], rtol=1e-9)
    assert np.abs(cg - co).max() < 1e-8
    for e in (0, 1):
        assert abs((cg - c)[elem == e].sum()) < 1e-11
