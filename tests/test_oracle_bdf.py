"""The oracle's DAEBDF (oracle/o_bdf.py) pinned by an independent stiff
solver: scipy's Radau on dc/dt = c_d(c) with the oracle's rates and Jacobian
products, positions and widths fixed. The variable-order steps must land on
the same concentrations to within the tolerances, keep every c in [0, 1] and
conserve each element."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
from o_bdf import OBdf  # noqa: E402

R_CRD, ZETA = [3.5, 3.5], [1.7, 1.7]


@pytest.fixture(scope="module")
def small():
    p = pot.johnson_cuh(nr=3000, nrho=3000)
    ff = O.OEam.from_knots(p, dmd=True)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((2, 2, 2), seed=3)
    x = x + np.random.default_rng(1).normal(0, 0.05, x.shape)
    s = x @ np.linalg.inv(H)
    s -= np.floor(s)
    return p, ff, s @ H, H, elem, alpha, c


def test_bdf_matches_an_independent_stiff_solver(small):
    from scipy.integrate import solve_ivp

    p, ff, x, H, elem, alpha, c = small
    d = O.ODMD(ff, H, x, elem, alpha, c, nquad=3, masses=p.masses)
    T = 3000.0
    b = OBdf(d, R_CRD, ZETA, p.masses, a_tol=1e-8, max_nsteps=2000, x=x, alpha=alpha)
    n = b.run(T)
    cb = d.state_c()
    assert b.t_cur == T and n < 2000 and b.counts["integ_acc"] == n
    assert max(b.q, 1) >= 1 and b.counts["nonlin_rej"] == 0

    def f(t, y):
        d.set_state(x, alpha, y.reshape(c.shape))
        return d.c_d(R_CRD, ZETA, p.masses)[0].ravel()

    def jac(t, y):
        d.set_state(x, alpha, y.reshape(c.shape))
        d.c_d(R_CRD, ZETA, p.masses)
        m = y.size
        Jm = np.zeros((m, m))
        for k in range(m):
            e = np.zeros(m)
            e[k] = 1.0
            Jm[:, k] = d.J(R_CRD, ZETA, p.masses, e.reshape(c.shape)).ravel()
        return Jm

    sol = solve_ivp(f, (0.0, T), c.ravel(), method="Radau", jac=jac, rtol=1e-11, atol=1e-13)
    cs = sol.y[:, -1].reshape(c.shape)
    moved = np.abs(cs - c).max()
    err = np.abs(cb - cs).max()
    print(f"BDF: {n} steps, max order {b.q}; max|c_BDF - c_Radau| {err:.2e} over a change of {moved:.3f}")
    assert moved > 0.05 and err < 1e-6
    assert cb.min() >= 0.0 and cb.max() <= 1.0
    for e in (0, 1):
        assert abs((cb - c)[elem == e].sum()) < 1e-12


def test_bdf_order_climbs_and_steps_grow(small):
    p, ff, x, H, elem, alpha, c = small
    d = O.ODMD(ff, H, x, elem, alpha, c, nquad=3, masses=p.masses)
    b = OBdf(d, R_CRD, ZETA, p.masses, a_tol=1e-8, max_nsteps=60, ntally=10, x=x, alpha=alpha)
    b.run(1e6)
    ts = [r[1] for r in b.tally]
    assert len(b.tally) == 7 and ts == sorted(ts)
    assert b.q >= 3  # the order rises from 1 as the solution smooths
    assert (ts[-1] - ts[-2]) > 2.0 * (ts[2] - ts[1])  # and the steps grow
