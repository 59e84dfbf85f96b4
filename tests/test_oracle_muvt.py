"""The oracle's MDMuVT (src/md_muvt.cpp:122-280): NVT with a grand-canonical
exchange of the gas element every nevery steps.

Pins:
* with no attempts the run is the NVT trajectory (the exchange steps only
  wrap the atoms into the box and rebuild the list);
* with atoms that do not move, the exchanges of one run are one chain: the
  same decisions, atom order and final state as a single SGCMC chain of all
  the attempts from the same seed, so the random stream, the reload of the
  atoms and their order carry across exchanges;
* the chain samples the grand-canonical ideal gas: with every interaction
  switched off the gas count is Poisson with mean z V, z = exp(mu / kT) /
  lambda^3, lambda = h / sqrt(2 pi m kT) (GCMC::init, src/gcmc.cpp:86-116),
  and raising mu by kT ln 2 doubles it.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot
from test_oracle_gcmc import GAS, MASSES, small_cell

KB = pot.KB_EV
HP = 4.13566766225 * 0.1 * math.sqrt(1.60217656535 / 1.66053904020)


@pytest.fixture(scope="module")
def ff():
    return O.OEam.read(pot.NIALH_JEA)


def test_without_attempts_the_run_is_nvt(ff):
    x, H, elem = small_cell(3, k=(4, 4, 4), h_frac=0.1)
    v = O.o_create_T(elem, MASSES, KB, 300.0, 5)
    o = O.OMD(ff, H, x, elem)
    vm, Tm, _, hist, counts, _ = o.muvt_run(v, MASSES, KB, HP, 300.0, pot.FS, -2.3, GAS, 1, 10, 0, 40)
    o2 = O.OMD(ff, H, x, elem)
    v2, _, T2, _ = o2.nvt_run(v, MASSES, KB, 300.0, pot.FS, 40)
    ds = (o.x - o2.x) @ np.linalg.inv(H)
    ds -= np.round(ds)
    assert np.abs(ds @ H).max() < 1e-11
    np.testing.assert_allclose(vm, v2, rtol=0, atol=1e-12)
    assert abs(Tm - T2) < 1e-10 * T2
    assert counts.sum() == 0 and np.all(hist[:, 2] == (elem == GAS).sum())


def test_exchanges_of_one_run_form_one_chain(ff):
    # dt so small that nothing moves: 20 exchanges of 50 attempts are the
    # single chain of 1000 attempts from the same seed
    x, H, elem = small_cell(5, k=(4, 4, 4), h_frac=0.1)
    v = O.o_create_T(elem, MASSES, KB, 300.0, 6)
    o = O.OMD(ff, H, x, elem)
    vm, Tm, _, hist, counts, nb = o.muvt_run(v, MASSES, KB, HP, 300.0, 1e-9 * pot.FS, -2.1, GAS, 77, 1, 50, 20)
    g = O.OGCMC(ff, H, x, elem, GAS, -2.1, 300.0, MASSES, 77, cap=len(x) + 4096, kB=KB, hP=HP)
    dec, du = g.attempts(1000)
    st = g.state()
    assert counts[0] == (dec == 1).sum() and counts[1] == (dec == 2).sum()
    assert counts[2] == ((dec == 1) | (dec == -1)).sum()
    assert counts[0] > 5 and counts[1] > 5
    assert o.n == len(st["elem"])
    np.testing.assert_array_equal(o.elem, st["elem"])
    ds = (o.x - st["x"]) @ np.linalg.inv(H)
    ds -= np.round(ds)
    assert np.abs(ds @ H).max() < 1e-9
    # the gas count after each exchange follows the chain
    ngas = (elem == GAS).sum() + np.cumsum((dec == 1).astype(int) - (dec == 2).astype(int))
    np.testing.assert_array_equal(hist[:, 2], ngas[49::50])


def ideal_gas_ff():
    # F = 0 and phi = 0: no atom interacts with any other
    p = pot.johnson_cu(nr=2000, nrho=2000)
    q = pot.EamKnots(p.names, p.masses, p.nr, p.dr, p.nrho, p.drho, np.zeros_like(p.F), p.rho,
                     np.zeros_like(p.r_phi), False)
    return O.OEam.from_knots(q)


def gas_counts(ffz, mu, nattempts, seed):
    x, H = pot.fcc(3.615, (4, 4, 4))
    x = x[:5]
    elem = np.zeros(len(x), np.uint8)
    g = O.OGCMC(ffz, H, x, elem, 0, mu, 300.0, [pot.MASS["Cu"]], seed, cap=400, kB=KB, hP=HP)
    dec, _ = g.attempts(nattempts)
    n = len(x) + np.cumsum((dec == 1).astype(int) - (dec == 2).astype(int))
    return n[nattempts // 10:], np.linalg.det(H)


def test_the_chain_samples_the_ideal_gas():
    ffz = ideal_gas_ff()
    kT = KB * 300.0
    lam = HP / math.sqrt(2.0 * math.pi * kT * pot.MASS["Cu"])
    x, H = pot.fcc(3.615, (4, 4, 4))
    vol = np.linalg.det(H)
    mu = kT * math.log(20.0 * lam ** 3 / vol)  # z V = 20
    n, _ = gas_counts(ffz, mu, 60000, 11)
    # Poisson(20): mean and variance 20; about 2,000 independent samples
    assert abs(n.mean() - 20.0) < 1.0, n.mean()
    assert 0.8 < n.var() / n.mean() < 1.25, (n.mean(), n.var())
    n2, _ = gas_counts(ffz, mu + kT * math.log(2.0), 60000, 12)
    assert abs(n2.mean() - 40.0) < 2.0, n2.mean()
