"""Pins the oracle's grand-canonical exchange restatement (single-cell SGCMC
with EAM exchange energies).

* delta_u of every accepted insertion equals E(after) - E(before) of two
  independent oracle energy calls (ghost half list, all periodic images),
  also in a box smaller than the H-H cutoff, where the trial point meets its
  own images; the same for accepted deletions (boxes without self images:
  there the reference counts the deleted atom's self-image pairs 1.5 times).
* the incrementally kept rho/F of every atom equals a fresh computation.
* the decision sequence and the final RNG seed follow from a separate
  Python restatement of the draw order (Park-Miller, polar Gaussian) and of
  the acceptance rules fed with the oracle's delta_u.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

GAS = 2  # H in NiAlH_jea.eam.alloy
MASSES = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
T = 300.0


@pytest.fixture(scope="module")
def ff():
    return O.OEam.read(pot.NIALH_JEA)


def energy(ff, H, x, elem):
    return O.OMD(ff, H, x, elem, skin=0.5).energy()


def small_cell(seed, k=(3, 3, 3), tilt=(0.0, 0.0, 0.0), h_frac=0.05):
    rng = np.random.default_rng(seed)
    x, H = pot.fcc(3.52, k)
    H = H.copy()
    H[1, 0], H[2, 0], H[2, 1] = tilt
    s = x @ np.linalg.inv(pot.fcc(3.52, k)[1])
    x = s @ H + rng.normal(0, 0.03, x.shape)
    octa = (x[: len(x) // 4] + np.array([1.76, 0.0, 0.0]))
    nh = max(1, int(h_frac * len(octa)))
    xh = octa[rng.choice(len(octa), nh, replace=False)]
    return np.concatenate([x, xh]), H, np.concatenate([np.zeros(len(x), np.uint8), np.full(nh, GAS, np.uint8)])


def step_and_check(ff, g, H, nsteps, want, tol=2e-9):
    checked = 0
    st = g.state()
    e0 = energy(ff, H, st["x"], st["elem"])
    for _ in range(nsteps):
        d, du = g.attempts(1)
        if d[0] in (1, 2):
            st = g.state()
            e1 = energy(ff, H, st["x"], st["elem"])
            delta = (e1 - e0) if d[0] == 1 else (e0 - e1)
            assert abs(delta - du[0]) < tol * max(1.0, abs(e0)), (d[0], delta, du[0])
            e0 = e1
            checked += d[0] == want
    return checked


@pytest.mark.parametrize("tilt", [(0.0, 0.0, 0.0), (1.1, -0.7, 0.9)])
def test_accepted_insertions_change_the_energy_by_delta_u(ff, tilt):
    x, H, elem = small_cell(1, tilt=tilt)
    g = O.OGCMC(ff, H, x, elem, GAS, 1.0, T, MASSES, seed=11)
    assert step_and_check(ff, g, H, 40, 1) >= 5


def test_insertion_meets_its_own_images_in_a_box_below_the_cutoff(ff):
    # a 2.5 A edge is below the 2.79 A H-H cutoff: the trial point sees its
    # own periodic images (next_jatm_self, one half of each pair)
    rng = np.random.default_rng(4)
    H = np.array([[2.5, 0, 0], [0.4, 6.0, 0], [-0.3, 0.5, 6.5]])
    x = rng.uniform(0, 1, (4, 3)) @ H
    elem = np.zeros(4, np.uint8)
    assert ff.cut()[GAS, GAS] > 2.5
    g = O.OGCMC(ff, H, x, elem, GAS, 3.0, T, MASSES, seed=5)
    assert g.info()["max_trials"] > 27
    assert step_and_check(ff, g, H, 30, 1) >= 3


def test_accepted_deletions_change_the_energy_by_delta_u(ff):
    x, H, elem = small_cell(2, h_frac=0.3)
    g = O.OGCMC(ff, H, x, elem, GAS, -6.0, T, MASSES, seed=3)
    assert step_and_check(ff, g, H, 40, 2) >= 5


def test_kept_density_matches_a_fresh_computation(ff):
    x, H, elem = small_cell(3, k=(3, 4, 3), tilt=(0.5, 0.2, -0.4), h_frac=0.1)
    g = O.OGCMC(ff, H, x, elem, GAS, -2.0, T, MASSES, seed=17)
    d, _ = g.attempts(400)
    assert (d == 1).sum() > 0 and (d == 2).sum() > 0
    st = g.state()
    fresh = O.OGCMC(ff, H, st["x"], st["elem"], GAS, -2.0, T, MASSES, seed=17).state()
    np.testing.assert_allclose(st["rho"], fresh["rho"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(st["F"], fresh["F"], rtol=1e-12, atol=1e-13)
    np.testing.assert_array_equal(st["elem"], fresh["elem"])


class PM:
    """Park-Miller minimal standard with the polar Gaussian of the reference."""

    def __init__(self, seed):
        self.seed, self.chk, self.keep = seed, False, 0.0

    def uniform(self):
        hi, lo = divmod(self.seed, 127773)
        t = 16807 * lo - 2836 * hi
        self.seed = t if t > 0 else t + 2147483647
        return self.seed * 1.0 / 2147483647

    def gaussian(self):
        if self.chk:
            self.chk = False
            return self.keep
        self.chk = True
        s = 1.0
        while s >= 1.0:
            u = 2.0 * self.uniform() - 1.0
            v = 2.0 * self.uniform() - 1.0
            s = u * u + v * v
        f = math.sqrt(-2.0 * math.log(s) / s)
        self.keep = v * f
        return u * f


def xexp(v):
    return math.exp(v) if v < 709.0 else math.inf


def test_decisions_follow_the_draw_order_and_acceptance_rules(ff):
    x, H, elem = small_cell(5, h_frac=0.1)
    mu = -2.2
    g = O.OGCMC(ff, H, x, elem, GAS, mu, T, MASSES, seed=123)
    kT = pot.KB_EV * T
    lam = pot.HP / math.sqrt(2.0 * math.pi * kT * MASSES[GAS])
    z = math.exp(mu / kT) / lam ** 3
    vol = H[0, 0] * H[1, 1] * H[2, 2]
    info = g.info()
    assert abs(info["z_fac"] - z) < 1e-13 * z and abs(info["vol"] - vol) < 1e-12 * vol
    d, du = g.attempts(600)
    rng = PM(123)
    ngas = int((elem == GAS).sum())
    for k in range(len(d)):
        if rng.uniform() < 0.5:
            for _ in range(3):
                rng.uniform()
            acc = rng.uniform() < z * vol / ((ngas + 1.0) * xexp(du[k] / kT))
            if acc:
                for _ in range(3):
                    rng.gaussian()
                ngas += 1
            assert d[k] == (1 if acc else -1), k
        else:
            if ngas == 0:
                assert d[k] == 0
                continue
            rng.uniform()
            acc = rng.uniform() < ngas * xexp(du[k] / kT) / (z * vol)
            ngas -= acc
            assert d[k] == (2 if acc else -2), k
    assert rng.seed == g.rng_seed
    assert ngas == g.ngas
    assert {1, -1, 2, -2} <= set(d.tolist())
