"""Pins the oracle's DMD restatement: closed-form Gaussian smearing, the
alpha -> 0 limit, finite differences of the free energy against f_x,
f_alpha and the chemical potential, and the list against brute force."""
import math

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

KB = pot.KB_EV


def closed_form(s, r, a):
    q = s * s + a * a
    g = (s * s / q) ** 1.5 * math.exp(-r * r / q)
    dr = g * (-2.0 / q)                      # (1/r) dg/dr
    da = g * (-3.0 / q + 2.0 * r * r / q / q)  # (1/a) dg/da
    return g, dr, da


@pytest.fixture(scope="module")
def gauss16():
    p = pot.gaussian_knots(s=1.5)
    ff = O.OEam.from_knots(p, dmd=True)
    x, H = pot.fcc(8.0, (2, 2, 2))
    n = len(x)
    d = O.ODMD(ff, H, x, np.zeros((n, 1), np.uint8), np.full((n, 1), 0.1), np.full((n, 1), 0.5), nquad=16,
               masses=[1.0])
    return d


@pytest.mark.parametrize("r,a", [(0.7, 0.05), (1.3, 0.2), (2.5, 0.3), (3.4, 0.15), (1.0, 0.45)])
def test_smearing_matches_3d_gaussian_convolution(gauss16, r, a):
    out = gauss16.pair_eval(0, 0, r, a)
    g, dr, da = closed_form(1.5, r, a)
    for k in range(3):  # phi, rho_ij, rho_ji all hold r*exp(-r^2/s^2)
        assert abs(out[k] - g) < 1e-8 * max(1.0, abs(g)), (k, out[k], g)
        assert abs(out[3 + k] - dr) < 1e-7 * max(1.0, abs(dr)), (k, out[3 + k], dr)
        assert abs(out[6 + k] - da) < 1e-6 * max(1.0, abs(da)), (k, out[6 + k], da)


def test_zero_width_recovers_the_table():
    p = pot.johnson_cuh(nr=2000, nrho=2000)
    ff_d = O.OEam.from_knots(p, dmd=True)
    x, H = pot.fcc(3.615, (2, 2, 2))
    n = len(x)
    d = O.ODMD(ff_d, H, x, np.zeros((n, 1), np.uint8), np.full((n, 1), 0.1), np.full((n, 1), 0.9), nquad=3,
               masses=p.masses)
    for r in (2.2, 2.9, 3.7):
        out = d.pair_eval(0, 1, r, 1e-7)
        phi = p.analytic["phi"][0][1](r)
        rho_01 = p.analytic["rho"][0][0](r)   # setfl: density of a Cu atom
        rho_10 = p.analytic["rho"][1][0](r)   # density of an H atom
        assert abs(out[0] - phi) < 1e-6
        assert abs(out[1] - rho_01) < 1e-6
        assert abs(out[2] - rho_10) < 1e-6


@pytest.fixture(scope="module")
def cuh_small():
    p = pot.johnson_cuh(nr=3000, nrho=3000)
    ff = O.OEam.from_knots(p, dmd=True)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((2, 2, 2), seed=3)
    x = x + np.random.default_rng(1).normal(0, 0.05, x.shape)
    # wrapped into the box, as a list build leaves them, so set_state keeps
    # the image shifts of the list valid
    s = x @ np.linalg.inv(H)
    s -= np.floor(s)
    x = s @ H
    return p, ff, x, H, elem, alpha, c


def make(ff, p, x, H, elem, alpha, c, nquad=12):
    # The reference differentiates the smeared integral (Gaussian
    # integration by parts turns d/dr and d/dalpha into moments), not the
    # N-point sum, so forces equal -grad FE only as far as the rule is exact:
    # the gradient checks use a 12-point rule.
    return O.ODMD(ff, H, x, elem, alpha, c, nquad=nquad, masses=p.masses)


def test_forces_are_free_energy_gradients(cuh_small):
    p, ff, x, H, elem, alpha, c = cuh_small
    d = make(ff, p, x, H, elem, alpha, c)
    f, fa, mu, fe, FH, S = d.force()
    assert abs(d.energy() - fe) < 1e-10 * abs(fe)
    h = 1e-6
    rng = np.random.default_rng(0)
    for k in rng.choice(len(x), 4, replace=False):
        for q in range(3):
            xp = x.copy(); xp[k, q] += h
            xm = x.copy(); xm[k, q] -= h
            d.set_state(xp, alpha, c); ep = d.energy()
            d.set_state(xm, alpha, c); em = d.energy()
            fd = -(ep - em) / (2 * h)
            assert abs(fd - f[k, q]) < 1e-5 * max(1.0, abs(f[k, q])), (k, q, fd, f[k, q])
        for q in range(2):
            ap = alpha.copy(); ap[k, q] += h
            am = alpha.copy(); am[k, q] -= h
            d.set_state(x, ap, c); ep = d.energy()
            d.set_state(x, am, c); em = d.energy()
            fd = -(ep - em) / (2 * h)
            assert abs(fd - fa[k, q]) < 1e-5 * max(1.0, abs(fa[k, q])), (k, q, fd, fa[k, q])
    d.set_state(x, alpha, c)


def test_raw_f_c_plus_site_terms_is_dFE_dc(cuh_small):
    # dFE/dc_i = f_c + mu_0 - 3 kT ln(alpha_i) + kT (ln c_i - ln c_v)
    p, ff, x, H, elem, alpha, c = cuh_small
    d = make(ff, p, x, H, elem, alpha, c)
    f, fa, mu, fe, FH, S = d.force()
    kT = KB * 300.0
    lam = pot.HP / np.sqrt(2 * np.pi * kT * np.asarray(p.masses))
    mu0 = kT * 1.5 * (2 * np.log(lam) - 1 - np.log(np.pi))
    h = 1e-7
    for k in (0, 5, 17):
        for q in range(2):
            cp = c.copy(); cp[k, q] += h
            cm = c.copy(); cm[k, q] -= h
            d.set_state(x, alpha, cp); ep = d.energy()
            d.set_state(x, alpha, cm); em = d.energy()
            fd = (ep - em) / (2 * h)
            cv = 1.0 - c[k].sum()
            an = mu[k, q] + mu0[elem[k, q]] - 3 * kT * np.log(alpha[k, q]) + kT * (np.log(c[k, q]) - np.log(cv))
            assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (k, q, fd, an)
    d.set_state(x, alpha, c)


def test_virial_is_minus_strain_derivative(cuh_small):
    p, ff, x, H, elem, alpha, c = cuh_small
    _, _, _, fe, FH, _ = make(ff, p, x, H, elem, alpha, c).force()
    h = 1e-6
    for a, b, k in [(0, 0, 0), (1, 1, 3), (2, 2, 5), (1, 0, 1)]:
        eps = np.zeros((3, 3))
        eps[a, b] = h
        ep = make(ff, p, x @ (np.eye(3) + eps), H @ (np.eye(3) + eps), elem, alpha, c).energy()
        em = make(ff, p, x @ (np.eye(3) - eps), H @ (np.eye(3) - eps), elem, alpha, c).energy()
        assert abs(-(ep - em) / (2 * h) - FH[k]) < 1e-5 * max(1.0, abs(FH[k]))


def test_dmd_half_list_matches_brute_force(cuh_small):
    p, ff, x, H, elem, alpha, c = cuh_small
    d = make(ff, p, x, H, elem, alpha, c, nquad=3)
    i, j, img = d.pairs()
    got = {(int(a), int(b), *map(int, m)) for a, b, m in zip(i, j, img)}
    assert len(got) == len(i)
    xg, wg = O.gauss_hermite(3)
    xmax = xg[-1]
    cut_sk = ff.cut() + 0.5
    B = np.linalg.inv(H)
    s = x @ B
    s -= np.floor(s)
    want = set()
    n = len(x)
    for a0 in range(-2, 3):
        for a1 in range(-2, 3):
            for a2 in range(-2, 3):
                xj = (s + np.array([a0, a1, a2])) @ H
                xi = s @ H
                r = np.sqrt(((xi[:, None, :] - xj[None, :, :]) ** 2).sum(-1))
                for si in range(n):
                    for sj in range(n):
                        if (a0, a1, a2) == (0, 0, 0) and si >= sj:
                            continue
                        for ic in range(2):
                            for jc in range(2):
                                aij = math.sqrt(alpha[si, ic] ** 2 + alpha[sj, jc] ** 2)
                                if r[si, sj] - xmax * aij < cut_sk[elem[si, ic], elem[sj, jc]]:
                                    want.add((si * 2 + ic, sj * 2 + jc, a0, a1, a2))
    assert got == want


# ---- concentration rates (ForceFieldEAMDMD::__c_d_calc) ----

R_CRD = [3.5, 3.5]  # first fcc shell only, as examples/diff/in_Cu_diff.py
ZETA = [1.7, 1.7]


def test_exchange_barrier_symmetric_closed_form():
    # equal widths and no chemical-potential difference: the barrier sits
    # half way (z0 = 1/2), Q = gamma/16, alpha_Q^2 = 1/gamma, theta = -1/2
    for g in (4.0, 64.0, 913.0):
        Q, aQ, th = O.calc_q(g, g, 0.0)
        assert abs(Q - g / 16.0) < 1e-14 * g and abs(aQ * g - 1.0) < 1e-14 and abs(th + 0.5) < 1e-14


def test_exchange_barrier_detailed_balance():
    # seen from the other site the barrier is Q - mu_ji (so forward and
    # backward rates carry the same prefactor exp(-Q + mu_ji/2)), the width
    # is the same and the two theta add up to -1
    rng = np.random.default_rng(5)
    for _ in range(200):
        gi, gj = np.exp(rng.uniform(np.log(20.0), np.log(3000.0), 2))
        mu = rng.uniform(-0.05, 0.05) * min(gi, gj)
        q1, a1, t1 = O.calc_q(gi, gj, mu)
        q2, a2, t2 = O.calc_q(gj, gi, -mu)
        assert abs(q2 - (q1 - mu)) < 1e-12 * (abs(q1) + abs(mu))
        assert abs(a2 - a1) < 1e-13 * a1
        assert abs(t1 + t2 + 1.0) < 1e-12


def test_concentration_rates_conserve_mass_per_element(cuh_small):
    p, ff, x, H, elem, alpha, c = cuh_small
    d = make(ff, p, x, H, elem, alpha, c, nquad=3)
    cd, norm = d.c_d(R_CRD, ZETA, p.masses)
    assert norm == np.abs(cd).max() and norm > 0.0
    for e in (0, 1):
        s = cd[elem == e]
        assert abs(s.sum()) < 1e-12 * np.abs(s).sum(), e


def test_concentration_rates_vanish_at_uniform_state(cuh_small):
    # every site equivalent (perfect lattice, one alpha and c per element):
    # equal chemical potentials and equal terms, no net exchange
    p, ff, _, _, _, _, _ = cuh_small
    x, H = pot.fcc(3.615, (3, 3, 3))
    n = len(x)
    elem = np.tile(np.array([0, 1], np.uint8), (n, 1))
    alpha = np.tile([0.12, 0.2], (n, 1))
    c = np.tile([0.7, 0.2], (n, 1))
    d = make(ff, p, x, H, elem, alpha, c, nquad=3)
    cd, norm = d.c_d(R_CRD, ZETA, p.masses)
    d2 = make(ff, p, x + np.random.default_rng(2).normal(0, 0.05, x.shape), H, elem, alpha, c, nquad=3)
    scale = d2.c_d(R_CRD, ZETA, p.masses)[1]
    assert scale > 0.0 and norm < 1e-9 * scale


def test_jacobian_product_is_the_derivative_of_the_rates(cuh_small):
    # __J(x) is d c_d / d c applied to x: central differences of c_d along x
    # (through the chemical potentials, F'' and the vacancy concentrations)
    # agree to the differencing error
    p, ff, x, H, elem, alpha, c = cuh_small
    d = make(ff, p, x, H, elem, alpha, c, nquad=3)
    v = np.random.default_rng(5).normal(0.0, 1.0, c.shape)
    Jv = d.J(R_CRD, ZETA, p.masses, v)
    h = 1e-6
    d.set_state(x, alpha, c + h * v)
    cp = d.c_d(R_CRD, ZETA, p.masses)[0]
    d.set_state(x, alpha, c - h * v)
    cm = d.c_d(R_CRD, ZETA, p.masses)[0]
    fd = (cp - cm) / (2.0 * h)
    assert np.abs(fd).max() > 1e-4
    assert np.abs(fd - Jv).max() < 1e-7 * np.abs(fd).max()
    # linear in x
    w = np.random.default_rng(6).normal(0.0, 1.0, c.shape)
    d.set_state(x, alpha, c)
    Jw = d.J(R_CRD, ZETA, p.masses, w)
    Jvw = d.J(R_CRD, ZETA, p.masses, 2.0 * v - 3.0 * w)
    assert np.abs(Jvw - (2.0 * Jv - 3.0 * Jw)).max() < 1e-12 * np.abs(Jvw).max()


def test_gmres_minimizes_the_residual_over_the_krylov_space():
    # __GMRES::solve from a zero guess: after k products x is the minimiser
    # of |b - A x| over span(b, Ab, ..., A^(k-1) b); with k = n it solves A x = b
    rng = np.random.default_rng(1)
    n = 10
    A = 3.0 * np.eye(n) + rng.normal(0.0, 0.5, (n, n))
    b = rng.normal(0.0, 1.0, n)
    ok, x, nrm = O.gmres_dense(A, b, n, 1e-13)
    assert ok and np.abs(x - np.linalg.solve(A, b)).max() < 1e-12
    assert abs(nrm - np.linalg.norm(x)) < 1e-13  # |x| from the orthonormal basis
    for m in (1, 2, 4, 6):
        ok, x, _ = O.gmres_dense(A, b, m, 1e-30)
        K = np.stack([np.linalg.matrix_power(A, k) @ b for k in range(m)], 1)
        xr = K @ np.linalg.lstsq(A @ K, b, rcond=None)[0]
        assert not ok and np.abs(x - xr).max() < 1e-11 * np.abs(xr).max(), m
    # the tolerance stops it early, at the first k whose residual is below it
    res = [np.linalg.norm(b - A @ O.gmres_dense(A, b, m, 1e-30)[1]) for m in range(1, n + 1)]
    k = next(i for i, r in enumerate(res) if r < 0.05) + 1
    ok, x, _ = O.gmres_dense(A, b, n, 0.05)
    assert ok and np.abs(x - O.gmres_dense(A, b, k, 1e-30)[1]).max() < 1e-14


@pytest.mark.parametrize("dt", [30.0, 300.0])
def test_implicit_step_solves_backward_euler(cuh_small, dt):
    # DAEImplicit::newton on c - c_n = dt c_d(c) (beta = dt, a = -c_n / dt):
    # the result is a root of F to well inside a_tol, keeps 0 <= c <= 1 and
    # conserves each element
    p, ff, x, H, elem, alpha, c = cuh_small
    d = make(ff, p, x, H, elem, alpha, c, nquad=3)
    a_tol = 1e-8
    ok, it, cn = d.newton(R_CRD, ZETA, p.masses, dt, -c / dt, a_tol=a_tol)
    assert ok and it >= 2
    assert np.abs(cn - c).max() > 1e-3
    d.set_state(x, alpha, cn)
    F = cn / dt - c / dt - d.c_d(R_CRD, ZETA, p.masses)[0]
    assert np.linalg.norm(F) < 0.01 * a_tol * np.sqrt(c.size)
    assert cn.min() >= 0.0 and cn.max() <= 1.0
    for e in (0, 1):
        assert abs((cn - c)[elem == e].sum()) < 1e-13


def test_implicit_step_that_does_not_converge_restores_c(cuh_small):
    p, ff, x, H, elem, alpha, c = cuh_small
    d = make(ff, p, x, H, elem, alpha, c, nquad=3)
    ok, it, cn = d.newton(R_CRD, ZETA, p.masses, 3000.0, -c / 3000.0, a_tol=1e-8)
    assert not ok
    np.testing.assert_array_equal(cn, c)


def one_component(seed=7, eps=None):
    p = pot.johnson_cuh(nr=3000, nrho=3000)
    ff = O.OEam.from_knots(p, dmd=True)
    x, H = pot.fcc(3.615, (2, 2, 2))
    rng = np.random.default_rng(seed)
    x = x + rng.normal(0, 0.05, x.shape)
    s = x @ np.linalg.inv(H)
    s -= np.floor(s)
    x = s @ H
    n = len(x)
    alpha = rng.uniform(0.08, 0.2, (n, 1))
    c = rng.uniform(0.3, 1.0, (n, 1))
    c[0, 0] = 0.0 if eps is None else eps  # a vacancy
    return p, ff, x, H, np.zeros((n, 1), np.uint8), alpha, c


def test_unit_gradient_is_the_force_per_unit_c():
    # calc_mu's f and f_alpha: with one component, c_i times them is minus the
    # free-energy force on that site; on the vacancy (c = 0), where the force
    # vanishes, they are the limit of -force / c as c -> 0
    p, ff, x, H, elem, alpha, c = one_component()
    d = O.ODMD(ff, H, x, elem, alpha, c, nquad=3, masses=p.masses)
    g, ga = d.unit_grad()
    f, fa, _, _, _, _ = d.force()
    live = c[:, 0] > 0.0
    np.testing.assert_allclose((g * c)[live], -f[live], rtol=0, atol=1e-12 * np.abs(f).max())
    np.testing.assert_allclose((ga * c)[live], -fa[live], rtol=0, atol=1e-12 * np.abs(fa).max())
    assert np.abs(g[0]).max() > 1e-3 and np.all(f[0] == 0.0)
    eps = 1e-7
    p, ff, x, H, elem, alpha, ce = one_component(eps=eps)
    de = O.ODMD(ff, H, x, elem, alpha, ce, nquad=3, masses=p.masses)
    fe_, fae, _, _, _, _ = de.force()
    assert np.abs(-fe_[0] / eps - g[0]).max() < 1e-5 * np.abs(g[0]).max()
    assert abs(-fae[0, 0] / eps - ga[0, 0]) < 1e-5 * abs(ga[0, 0])
