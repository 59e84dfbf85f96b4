"""Pins the oracle's neighbor list and two-pass EAM against brute force,
closed forms, finite differences and invariants."""
import numpy as np
import pytest

import eam_ref
import oracle_lib as O
from mapp4py_amd import potentials as pot


def jitter(x, sigma, seed):
    return x + np.random.default_rng(seed).normal(0.0, sigma, x.shape)


@pytest.fixture(scope="module")
def cu():
    p = pot.johnson_cu(nr=3000, nrho=3000)
    return p, O.OEam.from_knots(p)


@pytest.fixture(scope="module")
def feni():
    p = pot.johnson_feni_fs(nr=3000, nrho=3000)
    return p, O.OEam.from_knots(p)


def triclinic(H, tilt):
    H = np.array(H, float)
    H[1, 0] += tilt[0]
    H[2, 0] += tilt[1]
    H[2, 1] += tilt[2]
    return H


@pytest.mark.parametrize("cells,tilt", [((3, 3, 3), (0, 0, 0)), ((2, 2, 3), (0.4, -0.3, 0.5)),
                                        ((1, 2, 2), (0.2, 0.1, -0.2))])
def test_half_list_equals_brute_force(cu, cells, tilt):
    p, ff = cu
    x, H = pot.fcc(3.615, cells)
    H = triclinic(H, tilt)
    x = jitter(x, 0.1, 3)
    elem = np.zeros(len(x), np.uint8)
    md = O.OMD(ff, H, x, elem, skin=0.5)
    i, j, img = md.pairs()
    got = {(int(a), int(b), int(c[0]), int(c[1]), int(c[2])) for a, b, c in zip(i, j, img)}
    assert len(got) == len(i)
    cut_sk = ff.cut() + 0.5
    want = O.brute_half_pairs(H, x, elem, cut_sk, kmax=3)
    assert got == want


def test_fs_energy_forces_virial_match_closed_form(feni):
    p, ff = feni
    x, H = pot.bcc(2.86, (3, 3, 3))
    x = jitter(x, 0.08, 11)
    elem = (np.random.default_rng(5).random(len(x)) < 0.4).astype(np.uint8)
    md = O.OMD(ff, H, x, elem, skin=0.3)
    f, e, FH, rho = md.force()
    xw = md.x  # wrapped positions used by the list
    E, fa, W, rho_a = eam_ref.energy_forces(p.analytic, True, H, xw, elem, kmax=2)
    np.testing.assert_allclose(rho, rho_a, rtol=2e-7)
    assert abs(e - E) < 2e-7 * abs(E)
    np.testing.assert_allclose(f, fa, atol=2e-5)
    Wv = np.array([W[0, 0], W[0, 1], W[0, 2], W[1, 1], W[1, 2], W[2, 2]])
    np.testing.assert_allclose(FH, Wv, atol=2e-4 * np.abs(Wv).max())
    assert abs(md.energy() - e) < 1e-10 * abs(e)


def test_forces_are_minus_energy_gradient(cu):
    p, ff = cu
    x, H = pot.fcc(3.615, (3, 3, 3))
    x = jitter(x, 0.1, 1)
    elem = np.zeros(len(x), np.uint8)
    md = O.OMD(ff, H, x, elem, skin=0.5)
    f, e, FH, _ = md.force()
    x0 = md.x
    h = 1e-5
    rng = np.random.default_rng(2)
    for k in rng.choice(len(x0), 6, replace=False):
        for d in range(3):
            xp = x0.copy(); xp[k, d] += h
            xm = x0.copy(); xm[k, d] -= h
            md.x = xp; md.update(); ep = md.energy()
            md.x = xm; md.update(); em = md.energy()
            fd = -(ep - em) / (2 * h)
            assert abs(fd - f[k, d]) < 2e-6 * max(1.0, abs(f[k, d])), (k, d, fd, f[k, d])
    np.testing.assert_allclose(f.sum(0), 0.0, atol=1e-10)


def test_virial_is_minus_strain_derivative(cu):
    p, ff = cu
    x, H = pot.fcc(3.615, (3, 3, 3))
    x = jitter(x, 0.05, 4)
    elem = np.zeros(len(x), np.uint8)
    _, e0, FH, _ = O.OMD(ff, H, x, elem).force()
    h = 1e-6
    for a, b, k in [(0, 0, 0), (1, 1, 3), (2, 2, 5), (1, 0, 1), (2, 0, 2), (2, 1, 4)]:
        # lower-triangular deformation keeps H lower triangular; F_H is symmetric
        eps = np.zeros((3, 3))
        eps[a, b] = h
        ep = O.OMD(ff, H @ (np.eye(3) + eps), x @ (np.eye(3) + eps), elem).force()[1]
        em = O.OMD(ff, H @ (np.eye(3) - eps), x @ (np.eye(3) - eps), elem).force()[1]
        dE = (ep - em) / (2 * h)
        assert abs(-dE - FH[k]) < 1e-5 * max(1.0, abs(FH[k])), (a, b, -dE, FH[k])


def test_perfect_lattice_has_zero_force_and_cohesive_energy(cu):
    p, ff = cu
    x, H = pot.fcc(3.615, (4, 4, 4))
    elem = np.zeros(len(x), np.uint8)
    f, e, FH, rho = O.OMD(ff, H, x, elem).force()
    assert np.abs(f).max() < 1e-10
    assert np.ptp(rho) < 1e-10
    # diagonal virial, equal components, no shear
    assert abs(FH[0] - FH[3]) < 1e-8 and abs(FH[0] - FH[5]) < 1e-8
    assert max(abs(FH[1]), abs(FH[2]), abs(FH[4])) < 1e-8
    assert -4.0 < e / len(x) < -3.0


def test_pair_only_limit_is_plain_pair_sum(cu):
    p, _ = cu
    q = pot.EamKnots(p.names, p.masses, p.nr, p.dr, p.nrho, p.drho, np.zeros_like(p.F), p.rho, p.r_phi, False)
    ff = O.OEam.from_knots(q)
    x, H = pot.fcc(3.615, (2, 2, 2))
    x = jitter(x, 0.1, 9)
    elem = np.zeros(len(x), np.uint8)
    md = O.OMD(ff, H, x, elem)
    _, e, _, _ = md.force()
    xw = md.x
    i, j, d, r = eam_ref.pair_arrays(H, xw, p.analytic["rc"], kmax=3)
    E = 0.5 * p.analytic["phi"][0][0](r).sum()
    assert abs(e - E) < 1e-7 * abs(E)


def test_small_box_multiple_images(cu):
    # box 3.615 A across: every atom sees several images of every other atom and itself
    p, ff = cu
    x, H = pot.fcc(3.615, (1, 1, 1))
    x = jitter(x, 0.05, 6)
    elem = np.zeros(len(x), np.uint8)
    md = O.OMD(ff, H, x, elem)
    f, e, FH, rho = md.force()
    E, fa, W, rho_a = eam_ref.energy_forces(p.analytic, False, H, md.x, elem, kmax=3)
    assert abs(e - E) < 2e-7 * abs(E)
    np.testing.assert_allclose(f, fa, atol=2e-5)


# ---------------- the Nose-Hoover chain of the NVT driver ----------------
# ThermostatNHC::operator() (src/thermostat.cpp:47-104) integrates a chain of
# three thermostat velocities xi_1..3 (eta_d) acting on the reduced temperature
# T_r = T_part / T. One call is a reverse sweep (xi_3, xi_2, xi_1), the velocity
# scale exp(-ddt xi_1), and a forward sweep (xi_1, xi_2, xi_3), each chain
# update the exact flow of y' = -a y + b over ddt2 with a, b frozen. The two
# sweeps are not mirror images: the reverse sweep drives xi_2 by xi_1^2 - w^2
# (src/thermostat.cpp:79), the forward sweep by N_f xi_1^2 - w^2 (:94). A call
# is therefore consistent, at first order in ddt, with the chain
#     T_r'  = -2 xi_1 T_r
#     xi_1' = w^2 (T_r - 1) - xi_1 xi_2
#     xi_2' = c xi_1^2 - w^2 - xi_2 xi_3,     c = (1 + N_f) / 2
#     xi_3' = xi_2^2 - w^2
# with w = 1 / t_relax. These equations are integrated here independently
# (scipy DOP853) and the oracle is held to them. The same chain conserves
#     H' = K + U + Q1 xi_1^2 / 2 + Q2 (xi_2^2 + xi_3^2) / 2
#          + N_f kT eta_1 + (N_f kT / c) (eta_2 + eta_3)
# with Q1 = N_f kT t_relax^2, Q2 = Q1 / c and eta_j the time integral of xi_j
# (dH'/dt = 0 term by term, as for the textbook chain, where c = N_f). Each
# test also plants a mistake in the oracle's thermostat and shows it fails:
# ddt and ddt2 swapped, the forward coupling's sign flipped, and the textbook
# drive N_f xi_1^2 in both sweeps.


def chain_ode(n_f, w2, T_r0, ts, c):
    from scipy.integrate import solve_ivp

    def rhs(t, y):
        tr, a, b, g = y
        return [-2.0 * a * tr, w2 * (tr - 1.0) - a * b, c * a * a - w2 - b * g, b * b - w2]

    sol = solve_ivp(rhs, (0.0, ts[-1]), [T_r0, 0.0, 0.0, 0.0], method="DOP853", rtol=1e-13, atol=1e-15,
                    t_eval=ts)
    assert sol.success
    return sol.y


def test_thermostat_converges_to_the_chain_equations():
    # the thermostat alone on a reduced temperature, t_relax = 1, over six
    # t_relax at shrinking steps: first order towards c = (1 + N_f) / 2
    n_f, T_r0 = 24, 2.5

    def err(k, c, mutation=0):
        dt = 1.0 / k
        with O.nhc_mutation(mutation):
            _, h = O.nhc_run(T_r0, n_f, dt, 1.0, 6 * k)
        y = chain_ode(n_f, 1.0, T_r0, dt * np.arange(1, 6 * k + 1), c)
        return np.abs(h[:, 0] - y[0]).max()

    c = (1 + n_f) / 2.0
    e = [err(k, c) for k in (100, 200, 400)]
    # measured 4.3e-3, 2.2e-3, 1.1e-3
    assert e[2] < 2e-3
    for a, b in zip(e, e[1:]):
        assert 1.8 < a / b < 2.2, e
    # the textbook drive in both sweeps makes the call symmetric: second order
    # towards c = N_f, which is what fixes c for the reference's sweeps
    e3 = [err(k, float(n_f), 3) for k in (100, 200)]
    assert 3.6 < e3[0] / e3[1] < 4.4 and e3[1] < 1e-4, e3
    for mutation in (1, 2, 3):
        assert err(400, c, mutation) > 0.05, mutation


def zero_force_cu():
    # F = 0 and phi = 0 leave every force exactly zero; rho keeps the cut-off
    p = pot.johnson_cu(nr=2000, nrho=2000)
    q = pot.EamKnots(p.names, p.masses, p.nr, p.dr, p.nrho, p.drho, np.zeros_like(p.F), p.rho,
                     np.zeros_like(p.r_phi), False)
    return O.OEam.from_knots(q)


def test_zero_force_nvt_temperature_follows_the_chain_equations():
    # MDNVT::run (src/md_nvt.cpp:405-470) calls ThermostatNHC(dt/2, 100 dt, 3, 1)
    # twice per step; without forces nothing else changes T_part
    ff = zero_force_cu()
    x, H = pot.fcc(3.615, (2, 2, 2))
    n = len(x)
    elem = np.zeros(n, np.uint8)
    masses = [pot.MASS["Cu"]]
    v = O.o_create_T(elem, masses, pot.KB_EV, 900.0, 3)
    T0 = (masses[0] * (v ** 2).sum()) / (3 * n * pot.KB_EV)
    dt, T, nsteps = pot.FS, 300.0, 800  # eight t_relax
    ts = dt * np.arange(1, nsteps + 1)
    y = chain_ode(3 * n, 1.0 / (100.0 * dt) ** 2, T0 / T, ts, (1 + 3 * n) / 2.0)

    def run(mutation=0):
        with O.nhc_mutation(mutation):
            md = O.OMD(ff, H, x, elem, skin=0.5)
            v1, out, h, _ = md.nvt_run_hist(v, masses, pot.KB_EV, T, dt, nsteps)
        assert out[0] == 0.0
        return h

    h = run()
    # measured 2.5e-3 at w dt = 0.01 (T_r from 2.8 to 0.9)
    assert np.abs(h[:, 0] / T - y[0]).max() < 5e-3
    # the chain itself, in units of w: measured 2.2e-2 (xi_2 at the end)
    np.testing.assert_allclose(h[:, 2:5] * (100.0 * dt), y[1:].T * (100.0 * dt), atol=4e-2)
    for mutation in (1, 2, 3):
        assert np.abs(run(mutation)[:, 0] / T - y[0]).max() > 0.1, mutation


def test_nvt_conserves_the_chains_extended_energy(cu):
    # with forces, cooling from about 600 K to a 300 K target over 20 t_relax
    p, ff = cu
    x, H = pot.fcc(3.615, (3, 3, 3))
    x = jitter(x, 0.05, 12)
    n = len(x)
    n_f = 3 * n
    elem = np.zeros(n, np.uint8)
    masses = [pot.MASS["Cu"]]
    v = O.o_create_T(elem, masses, pot.KB_EV, 600.0, 7)
    dt, T, nsteps = pot.FS, 300.0, 2000
    kT, tau = pot.KB_EV * T, 100.0 * dt
    c = (1 + n_f) / 2.0
    Q1 = n_f * kT * tau ** 2

    def run(mutation=0):
        with O.nhc_mutation(mutation):
            md = O.OMD(ff, H, x, elem, skin=0.5)
            _, e0, _, _ = md.force()
            _, _, h, nreb = md.nvt_run_hist(v, masses, pot.KB_EV, T, dt, nsteps)
        k = 0.5 * n_f * pot.KB_EV * h[:, 0]
        ext = (k + h[:, 1] + 0.5 * Q1 * h[:, 2] ** 2 + 0.5 * (Q1 / c) * (h[:, 3] ** 2 + h[:, 4] ** 2)
               + n_f * kT * h[:, 5] + (n_f * kT / c) * (h[:, 6] + h[:, 7]))
        h0 = 0.5 * masses[0] * (v ** 2).sum() + e0
        return np.abs(ext - h0).max(), (k + h[:, 1])[-1] - h0, nreb

    drift, work, nreb = run()
    assert nreb > 10 and work < -3.0  # the chain removed several eV
    # measured 3.6e-3 eV against 3.8 eV of thermostat work
    assert drift < 1e-2, drift
    for mutation in (1, 2, 3):
        # measured 3.8, 196 and 0.52 eV
        assert run(mutation)[0] > 0.2, mutation


def test_threaded_oracle_is_the_same_computation(feni):
    # the CPU baseline's all-cores mode: identical half list (same rows, same
    # order) and the same forces, energy and virial to rounding of the sums
    p, ff = feni
    x, H = pot.bcc(2.86, (6, 6, 6))
    x = jitter(x, 0.08, 13)
    elem = (np.random.default_rng(14).random(len(x)) < 0.3).astype(np.uint8)
    one = O.OMD(ff, H, x, elem, skin=0.4)
    f1, e1, W1, rho1 = one.force()
    pairs1 = one.pairs()
    with O.threads(4) as k:
        assert k == 4
        many = O.OMD(ff, H, x, elem, skin=0.4)
        f4, e4, W4, rho4 = many.force()
        pairs4 = many.pairs()
        masses = [pot.MASS["Fe"], pot.MASS["Ni"]]
        v = O.o_create_T(elem, masses, pot.KB_EV, 800.0, 3)
        v4, out4, T4, nreb4 = many.nvt_run(v, masses, pot.KB_EV, 800.0, pot.FS, 60)
    for a, b in zip(pairs1, pairs4):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(f4, f1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(rho4, rho1, rtol=1e-14)
    assert abs(e4 - e1) < 1e-11 * abs(e1)
    np.testing.assert_allclose(W4, W1, rtol=0, atol=1e-10)
    v1, out1, T1, nreb1 = one.nvt_run(v, masses, pot.KB_EV, 800.0, pot.FS, 60)
    assert nreb1 == nreb4 >= 1
    np.testing.assert_allclose(v4, v1, rtol=0, atol=1e-9)
    assert abs(T4 - T1) < 1e-9 * T1
