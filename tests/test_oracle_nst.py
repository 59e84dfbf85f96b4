"""The oracle's NST integrator (MDNST, src/md_nst.cpp).

Pins:
* MDMath::calc's matrix functions equal scipy's: E = expm(Z0), F0 = dt2
  phi(Z0), F1 = dt phi(Z1) with phi(Z) the top-right block of expm of
  [[Z, I], [0, 0]] -- for distinct diagonals and zero off-diagonals exactly,
  and for equal or near-equal diagonals up to the reference's extra half
  term in entry (2, 0) (its coincident-node second divided difference);
* with box relaxation times so long that the box cannot move, the NST
  trajectory is the NVT one;
* the box follows the target: a tensile target swells the crystal and a
  compressive one shrinks it, by about the strain the bulk modulus gives; a
  shear target tilts the box the way the stress asks, and the time-averaged
  stress approaches the target; a planted sign error in the box kick makes
  the box run away from it.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle_lib as O
from mapp4py_amd import potentials as pot

KB = 8.617330350e-5


def phi_expm(Z):
    n = Z.shape[0]
    A = np.zeros((2 * n, 2 * n))
    A[:n, :n] = Z
    A[:n, n:] = np.eye(n)
    return expm(A)[:n, n:]


def phi2(x):
    """phi''(x) for phi(x) = (e^x - 1) / x, by its series near 0"""
    if abs(x) < 1e-3:
        return 1.0 / 3.0 + x / 4.0 + x * x / 10.0
    e = np.exp(x)
    return (e * (x * x - 2 * x + 2) - 2) / x ** 3


@pytest.mark.parametrize("A,coincident", [
    ([[0.1, 0, 0], [0.05, -0.2, 0], [0.03, 0.02, 0.15]], False),
    ([[0.1, 0, 0], [0.05, 0.1, 0], [0.03, 0.02, 0.1]], True),                # equal diagonals
    ([[0.1, 0, 0], [0.05, 0.1 + 1e-12, 0], [0.03, 0.02, 0.1 - 1e-12]], True),  # nearly equal
    ([[0.3, 0, 0], [0.0, -0.1, 0], [0.07, 0.0, 0.2]], False),                # zero off-diagonals
    ([[2e-4, 0, 0], [1e-5, -3e-4, 0], [4e-6, 2e-5, 1e-4]], False),           # a real box velocity
])
def test_matrix_functions_match_scipy(A, coincident):
    A = np.array(A)
    a, dt, dt2 = 0.013, 1.3, 0.65
    E, F0, F1 = O.nst_calc(a, dt, dt2, A)
    Z0 = -dt2 * (A + a * np.eye(3))
    Et, F0t, F1t = expm(Z0), dt2 * phi_expm(Z0), dt * phi_expm(dt * A)
    if coincident:
        # With three coincident diagonal values the reference's second
        # divided differences (src/md_nst.cpp:975-998) return g''(m) where the
        # divided difference is g''(m) / 2: entry (2, 0) carries one extra
        # half term, restated as written.
        m0, m1 = np.trace(Z0) / 3.0, dt * np.trace(A) / 3.0
        Et[2, 0] += dt2 * dt2 * A[2, 1] * A[1, 0] * np.exp(m0) / 2.0
        F0t[2, 0] += dt2 ** 3 * A[2, 1] * A[1, 0] * phi2(m0) / 2.0
        F1t[2, 0] += dt ** 3 * A[2, 1] * A[1, 0] * phi2(m1) / 2.0
    # phi is the reference's four-term Taylor series below |z| = 1/64: its
    # truncation (z^4 / 120 on phi, z^3 / 30 on a divided difference) stays
    # below 2e-7 of an entry
    np.testing.assert_allclose(E, Et, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(F0, F0t, rtol=2e-7, atol=1e-14)
    np.testing.assert_allclose(F1, F1t, rtol=2e-7, atol=1e-14)


def cu_cell(cells=4, seed=3):
    p = pot.johnson_cu(nr=3000, nrho=3000)
    x, H = pot.fcc(3.615, (cells,) * 3)
    rng = np.random.default_rng(seed)
    x = x + rng.normal(0, 0.02, x.shape)
    elem = np.zeros(len(x), np.uint8)
    v = O.o_create_T(elem, [63.546], KB, 300.0, 17)
    return p, H, x, elem, v


def test_immobile_box_gives_the_nvt_trajectory():
    p, H, x, elem, v = cu_cell()
    S = np.full((3, 3), np.nan)
    np.fill_diagonal(S, 0.0)
    o1 = O.OMD(O.OEam.from_knots(p), H, x, elem)
    v1, H1, T1, out1, hist, _ = o1.nst_run(v.copy(), [63.546], KB, 300.0, 1.0, S, 60, tau=np.full((3, 3), 1e12))
    o2 = O.OMD(O.OEam.from_knots(p), H, x, elem)
    v2, out2, T2, _ = o2.nvt_run(v.copy(), [63.546], KB, 300.0, 1.0, 60)
    np.testing.assert_allclose(H1, H, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v1, v2, rtol=0, atol=1e-10)
    np.testing.assert_allclose(o1.x, o2.x, rtol=0, atol=1e-9)
    assert abs(T1 - T2) < 1e-8 * T2


def run_target(S, nsteps=600, tau=40.0):
    p, H, x, elem, v = cu_cell()
    o = O.OMD(O.OEam.from_knots(p), H, x, elem)
    return o.nst_run(v, [63.546], KB, 300.0, 1.0, S, nsteps, tau=np.full((3, 3), tau)), H


def test_hydrostatic_target_changes_the_volume_the_right_way():
    vols = {}
    for s in (-0.02, 0.0, 0.02):
        S = np.full((3, 3), np.nan)
        np.fill_diagonal(S, s)
        (v, H1, T, out, hist, _), H0 = run_target(S)
        vol = hist[:, 2] * hist[:, 4] * hist[:, 7]
        vols[s] = vol[300:].mean()
        # the stress follows the target over the second half of the run
        mean_s = hist[300:, [8, 11, 13]].mean()
        assert abs(mean_s - s) < 0.006, (s, mean_s)
    assert vols[-0.02] < vols[0.0] < vols[0.02]
    # about the bulk modulus of the Johnson Cu model (0.86 eV/A^3 at 0 K)
    B = -vols[0.0] * 0.04 / (vols[0.02] - vols[-0.02]) * -1.0
    assert 0.4 < abs(B) < 1.6, B


def test_shear_target_tilts_the_box_the_way_the_stress_asks():
    tilt = {}
    for s in (-0.01, 0.01):
        S = np.full((3, 3), np.nan)
        np.fill_diagonal(S, 0.0)
        S[1, 0] = s
        (v, H1, T, out, hist, _), H0 = run_target(S)
        tilt[s] = hist[300:, 3].mean()
        assert abs(hist[300:, 9].mean() - s) < 0.004
    assert tilt[-0.01] < 0.0 < tilt[0.01]


def test_a_wrong_sign_in_the_box_kick_fails_the_stress_check():
    S = np.full((3, 3), np.nan)
    np.fill_diagonal(S, 0.02)
    with O.nst_mutation(1):
        (v, H1, T, out, hist, nr), H0 = run_target(S, nsteps=300)
    # the box runs away (the run stops) or the stress misses the target
    if nr != -1:
        assert abs(hist[150:, [8, 11, 13]].mean() - 0.02) > 0.006
