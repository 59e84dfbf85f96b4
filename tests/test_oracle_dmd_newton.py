"""The oracle's second-derivative pass and Newton stage of the DMD error
minimisation: ForceFieldEAMDMD::__prepJ_n_res / __J
(src/ff_eam_dmd.cpp:364-831) and DAE::min_error_false (src/dae.cpp:273-330).

Pins:
* the smeared second derivatives ((1/r d/dr)^2, (1/a d/da)(1/r d/dr),
  (1/a d/da)^2) equal the closed form of the 3-D Gaussian convolution of
  r exp(-r^2/s^2) (16 Hermite points);
* J (dx, dalpha) equals minus the central difference of the gradient per
  unit c (o_dmd_unit_grad) along (dx, dalpha), on two elements with
  fs-style densities, a curved embedding function, partly empty sites and
  16 Hermite points; two planted mistakes (no F'' pass, a flipped
  (1/a d/da)^2 sign) fail that check;
* from a relaxed cell with a vacancy (the reference's diffusion example in
  small), the Newton stage drives the gradient per unit c to a_tol
  sqrt(3 n + n c_dim), quadratically.
"""
import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

S = 1.5


def closed_form2(s, r, a):
    """g = (s^2/S)^(3/2) exp(-r^2/S), S = s^2 + a^2, and D_r = (1/r) d/dr,
    D_a = (1/a) d/da applied once and twice."""
    Sg = s * s + a * a
    g = (s * s / Sg) ** 1.5 * np.exp(-r * r / Sg)
    h = -3.0 / Sg + 2.0 * r * r / Sg ** 2
    dr = -2.0 * g / Sg
    da = g * h
    drr = 4.0 * g / Sg ** 2
    dra = -2.0 * g * h / Sg + 4.0 * g / Sg ** 2
    daa = g * (h * h + 6.0 / Sg ** 2 - 8.0 * r * r / Sg ** 3)
    return g, dr, da, drr, dra, daa


@pytest.fixture(scope="module")
def gauss16():
    p = pot.gaussian_knots(s=S)
    x, H = pot.fcc(8.0, (2, 2, 2))
    n = len(x)
    return O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, np.zeros((n, 1), np.uint8), np.full((n, 1), 0.1),
                  np.full((n, 1), 0.5), nquad=16, masses=[1.0])


@pytest.mark.parametrize("r,a", [(0.7, 0.05), (1.3, 0.2), (2.5, 0.3), (3.4, 0.15), (1.0, 0.45)])
def test_second_derivatives_match_the_gaussian_closed_form(gauss16, r, a):
    out = gauss16.pair_eval2(0, 0, r, a)
    ref = closed_form2(S, r, a)
    for q, v in enumerate(ref):
        scale = max(abs(v), 1e-3)
        for k in range(3):
            assert abs(out[3 * q + k] - v) < 2e-6 * scale, (q, k, out[3 * q + k], v)


def smooth_pair(nr=16000, dr=0.0005, nrho=4000, drho=0.005):
    """two elements, every function of the Gaussian family (exact under the
    Hermite rule up to its order), distinct rho_ab and rho_ba (fs), and an
    embedding function with F'' far from zero"""
    r = np.arange(nr) * dr
    q = np.arange(nrho) * drho
    g = lambda s, amp: np.where(r < 7.5, amp * np.exp(-r * r / (s * s)), 0.0)  # noqa: E731
    rho = np.zeros((2, 2, nr))
    rho[0, 0], rho[0, 1], rho[1, 0], rho[1, 1] = g(1.5, 1.0), g(1.3, 0.7), g(1.7, 0.5), g(1.4, 0.9)
    r_phi = np.zeros((2, 2, nr))
    r_phi[0, 0] = r * g(1.2, 0.4)
    r_phi[0, 1] = r_phi[1, 0] = r * g(1.1, 0.3)
    r_phi[1, 1] = r * g(1.0, 0.2)
    F = np.stack([-2.0 * np.sqrt(q + 0.3) + 0.4 * q * q, -1.5 * np.sqrt(q + 0.2) + 0.2 * q * q])
    return pot.EamKnots(["X", "Y"], [1.0, 2.0], nr, dr, nrho, drho, F, rho, r_phi, True)


@pytest.fixture(scope="module")
def cell():
    rng = np.random.default_rng(21)
    x, H = pot.fcc(3.1, (2, 2, 2))
    n = len(x)
    x = x + rng.normal(0, 0.08, x.shape)
    elem = np.tile(np.array([0, 1], np.uint8), (n, 1))
    alpha = rng.uniform(0.12, 0.3, (n, 2))
    c = np.stack([rng.uniform(0.3, 0.7, n), rng.uniform(0.0, 0.3, n)], axis=1)
    c[3] = [0.0, 0.0]        # an empty site: fcoef 1 for its zero components
    c[5, 1] = 0.0            # a zero component beside a positive one
    return smooth_pair(), H, x, elem, alpha, c


def make(cell):
    p, H, x, elem, alpha, c = cell
    return O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=16, masses=p.masses, T=500.0)


def grad_at(o, x, alpha, c):
    o.set_state(x, alpha, c)
    g, ga = o.unit_grad()
    return np.concatenate([g.ravel(), ga.ravel()])


def fd_check(cell, seed, eps=1e-5):
    p, H, x, elem, alpha, c = cell
    o = make(cell)
    rng = np.random.default_rng(seed)
    dx = rng.normal(0, 1.0, x.shape)
    da = rng.normal(0, 0.3, alpha.shape)
    o.set_state(x, alpha, c)
    jx, ja = o.Jxa(dx, da)
    J = np.concatenate([jx.ravel(), ja.ravel()])
    fp = grad_at(o, x + eps * dx, alpha + eps * da, c)
    fm = grad_at(o, x - eps * dx, alpha - eps * da, c)
    fd = -(fp - fm) / (2 * eps)
    o.set_state(x, alpha, c)
    return float(np.abs(J - fd).max() / np.abs(fd).max()), float(np.abs(fd).max())


@pytest.mark.parametrize("seed", [1, 2])
def test_jacobian_is_minus_the_derivative_of_the_unit_gradient(cell, seed):
    err, scale = fd_check(cell, seed)
    assert scale > 1.0
    assert err < 2e-6, err


@pytest.mark.parametrize("mutation", [1, 2])
def test_planted_jacobian_mistakes_fail_the_pin(cell, mutation):
    with O.jxa_mutation(mutation):
        err, _ = fd_check(cell, 1)
    assert err > 1e-3, (mutation, err)


def test_x_and_alpha_blocks_both_enter(cell):
    # directions along x only and alpha only give the two halves of the pin
    p, H, x, elem, alpha, c = cell
    o = make(cell)
    rng = np.random.default_rng(4)
    for dx, da in ((rng.normal(size=x.shape), np.zeros_like(alpha)), (np.zeros_like(x), rng.normal(size=alpha.shape))):
        o.set_state(x, alpha, c)
        jx, ja = o.Jxa(dx, da)
        J = np.concatenate([jx.ravel(), ja.ravel()])
        eps = 1e-5
        fd = -(grad_at(o, x + eps * dx, alpha + eps * da, c) - grad_at(o, x - eps * dx, alpha - eps * da, c)) / (2 * eps)
        assert np.abs(J - fd).max() < 2e-6 * np.abs(fd).max()
        assert np.abs(jx).max() > 0 and np.abs(ja).max() > 0


@pytest.fixture(scope="module")
def vacancy_cell(tmp_path_factory):
    """a 3x3x3 fcc Cu cell of the reference's diffusion example (Mishin
    tables, 5 Hermite points, 300 K, alpha 0.2, c = 1) with one vacancy,
    relaxed in x and alpha at fixed c by scipy's L-BFGS-B on the oracle's
    free energy; the vacancy's alpha (no free-energy gradient) stays at 0.2"""
    from scipy.optimize import minimize

    from test_cfg import HP_EXAMPLE, KB_EXAMPLE, unpack

    path = unpack("Cu_mishin.eam.alloy.gz", tmp_path_factory.mktemp("cu"))
    ff = O.OEam.read(path, 0, dmd=True)
    x, H = pot.fcc(3.6149, (3, 3, 3))
    n = len(x)
    elem = np.zeros((n, 1), np.uint8)
    alpha = np.full((n, 1), 0.2)
    c = np.ones((n, 1))
    c[7] = 0.0
    o = O.ODMD(ff, H, x, elem, alpha, c, nquad=5, kB=KB_EXAMPLE, T=300.0, hP=HP_EXAMPLE, masses=[63.546])

    def fun(z):
        o.set_state(z[:3 * n].reshape(n, 3), z[3 * n:].reshape(n, 1), c)
        f, fa, _, fe, _, _ = o.force()
        return fe, -np.concatenate([f.ravel(), fa.ravel()])

    z0 = np.concatenate([x.ravel(), alpha.ravel()])
    bounds = [(None, None)] * (3 * n) + [(1e-3, None)] * n
    r = minimize(fun, z0, jac=True, method="L-BFGS-B", bounds=bounds, options={"gtol": 1e-7, "ftol": 1e-15})
    return o, r.x[:3 * n].reshape(n, 3), r.x[3 * n:].reshape(n, 1), c


def test_newton_stage_converges_quadratically(vacancy_cell):
    o, x, alpha, c = vacancy_cell
    n = len(x)
    res = [np.linalg.norm(grad_at(o, x, alpha, c))]
    o.set_state(x, alpha, c)
    for _ in range(4):
        it, r = o.newton_xa(1e-30, max_newton=1, max_gmres=8)
        res.append(r)
    # quadratic: each residual below a constant times the square of the last
    assert res[0] > 1.0, res
    for k in range(1, 4):
        assert res[k + 1] < 10.0 * res[k] ** 2 or res[k + 1] < 1e-11, res
    assert res[4] < 1e-11, res
    # the stage's own loop stops at a_tol sqrt(3 n + n c_dim)
    o.set_state(x, alpha, c)
    a_tol = 1e-9
    it, r = o.newton_xa(a_tol, max_newton=10, max_gmres=8)
    assert r <= a_tol * np.sqrt(4 * n) and 3 <= it <= 4, (it, r)
    # the end state is a stationary point of the free energy; the vacancy's
    # alpha has moved to where its gradient per unit c vanishes
    x1, a1 = o.state_xa()
    f, fa, _, _, _, _ = o.force()
    assert np.abs(f).max() < 1e-8 and np.abs(fa).max() < 1e-8
    assert abs(a1[7, 0] - 0.2) > 0.05 and abs(a1[7, 0] - np.median(a1)) < 0.05
