"""The oracle's MD minimisers (oracle/o_min.py) pinned by what the
mathematics fixes:

* the back-tracking line search on a quadratic accepts the first step
  max_a rho^j inside the Armijo bound 2 (1 - c) |dfa| / h.A.h, with max_a the
  smaller of 1 and max_dx / max |h|;
* the L-BFGS two-loop direction equals H f for the dense BFGS inverse-Hessian
  update H <- (I - r s y^T) H (I - r y s^T) + r s s^T of gamma I over the
  stored pairs, oldest first;
* on a quadratic both minimisers reach A^-1 b; on the oracle EAM both reach
  the minimum scipy's L-BFGS-B finds from the same start.
"""
import os
import sys

import numpy as np
import pytest
from scipy.optimize import minimize

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import o_min  # noqa: E402
import oracle_lib as O  # noqa: E402
from mapp4py_amd import potentials as pot  # noqa: E402


class Quad:
    """E = x.A.x / 2 - b.x over [n][3] positions."""

    def __init__(self, A, b, x):
        self.A, self.b = A, b
        self.x = np.array(x, dtype=float)
        self.n = self.x.shape[0]

    def update(self):
        return False

    def energy(self):
        v = self.x.ravel()
        return 0.5 * v @ self.A @ v - self.b @ v

    def force(self):
        v = self.x.ravel()
        return (-(self.A @ v - self.b)).reshape(-1, 3), self.energy()


def spd(n, seed, cond=20.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    return Q @ np.diag(np.geomspace(1.0, cond, n)) @ Q.T


@pytest.mark.parametrize("max_dx,c,rho", [(10.0, 0.4, 0.5), (0.05, 0.4, 0.5), (10.0, 0.1, 0.3)])
def test_backtrack_first_step_is_the_armijo_closed_form(max_dx, c, rho):
    n = 12
    A = spd(n, 1) * 3.0
    b = np.random.default_rng(2).normal(size=n)
    q = Quad(A, b, np.zeros((n // 3, 3)))
    x0 = q.x.copy()
    r = o_min.min_lbfgs(q, 1, m=2, e_tol=0.0, max_dx=max_dx, ls=o_min.BackTrack(c=c, rho=rho))
    f = b.copy()  # force at x = 0
    dfa = -f @ f
    curv = f @ A @ f
    bound = 2.0 * (1.0 - c) * (-dfa) / curv
    a = min(1.0, max_dx / np.abs(f).max())
    while a > bound:
        a *= rho
    e_expect = Quad(A, b, (x0.ravel() + a * f).reshape(-1, 3)).energy()
    assert r["status"] == o_min.MIN_F_MAX_ITER and r["steps"] == 1
    assert r["e"] == pytest.approx(e_expect, rel=1e-13, abs=1e-13)
    np.testing.assert_allclose(q.x.ravel(), x0.ravel() + a * f, rtol=0, atol=1e-13)


def test_two_loop_direction_is_the_dense_bfgs_update():
    n = 9
    A = spd(n, 3)
    b = np.random.default_rng(4).normal(size=n)
    q = Quad(A, b, np.zeros((n // 3, 3)))
    seen = []

    def trace(f, h, s, y, rho, gamma):
        seen.append((f, h, s, y, rho, gamma))

    o_min.min_lbfgs(q, 7, m=3, e_tol=0.0, max_dx=10.0, trace=trace)
    checked = 0
    for f, h, s, y, rho, gamma in seen:
        k = len(s)
        H = gamma * np.eye(n)
        for i in range(k - 1, -1, -1):  # oldest pair first
            si, yi = s[i].ravel(), y[i].ravel()
            r = 1.0 / (si @ yi)
            assert rho[i] == pytest.approx(r, rel=1e-14)
            V = np.eye(n) - r * np.outer(yi, si)
            H = V.T @ H @ V + r * np.outer(si, si)
        np.testing.assert_allclose(h.ravel(), H @ f.ravel(), rtol=1e-9, atol=1e-12)
        checked += k > 0
    assert checked >= 4


@pytest.mark.parametrize("which", ["cg", "lbfgs"])
def test_quadratic_minimum(which):
    n = 15
    A = spd(n, 5)
    b = np.random.default_rng(6).normal(size=n)
    q = Quad(A, b, np.zeros((n // 3, 3)))
    if which == "cg":
        r = o_min.min_cg(q, 2000, e_tol=0.0, max_dx=10.0)
    else:
        r = o_min.min_lbfgs(q, 2000, m=5, e_tol=0.0, max_dx=10.0)
    print(which, r["steps"], r["status"])
    np.testing.assert_allclose(q.x.ravel(), np.linalg.solve(A, b), atol=1e-6)


def test_zero_steps_and_zero_gradient():
    A = np.eye(6)
    q = Quad(A, np.zeros(6), np.zeros((2, 3)))
    r = o_min.min_lbfgs(q, 0)
    assert r["status"] == o_min.MIN_F_MAX_ITER and r["steps"] == 0
    # zero force: the CG's first test counts one step (src/min_cg.h:490-494)
    r = o_min.min_cg(q, 10)
    assert r["status"] == o_min.LS_F_GRAD0 and r["steps"] == 1
    r = o_min.min_lbfgs(q, 10)
    assert r["status"] == o_min.LS_F_GRAD0 and r["steps"] == 0


def cu_system(seed, cells=3, sigma=0.08):
    p = pot.johnson_cu(nr=4000, nrho=4000)
    x, H = pot.fcc(3.615, (cells, cells, cells))
    x = np.delete(x, 5, axis=0)  # a vacancy
    x = x + np.random.default_rng(seed).normal(0.0, sigma, x.shape)
    elem = np.zeros(len(x), np.uint8)
    return p, H, x, elem


def scipy_minimum(p, H, x, elem):
    o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=1.0)

    def fg(v):
        o.x = v.reshape(-1, 3)
        o.update()
        f, e = o.force()[:2]
        return e, -f.ravel()

    r = minimize(fg, x.ravel(), jac=True, method="L-BFGS-B", options=dict(maxiter=5000, gtol=1e-10, ftol=1e-16))
    return r.fun, r.x.reshape(-1, 3)


@pytest.mark.parametrize("which", ["cg", "lbfgs"])
def test_eam_minimum_matches_scipy(which):
    p, H, x, elem = cu_system(7)
    e_ref, x_ref = scipy_minimum(p, H, x, elem)
    o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=0.5)
    if which == "cg":
        r = o_min.min_cg(o, 3000, e_tol=1e-13)
    else:
        r = o_min.min_lbfgs(o, 3000, m=5, e_tol=1e-13)
    f = o.force()[0]
    print(which, r["steps"], r["status"], r["e"], e_ref, np.abs(f).max())
    assert r["e"] == pytest.approx(e_ref, rel=1e-11)
    assert np.abs(f).max() < 1e-4
    L = np.diag(H)
    d = (o.x - x_ref + 0.5 * L) % L - 0.5 * L
    assert np.abs(d).max() < 1e-4


class Line:
    """A function of the step length alone, for the bracketing routine."""

    def __init__(self, fn):
        self.fn = fn
        self.calls = []

    def F(self, a):
        self.calls.append(float(a))
        return self.fn(float(a))


@pytest.mark.parametrize("amin,max_a", [(0.3, 1.0), (2.5, 10.0), (0.9, 1.0), (1e-4, 1.0), (7.0, 3.0)])
def test_bracket_encloses_the_minimum(amin, max_a):
    fn = lambda a: (a - amin) ** 2  # noqa: E731
    dfa = -2.0 * amin
    st, a, b, c, fa, fb, fc = o_min.bracket(Line(fn), dfa, max_a, 0.0, fn(0.0))
    assert st == o_min.B_S
    assert fa == fn(a) and fb == fn(b) and fc == fn(c)
    assert 0.0 <= a < b <= c <= max_a
    assert fb <= fa
    if c < max_a:  # a true bracket: the minimum lies inside
        assert fb <= fc and a <= amin <= c
    else:
        assert amin >= b or fb <= fc


def test_bracket_reports_no_lower_energy():
    st = o_min.bracket(Line(lambda a: 1.0 + a), -1.0, 1.0, 0.0, 1.0)[0]
    assert st == o_min.B_F_DOWNHILL


@pytest.mark.parametrize("ls", ["golden", "brent"])
def test_exact_line_minimum_on_a_quadratic(ls):
    """With enough iterations both searches land on the exact minimiser of the
    quadratic along h = f: alpha* = f.f / f.A.f."""
    n = 12
    A = spd(n, 8) * 2.0
    b = np.random.default_rng(9).normal(size=n)
    q = Quad(A, b, np.zeros((n // 3, 3)))
    f = b.copy()
    a_star = (f @ f) / (f @ A @ f)
    search = o_min.GoldenSection(max_iter=200) if ls == "golden" else o_min.Brent(max_iter=200)
    r = o_min.min_lbfgs(q, 1, m=2, e_tol=0.0, max_dx=10.0, ls=search)
    x = q.x.ravel()
    alpha = (x @ f) / (f @ f)
    e_star = Quad(A, b, (a_star * f).reshape(-1, 3)).energy()
    print(ls, alpha, a_star)
    assert alpha == pytest.approx(a_star, rel=1e-6)
    assert r["e"] == pytest.approx(e_star, rel=1e-12)


@pytest.mark.parametrize("ls", ["golden", "brent"])
def test_eam_minimum_with_golden_and_brent(ls):
    p, H, x, elem = cu_system(8)
    e_ref, _ = scipy_minimum(p, H, x, elem)
    o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=0.5)
    search = o_min.GoldenSection() if ls == "golden" else o_min.Brent()
    r = o_min.min_lbfgs(o, 3000, m=5, e_tol=1e-13, ls=search)
    print(ls, r["steps"], r["status"], r["e"], e_ref)
    assert r["e"] == pytest.approx(e_ref, rel=1e-10)
