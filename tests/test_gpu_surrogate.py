"""GPU checks of the learned surrogate of the DMD smeared functions.

* the training kernel's loss and gradient equal the NumPy reference that
  rounds to BF16 where the kernel feeds the matrix cores (FP32 accumulation
  order aside), and stay within BF16 accuracy of the exact gradient;
* the inference kernel (split BF16, three products per MFMA step) equals the
  exact float64 forward pass to ~1e-4 of the output scale;
* training lowers the loss, and DMD evaluations through the trained networks
  are compared with the quadrature they replace. The bars there (outputs
  within 8% of each output's range, forces within 0.1 eV/A, f_alpha 0.2,
  free energy 0.01 eV/site) are what a 64-wide BF16 network reaches; they
  are not the surrogate's target of 1e-4 eV/A at 5x the quadrature's speed,
  which it misses (bench.py's dmd.surrogate states the measured speed and
  force error against that target).
"""
import sys

import numpy as np
import pytest

import surr_ref as R
from mapp4py_amd import potentials as pot

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


def randomized(M, seed, net=1):
    s = M.Surrogate(2, box=(1.8, 6.1, 0.05, 0.30), seed=seed)
    rng = np.random.default_rng(seed)
    p = s.params(net).astype(np.float64)
    q = R.unpack(p)
    q["b2"] = rng.normal(0, 0.3, 64)
    q["b3"] = rng.normal(0, 0.3, 6)
    q["W1"][:, 3] = rng.normal(0, 0.3, 64)
    p = R.pack(q).astype(np.float32)
    s.set_params(net, p)
    return s, p.astype(np.float64), rng


def test_training_gradient_matches_the_numpy_reference(M):
    s, p, rng = randomized(M, 3)
    x = rng.uniform(-1, 1, (256, 3)).astype(np.float32)
    t = rng.normal(0, 1, (256, 6)).astype(np.float32)
    loss, g = s.loss_grad(1, x, t)
    l_em, g_em = R.loss_grad(p, x.astype(np.float64), t.astype(np.float64), emulate=True)
    l_ex, g_ex = R.loss_grad(p, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - l_em) < 1e-5 * l_em, (loss, l_em)
    for k, (a, b, _) in R.SL.items():
        err = np.linalg.norm(g[a:b] - g_em[a:b]) / np.linalg.norm(g_em[a:b])
        err_ex = np.linalg.norm(g[a:b] - g_ex[a:b]) / np.linalg.norm(g_ex[a:b])
        say(f"grad {k}: rel err vs BF16-emulating reference {err:.2e}, vs exact {err_ex:.2e}")
        assert err < 2e-3, k
        assert err_ex < 5e-2, k


def test_inference_matches_the_exact_forward_pass(M):
    s, p, rng = randomized(M, 4, net=2)
    sc = rng.uniform(0.5, 2.0, 6).astype(np.float32)
    sh = rng.normal(0, 1, 6).astype(np.float32)
    s.set_norm(2, sc, sh)
    n = 1000
    rin = np.stack([rng.uniform(1.8, 6.1, n), rng.uniform(0.05, 0.3, n), rng.uniform(0.05, 0.3, n)], 1)
    rin = rin.astype(np.float32)
    out = s.infer(2, rin)
    r64 = rin.astype(np.float64)
    x = np.stack([(r64[:, 0] - 1.8) * (2 / np.float32(6.1 - 1.8)) - 1,
                  (r64[:, 1] - 0.05) * (2 / np.float32(0.3 - 0.05)) - 1,
                  (r64[:, 2] - 0.05) * (2 / np.float32(0.3 - 0.05)) - 1], 1)
    y = R.forward(p, x)[0] * sc + sh
    err = np.abs(out - y).max(0) / np.abs(y).max(0)
    say(f"inference rel error per output {np.array2string(err, precision=2)}")
    assert err.max() < 5e-4


@pytest.fixture(scope="module")
def trained(M):
    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((4, 4, 4), seed=8)
    x = x + np.random.default_rng(9).normal(0, 0.05, x.shape)
    d = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    s = M.Surrogate(2, box=(2.2, 6.1, 0.05, 0.30), seed=1)
    l0, l1 = s.train(d, nsamples=400_000, steps=3000, batch=4096, lr=2e-3)
    say(f"training loss {l0:.3e} -> {l1:.3e}")
    return d, s, l0, l1


def test_training_lowers_the_loss(trained):
    _, _, l0, l1 = trained
    assert l1 < 0.05 * l0


def test_networks_against_the_quadrature(M, trained):
    d, s, _, _ = trained
    rng = np.random.default_rng(11)
    n = 4096
    r = rng.uniform(2.3, 6.0, n)
    a1, a2 = rng.uniform(0.05, 0.3, n), rng.uniform(0.05, 0.3, n)
    aij = np.sqrt(a1 * a1 + a2 * a2)
    for a in range(2):
        for b in range(2):
            q = d.smear(a, b, r, aij)
            want = np.stack([q[:, 0], q[:, 1], q[:, 3] * r, q[:, 4] * r, q[:, 6] * aij, q[:, 7] * aij], 1)
            got = s.infer(a * 2 + b, np.stack([r, a1, a2], 1))
            err = np.abs(got - want).max(0)
            rel = err / (np.abs(want).max(0) + 1e-12)
            say(f"net ({a},{b}) max abs error per output {np.array2string(err, precision=2)} "
                f"(relative to the output range {np.array2string(rel, precision=2)})")
            # measured at most 5% of an output's range after 3000 steps
            assert np.all(rel < 0.08)


def test_dmd_evaluation_through_the_networks(M, trained):
    d, s, _, _ = trained
    f0, fa0, mu0, fe0, FH0, S0 = d.derivative()
    d.use_surrogate(s)
    try:
        f1, fa1, mu1, fe1, FH1, S1 = d.derivative()
        nets, quad = d.surrogate_coverage()
    finally:
        d.use_surrogate(None)
    f2 = d.derivative()[0]
    np.testing.assert_array_equal(f2, f0)  # switching back restores the quadrature exactly
    df, dfa = np.abs(f1 - f0).max(), np.abs(fa1 - fa0).max()
    say(f"surrogate vs quadrature: max|df_x| {df:.2e} eV/A, max|df_alpha| {dfa:.2e} eV/A, "
        f"|dFE|/N {abs(fe1 - fe0) / len(f0):.2e} eV, network calls {nets}, quadrature slots {quad}")
    assert nets > 0
    # What this network achieves, not the target: forces 5.7e-2 eV/A and the
    # free energy 5.7e-3 eV per site from the quadrature (measured), about 570
    # times the 1e-4 eV/A the surrogate was meant to reach; the quadrature
    # stays the path that meets it
    assert df < 0.1 and dfa < 0.2 and abs(fe1 - fe0) / len(f0) < 0.01


def _time(fn, k=5):
    import time

    fn()
    t0 = time.perf_counter()
    for _ in range(k):
        fn()
    return (time.perf_counter() - t0) / k


def test_networks_gain_on_the_quadrature_as_its_order_grows(M):
    # What the networks buy: a fixed cost per pair whatever the number of
    # Gauss-Hermite points they were trained from, while the quadrature's
    # cost per pair grows with it. The training box covers the whole list
    # range for each rule (r_c + x_max sqrt(2) alpha_max grows with N). On
    # MI355X the FP64 quadrature stays the cheaper path at both orders; the
    # networks' relative speed improves with N.
    p = pot.johnson_cuh(nr=5000, nrho=5000, rc=5.5)
    x, H, elem, alpha, c = pot.dmd_cuh_sites((32, 32, 16), seed=4)
    x = x + np.random.default_rng(5).normal(0, 0.05, x.shape)
    ff = M.ForceFieldEAM.from_knots(p, dmd=True)
    t = {}
    for N in (3, 16):
        d = M.AtomsDMD(ff, H, x, elem, alpha, c, nquad=N, masses=p.masses)
        tq = _time(d.derivative)
        xmax = M.gauss_hermite(N)[0].max()
        s = M.Surrogate(2, box=(2.2, 5.5 + xmax * np.sqrt(2.0) * 0.30 + 0.05, 0.05, 0.30), seed=1)
        s.train(d, nsamples=100_000, steps=300, batch=4096, lr=2e-3)
        d.use_surrogate(s)
        t[N] = (tq, _time(d.derivative))
        assert d.surrogate_coverage()[1] == 0  # every pair went to the networks
        say(f"{len(x)} sites, N={N}: quadrature {t[N][0] * 1e3:.2f} ms, networks {t[N][1] * 1e3:.2f} ms per evaluation "
            f"(networks / quadrature speed {t[N][0] / t[N][1]:.2f})")
    assert t[16][0] > 2.0 * t[3][0]
    assert t[16][0] / t[16][1] > 1.2 * (t[3][0] / t[3][1])
