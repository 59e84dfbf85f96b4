"""Pins the NumPy reference of the surrogate MLP: its backward pass equals
central finite differences of its own loss (float64), and the BF16-emulating
variant stays within BF16 accuracy of the exact one."""
import numpy as np

import surr_ref as R


def make(seed, n=64):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, R.NPARAMS)
    x = rng.uniform(-1, 1, (n, 3))
    t = rng.normal(0, 1, (n, 6))
    return p, x, t


def test_backward_matches_finite_differences():
    p, x, t = make(0)
    loss, g = R.loss_grad(p, x, t)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b, _ in R.SL.values()])
    h = 1e-6
    for k in idx:
        pp = p.copy(); pp[k] += h
        pm = p.copy(); pm[k] -= h
        fd = (R.loss_grad(pp, x, t)[0] - R.loss_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_bf16_emulation_stays_close_to_the_exact_gradient():
    p, x, t = make(2, 256)
    l0, g0 = R.loss_grad(p, x, t)
    l1, g1 = R.loss_grad(p, x, t, emulate=True)
    assert abs(l1 - l0) < 1e-2 * l0
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)
    assert R.bf16(1.0 + 2.0 ** -9) == 1.0 and R.bf16(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7
