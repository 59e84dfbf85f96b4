"""The second-derivative pass and the Newton stage of the DMD error
minimisation on the GPU (mapp_dmd_Jxa, mapp_dmd_min_newton) against the
oracle's restatement of ForceFieldEAMDMD::__prepJ_n_res / __J
(src/ff_eam_dmd.cpp:364-831) and DAE::min_error_false (src/dae.cpp:273-330).

Bars: J products and the residual within 1e-9 of their largest entry (FP64
both sides, different summation order); the Newton stage takes the oracle's
number of steps and ends within 1e-9 A (x) and 1e-10 (alpha) of its state.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

pytestmark = pytest.mark.gpu
R_CRD, ZETA = [3.5], [1.7]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


def cuh_state(seed):
    x, H, elem, alpha, c = pot.dmd_cuh_sites((3, 3, 3), seed=seed)
    rng = np.random.default_rng(seed)
    x = x + rng.normal(0, 0.05, x.shape)
    c[2] = [0.0, 0.0]      # an empty site
    c[5, 1] = 0.0          # a zero component beside a positive one
    return x, H, elem, alpha, c


@pytest.mark.parametrize("seed", [3, 8])
def test_jacobian_products_match_the_oracle(M, seed):
    p = pot.johnson_cuh(nr=3000, nrho=3000)
    x, H, elem, alpha, c = cuh_state(seed)
    n = len(x)
    say(f"Jxa: Cu-H {n} sites x 2 components, seed {seed}")
    g = M.AtomsDMD(M.ForceFieldEAM.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    o = O.ODMD(O.OEam.from_knots(p, dmd=True), H, x, elem, alpha, c, nquad=3, masses=p.masses)
    rng = np.random.default_rng(seed + 100)
    for k in range(3):
        dx = rng.normal(0, 1.0, x.shape)
        da = rng.normal(0, 0.1, alpha.shape)
        if k == 1:
            dx[:] = 0.0
        if k == 2:
            da[:] = 0.0
        jx, ja, fx, fa = g.Jxa(dx, da, with_f=True)
        ox, oa = o.Jxa(dx, da)
        og, oga = o.unit_grad()
        sj = max(np.abs(ox).max(), np.abs(oa).max())
        sf = max(np.abs(og).max(), np.abs(oga).max())
        ej = max(np.abs(jx - ox).max(), np.abs(ja - oa).max()) / sj
        ef = max(np.abs(fx - og).max(), np.abs(fa - oga).max()) / sf
        say(f"  product {k}: rel |dJ| {ej:.1e} (scale {sj:.2e}), rel |df| {ef:.1e}")
        assert ej < 1e-9 and ef < 1e-9


@pytest.fixture(scope="module")
def vacancy(tmp_path_factory):
    from test_cfg import HP_EXAMPLE, KB_EXAMPLE, unpack

    path = unpack("Cu_mishin.eam.alloy.gz", tmp_path_factory.mktemp("cu"))
    x, H = pot.fcc(3.6149, (3, 3, 3))
    n = len(x)
    elem = np.zeros((n, 1), np.uint8)
    alpha = np.full((n, 1), 0.2)
    c = np.ones((n, 1))
    c[7] = 0.0
    return path, H, x, elem, alpha, c, KB_EXAMPLE, HP_EXAMPLE


def test_newton_stage_matches_the_oracle(M, vacancy):
    path, H, x, elem, alpha, c, kB, hP = vacancy
    n = len(x)
    say(f"min_newton: Cu vacancy cell, {n} sites, N=5, CG first")
    g = M.AtomsDMD(M.ForceFieldEAM.setfl(path, ["Cu"], dmd=True), H, x, elem, alpha, c, nquad=5, T=300.0, kB=kB,
                   hP=hP, masses=[63.546])
    r = g.min_cg(max_iter=2000, e_tol=0.0, grad_tol=1e-6)
    xs, as_, cs = g.state()
    o = O.ODMD(O.OEam.read(path, 0, dmd=True), H, xs, elem, as_, cs, nquad=5, kB=kB, T=300.0, hP=hP,
               masses=[63.546])
    a_tol = 1e-9
    k_g, res_g = g.min_newton(a_tol, 10, 8)
    k_o, res_o = o.newton_xa(a_tol, 10, 8)
    xg, ag, _ = g.state()
    xo, ao = o.state_xa()
    dxs = (xg - xo) @ np.linalg.inv(H)
    dxs -= np.round(dxs)
    dx = np.abs(dxs @ H).max()
    say(f"  CG {r['iters']} iterations; Newton: GPU {k_g} steps |f| {res_g:.2e}, oracle {k_o} steps |f| {res_o:.2e}; "
        f"max|dx| {dx:.1e} A, max|dalpha| {np.abs(ag - ao).max():.1e}")
    tol = a_tol * np.sqrt(4 * n)
    assert k_g == k_o and k_g >= 2
    assert res_g <= tol and res_o <= tol
    assert dx < 1e-9 and np.abs(ag - ao).max() < 1e-10
    # the vacancy's width moved: the free-energy CG leaves it where it was
    assert abs(as_[7, 0] - 0.2) < 1e-12 and abs(ag[7, 0] - 0.2) > 0.05
