"""GPU replay of the reference's diffusion example up to its first printed
line: examples/diff/configs/Cu-DMD.cfg read by the library's CFG importer,
x and alpha relaxed at fixed c by the GPU's CG (MinCGDMD) and finished by
Newton-Krylov on the GPU's gradient, then the free energy and stress
compared with step 0 of examples/diff/output_cu."""
import math
import sys

import numpy as np
import pytest

import mapp4py_amd as M
from test_cfg import HP_EXAMPLE, KB_EXAMPLE, golden, unpack

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def example(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("cu_gpu")
    cfg, pot = unpack("Cu-DMD.cfg.gz", tmp), unpack("Cu_mishin.eam.alloy.gz", tmp)
    d = M.import_cfg_dmd(cfg)
    ff = M.ForceFieldEAM.setfl(pot, ["Cu"], dmd=True)
    g = M.AtomsDMD(ff, d["H"], d["x"], d["elem"], d["alpha"], d["c"], nquad=5, T=300.0, kB=KB_EXAMPLE,
                   hP=HP_EXAMPLE, masses=d["masses"])
    return d, g, tmp


def test_relaxed_example_matches_the_reference_output(example):
    d, g, _ = example
    gd = golden()
    r = g.min_cg(max_iter=2000, e_tol=0.0, grad_tol=1e-7)
    say(f"CG: {r['iters']} iterations status {r['status']}, FE {r['fe0']:.6f} -> {r['fe']:.6f}, "
        f"max|grad| {r['grad_max']:.2e}")
    assert abs(r["fe"] - gd["FE"]) < 0.005
    # Newton-Krylov on the GPU gradient (the reference finishes with Newton)
    from scipy.optimize import newton_krylov

    xs, as_, cs = g.state()
    n = len(cs)

    def grad(v):
        g.set_state(v[:3 * n].reshape(n, 3), v[3 * n:].reshape(n, 1), cs)
        g.update()
        f, fa, _, _, _, _ = g.derivative()
        return -np.concatenate([f.ravel(), fa.ravel()])

    v = newton_krylov(grad, np.concatenate([xs.ravel(), as_.ravel()]), f_tol=1e-9, maxiter=30, inner_maxiter=20,
                      rdiff=1e-7)
    g.set_state(v[:3 * n].reshape(n, 3), v[3 * n:].reshape(n, 1), cs)
    g.update()
    f, fa, _, fe, FH, _ = g.derivative()
    S = -np.asarray(FH) / np.linalg.det(d["H"])
    say(f"relaxed on the GPU: FE {fe:.6f} (printed {gd['FE']:e}), S diag {S[0]:.8e} {S[3]:.8e} {S[5]:.8e} "
        f"(printed {gd['S_diag'][0]:e}), max|grad| {max(np.abs(f).max(), np.abs(fa).max()):.1e}")
    assert max(np.abs(f).max(), np.abs(fa).max()) < 1e-8
    assert abs(fe - gd["FE"]) < 0.005
    for k in (0, 3, 5):
        assert abs(S[k] - gd["S_diag"][0]) < 5e-9
    for k in (1, 2, 4):
        assert abs(S[k]) < 1e-12


def test_library_newton_stage_reaches_the_reference_output(example, tmp_path):
    # DAE::min_error_false as the example runs it: CG, then the Newton stage
    # with the example's a_tol 1e-7, 10 Newton steps and GMRES(4)
    # (examples/diff/in_Cu_diff.py), then the same stage to a tight tolerance
    d, _, _ = example
    gd = golden()
    pot = unpack("Cu_mishin.eam.alloy.gz", tmp_path)
    g = M.AtomsDMD(M.ForceFieldEAM.setfl(pot, ["Cu"], dmd=True), d["H"], d["x"], d["elem"], d["alpha"], d["c"],
                   nquad=5, T=300.0, kB=KB_EXAMPLE, hP=HP_EXAMPLE, masses=d["masses"])
    n = len(d["c"])
    r = g.min_cg(max_iter=2000, e_tol=0.0, grad_tol=1e-7)
    _, res0 = g.min_newton(1e-7, 0, 4)
    k1, res1 = g.min_newton(1e-7, 10, 4)
    k2, res2 = g.min_newton(1e-11, 20, 8)
    f, fa, _, fe, FH, _ = g.derivative()
    S = -np.asarray(FH) / np.linalg.det(d["H"])
    say(f"CG {r['iters']} iterations, max|grad| {r['grad_max']:.1e}, |f| {res0:.2e}; Newton (1e-7, 10, 4): {k1} "
        f"steps |f| {res1:.2e}; Newton (1e-11, 20, 8): {k2} steps |f| {res2:.2e}; FE {fe:.6f}, S diag {S[0]:.8e}")
    # GMRES(4) steps shrink the residual; within 10 of them the example's
    # stage either meets its tolerance or ends unconverged, as the reference's
    assert res1 < 0.5 * res0
    assert res1 <= 1e-7 * math.sqrt(4 * n) or k1 == 10
    assert res2 <= 1e-11 * math.sqrt(4 * n)
    assert max(np.abs(f).max(), np.abs(fa).max()) < 1e-8
    assert abs(fe - gd["FE"]) < 0.005
    for k in (0, 3, 5):
        assert abs(S[k] - gd["S_diag"][0]) < 5e-9


def test_export_then_import_returns_the_state(example):
    d, g, tmp = example
    p = tmp / "dump.cfg"
    g.export_cfg(p, ["Cu"])
    e = M.import_cfg_dmd(p)
    xs, as_, cs = g.state()
    # the same points modulo the box (the file holds s, wrapped on import)
    ds = (e["x"] - xs) @ np.linalg.inv(d["H"])
    ds -= np.round(ds)
    assert np.abs(ds @ d["H"]).max() < 1e-11
    s_in = e["x"] @ np.linalg.inv(d["H"])
    assert s_in.min() >= -1e-15 and s_in.max() < 1.0
    np.testing.assert_allclose(e["alpha"], as_, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(e["c"], cs)
    assert e["names"] == ["Cu"] and e["masses"][0] == pytest.approx(63.546)
