"""GPU parity of the batched grand-canonical exchange against the oracle.

Bars: accept/reject decisions bit-identical to the oracle's single-cell
SGCMC chains driven by the same per-cell seeds (the decision sequence is a
function of the random stream and the exchange energies, so one flipped
decision would derail the whole chain); exchange energies within 1e-9 eV
(FP64, summation order only); final atom order, elements and positions
identical; kept rho/F within 1e-10 relative. Every launch shape (one wave
or four waves per cell), the HIP graph replay and the tuner give the same
chains.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

pytestmark = pytest.mark.gpu

GAS = 2
MASSES = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]
T = 300.0
DU_TOL = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


@pytest.fixture(scope="module")
def ff(M):
    return M.ForceFieldEAM.setfl(pot.NIALH_JEA)


@pytest.fixture(scope="module")
def off():
    return O.OEam.read(pot.NIALH_JEA)


def special_cells():
    rng = np.random.default_rng(9)
    out = []
    # tilted 3x4x3 block with H
    x, H = pot.fcc(3.52, (3, 4, 3))
    Ht = H.copy()
    Ht[1, 0], Ht[2, 0], Ht[2, 1] = 0.9, -0.6, 0.7
    s = x @ np.linalg.inv(H)
    xt = s @ Ht + rng.normal(0, 0.03, x.shape)
    xh = xt[:20] + np.array([1.76, 0, 0])
    out.append((np.concatenate([xt, xh]), Ht, np.concatenate([np.zeros(len(xt), np.uint8), np.full(20, GAS, np.uint8)])))
    # a 2.5 A edge: the trial point meets its own images
    Hs = np.array([[2.5, 0, 0], [0.4, 6.0, 0], [-0.3, 0.5, 6.5]])
    out.append((rng.uniform(0, 1, (4, 3)) @ Hs, Hs, np.zeros(4, np.uint8)))
    # an empty cell: insertions only
    out.append((np.zeros((0, 3)), np.diag([7.04, 7.04, 7.04]), np.zeros(0, np.uint8)))
    return out


def batch_cells(n_random=10, seed=21, nmax=1600):
    cells = pot.gcmc_nih_cells(n_random, seed=seed, nmax=nmax, h_frac=0.05)
    return cells + special_cells()


def run_oracle(off, cells, mu, seeds, k_list):
    res = []
    for c, (x, H, e) in enumerate(cells):
        g = O.OGCMC(off, H, x, e, GAS, mu, T, MASSES, seed=int(seeds[c]))
        ds, us = [], []
        for k in k_list:
            d, u = g.attempts(k)
            ds.append(d)
            us.append(u)
        res.append((np.concatenate(ds), np.concatenate(us), g.state(), g.rng_seed, g.ngas))
    return res


def compare(M, gb, ores, dec, du, tag):
    n, ngas = gb.counts()
    nflip = 0
    for c, (d_o, u_o, st, seed_o, ngas_o) in enumerate(ores):
        if not np.array_equal(dec[c], d_o):
            k = int(np.nonzero(dec[c] != d_o)[0][0])
            raise AssertionError(f"{tag}: cell {c} decision {k}: gpu {dec[c][k]} oracle {d_o[k]} "
                                 f"dU gpu {du[c][k]!r} oracle {u_o[k]!r}")
        err = np.abs(du[c] - u_o).max() if len(u_o) else 0.0
        assert err < DU_TOL * max(1.0, np.abs(u_o).max() if len(u_o) else 1.0), (tag, c, err)
        cg = gb.cell(c)
        assert n[c] == len(st["elem"]) and ngas[c] == ngas_o, (tag, c)
        np.testing.assert_array_equal(cg["elem"], st["elem"])
        np.testing.assert_array_equal(cg["x"], st["x"])
        np.testing.assert_allclose(cg["rho"], st["rho"], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(cg["F"], st["F"], rtol=1e-10, atol=1e-12)
        assert cg["seed"] == seed_o, (tag, c)
    acc = {k: int((dec == k).sum()) for k in (1, -1, 2, -2, 0)}
    say(f"{tag}: {len(ores)} cells, decisions {acc}, max |dU - oracle| "
        f"{max(np.abs(du[c] - o[1]).max() for c, o in enumerate(ores)):.2e} eV")
    return nflip


@pytest.mark.parametrize("mu,waves,graph", [(-2.3, 1, False), (-1.9, 4, False), (-2.3, 0, True)])
def test_chains_bit_match_the_oracle(M, ff, off, mu, waves, graph):
    cells = batch_cells()
    seeds = np.arange(len(cells), dtype=np.int32) * 7919 + 1000
    gb = M.GCMCBatch(ff, cells, GAS, mu, T, MASSES, seeds, cap=[len(c[2]) + 96 for c in cells])
    if waves:
        gb.set_waves([waves] * 5)
    gb.use_graph(graph)
    k_list = [150, 250, 250] if graph else [150, 250]
    outs = [gb.attempts(k) for k in k_list]
    dec = np.concatenate([o[0] for o in outs], axis=1)
    du = np.concatenate([o[1] for o in outs], axis=1)
    ores = run_oracle(off, cells, mu, seeds, k_list)
    compare(M, gb, ores, dec, du, f"mu {mu} waves {waves or 'default'} graph {graph}")


def test_tuner_leaves_the_chains_alone(M, ff, off):
    cells = batch_cells(6, seed=5, nmax=3000)
    seeds = np.arange(len(cells), dtype=np.int32) + 77
    gb = M.GCMCBatch(ff, cells, GAS, -2.1, T, MASSES, seeds)
    w, ms = gb.tune(20)
    say(f"tuner: waves per class {w.tolist()} ms {np.round(ms, 3).tolist()}")
    assert set(w.tolist()) <= {1, 4}
    dec, du = gb.attempts(200)
    compare(M, gb, run_oracle(off, cells, -2.1, seeds, [200]), dec, du, "after tune")


def test_graph_replays_match_eager(M, ff):
    cells = batch_cells(8, seed=13)
    seeds = np.arange(len(cells), dtype=np.int32) * 31 + 5
    a = M.GCMCBatch(ff, cells, GAS, -1.9, T, MASSES, seeds)
    b = M.GCMCBatch(ff, cells, GAS, -1.9, T, MASSES, seeds)
    b.use_graph(True)
    for rep in range(4):
        da, ua = a.attempts(64)
        db, ub = b.attempts(64)
        np.testing.assert_array_equal(da, db)
        np.testing.assert_array_equal(ua, ub)
        assert np.array_equal(a.counts()[0], b.counts()[0]), rep


def test_single_trial_hooks_follow_the_oracle_chain(M, ff, off):
    # SGCMC::attmpt (src/sgcmc.cpp:255-327) kept on the caller's side -- its
    # random stream, draw order and acceptance -- with the exchange energy and
    # the commit of every trial through the single-trial hooks: the decisions
    # and energies of the oracle's own chain, trial by trial, and its final
    # atoms, densities and embedding energies
    from test_oracle_gcmc import PM, xexp

    cells = batch_cells(3, seed=17, nmax=900)
    seeds = np.arange(len(cells), dtype=np.int32) * 13 + 401
    gb = M.GCMCBatch(ff, cells, GAS, -2.0, T, MASSES, seeds, cap=[len(c[2]) + 96 for c in cells])
    k = 120
    for c, (x, H, e) in enumerate(cells):
        o = O.OGCMC(off, H, x, e, GAS, -2.0, T, MASSES, seed=int(seeds[c]))
        info = o.info()
        zv, beta, sigma = info["z_fac"] * info["vol"], info["beta"], info["sigma"]
        d_o, u_o = o.attempts(k)
        rng = PM(int(seeds[c]))
        ngas = int((e == GAS).sum())
        dec, du = np.zeros(k, np.int8), np.zeros(k)
        for a in range(k):
            if rng.uniform() < 0.5:
                sb = [rng.uniform() for _ in range(3)]
                du[a] = gb.xchng_energy(c, s=sb)
                acc = rng.uniform() < zv / ((ngas + 1.0) * xexp(beta * du[a]))
                if acc:
                    gb.xchng_commit([rng.gaussian() * sigma for _ in range(3)])
                    ngas += 1
                dec[a] = 1 if acc else -1
            else:
                if ngas == 0:
                    continue
                du[a] = gb.xchng_energy(c, igas=int(ngas * rng.uniform()))
                acc = rng.uniform() < ngas * xexp(beta * du[a]) / zv
                if acc:
                    gb.xchng_commit()
                    ngas -= 1
                dec[a] = 2 if acc else -2
        np.testing.assert_array_equal(dec, d_o)
        err = np.abs(du - u_o).max()
        assert err < DU_TOL * max(1.0, np.abs(u_o).max()), (c, err)
        st, cg = o.state(), gb.cell(c)
        np.testing.assert_array_equal(cg["elem"], st["elem"])
        np.testing.assert_array_equal(cg["x"], st["x"])
        np.testing.assert_allclose(cg["rho"], st["rho"], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(cg["F"], st["F"], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(cg["v"], st["v"], rtol=0, atol=0)
        assert rng.seed == o.rng_seed
        say(f"hooks, cell {c} ({len(e)} atoms): decisions {dict(zip(*np.unique(dec, return_counts=True)))}, "
            f"max |dU - oracle| {err:.1e} eV")
    # a trial commits once, and not after a batch launch
    gb.xchng_energy(0, s=[0.5, 0.5, 0.5])
    gb.xchng_commit([0.0, 0.0, 0.0])
    with pytest.raises(M.MappError, match="no trial to commit"):
        gb.xchng_commit([0.0, 0.0, 0.0])
    gb.xchng_energy(0, s=[0.25, 0.5, 0.5])
    gb.attempts(1)
    with pytest.raises(M.MappError, match="no trial to commit"):
        gb.xchng_commit([0.0, 0.0, 0.0])


def test_full_cell_stops_with_a_capacity_error(M, ff):
    x, H = pot.fcc(3.52, (3, 3, 3))
    cells = [(x, H, np.zeros(len(x), np.uint8))]
    gb = M.GCMCBatch(ff, cells, GAS, 5.0, T, MASSES, [3], cap=[len(x)])
    with pytest.raises(M.MappError, match="full"):
        gb.attempts(50)
