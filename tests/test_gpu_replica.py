"""Batched replica minimisation (mapp_replica_min_cg) against the oracle's
MinCG with LineSearchBackTrack (oracle/o_min.py over the oracle EAM), one
oracle run per replica.

Both sides are FP64, so each replica takes the oracle's steps: the same step
count and status, the final energy within 1e-10 relative and the positions
within 1e-7 A (modulo the box). The batch mixes sizes (107-500 atoms), one
and three elements, a tilted box and vacancies; the one-wave and four-wave
launch shapes, the tuner and graph replays give the same results.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import o_min  # noqa: E402
import oracle_lib as O  # noqa: E402
from mapp4py_amd import potentials as pot  # noqa: E402
from test_oracle_gcmc import small_cell  # noqa: E402

pytestmark = pytest.mark.gpu
NSTEPS = 15


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


def cu_cells():
    cells = []
    rng = np.random.default_rng(11)
    for k, (shape, tilt, vac) in enumerate([((3, 3, 3), 0.0, 0), ((4, 4, 4), 0.0, 3), ((3, 3, 5), 0.0, 1),
                                            ((3, 4, 3), 0.9, 0), ((5, 5, 4), 0.0, 2), ((3, 3, 3), -0.6, 1)]):
        x, H = pot.fcc(3.615, shape)
        if tilt:
            s = x @ np.linalg.inv(H)
            H = H.copy()
            H[1, 0] = tilt
            H[2, 1] = 0.5 * tilt
            x = s @ H
        if vac:
            x = np.delete(x, rng.choice(len(x), vac, replace=False), axis=0)
        x = x + rng.normal(0.0, 0.08, x.shape)
        cells.append((H, x, np.zeros(len(x), np.uint8)))
    return cells


def oracle_min(ff, H, x, elem):
    o = O.OMD(ff, H, x, elem, skin=0.5)
    r = o_min.min_cg(o, NSTEPS, e_tol=0.0, max_dx=1.0)
    return r, o.x


def wrapped_diff(a, b, H):
    ds = (a - b) @ np.linalg.inv(H)
    ds -= np.round(ds)
    return float(np.abs(ds @ H).max())


def check(batch, cells, res, oracle, label):
    worst_e, worst_x = 0.0, 0.0
    for r, (H, x, elem) in enumerate(cells):
        ro, xo = oracle[r]
        assert res["status"][r] == ro["status"], (label, r, res["status"][r], ro["status"])
        assert res["steps"][r] == ro["steps"], (label, r, res["steps"][r], ro["steps"])
        de = abs(res["e"][r] - ro["e"]) / abs(ro["e"])
        dx = wrapped_diff(batch.x(r), xo, H)
        worst_e, worst_x = max(worst_e, de), max(worst_x, dx)
        assert abs(res["e0"][r] - ro["e0"]) < 1e-10 * abs(ro["e0"])
        assert de < 1e-10 and dx < 1e-7, (label, r, de, dx)
    say(f"  {label}: {len(cells)} replicas, steps {list(res['steps'])}, worst rel dE {worst_e:.1e}, "
        f"worst dx {worst_x:.1e} A")


@pytest.fixture(scope="module")
def batch_inputs():
    p = pot.johnson_cu(nr=3000, nrho=3000)
    cells = cu_cells()
    ffo = O.OEam.from_knots(p)
    oracle = [oracle_min(ffo, *c) for c in cells]
    return p, cells, oracle


@pytest.mark.parametrize("waves", [1, 4])
def test_replicas_take_the_oracle_steps(M, batch_inputs, waves):
    p, cells, oracle = batch_inputs
    say(f"replica MinCG, Cu, {len(cells)} replicas, {waves} wave(s) each, {NSTEPS} steps")
    b = M.ReplicaBatch(M.ForceFieldEAM.from_knots(p), cells, skin=0.5)
    b.set_waves([waves] * 5)
    res = b.min_cg(NSTEPS, e_tol=0.0, max_dx=1.0)
    check(b, cells, res, oracle, f"{waves} wave(s)")
    assert np.all(res["e"] < res["e0"])


def test_three_element_replicas(M):
    ff = O.OEam.read(pot.NIALH_JEA)
    cells = []
    for s, f in ((3, 0.1), (4, 0.3), (5, 0.0)):
        x, H, elem = small_cell(s, k=(3, 3, 3), h_frac=f)
        cells.append((H, x, elem))
    oracle = [oracle_min(ff, *c) for c in cells]
    say(f"replica MinCG, Ni-H, {len(cells)} replicas")
    b = M.ReplicaBatch(M.ForceFieldEAM.setfl(pot.NIALH_JEA), cells, skin=0.5)
    res = b.min_cg(NSTEPS, e_tol=0.0, max_dx=1.0)
    check(b, cells, res, oracle, "Ni-H")


def test_tuner_and_graph_replays(M, batch_inputs):
    p, cells, oracle = batch_inputs
    b = M.ReplicaBatch(M.ForceFieldEAM.from_knots(p), cells, skin=0.5)
    waves, ms = b.tune(NSTEPS, e_tol=0.0, max_dx=1.0)
    say(f"  tuner: waves per class {list(waves)}, ms per threshold {[round(t, 2) for t in ms]}")
    b.use_graph(True)
    res = b.min_cg(NSTEPS, e_tol=0.0, max_dx=1.0)
    check(b, cells, res, oracle, "tuned, graph")
    # a second replay continues from the minimised positions: fewer steps,
    # no energy above the first run's end
    res2 = b.min_cg(NSTEPS, e_tol=0.0, max_dx=1.0)
    assert np.all(res2["e0"] == res["e"]) and np.all(res2["e"] <= res["e"])


def test_rejects_a_box_thinner_than_the_cutoff(M):
    p = pot.johnson_cu(nr=3000, nrho=3000)
    x, H = pot.fcc(3.615, (1, 3, 3))
    with pytest.raises(M.MappError, match="thinner than r_c"):
        M.ReplicaBatch(M.ForceFieldEAM.from_knots(p), [(H, x, np.zeros(len(x), np.uint8))])
