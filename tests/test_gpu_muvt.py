"""MDMuVT on the GPU (mapp_md_muvt_run) against the oracle's restatement of
MDMuVT::run (src/md_muvt.cpp:122-280): the same exchange decisions, atom
counts, element order and gas history, positions and velocities within the
FP32-force bar of the NVT parity tests (1e-6) after 60 steps with six
exchanges of 200 attempts; with no attempts, the GPU NVT trajectory to the
same bar."""
import math
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot
from test_oracle_gcmc import GAS, MASSES, small_cell

pytestmark = pytest.mark.gpu
KB = pot.KB_EV
HP = 4.13566766225 * 0.1 * math.sqrt(1.60217656535 / 1.66053904020)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


def wrapped_diff(a, b, H):
    ds = (a - b) @ np.linalg.inv(H)
    ds -= np.round(ds)
    return np.abs(ds @ H).max()


@pytest.mark.parametrize("mu,seed", [(-2.0, 2411895), (-2.3, 77)])
def test_muvt_matches_the_oracle(M, mu, seed):
    x, H, elem = small_cell(3, k=(4, 4, 4), h_frac=0.1)
    v = O.o_create_T(elem, MASSES, KB, 300.0, 5)
    nsteps, nevery, natt = 60, 10, 200
    say(f"muVT Ni-H: {len(x)} atoms ({(elem == GAS).sum()} H), mu {mu} eV, {nsteps} steps, exchange every "
        f"{nevery} steps x {natt} attempts, seed {seed}")
    g = M.AtomsMD(M.ForceFieldEAM.setfl(pot.NIALH_JEA), H, x, elem)
    g.x_d = v
    run = M.MuVT(mu, 300.0, pot.FS, GAS, seed, nevery=nevery, nattempts=natt, kB=KB, hP=HP)
    run.run(g, nsteps, MASSES)
    o = O.OMD(O.OEam.read(pot.NIALH_JEA), H, x, elem)
    vo, To, _, ho, co, _ = o.muvt_run(v, MASSES, KB, HP, 300.0, pot.FS, mu, GAS, seed, nevery, natt, nsteps)
    dx = wrapped_diff(g.x, o.x, H) if g.natms == o.n else np.inf
    dv = np.abs(g.x_d - vo).max() if g.natms == o.n else np.inf
    say(f"  GPU counts {list(run.counts)} atoms {g.natms}; oracle {list(co)} atoms {o.n}; max|dx| {dx:.1e} A, "
        f"max|dv| {dv:.1e}, T {run.T_part:.6f} vs {To:.6f}")
    np.testing.assert_array_equal(run.counts, co)
    assert co[0] + co[1] > 0
    assert g.natms == o.n
    np.testing.assert_array_equal(g.elem, o.elem)
    np.testing.assert_array_equal(run.hist[:, 1], ho[:, 2])
    assert dx < 1e-6 and dv < 1e-6
    np.testing.assert_allclose(run.hist[:, 0], ho[:, 0], rtol=1e-6)
    assert abs(run.T_part - To) < 1e-6 * To


def test_without_attempts_the_gpu_run_is_its_nvt(M):
    x, H, elem = small_cell(4, k=(4, 4, 4), h_frac=0.1)
    v = O.o_create_T(elem, MASSES, KB, 300.0, 9)
    ff = M.ForceFieldEAM.setfl(pot.NIALH_JEA)
    g = M.AtomsMD(ff, H, x, elem)
    g.x_d = v
    run = M.MuVT(-2.3, 300.0, pot.FS, GAS, 1, nevery=10, nattempts=0, kB=KB, hP=HP)
    run.run(g, 40, MASSES)
    g2 = M.AtomsMD(ff, H, x, elem)
    g2.x_d = v
    nvt = M.NVT(300.0, pot.FS, kB=KB)
    nvt.run(g2, 40, MASSES)
    dx = wrapped_diff(g.x, g2.x, H)
    say(f"muVT without attempts vs NVT, 40 steps: max|dx| {dx:.1e} A")
    assert g.natms == len(x) and np.all(run.counts == 0)
    # each exchange re-sorts the atoms from their caller order, so the FP32
    # pair sums run in another order than the NVT run's: the NVT parity bar
    assert dx < 1e-6 and np.abs(g.x_d - g2.x_d).max() < 1e-6
    assert abs(run.T_part - nvt.T_part) < 1e-6 * nvt.T_part


def test_rejects_a_system_beyond_the_exchange_capacity(M):
    x, H = pot.fcc(3.52, (16, 16, 17))  # 17,408 atoms
    elem = np.zeros(len(x), np.uint8)
    g = M.AtomsMD(M.ForceFieldEAM.setfl(pot.NIALH_JEA), H, x, elem)
    run = M.MuVT(-2.3, 300.0, pot.FS, GAS, 1, nevery=10, nattempts=10, kB=KB, hP=HP)
    with pytest.raises(M.MappError, match="at most 16384 atoms"):
        run.run(g, 10, MASSES)
