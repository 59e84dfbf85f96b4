"""The batched GCMC kernel source (mapp4py_amd/csrc/gcmc.hip) run on the CPU
under a SIMT emulation (tests/simt) and checked against the oracle's chains:
identical accept/reject decisions and final atom order, with one wave and with
four waves per cell and through graph replay. This pins the kernel's logic
without a GPU; the GPU tests pin the same chains on the device.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SIMT = os.path.join(HERE, "simt")
CSRC = os.path.join(ROOT, "mapp4py_amd", "csrc")
BUILD = os.path.join(SIMT, "_build")
GAS = 2
MASSES = [pot.MASS["Ni"], pot.MASS["Al"], pot.MASS["H"]]

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
u8p = C.POINTER(C.c_uint8)
i8p = C.POINTER(C.c_int8)


def P(a, t=dp):
    return a.ctypes.data_as(t)


@pytest.fixture(scope="module")
def emu():
    os.makedirs(BUILD, exist_ok=True)
    src = open(os.path.join(CSRC, "gcmc.hip")).read()
    # kernel launches become calls of the emulator's launcher
    src = re.sub(r"(\w+)<<<(.*?)>>>\(", r"::emu::launch(\1, \2, ", src)
    with open(os.path.join(BUILD, "gcmc_emu.cpp"), "w") as fh:
        fh.write(src)
    lib = os.path.join(BUILD, "libgcmc_emu.so")
    subprocess.run(["g++", "-std=c++20", "-O1", "-fPIC", "-shared", "-Wno-unknown-pragmas", "-I", SIMT, "-I", CSRC,
                    os.path.join(BUILD, "gcmc_emu.cpp"), os.path.join(SIMT, "emu.cpp"),
                    os.path.join(SIMT, "gcmc_harness.cpp"), os.path.join(CSRC, "tables.cpp"), "-o", lib,
                    "-lpthread"], check=True)
    L = C.CDLL(lib)
    L.emu_gcmc_create.restype = C.c_void_p
    L.emu_gcmc_create.argtypes = [C.c_char_p, C.c_int, ip, dp, dp, u8p, ip, C.c_int, C.c_double, C.c_double,
                                  C.c_double, C.c_double, dp, ip]
    L.emu_gcmc_free.argtypes = [C.c_void_p]
    L.emu_gcmc_attempts.argtypes = [C.c_void_p, C.c_int, i8p, dp]
    L.emu_gcmc_set_waves.argtypes = [C.c_void_p, ip]
    L.emu_gcmc_use_graph.argtypes = [C.c_void_p, C.c_int]
    L.emu_gcmc_counts.argtypes = [C.c_void_p, ip, ip]
    L.emu_gcmc_get_cell.argtypes = [C.c_void_p, C.c_int, dp, u8p, dp, dp, dp, ip]
    L.emu_gcmc_xchng_energy.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_int, dp]
    L.emu_gcmc_xchng_commit.argtypes = [C.c_void_p, dp]
    return L


def cells_for_test():
    rng = np.random.default_rng(9)
    cells = pot.gcmc_nih_cells(3, seed=3, nmax=400, h_frac=0.05)
    Hs = np.array([[2.5, 0, 0], [0.4, 6.0, 0], [-0.3, 0.5, 6.5]])  # the trial point meets its own images
    cells.append((rng.uniform(0, 1, (4, 3)) @ Hs, Hs, np.zeros(4, np.uint8)))
    cells.append((np.zeros((0, 3)), np.diag([7.04, 7.04, 7.04]), np.zeros(0, np.uint8)))
    return cells


def run(L, cells, mu, seeds, ks, waves, graph):
    natms = np.array([len(e) for _, _, e in cells], np.int32)
    x = np.ascontiguousarray(np.concatenate([np.asarray(c[0], float).reshape(-1, 3) for c in cells]))
    el = np.ascontiguousarray(np.concatenate([np.asarray(c[2], np.uint8) for c in cells]))
    H = np.ascontiguousarray(np.stack([np.asarray(c[1], float) for c in cells]))
    cap = natms + 96
    m = np.array(MASSES)
    g = L.emu_gcmc_create(pot.NIALH_JEA.encode(), len(cells), P(natms, ip), P(H), P(x), P(el, u8p), P(cap, ip),
                          GAS, mu, 300.0, pot.KB_EV, pot.HP, P(m), P(np.ascontiguousarray(seeds, np.int32), ip))
    assert g
    try:
        w = np.full(5, waves, np.int32)
        L.emu_gcmc_set_waves(g, P(w, ip))
        L.emu_gcmc_use_graph(g, int(graph))
        decs, dus = [], []
        for k in ks:
            d = np.zeros((len(cells), k), np.int8)
            u = np.zeros((len(cells), k))
            assert L.emu_gcmc_attempts(g, k, P(d, i8p), P(u)) == 0
            decs.append(d)
            dus.append(u)
        states = []
        for c in range(len(cells)):
            n = np.zeros(len(cells), np.int32)
            ng = np.zeros(len(cells), np.int32)
            L.emu_gcmc_counts(g, P(n, ip), P(ng, ip))
            xs, e, rho, F, v = (np.zeros((n[c], 3)), np.zeros(n[c], np.uint8), np.zeros(n[c]), np.zeros(n[c]),
                                np.zeros((n[c], 3)))
            seed = C.c_int()
            L.emu_gcmc_get_cell(g, c, P(xs), P(e, u8p), P(rho), P(F), P(v), C.byref(seed))
            states.append(dict(x=xs, elem=e, rho=rho, F=F, seed=seed.value, ngas=int(ng[c])))
        return np.concatenate(decs, 1), np.concatenate(dus, 1), states
    finally:
        L.emu_gcmc_free(g)


@pytest.mark.parametrize("mu,waves,graph,ks", [(-2.3, 1, False, [25, 35]), (-1.9, 4, True, [15, 20, 20])])
def test_emulated_kernel_reproduces_the_oracle_chains(emu, mu, waves, graph, ks):
    cells = cells_for_test()
    seeds = np.arange(len(cells), dtype=np.int32) * 7919 + 1000
    dec, du, states = run(emu, cells, mu, seeds, ks, waves, graph)
    off = O.OEam.read(pot.NIALH_JEA)
    for c, (x, H, e) in enumerate(cells):
        g = O.OGCMC(off, H, x, e, GAS, mu, 300.0, MASSES, seed=int(seeds[c]))
        d_o = np.concatenate([g.attempts(k)[0] for k in ks])
        np.testing.assert_array_equal(dec[c], d_o)
        st = g.state()
        np.testing.assert_array_equal(states[c]["elem"], st["elem"])
        np.testing.assert_array_equal(states[c]["x"], st["x"])
        np.testing.assert_allclose(states[c]["rho"], st["rho"], rtol=1e-10, atol=1e-12)
        assert states[c]["seed"] == g.rng_seed and states[c]["ngas"] == g.ngas


def test_emulated_one_cell_batch_reproduces_the_oracle_chain(emu):
    # a batch of one cell takes its set-up densities from one wave per atom
    # (the path of the MD exchange), the lanes splitting the other atoms
    cells = cells_for_test()
    x, H, e = max(cells, key=lambda c: len(c[2]))
    seed, mu, k = 4242, -2.3, 30
    dec, du, states = run(emu, [(x, H, e)], mu, np.array([seed], np.int32), [k], 4, False)
    g = O.OGCMC(O.OEam.read(pot.NIALH_JEA), H, x, e, GAS, mu, 300.0, MASSES, seed=seed)
    np.testing.assert_array_equal(dec[0], g.attempts(k)[0])
    st = g.state()
    np.testing.assert_array_equal(states[0]["elem"], st["elem"])
    np.testing.assert_array_equal(states[0]["x"], st["x"])
    np.testing.assert_allclose(states[0]["rho"], st["rho"], rtol=1e-10, atol=1e-12)
    assert states[0]["seed"] == g.rng_seed and states[0]["ngas"] == g.ngas


def test_emulated_single_trial_hooks_follow_the_oracle_chain(emu):
    # the hook kernels (one trial's exchange energy, its commit) driven by a
    # caller-side SGCMC::attmpt with the reference's draw order: the oracle's
    # decisions, energies and final cells
    from test_oracle_gcmc import PM, xexp

    cells = cells_for_test()
    mu, k = -2.0, 40
    seeds = np.arange(len(cells), dtype=np.int32) * 13 + 401
    natms = np.array([len(e) for _, _, e in cells], np.int32)
    x = np.ascontiguousarray(np.concatenate([np.asarray(c[0], float).reshape(-1, 3) for c in cells]))
    el = np.ascontiguousarray(np.concatenate([np.asarray(c[2], np.uint8) for c in cells]))
    H = np.ascontiguousarray(np.stack([np.asarray(c[1], float) for c in cells]))
    cap = natms + 96
    g = emu.emu_gcmc_create(pot.NIALH_JEA.encode(), len(cells), P(natms, ip), P(H), P(x), P(el, u8p), P(cap, ip),
                            GAS, mu, 300.0, pot.KB_EV, pot.HP, P(np.array(MASSES)), P(seeds, ip))
    assert g
    off = O.OEam.read(pot.NIALH_JEA)
    try:
        for c, (xc, Hc, e) in enumerate(cells):
            o = O.OGCMC(off, Hc, xc, e, GAS, mu, 300.0, MASSES, seed=int(seeds[c]))
            info = o.info()
            zv, beta, sigma = info["z_fac"] * info["vol"], info["beta"], info["sigma"]
            d_o, u_o = o.attempts(k)
            rng = PM(int(seeds[c]))
            ngas = int((e == GAS).sum())
            dec, du = np.zeros(k, np.int8), np.zeros(k)
            d = C.c_double()
            for a in range(k):
                if rng.uniform() < 0.5:
                    sb = np.array([rng.uniform() for _ in range(3)])
                    assert emu.emu_gcmc_xchng_energy(g, c, 1, P(sb), 0, C.byref(d)) == 0
                    du[a] = d.value
                    acc = rng.uniform() < zv / ((ngas + 1.0) * xexp(beta * du[a]))
                    if acc:
                        v = np.array([rng.gaussian() * sigma for _ in range(3)])
                        assert emu.emu_gcmc_xchng_commit(g, P(v)) == 0
                        ngas += 1
                    dec[a] = 1 if acc else -1
                else:
                    if ngas == 0:
                        continue
                    assert emu.emu_gcmc_xchng_energy(g, c, 2, None, int(ngas * rng.uniform()), C.byref(d)) == 0
                    du[a] = d.value
                    acc = rng.uniform() < ngas * xexp(beta * du[a]) / zv
                    if acc:
                        assert emu.emu_gcmc_xchng_commit(g, None) == 0
                        ngas -= 1
                    dec[a] = 2 if acc else -2
            np.testing.assert_array_equal(dec, d_o)
            assert np.abs(du - u_o).max() < 1e-9 * max(1.0, np.abs(u_o).max())
            n = np.zeros(len(cells), np.int32)
            ng = np.zeros(len(cells), np.int32)
            emu.emu_gcmc_counts(g, P(n, ip), P(ng, ip))
            xs, es, rho, F, v = (np.zeros((n[c], 3)), np.zeros(n[c], np.uint8), np.zeros(n[c]), np.zeros(n[c]),
                                 np.zeros((n[c], 3)))
            seed = C.c_int()
            emu.emu_gcmc_get_cell(g, c, P(xs), P(es, u8p), P(rho), P(F), P(v), C.byref(seed))
            st = o.state()
            np.testing.assert_array_equal(es, st["elem"])
            np.testing.assert_array_equal(xs, st["x"])
            np.testing.assert_allclose(rho, st["rho"], rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(F, st["F"], rtol=1e-10, atol=1e-12)
            assert ng[c] == o.ngas and rng.seed == o.rng_seed
        # a trial commits once; a second commit is refused (MAPP_ERR_STATE)
        sb, v, d = np.array([0.5, 0.5, 0.5]), np.zeros(3), C.c_double()
        assert emu.emu_gcmc_xchng_energy(g, 0, 1, P(sb), 0, C.byref(d)) == 0
        assert emu.emu_gcmc_xchng_commit(g, P(v)) == 0
        assert emu.emu_gcmc_xchng_commit(g, P(v)) == 5
    finally:
        emu.emu_gcmc_free(g)
