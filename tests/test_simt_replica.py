"""The batched replica-minimisation kernel source
(mapp4py_amd/csrc/replica.hip) run on the CPU under the SIMT emulation
(tests/simt) and checked against the oracle's MinCG with the back-tracking
search, replica by replica, with one wave and with four waves per replica and
through graph replay. This pins the kernel's logic without a GPU; the GPU
tests pin the same runs on the device.
"""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot
from test_oracle_gcmc import small_cell

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import o_min  # noqa: E402

SIMT = os.path.join(HERE, "simt")
CSRC = os.path.join(ROOT, "mapp4py_amd", "csrc")
BUILD = os.path.join(SIMT, "_build")
NSTEPS = 6

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
u8p = C.POINTER(C.c_uint8)


def P(a, t=dp):
    return a.ctypes.data_as(t)


@pytest.fixture(scope="module")
def emu():
    os.makedirs(BUILD, exist_ok=True)
    src = open(os.path.join(CSRC, "replica.hip")).read()
    src = re.sub(r"(\w+)<<<(.*?)>>>\(", r"::emu::launch(\1, \2, ", src)
    with open(os.path.join(BUILD, "replica_emu.cpp"), "w") as fh:
        fh.write(src)
    lib = os.path.join(BUILD, "libreplica_emu.so")
    subprocess.run(["g++", "-std=c++20", "-O1", "-fPIC", "-shared", "-Wno-unknown-pragmas", "-I", SIMT, "-I", CSRC,
                    os.path.join(BUILD, "replica_emu.cpp"), os.path.join(SIMT, "emu.cpp"),
                    os.path.join(SIMT, "replica_harness.cpp"), os.path.join(CSRC, "tables.cpp"), "-o", lib,
                    "-lpthread"], check=True)
    L = C.CDLL(lib)
    L.emu_replica_create.restype = C.c_void_p
    L.emu_replica_create.argtypes = [C.c_char_p, C.c_int, ip, dp, dp, u8p, C.c_double]
    L.emu_replica_free.argtypes = [C.c_void_p]
    L.emu_replica_min_cg.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, ip, ip, dp, dp]
    L.emu_replica_set_waves.argtypes = [C.c_void_p, ip]
    L.emu_replica_use_graph.argtypes = [C.c_void_p, C.c_int]
    L.emu_replica_get_x.argtypes = [C.c_void_p, C.c_int, dp]
    return L


@pytest.fixture(scope="module")
def cells():
    out = []
    for s, f in ((3, 0.1), (7, 0.0)):
        x, H, elem = small_cell(s, k=(3, 3, 3), h_frac=f)
        out.append((H, x, elem))
    ff = O.OEam.read(pot.NIALH_JEA)
    ref = []
    for H, x, elem in out:
        o = O.OMD(ff, H, x, elem, skin=0.5)
        ref.append((o_min.min_cg(o, NSTEPS, e_tol=0.0, max_dx=1.0), o.x))
    return out, ref


def run(emu, cells, waves, graph=False):
    natms = np.array([len(x) for _, x, _ in cells], np.int32)
    H = np.ascontiguousarray(np.concatenate([h.reshape(1, 9) for h, _, _ in cells]))
    x = np.ascontiguousarray(np.concatenate([xx for _, xx, _ in cells]))
    elem = np.ascontiguousarray(np.concatenate([e for _, _, e in cells]).astype(np.uint8))
    h = emu.emu_replica_create(pot.NIALH_JEA.encode(), len(cells), P(natms, ip), P(H), P(x), P(elem, u8p), 0.5)
    assert h
    w = np.full(5, waves, np.int32)
    emu.emu_replica_set_waves(h, P(w, ip))
    emu.emu_replica_use_graph(h, int(graph))
    n = len(cells)
    steps, status = np.zeros(n, np.int32), np.zeros(n, np.int32)
    e0, e = np.zeros(n), np.zeros(n)
    assert emu.emu_replica_min_cg(h, NSTEPS, 0.0, 1.0, P(steps, ip), P(status, ip), P(e0), P(e)) == 0
    xs = []
    for r in range(n):
        xr = np.zeros((natms[r], 3))
        emu.emu_replica_get_x(h, r, P(xr))
        xs.append(xr)
    emu.emu_replica_free(h)
    return steps, status, e0, e, xs


@pytest.mark.parametrize("waves,graph", [(1, False), (4, False), (4, True)])
def test_emulated_replicas_take_the_oracle_steps(emu, cells, waves, graph):
    cl, ref = cells
    steps, status, e0, e, xs = run(emu, cl, waves, graph)
    for r, (H, x, elem) in enumerate(cl):
        ro, xo = ref[r]
        assert steps[r] == ro["steps"] and status[r] == ro["status"]
        assert abs(e0[r] - ro["e0"]) < 1e-10 * abs(ro["e0"])
        assert abs(e[r] - ro["e"]) < 1e-10 * abs(ro["e"])
        ds = (xs[r] - xo) @ np.linalg.inv(H)
        ds -= np.round(ds)
        assert np.abs(ds @ H).max() < 1e-8
