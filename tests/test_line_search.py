"""The library's line searches (host code shared by the GPU minimisers,
reached here through mapp_line_search) against the oracle's restatement
(oracle/o_min.py), on energies along a line given as plain functions: the
same trial steps in the same order, the same status, step and energy.

Runs on the CPU: the line searches are host code and touch no GPU.
"""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import o_min  # noqa: E402
from mapp4py_amd import _lib  # noqa: E402

ENERGY = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)
RESET = C.CFUNCTYPE(None, C.c_void_p)

FUNCS = {
    "quad_inside": (lambda a: (a - 0.3) ** 2 - 0.09, -0.6, 1.0),
    "quad_far": (lambda a: (a - 2.5) ** 2 - 6.25, -5.0, 10.0),
    "quad_beyond_cap": (lambda a: (a - 7.0) ** 2 - 49.0, -14.0, 3.0),
    "quad_tiny": (lambda a: 1e4 * (a - 1e-4) ** 2 - 1e-4, -2.0, 1.0),
    "quartic": (lambda a: (a - 0.4) ** 4 - 0.0256 - 0.1 * a, -0.3560, 2.0),
    "double_well": (lambda a: (a * a - 1.0) ** 2 * 0.5 - 0.3 * a, -0.3, 3.0),
    "steep_then_flat": (lambda a: -math.tanh(20.0 * a), -20.0, 1.0),
    "uphill": (lambda a: a, -1.0, 1.0),
    "flat": (lambda a: 0.0, -1.0, 1.0),
}


class Line1D:
    def __init__(self, fn, dfa, max_a):
        self.fn, self.dfa, self.max_a = fn, dfa, max_a
        self.calls = []

    def ls_prep(self):
        return self.dfa, 1.0, self.max_a

    def F(self, a):
        self.calls.append(float(a))
        return self.fn(float(a))

    def F_reset(self):
        self.calls.append("reset")


def lib_search(kind, fn, dfa, max_a, nrgy, init_flag=0, prev_val=-1.0, **kw):
    p = _lib.MinParams(1, 2, 1e-8, 1.0, 10, kw.get("c", 0.4), kw.get("rho", 0.5), kw.get("min_alpha", 0.0), kind,
                       kw.get("max_iter", 5), int(kw.get("brack", True)), kw.get("tol", math.sqrt(2 * 2.2e-16)),
                       kw.get("zeps", 2.220446049250313e-19))
    calls = []
    ef = ENERGY(lambda ctx, a: (calls.append(a), fn(a))[1])
    rf = RESET(lambda ctx: calls.append("reset"))
    pv = C.c_double(prev_val)
    e = C.c_double(nrgy)
    alpha = C.c_double()
    st = C.c_int()
    _lib.check(_lib.lib().mapp_line_search(C.byref(p), dfa, max_a, init_flag, C.byref(pv), ef, rf, None, C.byref(e),
                                          C.byref(alpha), C.byref(st)))
    return st.value, e.value, alpha.value, calls, pv.value


def oracle_search(kind, fn, dfa, max_a, nrgy, init_flag=0, prev_val=-1.0, **kw):
    if kind == 0:
        ls = o_min.BackTrack(c=kw.get("c", 0.4), rho=kw.get("rho", 0.5), min_alpha=kw.get("min_alpha", 0.0))
    elif kind == 1:
        ls = o_min.GoldenSection(tol=kw.get("tol", math.sqrt(2 * 2.2e-16)), max_iter=kw.get("max_iter", 5),
                                 brack=kw.get("brack", True))
    else:
        ls = o_min.Brent(tol=kw.get("tol", math.sqrt(2 * 2.2e-16)), max_iter=kw.get("max_iter", 5),
                         zeps=kw.get("zeps", 2.220446049250313e-19), brack=kw.get("brack", True))
    ls.prev_val = prev_val
    h = Line1D(fn, dfa, max_a)
    st, e, a = ls.min(h, nrgy, init_flag)
    return st, e, a, h.calls, ls.prev_val


def same_calls(a, b):
    assert len(a) == len(b), (a, b)
    for u, v in zip(a, b):
        if u == "reset" or v == "reset":
            assert u == v
        else:
            assert u == pytest.approx(v, rel=1e-12, abs=1e-300)


CASES = [(0, {}), (0, {"c": 0.1, "rho": 0.3}), (0, {"min_alpha": 1e-3}), (1, {}), (1, {"brack": False}),
         (1, {"max_iter": 40}), (2, {}), (2, {"brack": False}), (2, {"max_iter": 40})]


@pytest.mark.parametrize("name", sorted(FUNCS))
@pytest.mark.parametrize("kind,kw", CASES)
def test_library_line_search_follows_the_oracle(name, kind, kw):
    fn, dfa, max_a = FUNCS[name]
    nrgy = fn(0.0)
    rl = lib_search(kind, fn, dfa, max_a, nrgy, **kw)
    ro = oracle_search(kind, fn, dfa, max_a, nrgy, **kw)
    assert rl[0] == ro[0], (rl[0], ro[0])
    same_calls(rl[3], ro[3])
    assert rl[1] == pytest.approx(ro[1], rel=1e-12, abs=1e-15)
    if rl[0] == o_min.LS_S:
        assert rl[2] == pytest.approx(ro[2], rel=1e-12)
        # the memory each search leaves (src/ls.h:516,584,705,765)
        assert rl[4] == pytest.approx(ro[4], rel=1e-12)


def test_prep_failures_and_the_backtrack_memory():
    fn = FUNCS["quad_inside"][0]
    for dfa, max_a, want in [(0.0, 1.0, o_min.LS_F_GRAD0), (-1.0, 0.0, o_min.LS_MIN_ALPHA),
                             (1.0, 1.0, o_min.LS_F_DOWNHILL)]:
        for kind in (0, 1, 2):
            assert lib_search(kind, fn, dfa, max_a, 0.0)[0] == want
    # init_flag 1 (the CG's): the previous accepted slope times step caps the next start
    pv_l = pv_o = -1.0
    for dfa in (-0.6, -0.2, -0.05):
        rl = lib_search(0, fn, dfa, 1.0, fn(0.0), init_flag=1, prev_val=pv_l)
        ro = oracle_search(0, fn, dfa, 1.0, fn(0.0), init_flag=1, prev_val=pv_o)
        same_calls(rl[3], ro[3])
        pv_l, pv_o = rl[4], ro[4]
        assert pv_l == pytest.approx(pv_o, rel=1e-14)
    # init_flag 2: the previous energy caps the next start
    # (src/ls.h:745-750); prev_val keeps minus the energy
    pv_l = pv_o = -1.0
    for dfa, e0 in ((-0.6, 0.5), (-0.2, 0.3), (-0.05, 0.29)):
        rl = lib_search(0, fn, dfa, 1.0, e0, init_flag=2, prev_val=pv_l)
        ro = oracle_search(0, fn, dfa, 1.0, e0, init_flag=2, prev_val=pv_o)
        assert rl[0] == ro[0]
        same_calls(rl[3], ro[3])
        pv_l, pv_o = rl[4], ro[4]
        assert pv_l == pv_o == -e0
    # a flag outside 0..2 leaves max_a and the memory as they are
    rl = lib_search(0, fn, -0.6, 3.0, fn(0.0), init_flag=5, prev_val=0.7)
    ro = oracle_search(0, fn, -0.6, 3.0, fn(0.0), init_flag=5, prev_val=0.7)
    same_calls(rl[3], ro[3])
    assert rl[4] == ro[4] == 0.7


BAD_PARAMS = [  # (ls, c, rho, min_alpha, max_iter, tol, zeps), each with one value out of range
    (3, 0.4, 0.5, 0.0, 5, 1e-8, 1e-19), (0, 1.0, 0.5, 0.0, 5, 1e-8, 1e-19), (0, 0.0, 0.5, 0.0, 5, 1e-8, 1e-19),
    (0, 0.4, 1.0, 0.0, 5, 1e-8, 1e-19), (0, 0.4, math.inf, 0.0, 5, 1e-8, 1e-19), (0, 0.4, 0.5, -1.0, 5, 1e-8, 1e-19),
    (1, 0.4, 0.5, 0.0, -1, 1e-8, 1e-19), (2, 0.4, 0.5, 0.0, 5, -1e-8, 1e-19), (2, 0.4, 0.5, 0.0, 5, 1e-8, math.nan),
    (0, math.nan, 0.5, 0.0, 5, 1e-8, 1e-19)]


@pytest.mark.parametrize("bad", BAD_PARAMS)
def test_rejects_parameters_under_which_a_search_may_not_end(bad):
    ls, c, rho, min_alpha, max_iter, tol, zeps = bad
    p = _lib.MinParams(1, 2, 1e-8, 1.0, 10, c, rho, min_alpha, ls, max_iter, 1, tol, zeps)
    pv, e, a, st = C.c_double(-1.0), C.c_double(0.0), C.c_double(), C.c_int()
    ef = ENERGY(lambda ctx, x: x * x)
    rf = RESET(lambda ctx: None)
    with pytest.raises(_lib.MappError, match="line search"):
        _lib.check(_lib.lib().mapp_line_search(C.byref(p), -1.0, 1.0, 0, C.byref(pv), ef, rf, None, C.byref(e),
                                              C.byref(a), C.byref(st)))
    # the DMD minimisers check the same fields before they touch the system
    it, stat = C.c_int(), C.c_int()
    out4 = (C.c_double * 4)()
    for lbfgs in (0, 1):
        with pytest.raises(_lib.MappError, match="line search"):
            _lib.check(_lib.lib().mapp_dmd_min_run(None, lbfgs, 2, 10, 1e-8, 0.1, 0.01, 0.0, C.byref(p),
                                                  C.byref(it), C.byref(stat), out4))
    # valid parameters reach the handle check
    good = _lib.MinParams(1, 2, 1e-8, 1.0, 10, 0.4, 0.5, 0.0, 0, 5, 1, 1e-8, 1e-19)
    with pytest.raises(_lib.MappError, match="d must not be NULL"):
        _lib.check(_lib.lib().mapp_dmd_min_run(None, 0, 2, 10, 1e-8, 0.1, 0.01, 0.0, C.byref(good), C.byref(it),
                                              C.byref(stat), out4))
