"""mapp_md_value_noalloc, the energy call that allocates no device memory,
against mapp_md_value (the entry point it varies) and the oracle, through the
C ABI: equal energies on two calls in a row (the second on a smaller system),
the last force call's densities left readable, sizes beyond its scratch rows
rejected with MAPP_ERR_CAPACITY.

Bar: the energy of mapp_md_value to 5e-8 eV per atom + 1e-6 eV (the EAM
energy bar); both calls run the same deterministic passes on the same inputs.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

pytestmark = pytest.mark.gpu

E_ATOM = 5e-8
MAX_ATOMS = 1 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


@pytest.fixture(scope="module")
def cu_pot():
    return pot.johnson_cu(nr=10000, nrho=10000)


def cu_system(M, p, cells, seed):
    x, H = pot.fcc(3.615, cells)
    x = x + np.random.default_rng(seed).normal(0.0, 0.05, x.shape)
    elem = np.zeros(len(x), np.uint8)
    return M.AtomsMD(M.ForceFieldEAM.from_knots(p), H, x, elem, skin=0.5), H, x, elem


def value_noalloc(M, g):
    e = C.c_double()
    st = M.lib().mapp_md_value_noalloc(g._h, C.byref(e))
    return st, e.value


def test_value_without_allocation_equals_value(M, cu_pot):
    big, H, x, elem = cu_system(M, cu_pot, (10, 10, 10), 41)
    small, Hs, xs, elems = cu_system(M, cu_pot, (6, 6, 6), 42)
    say(f"no-alloc value: {len(x)} then {len(xs)} atoms")
    f, e_force, _ = big.force_calc()
    rho = big.rho()
    e_val = big.value()
    eo = O.OMD(O.OEam.from_knots(cu_pot), H, x, elem, skin=0.5).energy()
    for g, n, e_ref, tag in ((big, len(x), e_val, "first call"), (small, len(xs), small.value(), "smaller system")):
        st, e = value_noalloc(M, g)
        assert st == 0, M.lib().mapp_last_error()
        say(f"  {tag}: value {e_ref:.10f} no-alloc {e:.10f} (difference {e - e_ref:.1e}, "
            f"{'bit-equal' if e == e_ref else 'not bit-equal'})")
        assert abs(e - e_ref) < E_ATOM * n + 1e-6
    # the first system again, after the smaller one used the rows
    st, e = value_noalloc(M, big)
    assert st == 0 and abs(e - e_val) < E_ATOM * len(x) + 1e-6
    assert abs(e - eo) < E_ATOM * len(x) + 1e-6
    assert abs(e_force - e_val) < E_ATOM * len(x) + 1e-6
    # like value(), it leaves the densities of the last force call readable
    np.testing.assert_array_equal(big.rho(), rho)


def test_value_without_allocation_rejects_larger_systems(M, cu_pot):
    g, H, x, elem = cu_system(M, cu_pot, (65, 64, 64), 43)
    assert len(x) > MAX_ATOMS
    say(f"no-alloc value: {len(x)} atoms, above the {MAX_ATOMS}-atom scratch rows")
    st, _ = value_noalloc(M, g)
    assert st == 4  # MAPP_ERR_CAPACITY
    assert b"at most 1048576 atoms" in M.lib().mapp_last_error()
    # the system is untouched and the allocating call still works
    e = g.value()
    assert np.isfinite(e) and e < 0.0
