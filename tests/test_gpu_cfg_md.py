"""md.export_cfg of a device handle (mapp_md_export_cfg) and md.import_cfg of
the reference's Fe cell replicated as examples/scalability-test does: the
file written from the GPU state reads back to the handle's positions,
velocities and elements, and an EAM force call on the imported, replicated
and diluted cell matches the oracle (forces 1e-5 eV/A)."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import mapp4py_amd

    return mapp4py_amd


def test_export_of_a_device_state_reads_back(M, tmp_path):
    p = pot.johnson_feni_fs(nr=3000, nrho=3000)
    d = M.replicate(M.import_cfg_md(os.path.join(GOLDEN, "Fe_300K.cfg")), [6, 6, 6])
    rng = np.random.default_rng(5)
    n = len(d["x"])
    elem = (rng.uniform(size=n) < 0.2).astype(np.uint8)  # 20 % Ni on the bcc Fe lattice
    x = d["x"] + rng.normal(0, 0.05, (n, 3))
    say(f"export_cfg: {n} atoms from the replicated Fe cell, 20% Ni")
    g = M.AtomsMD(M.ForceFieldEAM.from_knots(p), d["H"], x, elem)
    v = rng.normal(0, 0.02, (n, 3))
    g.x_d = v
    f, e, _ = g.force_calc()
    fo, eo, _, _ = O.OMD(O.OEam.from_knots(p), d["H"], x, elem).force()
    df = float(np.abs(f - fo).max())
    say(f"imported cell: max|df| {df:.2e} eV/A, dE/N {abs(e - eo) / n:.2e} eV")
    assert df < 1e-5 and abs(e - eo) < 5e-8 * n + 1e-6
    path = tmp_path / "state.cfg"
    g.export_cfg(path, ["Fe", "Ni"], [55.845, 58.6934], with_vel=True)
    r = M.import_cfg_md(path)
    names = np.array(r["names"])[r["elem"]]
    np.testing.assert_array_equal(names, np.array(["Fe", "Ni"])[elem])
    np.testing.assert_allclose(r["H"], d["H"], rtol=0, atol=1e-13)
    s = g.x @ np.linalg.inv(d["H"])
    np.testing.assert_allclose(r["x"], (s - np.floor(s)) @ d["H"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(r["v"], v, rtol=0, atol=1e-13)
    g.export_cfg(path, ["Fe", "Ni"], [55.845, 58.6934])
    assert M.import_cfg_md(path)["v"] is None
