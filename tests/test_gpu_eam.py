"""GPU parity of the EAM path against the oracle, through the C ABI.

Bars: neighbor sets bit-exact; forces within 1e-5 eV/A of the FP64 oracle
(FP32 pair/force arithmetic). Energy and virial carry the rounding of FP32
pair spline values (about 1e-7 relative per pair) summed in FP64, so their
bars are per atom plus a floor for cells too small for the pair errors to
average out: |dE| < 5e-8 eV * N + 1e-6 eV and max |dF_H| < 2e-7 eV * N +
5e-6 eV. The
virial is not bounded relative to itself because near zero pressure its
components are small differences of large pair terms.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from mapp4py_amd import potentials as pot

pytestmark = pytest.mark.gpu

F_TOL = 1e-5       # eV/A, max abs force error vs FP64
E_ATOM = 5e-8      # eV per atom, energy
W_ATOM = 2e-7      # eV per atom, virial components


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


@pytest.fixture(scope="module")
def feni_pot():
    return pot.johnson_feni_fs(nr=5000, nrho=5000)


def jitter(x, s, seed):
    return x + np.random.default_rng(seed).normal(0.0, s, x.shape)


def triples(i, j, img):
    return {(int(a), int(b), int(c[0]), int(c[1]), int(c[2])) for a, b, c in zip(i, j, img)}


def both(M, p, H, x, elem, skin=0.5):
    g = M.AtomsMD(M.ForceFieldEAM.from_knots(p), H, x, elem, skin=skin)
    o = O.OMD(O.OEam.from_knots(p), H, x, elem, skin=skin)
    return g, o


def check_forces(g, o, tag):
    f, e, FH = g.force_calc()
    fo, eo, FHo, rho_o = o.force()
    n = len(f)
    df = np.abs(f - fo).max()
    de = abs(e - eo) / n
    dW = np.abs(FH - FHo).max() / n
    say(f"{tag}: max|df|={df:.2e} eV/A  |dE|/N={de:.2e} eV (dE/E={abs(e - eo) / abs(eo):.1e})  "
        f"max|dF_H|/N={dW:.2e} eV (|F_H|max/N={np.abs(FHo).max() / n:.2e})")
    assert df < F_TOL, (tag, df)
    assert abs(e - eo) < E_ATOM * n + 1e-6, (tag, de)
    assert np.abs(FH - FHo).max() < W_ATOM * n + 5e-6, (tag, dW)
    np.testing.assert_allclose(g.rho(), rho_o, rtol=1e-7)  # FP32 spline values, FP64 sum
    return f, e, FH


@pytest.mark.parametrize("cells,tilt,seed", [((4, 4, 4), (0, 0, 0), 1), ((3, 4, 5), (0.7, -0.4, 0.9), 2),
                                             ((1, 2, 2), (0.3, 0.1, -0.2), 3), ((6, 6, 6), (0, 0, 0), 4)])
def test_neighbor_sets_bit_exact(M, cu_pot, cells, tilt, seed):
    say(f"neighbor parity: fcc {cells} tilt {tilt}")
    x, H = pot.fcc(3.615, cells)
    H = H.copy()
    H[1, 0], H[2, 0], H[2, 1] = tilt
    x = jitter(x, 0.1, seed)
    elem = np.zeros(len(x), np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    gs = triples(*g.half_pairs())
    os_ = triples(*o.pairs())
    assert len(gs) == len(os_) and gs == os_
    np.testing.assert_array_equal(g.x, o.x)  # wrapped positions re-expressed identically


def test_neighbor_sets_on_the_cutoff_sphere(M, cu_pot):
    # choose the lattice so the 3rd fcc shell sits exactly on r_c + skin: the
    # FP64 predicate's rounding alone decides membership of those pairs
    rc = float(O.OEam.from_knots(cu_pot).cut()[0, 0]) + 0.5
    a = rc / np.sqrt(1.5)
    x, H = pot.fcc(a, (5, 5, 5))
    elem = np.zeros(len(x), np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    gs, os_ = triples(*g.half_pairs()), triples(*o.pairs())
    say(f"cutoff-sphere lattice a={a:.6f}: {len(gs)} pairs")
    assert gs == os_


def test_forces_energy_virial_cu_32000(M, cu_pot):
    # config 1 geometry: fcc Cu 20^3 cells, jitter N(0, 0.05 A)
    x, H = pot.fcc(3.615, (20, 20, 20))
    x = jitter(x, 0.05, 2024)
    elem = np.zeros(len(x), np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    f, e, FH = check_forces(g, o, "Cu 32000")
    assert abs(g.value() - e) < E_ATOM * len(f)
    np.testing.assert_allclose(f.sum(0), 0.0, atol=2e-3)


def test_forces_fs_two_elements_triclinic(M, feni_pot):
    x, H = pot.bcc(2.86, (8, 8, 8))
    H = H.copy()
    H[1, 0], H[2, 1] = 0.8, -0.5
    x = jitter(x, 0.06, 7)
    elem = (np.random.default_rng(3).random(len(x)) < 0.2).astype(np.uint8)
    g, o = both(M, feni_pot, H, x, elem, skin=0.4)
    assert triples(*g.half_pairs()) == triples(*o.pairs())
    check_forces(g, o, "FeNi fs 1024")


def test_small_box_images(M, cu_pot):
    x, H = pot.fcc(3.615, (1, 1, 1))
    x = jitter(x, 0.05, 5)
    elem = np.zeros(len(x), np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    assert triples(*g.half_pairs()) == triples(*o.pairs())
    check_forces(g, o, "Cu 4 atoms, box < 2 r_c")


def test_shipped_setfl_and_fs_tables(M):
    import os

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(root, "mapp4py_amd", "data", "NiAlH_jea.eam.alloy")
    ffg = M.ForceFieldEAM.setfl(path, ["Ni", "H"])
    ffo = O.OEam.read(path, 0)
    # element subset: file order Ni Al H -> Ni, H
    assert ffg.ne == 2
    x, H = pot.fcc(3.52, (4, 4, 4))
    x = jitter(x, 0.05, 8)
    elem = (np.random.default_rng(1).random(len(x)) < 0.1).astype(np.uint8)
    g = M.AtomsMD(ffg, H, x, elem)
    # oracle with the full three-element file: map H to index 2
    o = O.OMD(ffo, H, x, np.where(elem == 1, 2, 0).astype(np.uint8))
    check_forces(g, o, "NiAlH_jea (Ni, H subset)")


def test_update_decides_like_reference(M, cu_pot):
    x, H = pot.fcc(3.615, (4, 4, 4))
    x = jitter(x, 0.05, 9)
    elem = np.zeros(len(x), np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    xw = g.x
    # below skin/2 everywhere: no rebuild; one atom past skin/2: rebuild
    small = xw.copy()
    small[:, 0] += 0.2
    g.x = small
    o.x = small
    assert g.update() is False and o.update() is False
    check_forces(g, o, "after small move")
    big = small.copy()
    big[7, 1] += 0.3
    g.x = big
    o.x = big
    assert g.update() is True and o.update() is True
    assert triples(*g.half_pairs()) == triples(*o.pairs())
    check_forces(g, o, "after rebuild")


def test_nvt_matches_oracle_short_run(M, cu_pot):
    x, H = pot.fcc(3.615, (6, 6, 6))
    x = jitter(x, 0.05, 10)
    elem = np.zeros(len(x), np.uint8)
    masses = [pot.MASS["Cu"]]
    v = O.o_create_T(elem, masses, pot.KB_EV, 300.0, 3)
    g, o = both(M, cu_pot, H, x, elem)
    g.x_d = v
    nvt = M.NVT(300.0, 1.0 * pot.FS)
    out = nvt.run(g, 40, masses)
    vo, outo, Tpo, nreb_o = o.nvt_run(v, masses, pot.KB_EV, 300.0, 1.0 * pot.FS, 40)
    dx = np.abs(g.x - o.x).max()
    dv = np.abs(g.x_d - vo).max()
    say(f"NVT 40 steps: max|dx|={dx:.2e} A  max|dv|={dv:.2e}  T {nvt.T_part:.6f} vs {Tpo:.6f}  "
        f"rebuilds {nvt.rebuilds} vs {nreb_o}")
    assert nvt.rebuilds == nreb_o
    assert dx < 1e-6 and dv < 1e-6
    assert abs(nvt.T_part - Tpo) < 1e-6 * Tpo
    assert abs(out[0] - outo[0]) < E_ATOM * len(x)


def test_box_change_gives_the_oracle_of_the_new_box(M, cu_pot):
    # what md_nst or a strain step does to atoms->H (src/atoms.cpp:187-194):
    # the handle takes the new box and positions, and its forces, energy and
    # virial are those of a fresh oracle in that box; then a box change that
    # keeps the Cartesian positions; then one the grid cannot hold, which
    # leaves the system as it was
    x, H = pot.fcc(3.615, (5, 5, 5))
    x = jitter(x, 0.05, 41)
    elem = np.zeros(len(x), np.uint8)
    g = M.AtomsMD(M.ForceFieldEAM.from_knots(cu_pot), H, x, elem)
    check_forces(g, O.OMD(O.OEam.from_knots(cu_pot), H, x, elem), "before the box change")
    strain = np.eye(3) + np.array([[0.02, 0, 0], [0.05, -0.015, 0], [-0.04, 0.03, 0.01]])
    H1, x1 = H @ strain, x @ strain
    g.set_box(H1, x1)
    check_forces(g, O.OMD(O.OEam.from_knots(cu_pot), H1, x1, elem), "strained triclinic box, new positions")
    H2 = H1.copy()
    H2[2, 2] *= 1.01
    g.set_box(H2)
    check_forces(g, O.OMD(O.OEam.from_knots(cu_pot), H2, g.x, elem), "box change keeping the positions")
    with pytest.raises(M.MappError):
        g.set_box(np.diag([2.0, 40.0, 40.0]))
    check_forces(g, O.OMD(O.OEam.from_knots(cu_pot), H2, g.x, elem), "after a rejected box")


def test_feni_nvt_matches_oracle(M, feni_pot):
    # the two-element eam/fs path (global-table passes, the F' array, the
    # fused second half-kick) through 250 NVT steps at 600 K in a sheared box,
    # with on-device rebuilds: MDNVT::run (src/md_nvt.cpp:383-480) over
    # ForceFieldEAM::__force_calc (src/ff_eam.cpp:99-239) with the FS density
    # convention rho[e_j][e_i] (src/import_eam.h:836-851)
    x, H = pot.bcc(2.86, (8, 8, 8))
    shear = np.eye(3)
    shear[1, 0], shear[2, 0], shear[2, 1] = 0.04, -0.03, 0.05
    x, H = jitter(x @ shear, 0.05, 31), H @ shear
    elem = (np.random.default_rng(32).random(len(x)) < 0.2).astype(np.uint8)
    masses = [pot.MASS["Fe"], pot.MASS["Ni"]]
    v = O.o_create_T(elem, masses, pot.KB_EV, 600.0, 5)
    g, o = both(M, feni_pot, H, x, elem)
    g.x_d = v
    nvt = M.NVT(600.0, 1.0 * pot.FS)
    out = nvt.run(g, 250, masses)
    vo, outo, Tpo, nreb_o = o.nvt_run(v, masses, pot.KB_EV, 600.0, 1.0 * pot.FS, 250)
    dx = np.abs(g.x - o.x).max()
    dv = np.abs(g.x_d - vo).max()
    say(f"Fe-Ni NVT 250 steps, {len(x)} atoms ({int(elem.sum())} Ni): max|dx|={dx:.2e} A  max|dv|={dv:.2e}  "
        f"T {nvt.T_part:.6f} vs {Tpo:.6f}  rebuilds {nvt.rebuilds} vs {nreb_o}  dE {out[0] - outo[0]:.2e} eV")
    assert nvt.rebuilds == nreb_o >= 2
    assert dx < 1e-6 and dv < 1e-6
    assert abs(nvt.T_part - Tpo) < 1e-6 * Tpo
    assert abs(out[0] - outo[0]) < E_ATOM * len(x) + 1e-6
    # forces at the GPU's final positions against a fresh oracle
    check_forces(g, O.OMD(O.OEam.from_knots(feni_pot), H, g.x, elem), "Fe-Ni after 250 NVT steps")


def test_nvt_long_run_forces_stay_exact(M, cu_pot):
    # after many steps and rebuilds the GPU list must still cover every pair
    # within r_c: compare with a fresh oracle at the GPU's positions
    x, H = pot.fcc(3.615, (8, 8, 8))
    x = jitter(x, 0.05, 11)
    elem = np.zeros(len(x), np.uint8)
    masses = [pot.MASS["Cu"]]
    g = M.AtomsMD(M.ForceFieldEAM.from_knots(cu_pot), H, x, elem)
    g.x_d = O.o_create_T(elem, masses, pot.KB_EV, 1200.0, 4)
    nvt = M.NVT(1200.0, 2.0 * pot.FS)
    nvt.run(g, 300, masses)
    say(f"NVT 300 steps at 1200 K: {nvt.rebuilds} rebuilds")
    assert nvt.rebuilds >= 2
    o = O.OMD(O.OEam.from_knots(cu_pot), H, g.x, elem)
    check_forces(g, o, "after 300 NVT steps")


def test_every_batch_shape_gives_the_oracle_forces(M, cu_pot):
    x, H = pot.fcc(3.615, (10, 10, 10))
    x = jitter(x, 0.05, 21)
    elem = np.zeros(len(x), np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    assert g.set_lds(True), "an MI355X holds the one-element spline rows in LDS"
    for lds in (True, False):
        assert g.set_lds(lds) == lds
        for rb in (1, 2, 4, 8):
            for fb in (1, 2, 4, 8):
                g.set_batch(rb, fb)
                check_forces(g, o, f"lds={lds} batch rho={rb} force={fb}")


def test_spline_rows_below_the_lds_window(M, cu_pot):
    # The LDS force pass holds the rows of the largest r that fit (about
    # 5,700 of the 9,999 intervals of this table, r > ~2.15 A); a compressed
    # lattice (nearest neighbors near 2.05 A) sends the first-shell pairs to
    # the global-table rows below the window. Both modes give the oracle's
    # forces, and the same forces to FP32 rounding. At this density the pair
    # terms are large and cancel, so energy and virial are held to FP32
    # rounding relative to the largest pair term rather than per atom.
    x, H = pot.fcc(2.9, (6, 6, 6))
    x = jitter(x, 0.05, 22)
    elem = np.zeros(len(x), np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    rho_lo, f_lo = g.lds_window()
    say(f"LDS windows: density rows from {rho_lo}, force rows from {f_lo} (r > {f_lo * cu_pot.dr:.3f} A)")
    assert f_lo * cu_pot.dr > 2.9 / np.sqrt(2.0)  # the first shell (2.05 A) starts below the window
    fo, eo, FHo, rho_o = o.force()
    out = {}
    for lds in (True, False):
        assert g.set_lds(lds) == lds
        f, e, FH = g.force_calc()
        df = np.abs(f - fo).max()
        say(f"compressed fcc a=2.9, lds={lds}: max|df|={df:.2e} eV/A  dE/E={abs(e - eo) / abs(eo):.1e}  "
            f"max|dF_H|/|F_H|max={np.abs(FH - FHo).max() / np.abs(FHo).max():.1e}")
        assert df < F_TOL
        assert abs(e - eo) < 1e-6 * abs(eo)
        assert np.abs(FH - FHo).max() < 1e-6 * np.abs(FHo).max()
        np.testing.assert_allclose(g.rho(), rho_o, rtol=2e-7)
        out[lds] = f
    assert np.abs(out[True] - out[False]).max() < 1e-5


@pytest.mark.parametrize("natoms", [1, 7])
def test_isolated_atoms_have_empty_rows(M, cu_pot, natoms):
    # atoms further apart than r_c + skin in a large periodic box: every row
    # is empty, forces vanish and the energy is n F(0), in both pass modes
    H = np.diag([40.0, 41.0, 42.0])
    x = np.array([[2.0 + 9.5 * k, 3.0 + 7.0 * (k % 2), 5.0 + 8.0 * (k % 3)] for k in range(natoms)])
    elem = np.zeros(natoms, np.uint8)
    g, o = both(M, cu_pot, H, x, elem)
    assert g.list_stats()[1] == 0
    fo, eo, FHo, rho_o = o.force()
    for lds in (True, False):
        g.set_lds(lds)
        f, e, FH = g.force_calc()
        say(f"{natoms} isolated atoms, lds={lds}: E {e:.9f} vs {eo:.9f}, max|f| {np.abs(f).max():.1e}")
        assert np.abs(f).max() == 0.0 and np.abs(FH).max() == 0.0
        assert abs(e - eo) < 1e-9 * max(1.0, abs(eo))
        np.testing.assert_array_equal(g.rho(), np.zeros(natoms))
