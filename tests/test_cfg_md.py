"""CFG files of MD systems: ImportCFGMD (src/import_cfg.cpp:32-604) and
ExportCFGMD (src/export_cfg_md.cpp:26-216) in the library, on the host.

Pinned by the reference's own configuration file (golden/Fe_300K.cfg, the
two-atom bcc Fe cell of examples/beginner-tutorial, which its input.py
replicates 64 x 64 x 64), by a box built independently in numpy (H0
Transform sqrt(1 + 2 eta), then the lower-triangular box of the same row
Gram matrix by Cholesky), and by write -> read round trips."""
import os

import numpy as np
import pytest

import mapp4py_amd as M

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def sqrtm_spd(A):
    w, V = np.linalg.eigh(A)
    return (V * np.sqrt(w)) @ V.T


def lower_box(H0, T, eta, A):
    Hx = H0 @ T @ sqrtm_spd(np.eye(3) + 2.0 * eta)
    return np.linalg.cholesky(Hx @ Hx.T) * A


def test_reference_fe_cell():
    d = M.import_cfg_md(os.path.join(GOLDEN, "Fe_300K.cfg"))
    a = 2.86016
    assert d["names"] == ["Fe"] and np.allclose(d["masses"], [55.845])
    assert d["v"] is None
    np.testing.assert_array_equal(d["elem"], [0, 0])
    np.testing.assert_allclose(d["H"], a * np.eye(3), rtol=0, atol=1e-15)
    np.testing.assert_allclose(d["x"], [[0, 0, 0], [a / 2] * 3], rtol=0, atol=1e-15)


def test_replicated_fe_cell_as_the_scalability_example():
    d = M.replicate(M.import_cfg_md(os.path.join(GOLDEN, "Fe_300K.cfg")), [64, 64, 64])
    a = 2.86016
    assert len(d["x"]) == 2 * 64 ** 3
    np.testing.assert_allclose(d["H"], 64 * a * np.eye(3), rtol=1e-15)
    # copy k = (k0 * 64 + k1) * 64 + k2 holds atoms 2k and 2k + 1
    k = 64 * 64 * 3 + 64 * 5 + 7
    np.testing.assert_allclose(d["x"][2 * k], a * np.array([3, 5, 7]), rtol=0, atol=1e-12)
    np.testing.assert_allclose(d["x"][2 * k + 1], a * np.array([3.5, 5.5, 7.5]), rtol=0, atol=1e-12)
    # a perfect bcc lattice: every fractional coordinate is a multiple of 1/128
    s = d["x"] @ np.linalg.inv(d["H"])
    assert np.abs(s * 128 - np.round(s * 128)).max() < 1e-9


def write_std(path, H0, T, eta, A, R, rows):
    with open(path, "w") as f:
        f.write(f"Number of particles = {len(rows)}\n# a comment line\nA = {A!r} Angstrom (basic length-scale)\n")
        for i in range(3):
            for j in range(3):
                f.write(f"H0({i + 1},{j + 1}) = {float(H0[i, j])!r} A\n")
                f.write(f"Transform({i + 1},{j + 1}) = {float(T[i, j])!r}\n")
                if eta[i, j] != 0.0:
                    f.write(f"eta({i + 1},{j + 1}) = {float(eta[i, j])!r}\n")
        f.write(f"R = {R!r} [ns^-1]\n")
        for r in rows:
            f.write(" ".join(repr(float(v)) if not isinstance(v, str) else v for v in r) + "\n")


def test_standard_file_with_velocities_and_a_strained_box(tmp_path):
    rng = np.random.default_rng(3)
    H0 = np.array([[3.1, 0.2, -0.1], [0.3, 2.9, 0.15], [-0.2, 0.1, 3.3]])
    T = np.array([[1.0, 0.05, 0.0], [0.0, 1.0, 0.02], [0.01, 0.0, 1.0]])
    eta = np.array([[0.01, 0.002, 0.0], [0.002, -0.005, 0.001], [0.0, 0.001, 0.02]])
    A, R = 1.7, 2.5
    n = 40
    s = rng.uniform(-0.5, 1.5, (n, 3))
    sd = rng.normal(0, 0.1, (n, 3))
    sym = np.where(rng.uniform(size=n) < 0.3, "Ni", "Fe")
    mass = {"Fe": 55.845, "Ni": 58.6934}
    rows = [[mass[e], e, *s[i], *sd[i]] for i, e in enumerate(sym)]
    rows[5][0] = 99.0  # a later mass of a known symbol is ignored (Elements::add_type)
    write_std(tmp_path / "std.cfg", H0, T, eta, A, R, rows)
    d = M.import_cfg_md(tmp_path / "std.cfg")
    H = lower_box(H0, T, eta, A)
    np.testing.assert_allclose(d["H"], H, rtol=0, atol=1e-13)
    assert d["H"][0, 1] == 0 and d["H"][0, 2] == 0 and d["H"][1, 2] == 0
    first = sym[0]
    assert d["names"][0] == first and set(d["names"]) == {"Fe", "Ni"}
    np.testing.assert_allclose(d["masses"], [mass[v] for v in d["names"]])
    np.testing.assert_array_equal(np.array(d["names"])[d["elem"]], sym)
    sw = s - np.floor(s)
    np.testing.assert_allclose(d["x"], sw @ H, rtol=0, atol=1e-12)
    np.testing.assert_allclose(d["v"], R * sd @ H, rtol=0, atol=1e-12)


def test_extended_file_with_velocity_columns_and_auxiliaries(tmp_path):
    lines = ["Number of particles = 3", "A = 1.0 Angstrom", "H0(1,1) = 4.0 A", "H0(2,2) = 5.0 A",
             "H0(3,3) = 6.0 A", "H0(2,1) = 1.0 A", "R = 2.0", "entry_count = 8",
             "auxiliary[0] = q [e]", "auxiliary[1] = w [1]", "12.0", "C", "0.1 0.2 0.3 1 0 0 7 8",
             "1.0", "H", "0.5 0.5 0.5 0 1 0 7 8", "12.0", "C", "1.25 -0.25 0.0 0 0 1 7 8"]
    (tmp_path / "e.cfg").write_text("\n".join(lines) + "\n")
    d = M.import_cfg_md(tmp_path / "e.cfg")
    H = np.array([[4.0, 0, 0], [1.0, 5.0, 0], [0, 0, 6.0]])
    np.testing.assert_allclose(d["H"], H, atol=1e-14)
    assert d["names"] == ["C", "H"]
    np.testing.assert_array_equal(d["elem"], [0, 1, 0])
    s = np.array([[0.1, 0.2, 0.3], [0.5, 0.5, 0.5], [0.25, 0.75, 0.0]])
    np.testing.assert_allclose(d["x"], s @ H, atol=1e-14)
    np.testing.assert_allclose(d["v"], 2.0 * np.eye(3) @ H, atol=1e-14)


@pytest.mark.parametrize("with_v", [False, True])
def test_write_then_read_round_trip(tmp_path, with_v):
    rng = np.random.default_rng(11)
    H = np.array([[10.0, 0, 0], [1.5, 9.0, 0], [-0.7, 0.4, 11.0]])
    n = 200
    s = rng.uniform(-0.2, 1.2, (n, 3))  # some outside the box: written wrapped
    x = s @ H
    v = rng.normal(0, 3.0, (n, 3)) if with_v else None
    elem = (rng.uniform(size=n) < 0.25).astype(np.uint8)
    aux = rng.normal(size=(n, 2))
    M.write_cfg_md(tmp_path / "w.cfg", H, x, elem, ["Fe", "Ni"], [55.845, 58.6934], v=v, aux=aux,
                   aux_names=["f_0", "f_1"])
    text = (tmp_path / "w.cfg").read_text()
    assert (".NO_VELOCITY." in text) != with_v and ("R = 1.0" in text) == with_v
    assert f"entry_count = {3 + (3 if with_v else 0) + 2}" in text
    d = M.import_cfg_md(tmp_path / "w.cfg")
    np.testing.assert_allclose(d["H"], H, rtol=0, atol=1e-14)
    # the reader numbers elements in order of first appearance
    assert d["names"] == (["Fe", "Ni"] if elem[0] == 0 else ["Ni", "Fe"])
    names = np.array(d["names"])[d["elem"]]
    np.testing.assert_array_equal(names, np.array(["Fe", "Ni"])[elem])
    np.testing.assert_allclose(d["x"], (s - np.floor(s)) @ H, rtol=0, atol=1e-13)
    if with_v:
        np.testing.assert_allclose(d["v"], v, rtol=0, atol=1e-13)
    else:
        assert d["v"] is None


@pytest.mark.parametrize("body,msg", [
    ("55.8 Fe 0 0 0 0 0\n", "invalid line"),
    ("-1.0 Fe 0 0 0 0 0 0\n", "mass"),
    ("55.8 Fe 0 0 0 0 0 0\n55.8 Fe 0.5 0.5 0.5 0 0 0\n55.8 Fe 0.1 0.1 0.1 0 0 0\n", "expected 2 atoms"),
])
def test_import_rejects_what_the_reference_rejects(tmp_path, body, msg):
    (tmp_path / "bad.cfg").write_text("Number of particles = 2\nA = 1.0\nH0(1,1) = 3.0\nH0(2,2) = 3.0\n"
                                      "H0(3,3) = 3.0\n" + body)
    with pytest.raises(Exception, match=msg):
        M.import_cfg_md(tmp_path / "bad.cfg")


def test_extended_body_needs_a_symbol_after_the_mass(tmp_path):
    (tmp_path / "bad.cfg").write_text("Number of particles = 1\nA = 1.0\nH0(1,1) = 3.0\nH0(2,2) = 3.0\n"
                                      "H0(3,3) = 3.0\n.NO_VELOCITY.\nentry_count = 3\n55.8\n0 0 0\n")
    with pytest.raises(Exception, match="chemical symbol"):
        M.import_cfg_md(tmp_path / "bad.cfg")
