"""funcfl (single-element DYNAMO) import, ff_eam_funcfl / ImportEAM::funcfl
(src/import_eam.h:68-344), in the oracle: the tables from a file written from
analytic Z(r), rho(r), F(rho) give the closed-form EAM energy and forces with
phi = 27.2 * 0.529 Z^2 / r, and on one grid the import is the identity on the
knots (minus the last r and rho knot, as the reference's grid rounding
gives)."""
import numpy as np
import pytest

import eam_ref
import oracle_lib as O
from mapp4py_amd import potentials as pot
from mapp4py_amd.potentials import _Exp

HB = 27.2 * 0.529  # Hartree * Bohr in eV * A


class _Pair:
    """phi(r) = HB Z(r)^2 / r with derivative, from Z with value and derivative."""

    def __init__(self, Z):
        self.Z = Z

    def __call__(self, r):
        return HB * self.Z(r) ** 2 / r

    def d(self, r):
        z = self.Z(r)
        return HB * (2.0 * z * self.Z.d(r) * r - z * z) / (r * r)


def funcs(rc=5.3, scale=1.0):
    Z = _Exp(0.3 * scale, 3.2, 2.55, rc, 0.4)
    rho = _Exp(1.1 * scale, 5.0, 2.55, rc, 0.4)
    F = lambda x: -2.1 * np.sqrt(x)  # noqa: E731
    dF = lambda x: -1.05 / np.sqrt(np.maximum(x, 1e-300))  # noqa: E731
    return Z, rho, F, dF


def write(path, Z, rho, F, nr, dr, nrho, drho, mass=63.546):
    r = np.arange(nr) * dr
    q = np.arange(nrho) * drho
    pot.write_funcfl(path, F(q), Z(r), rho(r), drho, dr, mass)


# The grids below are binary fractions. On a spacing such as 0.002 or 0.01,
# fl(k dr) / dr falls just below k at some knots (11 of 3000 at 0.002, 148 at
# 0.01); there the reference reads Z on interval k at an offset near 1
# (src/import_eam.h:291-305 passes k where rho and F pass l), so that knot of
# r phi holds about HB Z(r_{k+1})^2. The oracle and the library reproduce this
# as written; those knots are not compared with the closed form.


def test_funcfl_gives_the_closed_form_eam(tmp_path):
    Z, rho, F, dF = funcs()
    p = tmp_path / "a.funcfl"
    write(p, Z, rho, F, 2900, 2.0 ** -9, 3000, 2.0 ** -6)
    ff = O.OEam.funcfl([p])
    x, H = pot.fcc(3.615, (3, 3, 3))
    x = x + np.random.default_rng(2).normal(0, 0.08, x.shape)
    elem = np.zeros(len(x), np.uint8)
    m = O.OMD(ff, H, x, elem)
    f, E, _, _ = m.force()
    an = {"F": [F], "dF": [dF], "rho": [[rho]], "phi": [[_Pair(Z)]], "rc": 5.3}
    Er, fr = eam_ref.energy_forces(an, False, H, x, elem)[:2]
    print(f"funcfl vs closed form: E {E:.10f} / {Er:.10f}, max|df| {np.abs(f - fr).max():.2e}")
    assert abs(E - Er) < 1e-7 * abs(Er)
    assert np.abs(f - fr).max() < 1e-6


def test_one_grid_import_is_the_identity_on_the_knots(tmp_path):
    # binary spacings: r / dr lands on the knots exactly; the common grid
    # keeps n - 1 knots ((n - 1) dr / dr rounded)
    Z, rho, F, dF = funcs()
    nr, dr, nrho, drho = 1400, 2.0 ** -8, 1200, 2.0 ** -5
    p = tmp_path / "b.funcfl"
    write(p, Z, rho, F, nr, dr, nrho, drho)
    a = O.OEam.funcfl([p])
    assert a.info() == (1, nr - 1, nrho - 1, dr, drho)
    r = np.arange(nr - 1) * dr
    q = np.arange(nrho - 1) * drho
    zk = Z(r)  # written with repr: read back exactly
    b = O.OEam.from_knots(pot.EamKnots(["Cu"], [63.546], nr - 1, dr, nrho - 1, drho, F(q)[None, :],
                                       rho(r)[None, None, :], (HB * zk * zk)[None, None, :], False))
    x, H = pot.fcc(3.615, (3, 3, 3))
    x = x + np.random.default_rng(3).normal(0, 0.08, x.shape)
    elem = np.zeros(len(x), np.uint8)
    fa, Ea, _, _ = O.OMD(a, H, x, elem).force()
    fb, Eb, _, _ = O.OMD(b, H, x, elem).force()
    assert Ea == Eb
    np.testing.assert_array_equal(fa, fb)


def test_two_files_on_one_grid_mix_as_the_geometric_mean(tmp_path):
    # phi_ab = HB Z_a Z_b / r: the cross pair equals the geometric mean of the
    # two like pairs, and each element keeps its own density and F
    Za, rhoa, Fa, dFa = funcs()
    Zb, rhob, Fb, dFb = funcs(scale=1.3)
    pa, pb = tmp_path / "a.funcfl", tmp_path / "b.funcfl"
    write(pa, Za, rhoa, Fa, 2900, 2.0 ** -9, 3000, 2.0 ** -6)
    write(pb, Zb, rhob, Fb, 2900, 2.0 ** -9, 3000, 2.0 ** -6)
    ff = O.OEam.funcfl([pa, pb])
    x, H = pot.fcc(3.615, (3, 3, 3))
    x = x + np.random.default_rng(4).normal(0, 0.08, x.shape)
    elem = (np.random.default_rng(5).uniform(size=len(x)) < 0.4).astype(np.uint8)
    f, E, _, _ = O.OMD(ff, H, x, elem).force()

    class _Cross(_Pair):
        def __call__(self, r):
            return HB * Za(r) * Zb(r) / r

        def d(self, r):
            return HB * ((Za.d(r) * Zb(r) + Za(r) * Zb.d(r)) * r - Za(r) * Zb(r)) / (r * r)

    an = {"F": [Fa, Fb], "dF": [dFa, dFb], "rho": [[rhoa], [rhob]],
          "phi": [[_Pair(Za), _Cross(Za)], [_Cross(Za), _Pair(Zb)]], "rc": 5.3}
    Er, fr = eam_ref.energy_forces(an, False, H, x, elem)[:2]
    assert abs(E - Er) < 1e-7 * abs(Er)
    assert np.abs(f - fr).max() < 1e-6


def test_funcfl_rejects_a_file_shorter_than_the_common_grid(tmp_path):
    Z, rho, F, _ = funcs()
    pa, pb = tmp_path / "a.funcfl", tmp_path / "b.funcfl"
    write(pa, Z, rho, F, 2800, 0.002, 3000, 0.02)
    write(pb, Z, rho, F, 1000, 0.004, 3000, 0.02)  # coarser and shorter: the reference reads past it
    with pytest.raises(RuntimeError, match="fewer r knots"):
        O.OEam.funcfl([pa, pb])
