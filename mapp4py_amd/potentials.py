"""Synthetic EAM potentials and lattices for the benchmark workloads.

The workloads are defined on analytic embedded-atom models written out as
DYNAMO tables, because no Cu, Fe-Ni or Cu-H table with the required grids
ships with the reference:

* ``johnson_cu``  Johnson-form EAM for fcc Cu (R. A. Johnson, Phys. Rev. B 37,
  3924 (1988) functional form), written as setfl (eam/alloy);
* ``johnson_feni_fs``  a two-element Johnson-form model for bcc Fe-Ni written
  in the Finnis-Sinclair layout (eam/fs): every (host, neighbor) pair has its
  own density function;
* ``johnson_cuh``  Cu plus an interstitial-H model for the DMD workload.

Every function is multiplied by a smooth cutoff psi(x) = x^4/(1+x^4),
x = (r_c - r)/h, which vanishes with its derivatives at r_c; knots at r >= r_c
are exactly zero, so the reference's rule r_c = (last non-zero knot) * dr
(src/import_eam.h:52-64) recovers the intended cutoff within one grid step.
The functions also expose the analytic values so tests can compare tabulated
forces against closed forms.
"""
from __future__ import annotations

import os

import numpy as np

__all__ = [
    "EamKnots", "johnson_cu", "johnson_feni_fs", "johnson_cuh", "write_setfl", "write_fs",
    "fcc", "bcc", "dmd_cuh_sites", "gaussian_knots", "gcmc_nih_cells", "NIALH_JEA", "NIALH_ELEMS",
    "MASS", "KB_EV", "HP", "FS",
]

MASS = {"Cu": 63.546, "Fe": 55.845, "Ni": 58.6934, "H": 1.00794, "Al": 26.9815}
KB_EV = 8.617330350e-5
# Planck constant in eV * (Angstrom * sqrt(amu/eV)) time units
HP = 4.13566766225 * 0.1 * np.sqrt(1.60217656535 / 1.66053904020)
# one femtosecond in the time unit Angstrom*sqrt(amu/eV)
FS = 1.0e-15 / (1.0e-10 * np.sqrt(1.66053904020e-27 / 1.60217656535e-19))


class EamKnots:
    """Knot arrays of a tabulated EAM model.

    F[ne][nrho]; rho[ne][ne][nr] with rho[a][b] the density function listed in
    element a's block at position b (the density at a b-atom due to an
    a-atom for fs; setfl has one function per element); r_phi[ne][ne][nr].
    """

    def __init__(self, names, masses, nr, dr, nrho, drho, F, rho, r_phi, fs, analytic=None):
        self.names = list(names)
        self.masses = np.asarray(masses, dtype=np.float64)
        self.nr, self.dr, self.nrho, self.drho = int(nr), float(dr), int(nrho), float(drho)
        self.F = np.ascontiguousarray(F, dtype=np.float64)
        self.rho = np.ascontiguousarray(rho, dtype=np.float64)
        self.r_phi = np.ascontiguousarray(r_phi, dtype=np.float64)
        self.fs = bool(fs)
        self.analytic = analytic

    @property
    def ne(self):
        return len(self.names)


def _psi(r, rc, h):
    x = np.clip((rc - r) / h, 0.0, None)
    x4 = x ** 4
    return np.where(r < rc, x4 / (1.0 + x4), 0.0)


def _dpsi(r, rc, h):
    x = np.clip((rc - r) / h, 0.0, None)
    x3 = x ** 3
    d = 4.0 * x3 / (1.0 + x ** 4) ** 2 * (-1.0 / h)
    return np.where(r < rc, d, 0.0)


def _johnson_F(rho, Ec, alpha, beta, gamma, phi_e, rho_e):
    """Johnson embedding energy, zero at rho = 0."""
    rho = np.asarray(rho, dtype=np.float64)
    out = np.zeros_like(rho)
    m = rho > 0.0
    q = rho[m] / rho_e
    ab, gb = alpha / beta, gamma / beta
    out[m] = -Ec * (1.0 - ab * np.log(q)) * q ** ab - 6.0 * phi_e * q ** gb
    return out


def _johnson_dF(rho, Ec, alpha, beta, gamma, phi_e, rho_e):
    rho = np.asarray(rho, dtype=np.float64)
    out = np.zeros_like(rho)
    m = rho > 0.0
    q = rho[m] / rho_e
    ab, gb = alpha / beta, gamma / beta
    out[m] = (Ec * ab * ab * np.log(q) * q ** ab - 6.0 * phi_e * gb * q ** gb) / rho[m]
    return out


class _Exp:
    """A exp(-k (r/re - 1)) psi(r); value and derivative."""

    def __init__(self, A, k, re, rc, h):
        self.A, self.k, self.re, self.rc, self.h = A, k, re, rc, h

    def __call__(self, r):
        r = np.asarray(r, dtype=np.float64)
        return self.A * np.exp(-self.k * (r / self.re - 1.0)) * _psi(r, self.rc, self.h)

    def d(self, r):
        r = np.asarray(r, dtype=np.float64)
        e = self.A * np.exp(-self.k * (r / self.re - 1.0))
        return e * (-self.k / self.re) * _psi(r, self.rc, self.h) + e * _dpsi(r, self.rc, self.h)


def _tabulate(names, masses, nr, dr, nrho, drho, Ffun, rhofun, phifun, fs, analytic):
    ne = len(names)
    r = np.arange(nr, dtype=np.float64) * dr
    rr = np.arange(nrho, dtype=np.float64) * drho
    F = np.stack([Ffun[a](rr) for a in range(ne)])
    rho = np.zeros((ne, ne, nr))
    r_phi = np.zeros((ne, ne, nr))
    for a in range(ne):
        for b in range(ne):
            rho[a, b] = rhofun[a][b](r) if fs else rhofun[a][0](r)
            r_phi[a, b] = r * phifun[min(a, b)][max(a, b)](r)
    return EamKnots(names, masses, nr, dr, nrho, drho, F, rho, r_phi, fs, analytic)


def johnson_cu(nr=10000, nrho=10000, rc=4.95, dr=None, rho_max=40.0):
    """Johnson-form fcc Cu (a = 3.615 A), cutoff rc, grid nr x nrho."""
    re = 3.615 / np.sqrt(2.0)
    Ec, alpha, beta, gamma, phi_e = 3.54, 5.09, 5.85, 8.00, 0.59
    h = 0.6
    dr = float(dr) if dr is not None else 0.0005 * 10000 / nr
    f = _Exp(1.0, beta, re, rc, h)
    phi = _Exp(phi_e, gamma, re, rc, h)
    rho_e = 12.0
    par = (Ec, alpha, beta, gamma, phi_e, rho_e)
    Ff = lambda x: _johnson_F(x, *par)
    dFf = lambda x: _johnson_dF(x, *par)
    drho = rho_max / (nrho - 1)
    an = {"F": [Ff], "dF": [dFf], "rho": [[f]], "phi": [[phi]], "rc": rc}
    return _tabulate(["Cu"], [MASS["Cu"]], nr, dr, nrho, drho, [Ff], [[f]], [[phi]], False, an)


def johnson_feni_fs(nr=5000, nrho=5000, rc=5.3, rho_max=40.0):
    """Two-element (Fe, Ni) Johnson-form model in the eam/fs layout for bcc a = 2.86 A."""
    re = 2.86 * np.sqrt(3.0) / 2.0
    h = 0.6
    dr = rc * 1.0 / (nr - 1) * 1.0  # last knot lands at rc (zero), last non-zero one below
    dr = float(np.ceil(dr * 1e7) / 1e7)
    names = ["Fe", "Ni"]
    # density of element a seen by element b: rho[a][b] (fs block a, entry b)
    fr = [[_Exp(1.00, 6.0, re, rc, h), _Exp(1.05, 6.1, re, rc, h)],
          [_Exp(0.95, 6.1, re, rc, h), _Exp(1.10, 6.2, re, rc, h)]]
    ph = [[_Exp(0.35, 8.5, re, rc, h), _Exp(0.37, 8.35, re, rc, h)],
          [None, _Exp(0.40, 8.2, re, rc, h)]]
    pars = [(4.28, 5.0, 6.0, 8.5, 0.35, 8.0), (4.44, 5.0, 6.2, 8.2, 0.40, 8.8)]
    Ff = [lambda x, p=p: _johnson_F(x, *p) for p in pars]
    dFf = [lambda x, p=p: _johnson_dF(x, *p) for p in pars]
    drho = rho_max / (nrho - 1)
    an = {"F": Ff, "dF": dFf, "rho": fr, "phi": ph, "rc": rc}
    return _tabulate(names, [MASS["Fe"], MASS["Ni"]], nr, dr, nrho, drho, Ff, fr, ph, True, an)


def johnson_cuh(nr=5000, nrho=5000, rc=5.5, rho_max=40.0):
    """Cu-H model (setfl layout) for the diffusive-MD workload."""
    re = 3.615 / np.sqrt(2.0)
    h = 0.6
    dr = float(np.ceil(rc / (nr - 1) * 1e7) / 1e7)
    f_cu = _Exp(1.0, 5.85, re, rc, h)
    f_h = _Exp(0.5, 4.0, 2.0, rc, h)
    ph = [[_Exp(0.59, 8.0, re, rc, h), _Exp(0.30, 5.0, 1.8, rc, h)],
          [None, _Exp(0.05, 3.0, 2.5, rc, h)]]
    Ff = [lambda x: _johnson_F(x, 3.54, 5.09, 5.85, 8.0, 0.59, 12.0),
          lambda x: -1.5 * np.sqrt(np.clip(x, 0.0, None) / 6.0) + 0.02 * x]
    dFf = [lambda x: _johnson_dF(x, 3.54, 5.09, 5.85, 8.0, 0.59, 12.0),
           lambda x: np.where(x > 0, -0.75 / np.sqrt(6.0 * np.clip(x, 1e-300, None)), 0.0) + 0.02]
    drho = rho_max / (nrho - 1)
    an = {"F": Ff, "dF": dFf, "rho": [[f_cu], [f_h]], "phi": ph, "rc": rc}
    return _tabulate(["Cu", "H"], [MASS["Cu"], MASS["H"]], nr, dr, nrho, drho, Ff, [[f_cu], [f_h]], ph, False, an)


def _fmt_block(vals, per_line=5):
    lines = []
    for k in range(0, len(vals), per_line):
        lines.append(" ".join(repr(float(v)) for v in vals[k:k + per_line]))
    return "\n".join(lines)


def write_setfl(path, pot: EamKnots, rc_header=None, lattice=(3.615, "FCC")):
    """DYNAMO setfl (eam/alloy): 3 comments, elements, grid, per-element F and rho, r*phi lower triangle."""
    ne = pot.ne
    rch = pot.nr * pot.dr if rc_header is None else rc_header
    with open(path, "w") as fh:
        fh.write("# synthetic Johnson-form EAM written by mapp4py_amd.potentials\n# setfl\n#\n")
        fh.write(f"{ne} " + " ".join(pot.names) + "\n")
        fh.write(f"{pot.nrho} {pot.drho!r} {pot.nr} {pot.dr!r} {rch!r}\n")
        for a in range(ne):
            fh.write(f"0 {pot.masses[a]!r} {lattice[0]} {lattice[1]}\n")
            fh.write(_fmt_block(pot.F[a]) + "\n")
            fh.write(_fmt_block(pot.rho[a, 0]) + "\n")
        for a in range(ne):
            for b in range(a + 1):
                fh.write(_fmt_block(pot.r_phi[a, b]) + "\n")


def write_funcfl(path, F, Z, rho, drho, dr, mass, Znum=29, lattice=(3.615, "FCC"), comment="synthetic funcfl"):
    """DYNAMO funcfl (one element): comment, 'Z mass a lattice', 'nrho drho nr dr rc', then F [nrho],
    Z [nr] and rho [nr]; r*phi = 27.2*0.529*Z(r)^2 (Hartree*Bohr to eV*A)."""
    F, Z, rho = np.asarray(F, float), np.asarray(Z, float), np.asarray(rho, float)
    nr = len(Z)
    with open(path, "w") as fh:
        fh.write(f"# {comment}\n")
        fh.write(f"{Znum} {mass!r} {lattice[0]} {lattice[1]}\n")
        fh.write(f"{len(F)} {drho!r} {nr} {dr!r} {(nr - 1) * dr!r}\n")
        fh.write(_fmt_block(F) + "\n")
        fh.write(_fmt_block(Z) + "\n")
        fh.write(_fmt_block(rho) + "\n")


def write_fs(path, pot: EamKnots, rc_header=None, lattice=(2.86, "BCC")):
    """DYNAMO eam/fs: like setfl but element a's block lists ne density functions."""
    ne = pot.ne
    rch = pot.nr * pot.dr if rc_header is None else rc_header
    with open(path, "w") as fh:
        fh.write("# synthetic Johnson-form EAM written by mapp4py_amd.potentials\n# eam/fs\n#\n")
        fh.write(f"{ne} " + " ".join(pot.names) + "\n")
        fh.write(f"{pot.nrho} {pot.drho!r} {pot.nr} {pot.dr!r} {rch!r}\n")
        for a in range(ne):
            fh.write(f"0 {pot.masses[a]!r} {lattice[0]} {lattice[1]}\n")
            fh.write(_fmt_block(pot.F[a]) + "\n")
            for b in range(ne):
                fh.write(_fmt_block(pot.rho[a, b]) + "\n")
        for a in range(ne):
            for b in range(a + 1):
                fh.write(_fmt_block(pot.r_phi[a, b]) + "\n")


def _lattice(basis, a, n):
    n0, n1, n2 = n
    g = np.stack(np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij"), -1).reshape(-1, 3)
    pts = (g[:, None, :] + np.asarray(basis)[None, :, :]).reshape(-1, 3) * a
    H = np.diag([n0 * a, n1 * a, n2 * a]).astype(np.float64)
    return pts.astype(np.float64), H


def dmd_cuh_sites(cells, seed=4, a=3.615, alpha_lo=0.05, alpha_hi=0.30, vacancy=0.01):
    """Cu-H diffusive-MD sites on an fcc lattice, two components per site
    (component 0 = Cu, 1 = H): alpha log-uniform in [alpha_lo, alpha_hi] per
    component, c_H ~ Beta(2, 5), c_Cu = max(1 - vacancy - c_H, 0)."""
    x, H = fcc(a, cells)
    rng = np.random.default_rng(seed)
    n = len(x)
    elem = np.tile(np.array([0, 1], np.uint8), (n, 1))
    alpha = np.exp(rng.uniform(np.log(alpha_lo), np.log(alpha_hi), (n, 2)))
    cH = rng.beta(2.0, 5.0, n)
    c = np.stack([np.maximum(1.0 - vacancy - cH, 0.0), cH], axis=1)
    return x, H, elem, alpha, c


def gaussian_knots(s=1.5, nr=20000, dr=0.0005, nrho=2000, drho=0.01, r_zero=9.0):
    """Single-element table whose density and r*phi are r*exp(-r^2/s^2) after
    the DMD r-scaling, so the 3-D Gaussian smearing has a closed form:
    g(r, a) = (s^2/(s^2+a^2))^(3/2) exp(-r^2/(s^2+a^2))."""
    r = np.arange(nr) * dr
    f = np.where(r < r_zero, np.exp(-r * r / (s * s)), 0.0)
    F = np.zeros((1, nrho))
    rho = f[None, None, :].copy()
    r_phi = (r * f)[None, None, :].copy()
    return EamKnots(["X"], [1.0], nr, dr, nrho, drho, F, rho, r_phi, False)


def fcc(a, n):
    """fcc positions [4*n0*n1*n2][3] and the box H (rows = box vectors)."""
    return _lattice([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], a, n)


def bcc(a, n):
    return _lattice([[0, 0, 0], [0.5, 0.5, 0.5]], a, n)


NIALH_JEA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "NiAlH_jea.eam.alloy")
NIALH_ELEMS = ("Ni", "Al", "H")


def _fcc_shapes(nmin=108, nmax=4000, kmin=3):
    """Every fcc block k0 <= k1 <= k2 (k >= kmin) with 4 k0 k1 k2 atoms in
    [nmin, nmax], as (natoms, spread, (k0, k1, k2)) sorted by size."""
    out = []
    k = kmin
    while 4 * k ** 3 <= nmax:
        for k1 in range(k, nmax):
            if 4 * k * k1 * k1 > nmax:
                break
            for k2 in range(k1, nmax):
                n = 4 * k * k1 * k2
                if n > nmax:
                    break
                if n >= nmin:
                    out.append((n, k2 - k, (k, k1, k2)))
        k += 1
    return sorted(out)


def gcmc_nih_cells(ncells, seed=6, nmin=108, nmax=4000, a=3.52, jitter=0.03, h_frac=0.02):
    """Batch of periodic fcc Ni cells for grand-canonical H exchange: sizes
    drawn log-uniform in [nmin, nmax] atoms (each rounded to the nearest fcc
    block of at least 3 cells a side, most cubic first), Ni positions
    jittered N(0, jitter), H on a random h_frac of the octahedral sites.
    Element codes follow NiAlH_jea.eam.alloy: Ni = 0, H = 2.
    Returns a list of (x [n][3], H [3][3], elem [n] uint8)."""
    rng = np.random.default_rng(seed)
    shapes = _fcc_shapes(nmin, nmax)
    sizes = np.array([s[0] for s in shapes], dtype=np.float64)
    cells = []
    for _ in range(ncells):
        target = np.exp(rng.uniform(np.log(nmin), np.log(nmax)))
        best = min(range(len(shapes)), key=lambda q: (abs(sizes[q] - target), shapes[q][1]))
        k = shapes[best][2]
        x, H = fcc(a, k)
        x = x + rng.normal(0.0, jitter, x.shape)
        octa, _ = _lattice([[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5], [0.5, 0.5, 0.5]], a, k)
        nh = int(round(h_frac * len(octa)))
        pick = rng.choice(len(octa), nh, replace=False) if nh else np.zeros(0, np.int64)
        xh = octa[np.sort(pick)] + rng.normal(0.0, jitter, (nh, 3))
        xa = np.concatenate([x, xh])
        elem = np.concatenate([np.zeros(len(x), np.uint8), np.full(nh, 2, np.uint8)])
        cells.append((xa, H, elem))
    return cells
