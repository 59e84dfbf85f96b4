"""mapp4py_amd -- MI355X (gfx950) implementation of MAPP's EAM / DMD hot path.

The Python layer is a thin mirror of the reference's objects over the C ABI
(include/mapp_amd.h); all computation runs in the HIP library
``libmapp_amd.so``:

* :class:`ForceFieldEAM`  -- sim.ff_eam_setfl / ff_eam_fs (src/ff_eam.cpp:594-801)
* :class:`AtomsMD`        -- md.atoms with its ForceField, NeighborMD and DynamicMD
  (force_calc / value / update of src/ff_md.cpp:125-179, src/dynamic_md.h:38-68)
* :class:`NVT`            -- md.nvt(T, dt).run(sim, nsteps) (src/md_nvt.cpp:383-480)
* :class:`MinCG`, :class:`MinLBFGS` with :class:`LineSearchBackTrack`, :class:`LineSearchGoldenSection`,
  :class:`LineSearchBrent` -- md.min_cg, md.min_lbfgs, ls_bt, ls_golden, ls_brent (src/min_cg.h,
  src/min_l-bfgs.h, src/ls.h)
* :class:`AtomsDMD`       -- dmd.atoms with ForceFieldEAMDMD (src/ff_eam_dmd.cpp, src/ff_dmd.cpp)
* :class:`GCMCBatch`      -- many independent cells under SGCMC exchange (src/sgcmc.cpp,
  ForceFieldEAM's exchange hooks at src/ff_eam.cpp:323-578)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import MappError, check, f64, fp, i8p, i32p, i64p, ip, lib, ptr, u8p

__all__ = ["ForceFieldEAM", "AtomsMD", "AtomsDMD", "NVT", "NST", "MuVT", "MinCG", "MinLBFGS", "LineSearchBackTrack",
           "LineSearchGoldenSection", "LineSearchBrent", "GCMCBatch", "ReplicaBatch", "Surrogate", "MappError", "create_temp",
           "gauss_hermite", "lib"]


class ForceFieldEAM:
    """Tabulated EAM potential resident on the GPU."""

    def __init__(self, handle, ne):
        self._h = handle
        self.ne = ne

    @classmethod
    def setfl(cls, path, elems=None, dmd=False):
        return cls._from_file(lib().mapp_eam_create_setfl, path, elems, dmd)

    @classmethod
    def fs(cls, path, elems=None, dmd=False):
        return cls._from_file(lib().mapp_eam_create_fs, path, elems, dmd)

    @classmethod
    def _from_file(cls, fn, path, elems, dmd):
        h = C.c_void_p()
        if elems:
            arr = (C.c_char_p * len(elems))(*[e.encode() for e in elems])
            check(fn(str(path).encode(), arr, len(elems), int(dmd), C.byref(h)))
        else:
            check(fn(str(path).encode(), None, 0, int(dmd), C.byref(h)))
        ne = C.c_int()
        check(lib().mapp_eam_cutoff(h, C.byref(ne), None))
        return cls(h, ne.value)

    @classmethod
    def funcfl(cls, paths, dmd=False):
        """ff_eam_funcfl: one DYNAMO funcfl file per element, in element order."""
        paths = [paths] if isinstance(paths, (str, bytes, os.PathLike)) else list(paths)
        arr = (C.c_char_p * len(paths))(*[os.fsencode(p) for p in paths])
        h = C.c_void_p()
        check(lib().mapp_eam_create_funcfl(arr, len(paths), int(dmd), C.byref(h)))
        return cls(h, len(paths))

    @classmethod
    def from_knots(cls, pot, dmd=False):
        """pot: mapp4py_amd.potentials.EamKnots."""
        h = C.c_void_p()
        F, rho, rphi = f64(pot.F), f64(pot.rho), f64(pot.r_phi)
        check(lib().mapp_eam_create_knots(pot.ne, pot.nr, pot.nrho, pot.dr, pot.drho, ptr(F), ptr(rho),
                                          ptr(rphi), int(pot.fs), int(dmd), C.byref(h)))
        return cls(h, pot.ne)

    def cutoff(self):
        rc = np.zeros((self.ne, self.ne))
        check(lib().mapp_eam_cutoff(self._h, None, ptr(rc)))
        return rc

    def coefs(self, which, a, b=0, nrows=None, ncols=4):
        out = np.zeros((nrows, ncols))
        check(lib().mapp_eam_coefs(self._h, which, a, b, ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mapp_eam_destroy(self._h)
            self._h = None


class AtomsMD:
    """Atoms, neighbor list and EAM force field on the GPU (caller atom order)."""

    def __init__(self, ff: ForceFieldEAM, H, x, elem, skin=0.5):
        self.ff = ff
        x = f64(x).reshape(-1, 3)
        self.natms = x.shape[0]
        self.H = f64(H).reshape(3, 3)
        elem = np.ascontiguousarray(elem, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().mapp_md_create(ff._h, self.natms, ptr(self.H), ptr(x), ptr(elem, u8p), float(skin), C.byref(h)))
        self._h = h
        self.elem = elem
        self.pe = None
        self.F_H = None

    @property
    def vol(self):
        return float(self.H[0, 0] * self.H[1, 1] * self.H[2, 2])

    def _refresh_atoms(self):
        """the atom count and elements after a run that exchanges atoms"""
        n = C.c_int()
        check(lib().mapp_md_natms(self._h, C.byref(n)))
        self.natms = n.value
        el = np.zeros(self.natms, np.uint8)
        check(lib().mapp_md_get_elem(self._h, el.ctypes.data_as(C.c_void_p)))
        self.elem = el

    @property
    def x(self):
        out = np.zeros((self.natms, 3))
        check(lib().mapp_md_get_x(self._h, ptr(out)))
        return out

    @x.setter
    def x(self, val):
        val = f64(val).reshape(self.natms, 3)
        check(lib().mapp_md_set_x(self._h, ptr(val)))

    @property
    def x_d(self):
        out = np.zeros((self.natms, 3))
        check(lib().mapp_md_get_v(self._h, ptr(out)))
        return out

    @x_d.setter
    def x_d(self, val):
        val = f64(val).reshape(self.natms, 3)
        check(lib().mapp_md_set_v(self._h, ptr(val)))

    def update(self):
        r = C.c_int()
        check(lib().mapp_md_update(self._h, C.byref(r)))
        return bool(r.value)

    def set_box(self, H, x=None):
        """a new box H (lower triangular, rows = box vectors) and optionally
        new positions; the grid and the neighbor list are rebuilt"""
        Hn = f64(H).reshape(3, 3)
        xv = None if x is None else f64(x).reshape(self.natms, 3)
        check(lib().mapp_md_set_box(self._h, ptr(Hn), ptr(xv) if xv is not None else None))
        self.H = Hn

    def export_cfg(self, path, names, masses, with_vel=False):
        """md.export_cfg: the current positions (and velocities) as a CFG file,
        atoms in caller order; names and masses of the force field's elements."""
        check(lib().mapp_md_export_cfg(self._h, str(path).encode(), len(names), ptr(f64(masses)), _names(names),
                                       int(bool(with_vel))))

    def force_calc(self, want_f=True):
        """Returns (f [N][3], energy, F_H Voigt[6])."""
        f = np.zeros((self.natms, 3)) if want_f else None
        out = np.zeros(7)
        check(lib().mapp_md_force_calc(self._h, ptr(f) if want_f else None, ptr(out)))
        self.pe = out[0]
        self.F_H = out[1:].copy()
        return f, out[0], out[1:].copy()

    def value(self):
        e = C.c_double()
        check(lib().mapp_md_value(self._h, C.byref(e)))
        return e.value

    def rho(self):
        out = np.zeros(self.natms)
        check(lib().mapp_md_get_rho(self._h, ptr(out)))
        return out

    def half_pairs(self):
        n = C.c_int64()
        check(lib().mapp_md_pairs_count(self._h, C.byref(n)))
        i = np.zeros(n.value, np.int32)
        j = np.zeros(n.value, np.int32)
        img = np.zeros((n.value, 3), np.int8)
        check(lib().mapp_md_pairs(self._h, ptr(i, i32p), ptr(j, i32p), ptr(img, i8p)))
        return i, j, img

    def list_stats(self):
        m, s, r = C.c_int(), C.c_int64(), C.c_int()
        check(lib().mapp_md_list_stats(self._h, C.byref(m), C.byref(s), C.byref(r)))
        return m.value, s.value, r.value

    def set_batch(self, rho, force):
        check(lib().mapp_md_set_batch(self._h, int(rho), int(force)))

    def lds_window(self):
        """First spline interval held in LDS by the density and force passes (-1: none)."""
        a, b = C.c_int(), C.c_int()
        check(lib().mapp_md_lds_window(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_lds(self, on):
        """Spline rows of a one-element table in LDS (True) or read from the
        global tables (False); returns the mode in effect."""
        used = C.c_int(0)
        check(lib().mapp_md_set_lds(self._h, int(bool(on)), C.byref(used)))
        return bool(used.value)

    def profile(self, enable=True):
        check(lib().mapp_md_profile(self._h, int(enable)))

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_int64()
        check(lib().mapp_md_profile_read(self._h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mapp_md_destroy(self._h)
            self._h = None


class NVT:
    """md.nvt(T, dt): Nose-Hoover-chain NVT run on the GPU."""

    def __init__(self, T, dt, kB=8.617330350e-5):
        self.T, self.dt, self.kB = float(T), float(dt), float(kB)
        self.T_part = None
        self.rebuilds = 0

    def run(self, atoms: AtomsMD, nsteps, masses):
        masses = f64(masses)
        out = np.zeros(7)
        Tp = C.c_double()
        nr = C.c_int()
        check(lib().mapp_md_nvt_run(atoms._h, ptr(masses), self.kB, self.T, self.dt, int(nsteps), ptr(out),
                                    C.byref(Tp), C.byref(nr)))
        self.T_part = Tp.value
        self.rebuilds = nr.value
        atoms.pe = out[0]
        atoms.F_H = out[1:].copy()
        return out


class NST:
    """md.nst(S, T, dt): NVT with a box driven toward the target stress S
    (3x3, NaN for free components; src/md_nst.cpp). tau: box relaxation times
    (3x3, default 1000 dt); nreset: steps between re-takes of the reference
    box (0: never)."""

    def __init__(self, S, T, dt, kB=8.617330350e-5, tau=None, nreset=0):
        self.S = np.array(S, dtype=np.float64).reshape(3, 3)
        self.T, self.dt, self.kB = float(T), float(dt), float(kB)
        self.tau = None if tau is None else np.array(tau, dtype=np.float64).reshape(3, 3)
        self.nreset = int(nreset)
        self.T_part = None
        self.rebuilds = 0
        self.hist = None

    def run(self, atoms: AtomsMD, nsteps, masses):
        masses = f64(masses)
        out = np.zeros(7)
        H = np.zeros((3, 3))
        hist = np.zeros((max(int(nsteps), 1), 14))
        Tp = C.c_double()
        nr = C.c_int()
        tau = f64(self.tau) if self.tau is not None else None
        check(lib().mapp_md_nst_run(atoms._h, ptr(masses), self.kB, self.T, self.dt, ptr(f64(self.S)),
                                    ptr(tau) if tau is not None else None, self.nreset, int(nsteps), ptr(out),
                                    C.byref(Tp), ptr(H), ptr(hist), C.byref(nr)))
        self.T_part = Tp.value
        self.rebuilds = nr.value
        self.hist = hist[:int(nsteps)]
        atoms.H = H
        atoms.pe = out[0]
        atoms.F_H = out[1:].copy()
        return out


class MuVT:
    """md.muvt(mu, T, dt, gas_elem, seed): NVT whose gas element is exchanged
    with a reservoir at chemical potential mu every nevery steps (nattempts
    insertion/deletion attempts each; src/md_muvt.cpp). gas is the element's
    index in the force field; hP the Planck constant in the table's units
    (eV, A, amu: 4.13566766225e-1 * sqrt(1.60217656535 / 1.66053904020))."""

    HP = 4.13566766225 * 0.1 * (1.60217656535 / 1.66053904020) ** 0.5

    def __init__(self, mu, T, dt, gas, seed, nevery=1000, nattempts=1000, kB=8.617330350e-5, hP=None):
        self.mu, self.T, self.dt = float(mu), float(T), float(dt)
        self.gas, self.seed = int(gas), int(seed)
        self.nevery, self.nattempts = int(nevery), int(nattempts)
        self.kB = float(kB)
        self.hP = float(hP) if hP is not None else self.HP
        self.T_part = None
        self.counts = None
        self.hist = None
        self.builds = 0

    def run(self, atoms: AtomsMD, nsteps, masses):
        masses = f64(masses)
        out = np.zeros(7)
        Tp = C.c_double()
        counts = np.zeros(4, np.int32)
        hist = np.zeros((max(int(nsteps), 1), 2))
        nb = C.c_int()
        check(lib().mapp_md_muvt_run(atoms._h, ptr(masses), self.kB, self.hP, self.T, self.dt, self.mu, self.gas,
                                     self.seed, self.nevery, self.nattempts, int(nsteps), ptr(out), C.byref(Tp),
                                     counts.ctypes.data_as(C.POINTER(C.c_int)), ptr(hist), C.byref(nb)))
        self.T_part = Tp.value
        self.counts = counts
        self.hist = hist[:int(nsteps)]
        self.builds = nb.value
        atoms._refresh_atoms()
        atoms.pe = out[0]
        atoms.F_H = out[1:].copy()
        return out


class LineSearchBackTrack:
    """md.ls_bt: back-tracking line search (src/ls.cpp:456-461 defaults,
    LineSearchBackTrack::min at src/ls.h:711-769)."""

    def __init__(self, c=0.4, rho=0.5, min_alpha=0.0):
        if not 0.0 < c < 1.0 or not 0.0 < rho < 1.0 or min_alpha < 0.0:
            raise ValueError("ls_bt needs 0 < c < 1, 0 < rho < 1 and min_alpha >= 0")
        self.c, self.rho, self.min_alpha = float(c), float(rho), float(min_alpha)


class LineSearchGoldenSection:
    """mapp4py.ls_golden: golden-section line search over the bracketing
    routine (src/ls.cpp:95-101 defaults, src/ls.h:213-335,341-520)."""

    _kind = 1

    def __init__(self, tol=(2.0 * 2.220446049250313e-16) ** 0.5, max_iter=5, brack=True):
        if not tol >= 0.0 or int(max_iter) < 0:
            raise ValueError("line search needs tol >= 0 and max_iter >= 0")
        self.tol, self.max_iter, self.brack = float(tol), int(max_iter), bool(brack)
        self.zeps = 2.220446049250313e-16 * 1.0e-3


class LineSearchBrent(LineSearchGoldenSection):
    """mapp4py.ls_brent: Brent's line search over the bracketing routine
    (src/ls.cpp:262-268 defaults, src/ls.h:525-706)."""

    _kind = 2

    def __init__(self, tol=(2.0 * 2.220446049250313e-16) ** 0.5, max_iter=5, zeps=2.220446049250313e-16 * 1.0e-3,
                 brack=True):
        super().__init__(tol, max_iter, brack)
        if not zeps >= 0.0:
            raise ValueError("zeps must be >= 0")
        self.zeps = float(zeps)


_MIN_MSGS = {1: "line search failed: not downhill direction",
             2: "line search failed: gradient is zero",
             3: "line search failed: minimum alpha reached",
             4: "minimization finished: energy tolerance reached",
             5: "minimization finished: maximum iteration reached",
             7: "bracketing failed: maximum alpha reached",
             8: "bracketing failed: not downhill direction"}


class MinCG:
    """md.min_cg(e_tol, H_dof, affine, max_dx, ls).run(atoms, max_nsteps)
    (src/min_cg.cpp:180, MinCG::__run at src/min_cg.h:453-533) on the GPU.

    Only the positions are degrees of freedom: H_dof must be all False and
    affine False (box relaxation is not built)."""

    _lbfgs = 0

    def __init__(self, e_tol=1.0e-8, H_dof=None, affine=False, max_dx=1.0, ls=None):
        if H_dof is not None and np.any(np.asarray([v for row in H_dof for v in row], dtype=bool)):
            raise NotImplementedError("box degrees of freedom (H_dof) are not built")
        if affine:
            raise NotImplementedError("affine displacements are not built")
        if ls is None:
            ls = LineSearchBackTrack()
        if not isinstance(ls, (LineSearchBackTrack, LineSearchGoldenSection)):
            raise TypeError("ls must be LineSearchBackTrack, LineSearchGoldenSection or LineSearchBrent")
        if not e_tol >= 0.0 or not max_dx > 0.0:
            raise ValueError("e_tol must be >= 0 and max_dx > 0")
        self.e_tol, self.max_dx, self.ls, self.m = float(e_tol), float(max_dx), ls, 0
        self.ntally = 0
        self.report = None
        self.tally = None

    def run(self, atoms: "AtomsMD", max_nsteps):
        from ._lib import MinParams, MinReport
        ls = self.ls
        if isinstance(ls, LineSearchBackTrack):
            p = MinParams(self._lbfgs, int(self.m), self.e_tol, self.max_dx, int(max_nsteps), ls.c, ls.rho,
                          ls.min_alpha, 0, 5, 1, (2.0 * 2.220446049250313e-16) ** 0.5, 2.220446049250313e-19)
        else:
            p = MinParams(self._lbfgs, int(self.m), self.e_tol, self.max_dx, int(max_nsteps), 0.4, 0.5, 0.0,
                          ls._kind, ls.max_iter, int(ls.brack), ls.tol, ls.zeps)
        r = MinReport()
        cap = int(max_nsteps) + 1
        tally = np.zeros((cap, 2))
        rows = C.c_int()
        check(lib().mapp_md_minimize(atoms._h, C.byref(p), C.byref(r), ptr(tally), cap, C.byref(rows)))
        self.tally = tally[:rows.value]
        self.report = {k: getattr(r, k) for k, _ in MinReport._fields_}
        self.report["message"] = _MIN_MSGS.get(r.status, "")
        atoms.pe = r.e
        return self.report


class MinLBFGS(MinCG):
    """md.min_lbfgs(m, e_tol, H_dof, affine, max_dx, ls).run(atoms, max_nsteps)
    (src/min_l-bfgs.cpp:162, MinLBFGS::__run at src/min_l-bfgs.h:114-217)."""

    _lbfgs = 1

    def __init__(self, m=2, e_tol=1.0e-8, H_dof=None, affine=False, max_dx=1.0, ls=None):
        super().__init__(e_tol, H_dof, affine, max_dx, ls)
        if not 0 <= int(m) <= 64:
            raise ValueError("m must lie in [0, 64]")
        self.m = int(m)


def create_temp(elem, masses, T, seed, kB=8.617330350e-5):
    """AtomsMD::create_T: Maxwell velocities from the reference's random stream."""
    elem = np.ascontiguousarray(elem, dtype=np.uint8)
    masses = f64(masses)
    v = np.zeros((elem.shape[0], 3))
    check(lib().mapp_create_T(elem.shape[0], ptr(elem, u8p), ptr(masses), float(kB), float(T), int(seed), ptr(v)))
    return v


class AtomsDMD:
    """dmd.atoms with ForceFieldEAMDMD, NeighborDMD and DynamicDMD on the GPU.

    elem / alpha / c are [nsites][c_dim]; c < 0 marks an absent component.
    """

    def __init__(self, ff: ForceFieldEAM, H, x, elem, alpha, c, nquad=3, skin=0.5, T=300.0,
                 kB=8.617330350e-5, hP=None, masses=None):
        from .potentials import HP

        self.ff = ff
        x = f64(x).reshape(-1, 3)
        self.nsites = x.shape[0]
        elem = np.ascontiguousarray(elem, dtype=np.uint8).reshape(self.nsites, -1)
        self.c_dim = elem.shape[1]
        alpha = f64(alpha).reshape(self.nsites, self.c_dim)
        c = f64(c).reshape(self.nsites, self.c_dim)
        self.H = f64(H).reshape(3, 3)
        self.masses = f64(masses)
        self.kB, self.T = float(kB), float(T)
        h = C.c_void_p()
        check(lib().mapp_dmd_create(ff._h, self.nsites, self.c_dim, int(nquad), ptr(self.H), ptr(x), ptr(elem, u8p),
                                    ptr(alpha), ptr(c), float(skin), self.kB, self.T,
                                    float(HP if hP is None else hP), ptr(self.masses), C.byref(h)))
        self._h = h

    def state(self):
        x = np.zeros((self.nsites, 3))
        a = np.zeros((self.nsites, self.c_dim))
        c = np.zeros((self.nsites, self.c_dim))
        check(lib().mapp_dmd_get_state(self._h, ptr(x), ptr(a), ptr(c)))
        return x, a, c

    def set_state(self, x, alpha, c):
        x, alpha, c = f64(x), f64(alpha), f64(c)
        check(lib().mapp_dmd_set_state(self._h, ptr(x), ptr(alpha), ptr(c)))

    def update(self):
        r = C.c_int()
        check(lib().mapp_dmd_update(self._h, C.byref(r)))
        return bool(r.value)

    def derivative(self, with_c=False, ext_mu=None):
        """Returns (f, f_alpha, mu, fe, F_H, S)."""
        f = np.zeros((self.nsites, 3))
        fa = np.zeros((self.nsites, self.c_dim))
        mu = np.zeros((self.nsites, self.c_dim))
        out = np.zeros(8)
        em = None if ext_mu is None else f64(ext_mu)
        check(lib().mapp_dmd_derivative(self._h, int(with_c), ptr(em) if em is not None else None, ptr(f), ptr(fa),
                                        ptr(mu), ptr(out)))
        return f, fa, mu, out[0], out[1:7].copy(), out[7]

    def evaluate(self, with_c=False, ext_mu=None):
        """derivative() leaving f, f_alpha and mu on the device; returns (fe, F_H, S)."""
        out = np.zeros(8)
        em = None if ext_mu is None else f64(ext_mu)
        check(lib().mapp_dmd_derivative(self._h, int(with_c), ptr(em) if em is not None else None, None, None, None,
                                        ptr(out)))
        return out[0], out[1:7].copy(), out[7]

    def c_d(self, r_crd, zeta, fetch=True):
        """ForceFieldDMD::c_d_calc: (c_d [nsites][c_dim] or None, max |c_d|) for the
        exchange radii r_crd and widths zeta (the reference's C) per element."""
        out = np.zeros((self.nsites, self.c_dim)) if fetch else None
        norm = C.c_double()
        rc, zt = f64(r_crd), f64(zeta)
        check(lib().mapp_dmd_c_d(self._h, ptr(rc), ptr(zt), ptr(out) if fetch else None, C.byref(norm)))
        return out, norm.value

    def J(self, x):
        """ForceFieldEAMDMD::__J: (d c_d / d c) x [nsites][c_dim] at the state and
        exchange parameters of the last c_d call."""
        xv = f64(np.asarray(x).reshape(self.nsites, self.c_dim))
        out = np.zeros((self.nsites, self.c_dim))
        check(lib().mapp_dmd_J(self._h, ptr(xv), ptr(out)))
        return out

    def newton(self, r_crd, zeta, beta, a, a_tol=1e-8, err_fac=1.0, max_nnewton_iters=5, max_ngmres_iters=5):
        """DAEImplicit::newton: solve c / beta + a - c_d(c) = 0 from the current c.
        Returns (converged, Newton iterations); c stays on the device (state())."""
        av = f64(np.asarray(a).reshape(self.nsites, self.c_dim))
        rc, zt = f64(r_crd), f64(zeta)
        ok, it = C.c_int(), C.c_int()
        check(lib().mapp_dmd_newton(self._h, ptr(rc), ptr(zt), float(beta), ptr(av), float(a_tol), float(err_fac),
                                    int(max_nnewton_iters), int(max_ngmres_iters), C.byref(ok), C.byref(it)))
        return bool(ok.value), it.value

    def export_cfg(self, path, names, masses=None):
        """dmd.export_cfg: the current state as an extended CFG file (caller order);
        names and masses of the force field's elements (masses default to the
        system's)."""
        m = f64(self.masses if masses is None else masses)
        arr = (C.c_char_p * len(names))(*[str(n).encode() for n in names])
        check(lib().mapp_dmd_export_cfg(self._h, str(path).encode(), len(names), ptr(m), arr))

    def bdf_run(self, r_crd, zeta, t_tot, a_tol=1e-8, min_dt=2.220446049250313e-16, max_nsteps=1000,
                max_nnewton_iters=5, max_ngmres_iters=5, nreset=0, ntally=0, relax_max_iter=1000):
        """dmd.bdf().run: variable-order BDF of the concentrations over t_tot.
        Returns (report dict, tally rows [k][3] of (step, time, free energy))."""
        from ._lib import BdfParams, BdfReport

        p = BdfParams(float(a_tol), float(min_dt), int(max_nsteps), int(max_nnewton_iters), int(max_ngmres_iters),
                      int(nreset), int(ntally), int(relax_max_iter))
        r = BdfReport()
        cap = (int(max_nsteps) // int(ntally) + 2) if ntally else 0
        tally = np.zeros((max(cap, 1), 3))
        rows = C.c_int()
        rc, zt = f64(r_crd), f64(zeta)
        check(lib().mapp_dmd_bdf_run(self._h, ptr(rc), ptr(zt), float(t_tot), C.byref(p), C.byref(r), ptr(tally),
                                     cap, C.byref(rows)))
        rep = {k: getattr(r, k) for k, _ in BdfReport._fields_}
        return rep, tally[:rows.value]

    def Jxa(self, dx, dalpha, with_f=False):
        """ForceFieldEAMDMD::__prepJ_n_res + __J at the current state: returns
        (Jdx [nsites][3], Jdalpha [nsites][c_dim]) = -(d f / d(x, alpha)) (dx,
        dalpha) at fixed c, f the gradient per unit c; with_f adds (f, f_alpha)."""
        dx = f64(dx).reshape(self.nsites, 3)
        da = f64(dalpha).reshape(self.nsites, self.c_dim)
        jx, ja = np.zeros((self.nsites, 3)), np.zeros((self.nsites, self.c_dim))
        fx, fa = (np.zeros((self.nsites, 3)), np.zeros((self.nsites, self.c_dim))) if with_f else (None, None)
        check(lib().mapp_dmd_Jxa(self._h, ptr(dx), ptr(da), ptr(jx), ptr(ja), ptr(fx) if with_f else None,
                                 ptr(fa) if with_f else None))
        return (jx, ja, fx, fa) if with_f else (jx, ja)

    def min_newton(self, a_tol, max_nnewton_iters=5, max_ngmres_iters=5):
        """the Newton stage of DAE::min_error_false over x and alpha at fixed c:
        (steps taken, |f| at the end)"""
        k, r = C.c_int(), C.c_double()
        check(lib().mapp_dmd_min_newton(self._h, float(a_tol), int(max_nnewton_iters), int(max_ngmres_iters),
                                        C.byref(k), C.byref(r)))
        return k.value, r.value

    def unit_gradient(self):
        """calc_mu's gradient per unit c at the state of the last c_d call:
        (g [nsites][3], g_alpha [nsites][c_dim])."""
        g = np.zeros((self.nsites, 3))
        ga = np.zeros((self.nsites, self.c_dim))
        check(lib().mapp_dmd_unit_gradient(self._h, ptr(g), ptr(ga)))
        return g, ga

    def J_entries(self):
        """(list entries, second-list entries) of the last J state."""
        a, b = C.c_int64(), C.c_int64()
        check(lib().mapp_dmd_J_entries(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_batch(self, b):
        """List entries per batch of the quadrature passes (1, 2 or 4)."""
        check(lib().mapp_dmd_set_batch(self._h, int(b)))

    def set_box(self, H, x=None):
        """a new box H and optionally new positions (alpha and c kept); the
        grid and the component lists are rebuilt"""
        Hn = f64(H).reshape(3, 3)
        xv = None if x is None else f64(x).reshape(self.nsites, 3)
        check(lib().mapp_dmd_set_box(self._h, ptr(Hn), ptr(xv) if xv is not None else None))
        self.H = Hn

    def force_calc_static(self):
        """ForceFieldDMD::force_calc_static: (fe, F_H[6], S) of the state of the
        last c_d call, without new passes"""
        out = np.zeros(8)
        check(lib().mapp_dmd_force_calc_static(self._h, ptr(out)))
        return out[0], out[1:7].copy(), out[7]

    def set_fp32(self, on=True, rows=False):
        """FP32 pair splines (FP64 moments and sums) in the free energy and its
        gradient: from the knots held in LDS, or with rows=True from FP32
        copies of the coefficient rows; the concentration-rate path stays FP64"""
        check(lib().mapp_dmd_set_fp32(self._h, 0 if not on else (2 if rows else 1)))

    def knot_window(self):
        """first knot of the LDS window of the FP32 knots, or -1 where they
        are not in LDS"""
        k = C.c_int()
        check(lib().mapp_dmd_knot_window(self._h, C.byref(k)))
        return k.value

    def value(self, with_c=False, ext_mu=None):
        e = C.c_double()
        em = None if ext_mu is None else f64(ext_mu)
        check(lib().mapp_dmd_value(self._h, int(with_c), ptr(em) if em is not None else None, C.byref(e)))
        return e.value

    def half_pairs(self):
        n = C.c_int64()
        check(lib().mapp_dmd_pairs_count(self._h, C.byref(n)))
        i = np.zeros(n.value, np.int32)
        j = np.zeros(n.value, np.int32)
        img = np.zeros((n.value, 3), np.int8)
        check(lib().mapp_dmd_pairs(self._h, ptr(i, i32p), ptr(j, i32p), ptr(img, i8p)))
        return i, j, img

    def smear(self, ei, ej, r, a):
        r, a = f64(r), f64(a)
        out = np.zeros((len(r), 9))
        check(lib().mapp_dmd_smear(self._h, int(ei), int(ej), len(r), ptr(r), ptr(a), ptr(out)))
        return out

    def _min_run(self, lbfgs, m, max_iter, e_tol, max_dx, max_dalpha, grad_tol, ls):
        from ._lib import MinParams
        it, st = C.c_int(), C.c_int()
        out = np.zeros(4)
        if isinstance(ls, LineSearchBackTrack):
            p = MinParams(lbfgs, m, e_tol, max_dx, max_iter, ls.c, ls.rho, ls.min_alpha, 0, 5, 1,
                          (2.0 * 2.220446049250313e-16) ** 0.5, 2.220446049250313e-19)
        elif isinstance(ls, LineSearchGoldenSection):
            p = MinParams(lbfgs, m, e_tol, max_dx, max_iter, 0.4, 0.5, 0.0, ls._kind, ls.max_iter, int(ls.brack),
                          ls.tol, ls.zeps)
        else:
            raise TypeError("ls must be LineSearchBackTrack, LineSearchGoldenSection or LineSearchBrent")
        check(lib().mapp_dmd_min_run(self._h, int(lbfgs), int(m), int(max_iter), float(e_tol), float(max_dx),
                                     float(max_dalpha), float(grad_tol), C.byref(p), C.byref(it), C.byref(st),
                                     ptr(out)))
        return {"iters": it.value, "status": st.value, "fe0": out[0], "fe": out[1], "grad_norm": out[2],
                "grad_max": out[3]}

    def min_cg(self, max_iter=1000, e_tol=1e-8, max_dx=1.0, max_dalpha=0.1, grad_tol=0.0, ls=None):
        """dmd.min_cg(e_tol, max_dx, max_dalpha, ls).run(atoms, max_iter) over x and
        alpha (src/min_cg_dmd.h:633-700); ls defaults to the back-tracking search."""
        if ls is not None:
            return self._min_run(0, 0, max_iter, e_tol, max_dx, max_dalpha, grad_tol, ls)
        it, st = C.c_int(), C.c_int()
        out = np.zeros(4)
        check(lib().mapp_dmd_min_cg(self._h, int(max_iter), float(e_tol), float(max_dx), float(max_dalpha),
                                    float(grad_tol), C.byref(it), C.byref(st), ptr(out)))
        return {"iters": it.value, "status": st.value, "fe0": out[0], "fe": out[1], "grad_norm": out[2],
                "grad_max": out[3]}

    def min_lbfgs(self, m=2, max_iter=1000, e_tol=1e-8, max_dx=1.0, max_dalpha=0.1, grad_tol=0.0, ls=None):
        """dmd.min_lbfgs(m, e_tol, max_dx, max_dalpha, ls).run(atoms, max_iter) over
        x and alpha (src/min_l-bfgs_dmd.h:113-210); report as min_cg."""
        if ls is not None:
            return self._min_run(1, m, max_iter, e_tol, max_dx, max_dalpha, grad_tol, ls)
        it, st = C.c_int(), C.c_int()
        out = np.zeros(4)
        check(lib().mapp_dmd_min_lbfgs(self._h, int(m), int(max_iter), float(e_tol), float(max_dx),
                                       float(max_dalpha), float(grad_tol), C.byref(it), C.byref(st), ptr(out)))
        return {"iters": it.value, "status": st.value, "fe0": out[0], "fe": out[1], "grad_norm": out[2],
                "grad_max": out[3]}

    def use_surrogate(self, s=None):
        """Evaluate with the networks of s inside their training box (None: quadrature only)."""
        self._surr = s
        check(lib().mapp_dmd_use_surrogate(self._h, s._h if s is not None else None))

    def surrogate_coverage(self):
        """(network evaluations, slots left to the quadrature) of the last evaluation"""
        a, b = C.c_int64(), C.c_int64()
        check(lib().mapp_dmd_surrogate_coverage(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def profile(self, enable=True):
        check(lib().mapp_dmd_profile(self._h, int(enable)))

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_int64()
        check(lib().mapp_dmd_profile_read(self._h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mapp_dmd_destroy(self._h)
            self._h = None


def _names(names):
    return (C.c_char_p * len(names))(*[str(v).encode() for v in names])


def import_cfg_md(path):
    """md.atoms.import_cfg: a standard or extended CFG file of an MD system as a
    dict with H [3][3] (lower triangular, rows = box vectors), x [natms][3]
    (Cartesian), v [natms][3] (Cartesian; None for .NO_VELOCITY.), elem
    [natms], masses and names (element symbols in order of first
    appearance)."""
    h = C.c_void_p()
    check(lib().mapp_cfg_md_read(str(path).encode(), C.byref(h)))
    try:
        n, ne, hv = C.c_int(), C.c_int(), C.c_int()
        check(lib().mapp_cfg_md_info(h, C.byref(n), C.byref(ne), C.byref(hv)))
        n, ne, hv = n.value, ne.value, bool(hv.value)
        out = {"H": np.zeros((3, 3)), "x": np.zeros((n, 3)), "v": np.zeros((n, 3)) if hv else None,
               "elem": np.zeros(n, np.uint8), "masses": np.zeros(max(ne, 1))}
        check(lib().mapp_cfg_md_get(h, ptr(out["H"]), ptr(out["x"]), ptr(out["v"]) if hv else None,
                                    ptr(out["elem"], u8p), ptr(out["masses"])))
        out["masses"] = out["masses"][:ne]
        out["names"] = [lib().mapp_cfg_md_element(h, i).decode() for i in range(ne)]
        return out
    finally:
        lib().mapp_cfg_md_free(h)


def replicate(cfg, N):
    """atoms *= N (Atoms::mul, src/atoms.cpp:335-390) on a dict from
    import_cfg_md: copy k = (k0 N1 + k1) N2 + k2 of every atom shifted by
    (k0, k1, k2) H, copies in order of k, box row d scaled by N[d]."""
    N = [int(v) for v in N]
    H = np.asarray(cfg["H"], dtype=np.float64)
    k = np.array([(a, b, c) for a in range(N[0]) for b in range(N[1]) for c in range(N[2])], dtype=np.float64)
    shift = k @ H
    out = dict(cfg)
    out["x"] = (cfg["x"][None, :, :] + shift[:, None, :]).reshape(-1, 3)
    out["elem"] = np.tile(cfg["elem"], len(k))
    if cfg.get("v") is not None:
        out["v"] = np.tile(cfg["v"], (len(k), 1))
    out["H"] = H * np.array(N, dtype=np.float64)[:, None]
    return out


def write_cfg_md(path, H, x, elem, names, masses, v=None, aux=None, aux_names=None):
    """md.export_cfg for host arrays: x (and v) Cartesian [natms][3], H lower
    triangular, elem [natms]; aux [natms][naux] extra columns named aux_names."""
    x = f64(x).reshape(-1, 3)
    n = x.shape[0]
    elem = np.ascontiguousarray(elem, dtype=np.uint8).reshape(n)
    H, m = f64(H).reshape(3, 3), f64(masses)
    vv = None if v is None else f64(v).reshape(n, 3)
    naux = 0 if aux is None else int(np.asarray(aux).reshape(n, -1).shape[1])
    av = None if aux is None else f64(aux).reshape(n, naux)
    an = _names(aux_names or []) if naux else None
    check(lib().mapp_cfg_md_write(str(path).encode(), n, ptr(H), ptr(x), ptr(elem, u8p),
                                  ptr(vv) if vv is not None else None, naux, ptr(av) if av is not None else None, an,
                                  len(names), ptr(m), _names(names)))


def import_cfg_dmd(path):
    """dmd.atoms.import_cfg: an extended CFG file of a DMD system as a dict with
    H [3][3], x [natms][3], elem / alpha / c [natms][c_dim], masses and names
    (the element symbols in file order)."""
    h = C.c_void_p()
    check(lib().mapp_cfg_dmd_read(str(path).encode(), C.byref(h)))
    try:
        n, cd, ne = C.c_int(), C.c_int(), C.c_int()
        check(lib().mapp_cfg_dmd_info(h, C.byref(n), C.byref(cd), C.byref(ne)))
        n, cd, ne = n.value, max(cd.value, 1), ne.value
        out = {"H": np.zeros((3, 3)), "x": np.zeros((n, 3)), "elem": np.zeros((n, cd), np.uint8),
               "alpha": np.zeros((n, cd)), "c": np.zeros((n, cd)), "masses": np.zeros(max(ne, 1))}
        check(lib().mapp_cfg_dmd_get(h, ptr(out["H"]), ptr(out["x"]), ptr(out["elem"], u8p), ptr(out["alpha"]),
                                     ptr(out["c"]), ptr(out["masses"])))
        out["masses"] = out["masses"][:ne]
        out["names"] = [lib().mapp_cfg_dmd_element(h, i).decode() for i in range(ne)]
        return out
    finally:
        lib().mapp_cfg_dmd_free(h)


def write_cfg_dmd(path, H, x, elem, alpha, c, names, masses):
    """dmd.export_cfg for host arrays: x Cartesian [natms][3], H lower triangular,
    elem / alpha / c [natms][c_dim]."""
    x = f64(x).reshape(-1, 3)
    n = x.shape[0]
    elem = np.ascontiguousarray(elem, dtype=np.uint8).reshape(n, -1)
    cd = elem.shape[1]
    alpha, c, H, m = f64(alpha).reshape(n, cd), f64(c).reshape(n, cd), f64(H).reshape(3, 3), f64(masses)
    arr = (C.c_char_p * len(names))(*[str(v).encode() for v in names])
    check(lib().mapp_cfg_dmd_write(str(path).encode(), n, cd, ptr(H), ptr(x), ptr(elem, u8p), ptr(alpha), ptr(c),
                                   len(names), ptr(m), arr))


def gauss_hermite(n):
    x = np.zeros(n)
    w = np.zeros(n)
    check(lib().mapp_gauss_hermite(int(n), ptr(x), ptr(w)))
    return x, w


class ReplicaBatch:
    """Many small periodic cells ("replicas") minimised together: each runs
    md.min_cg with LineSearchBackTrack over its atom positions
    (src/min_cg.h:453-533, src/ls.h:711-769) inside one persistent launch.
    cells: a list of (H, x, elem) per replica."""

    def __init__(self, ff: ForceFieldEAM, cells, skin=0.5):
        self.ff = ff
        natms = np.array([len(f64(x).reshape(-1, 3)) for _, x, _ in cells], np.int32)
        H = np.ascontiguousarray(np.concatenate([f64(h).reshape(1, 9) for h, _, _ in cells]))
        x = np.ascontiguousarray(np.concatenate([f64(xx).reshape(-1, 3) for _, xx, _ in cells]))
        elem = np.ascontiguousarray(np.concatenate([np.asarray(e, np.uint8).ravel() for _, _, e in cells]))
        self.natms = natms
        self.H = H.reshape(-1, 3, 3)
        h = C.c_void_p()
        check(lib().mapp_replica_create(ff._h, len(natms), natms.ctypes.data_as(i32p), ptr(H), ptr(x),
                                        ptr(elem, u8p), float(skin), C.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.natms)

    @staticmethod
    def _ls(ls):
        ls = ls or LineSearchBackTrack()
        if not isinstance(ls, LineSearchBackTrack):
            raise ValueError("replica minimisation runs the back-tracking line search")
        return ls

    def min_cg(self, max_nsteps=1000, e_tol=1e-8, max_dx=1.0, ls=None):
        """per replica: steps, status (MinCG's err), e0, e, evals"""
        ls = self._ls(ls)
        n = len(self.natms)
        steps, status, evals = (np.zeros(n, np.int32) for _ in range(3))
        e0, e = np.zeros(n), np.zeros(n)
        check(lib().mapp_replica_min_cg(self._h, int(max_nsteps), float(e_tol), float(max_dx), ls.c, ls.rho,
                                        ls.min_alpha, steps.ctypes.data_as(i32p), status.ctypes.data_as(i32p),
                                        ptr(e0), ptr(e), evals.ctypes.data_as(i32p)))
        return {"steps": steps, "status": status, "e0": e0, "e": e, "evals": evals}

    def tune(self, max_nsteps=1000, e_tol=1e-8, max_dx=1.0, ls=None):
        ls = self._ls(ls)
        w = np.zeros(5, np.int32)
        ms = np.zeros(6)
        check(lib().mapp_replica_tune(self._h, int(max_nsteps), float(e_tol), float(max_dx), ls.c, ls.rho,
                                      ls.min_alpha, w.ctypes.data_as(i32p), ptr(ms)))
        return w, ms

    def set_waves(self, waves):
        w = np.ascontiguousarray(waves, dtype=np.int32)
        check(lib().mapp_replica_set_waves(self._h, w.ctypes.data_as(i32p)))

    def use_graph(self, on=True):
        check(lib().mapp_replica_use_graph(self._h, int(bool(on))))

    def x(self, r):
        out = np.zeros((int(self.natms[r]), 3))
        check(lib().mapp_replica_get_x(self._h, int(r), ptr(out)))
        return out

    def profile(self, enable=True):
        check(lib().mapp_replica_profile(self._h, int(enable)))

    def profile_read(self):
        ms, n = C.c_double(), C.c_int64()
        check(lib().mapp_replica_profile_read(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mapp_replica_destroy(self._h)
            self._h = None


class GCMCBatch:
    """Independent periodic cells, each with its own SGCMC chain of insert /
    delete attempts of one gas element (md.muvt's exchange step,
    src/sgcmc.cpp:145-327), all advanced together on the GPU.

    cells: list of (x [n][3], H [3][3], elem [n]); seeds: one stream seed per
    cell. Decisions: 1 / -1 accepted / rejected insertion, 2 / -2 deletion,
    0 deletion drawn with no gas atom.
    """

    def __init__(self, ff: ForceFieldEAM, cells, gas, mu, T, masses, seeds, cap=None, kB=8.617330350e-5, hP=None):
        from .potentials import HP

        self.ff = ff
        self.ncells = len(cells)
        natms = np.array([len(e) for _, _, e in cells], dtype=np.int32)
        x = f64(np.concatenate([np.asarray(c[0], dtype=np.float64).reshape(-1, 3) for c in cells]))
        elem = np.ascontiguousarray(np.concatenate([np.asarray(c[2], dtype=np.uint8) for c in cells]))
        H = f64(np.stack([np.asarray(c[1], dtype=np.float64).reshape(3, 3) for c in cells]))
        seeds = np.ascontiguousarray(seeds, dtype=np.int32)
        if len(seeds) != self.ncells:
            raise ValueError("one seed per cell")
        capp = None
        if cap is not None:
            cap = np.ascontiguousarray(cap, dtype=np.int32)
            capp = ptr(cap, i32p)
        self.masses = f64(masses)
        h = C.c_void_p()
        check(lib().mapp_gcmc_create(ff._h, self.ncells, ptr(natms, i32p), ptr(H), ptr(x), ptr(elem, u8p), capp,
                                     int(gas), float(mu), float(T), float(kB), float(HP if hP is None else hP),
                                     ptr(self.masses), ptr(seeds, i32p), C.byref(h)))
        self._h = h

    def attempts(self, k):
        d = np.zeros((self.ncells, k), np.int8)
        du = np.zeros((self.ncells, k))
        check(lib().mapp_gcmc_attempts(self._h, int(k), ptr(d, i8p), ptr(du)))
        return d, du

    def run(self, k):
        """k attempts per cell without copying decisions back"""
        check(lib().mapp_gcmc_attempts(self._h, int(k), None, None))

    def tune(self, k=16):
        w = np.zeros(5, np.int32)
        ms = np.zeros(6)
        check(lib().mapp_gcmc_tune(self._h, int(k), ptr(w, i32p), ptr(ms)))
        return w, ms

    def set_waves(self, waves):
        w = np.ascontiguousarray(waves, dtype=np.int32)
        check(lib().mapp_gcmc_set_waves(self._h, ptr(w, i32p)))

    def use_graph(self, on=True):
        check(lib().mapp_gcmc_use_graph(self._h, int(on)))

    def xchng_energy(self, cell, s=None, igas=None):
        """ForceFieldEAM::pre_xchng_energy + xchng_energy of one trial in one
        cell: an insertion of the gas element at fractional s, or (igas given)
        the deletion of the igas-th gas atom. Returns the exchange energy; the
        trial is kept for xchng_commit."""
        du = C.c_double()
        if igas is None:
            sv = f64(s).reshape(3)
            check(lib().mapp_gcmc_xchng_energy(self._h, int(cell), 1, ptr(sv), 0, C.byref(du)))
        else:
            check(lib().mapp_gcmc_xchng_energy(self._h, int(cell), 2, None, int(igas), C.byref(du)))
        return du.value

    def xchng_commit(self, v=None):
        """the kept trial accepted: post_xchng_energy after ins_succ / del_succ
        (v: the inserted atom's velocity)"""
        vv = None if v is None else f64(v).reshape(3)
        check(lib().mapp_gcmc_xchng_commit(self._h, ptr(vv) if vv is not None else None))

    def counts(self):
        n = np.zeros(self.ncells, np.int32)
        g = np.zeros(self.ncells, np.int32)
        check(lib().mapp_gcmc_counts(self._h, ptr(n, i32p), ptr(g, i32p)))
        return n, g

    def cell(self, c):
        n = int(self.counts()[0][c])
        x, rho, F, v = np.zeros((n, 3)), np.zeros(n), np.zeros(n), np.zeros((n, 3))
        elem = np.zeros(n, np.uint8)
        seed = C.c_int32()
        check(lib().mapp_gcmc_get_cell(self._h, int(c), ptr(x), ptr(elem, u8p), ptr(rho), ptr(F), ptr(v),
                                       C.byref(seed)))
        return dict(x=x, elem=elem, rho=rho, F=F, v=v, seed=seed.value)

    def profile(self, enable=True):
        check(lib().mapp_gcmc_profile(self._h, int(enable)))

    def profile_read(self, kernel="gcmc"):
        ms, n = C.c_double(), C.c_int64()
        check(lib().mapp_gcmc_profile_read(self._h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mapp_gcmc_destroy(self._h)
            self._h = None


class Surrogate:
    """Per-element-pair MLPs (3 -> 64 -> 64 -> 6, tanh) standing in for the
    DMD Gauss-Hermite smearing; trained on the GPU from the quadrature of an
    AtomsDMD's force field. box = (r_lo, r_hi, alpha_lo, alpha_hi)."""

    NPARAMS = 4806

    def __init__(self, ne, box=(1.8, 6.7, 0.05, 0.30), seed=1):
        self.ne = ne
        self.box = tuple(float(v) for v in box)
        b = np.ascontiguousarray(self.box, dtype=np.float32)
        h = C.c_void_p()
        check(lib().mapp_surr_create(int(ne), ptr(b, fp), int(seed), C.byref(h)))
        self._h = h

    def train(self, dmd, nsamples=2_000_000, steps=4000, batch=4096, lr=2e-3):
        out = np.zeros(2)
        check(lib().mapp_surr_train(self._h, dmd._h, int(nsamples), int(steps), int(batch), float(lr), ptr(out)))
        return out[0], out[1]

    def loss_grad(self, net, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.NPARAMS, np.float32)
        loss = C.c_double()
        check(lib().mapp_surr_loss_grad(self._h, int(net), len(x), ptr(x, fp), ptr(t, fp), ptr(g, fp),
                                        C.byref(loss)))
        return loss.value, g

    def params(self, net):
        p = np.zeros(self.NPARAMS, np.float32)
        check(lib().mapp_surr_get_params(self._h, int(net), ptr(p, fp)))
        return p

    def set_params(self, net, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().mapp_surr_set_params(self._h, int(net), ptr(p, fp)))

    def norm(self, net):
        sc, sh = np.zeros(6, np.float32), np.zeros(6, np.float32)
        check(lib().mapp_surr_get_norm(self._h, int(net), ptr(sc, fp), ptr(sh, fp)))
        return sc, sh

    def set_norm(self, net, scale, shift):
        sc = np.ascontiguousarray(scale, dtype=np.float32)
        sh = np.ascontiguousarray(shift, dtype=np.float32)
        check(lib().mapp_surr_set_norm(self._h, int(net), ptr(sc, fp), ptr(sh, fp)))

    def infer(self, net, rin):
        """outputs [n][6] for raw inputs rin[n][3] = (r, alpha_a, alpha_b)"""
        rin = np.ascontiguousarray(rin, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((len(rin), 6), np.float32)
        check(lib().mapp_surr_infer(self._h, int(net), len(rin), ptr(rin, fp), ptr(out, fp)))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mapp_surr_destroy(self._h)
            self._h = None
