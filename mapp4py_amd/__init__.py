"""mapp4py_amd -- MI355X (gfx950) implementation of MAPP's EAM / DMD hot path.

The Python layer is a thin mirror of the reference's objects over the C ABI
(include/mapp_amd.h); all computation runs in the HIP library
``libmapp_amd.so``:

* :class:`ForceFieldEAM`  -- sim.ff_eam_setfl / ff_eam_fs (src/ff_eam.cpp:594-801)
* :class:`AtomsMD`        -- md.atoms with its ForceField, NeighborMD and DynamicMD
  (force_calc / value / update of src/ff_md.cpp:125-179, src/dynamic_md.h:38-68)
* :class:`NVT`            -- md.nvt(T, dt).run(sim, nsteps) (src/md_nvt.cpp:383-480)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import MappError, check, f64, i8p, i32p, i64p, ip, lib, ptr, u8p

__all__ = ["ForceFieldEAM", "AtomsMD", "NVT", "MappError", "create_temp", "lib"]


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


def create_temp(elem, masses, T, seed, kB=8.617330350e-5):
    """AtomsMD::create_T: Maxwell velocities from the reference's random stream."""
    elem = np.ascontiguousarray(elem, dtype=np.uint8)
    masses = f64(masses)
    v = np.zeros((elem.shape[0], 3))
    check(lib().mapp_create_T(elem.shape[0], ptr(elem, u8p), ptr(masses), float(kB), float(T), int(seed), ptr(v)))
    return v
