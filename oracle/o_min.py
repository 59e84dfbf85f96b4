"""Oracle restatement of the reference's MD energy minimisers over the
positions (box fixed, H_dof all False; per-atom displacements, affine =
False): MinCG::__run (src/min_cg.h:453-533), MinLBFGS::__run
(src/min_l-bfgs.h:114-217), the back-tracking line search
LineSearchBackTrack::min (src/ls.h:711-769, defaults src/ls.cpp:456-461)
and MinMDHandler's ls_prep / F / F_reset (src/min_cg.h:231-278).

Test infrastructure: only tests/ and the benchmark's CPU baseline use it.

``system`` is any object with a position property ``x`` ([n][3]; setting it
moves the atoms), ``update()`` (the neighbor-list skin check and rebuild of
DynamicMD::update), ``energy()`` (ForceFieldMD::value) and ``force()``
returning (forces [n][3], energy, ...) as ForceFieldMD::derivative.
Vectors are numpy arrays; inner products are plain sums over all components.
"""
import numpy as np

LS_S, LS_F_DOWNHILL, LS_F_GRAD0, LS_MIN_ALPHA, MIN_S_TOLERANCE, MIN_F_MAX_ITER, B_S, B_F_MAX_ALPHA, B_F_DOWNHILL = \
    range(9)
EPS = float(np.finfo(np.float64).eps)
SQRT_EPS = float(np.sqrt(EPS))
GOLDEN = 0.5 + 0.5 * float(np.sqrt(5.0))
F64 = np.float64


def _dot(a, b):
    """Inner products as numpy scalars: a zero denominator gives inf or nan
    as in C++ instead of raising."""
    return np.float64(np.sum(a * b))


class Handler:
    """MinMDHandler<BC=false, X=true> (src/min_cg.h:144-278)."""

    def __init__(self, system, max_dx):
        self.s = system
        self.max_dx = max_dx
        self.x0 = None
        self.h = None
        self.f = None
        self.f_h = 0.0

    def force_calc(self):
        out = self.s.force()
        self.f = np.array(out[0], dtype=float)
        return float(out[1])

    def ls_prep(self):
        """src/min_cg.h:231-257: (dfa, h_norm, max_a); the X branch bounds
        every component of the step by max_dx (src/min_cg.h:214-225)."""
        h_norm = _dot(self.h, self.h)
        dfa = -self.f_h
        if h_norm == 0.0:
            return 0.0, h_norm, 0.0
        if dfa >= 0.0:
            return 1.0, h_norm, 0.0
        h_norm = np.sqrt(h_norm)
        max_a = abs(self.max_dx / float(np.max(np.abs(self.h))))
        return dfa, h_norm, max_a

    def F(self, alpha):
        self.s.x = self.x0 + alpha * self.h
        self.s.update()
        return float(self.s.energy())

    def F_reset(self):
        self.s.x = self.x0
        self.s.update()


class BackTrack:
    """LineSearchBackTrack (src/ls.cpp:456-461, src/ls.h:711-769)."""

    def __init__(self, c=0.4, rho=0.5, min_alpha=0.0):
        self.c, self.rho, self.min_alpha = c, rho, min_alpha
        self.prev_val = -1.0

    def reset(self):
        self.prev_val = -1.0

    def min(self, func, nrgy, init_flag):
        """Returns (status, energy, alpha)."""
        init_energy = nrgy
        dfa, _h_norm, max_a = func.ls_prep()
        if dfa == 0.0:
            return LS_F_GRAD0, nrgy, 0.0
        if max_a == 0.0:
            return LS_MIN_ALPHA, nrgy, 0.0
        if dfa >= 0.0:
            return LS_F_DOWNHILL, nrgy, 0.0
        if init_flag == 0:
            max_a = min(max_a, 1.0)
        elif init_flag == 1:
            if self.prev_val > 0.0 and dfa < 0.0:
                max_a = min(max_a, -self.prev_val / dfa)
            self.prev_val = -dfa
        elif init_flag == 2:
            if self.prev_val > 0.0 and dfa < 0.0:
                max_a = min(max_a, min(1.0, 2.02 * (self.prev_val + nrgy) / dfa))
            self.prev_val = -nrgy
        if max_a <= self.min_alpha:
            return LS_MIN_ALPHA, nrgy, 0.0
        alpha = max_a
        while True:
            ideal = nrgy + alpha * self.c * dfa
            cur = func.F(alpha)
            if cur <= ideal:
                if init_flag == 1:
                    self.prev_val *= alpha
                return LS_S, cur, alpha
            alpha *= self.rho
            # a nan step (from a 0/0 curvature pair once the forces stop
            # changing) never meets the reference's exit test; it ends here
            if alpha <= self.min_alpha or alpha != alpha:
                func.F_reset()
                return LS_MIN_ALPHA, init_energy, alpha


def bracket(func, dfa, max_a, a, fa):
    """LineSearch::bracket (src/ls.h:213-335). Returns (status, a, b, c, fa,
    fb, fc). Arithmetic in numpy float64 scalars so that a zero divisor gives
    inf or nan as in C++."""
    dfa, max_a, a, fa = F64(dfa), F64(max_a), F64(a), F64(fa)
    with np.errstate(all="ignore"):
        q = F64(1.0e-14) / (max_a * abs(dfa))
        if 0.01 < q < 1.0:
            b = q * max_a
        elif q >= 1.0:
            b = max_a
        else:
            b = F64(0.01) * max_a
        fb = F64(func.F(b))
        c = fc = F64(0.0)
        if fb >= fa:
            c, fc = b, fb
            it = 20
            r = (fa - fc) / (c * dfa)
            b = c * 0.5 / (1.0 + r)
            while fb >= fa and it and b > EPS:
                fb = F64(func.F(b))
                if fb < fa:
                    continue
                c, fc = b, fb
                it -= 1
                r = (fa - fc) / (c * dfa)
                b = c * 0.5 / (1.0 + r)
            if fb >= fa:
                b = F64(SQRT_EPS) if SQRT_EPS < max_a else max_a
                fb = F64(func.F(b))
                if fb >= fa:
                    return B_F_DOWNHILL, a, b, c, fa, fb, fc
            return B_S, a, b, c, fa, fb, fc
        fc = fb
        while fb >= fc:
            c = b + GOLDEN * (b - a)
            if c >= max_a:
                c = max_a
                fc = F64(func.F(c))
                return B_S, a, b, c, fa, fb, fc
            fc = F64(func.F(c))
            if fc > fb:
                continue
            ulim = max_a
            r = (b - a) * (fb - fc)
            q = (b - c) * (fb - fa)
            u = b - ((b - c) * q - (b - a) * r) / (2.0 * (q - r))
            if b < u < c:
                fu = F64(func.F(u))
                if fu < fc:
                    return B_S, b, u, c, fb, fu, fc
                elif fu > fb:
                    return B_S, a, b, u, fa, fb, fu
                a, b, fa, fb = b, c, fb, fc
            elif u > c:
                u = u if u < ulim else ulim
                fu = F64(func.F(u))
                a, b, c, fa, fb, fc = b, c, u, fb, fc, fu
            else:
                a, b, fa, fb = b, c, fb, fc
        return B_S, a, b, c, fa, fb, fc


def _prep_checks(func):
    dfa, _h_norm, max_a = func.ls_prep()
    if dfa == 0.0:
        return LS_F_GRAD0, dfa, max_a
    if max_a == 0.0:
        return LS_MIN_ALPHA, dfa, max_a
    if dfa >= 0.0:
        return LS_F_DOWNHILL, dfa, max_a
    return LS_S, dfa, max_a


class GoldenSection:
    """LineSearchGoldenSection (src/ls.cpp:95-101, src/ls.h:341-520)."""

    def __init__(self, tol=float(np.sqrt(2.0 * EPS)), max_iter=5, brack=True):
        self.tol, self.max_iter, self.brack = tol, max_iter, brack
        self.prev_val = -1.0

    def reset(self):
        self.prev_val = -1.0

    def min(self, func, nrgy, init_flag):
        gold, cgold = 0.61803399, 0.38196601
        st, dfa, max_a = _prep_checks(func)
        if st != LS_S:
            return st, nrgy, 0.0
        calc = 0
        x0 = x1 = x2 = x3 = f1 = f2 = F64(0.0)
        if self.brack:
            st, a, b, c, fa, fb, fc = bracket(func, dfa, max_a, 0.0, nrgy)
            if st != B_S:
                func.F_reset()
                return st, nrgy, 0.0
            x0, x3 = a, c
            if c - b > b - a:
                x1, x2 = b, b + cgold * (c - b)
                f1, f2 = fb, F64(func.F(x2))
                calc = 2
            else:
                x1, x2 = b + cgold * (a - b), b
                f1, f2 = F64(func.F(x1)), fb
                calc = 1
            it = self.max_iter
            while x3 - x0 > self.tol * (x1 + x2) and x3 - x0 > EPS and it:
                if f2 < f1:
                    x0, x1, x2 = x1, x2, gold * x2 + cgold * x3
                    f1, f2 = f2, F64(func.F(x2))
                    calc = 2
                else:
                    x3, x2, x1 = x2, x1, gold * x1 + cgold * x0
                    f2, f1 = f1, F64(func.F(x1))
                    calc = 1
                it -= 1
        else:
            f0 = F64(nrgy)
            x3 = F64(max_a)
            f3 = F64(func.F(x3))
            it = self.max_iter
            set_left = set_right = False
            while x3 - x0 > self.tol * (x1 + x2) and x3 - x0 > EPS and it:
                delta = cgold * (x3 - x0)
                if not set_left:
                    x1 = x0 + delta
                    f1 = F64(func.F(x1))
                    calc = 1
                    set_left = True
                if not set_right:
                    x2 = x3 - delta
                    f2 = F64(func.F(x2))
                    calc = 2
                    set_right = True
                if (f0 < f2 and f0 < f3) or (f1 < f2 and f1 < f3):
                    set_left = False
                    x3, x2, f3, f2 = x2, x1, f2, f1
                else:
                    set_right = False
                    x0, x1, f0, f1 = x1, x2, f1, f2
                it -= 1
            if not set_left and f0 < f1:
                if x0 == 0.0:
                    func.F_reset()
                    return LS_MIN_ALPHA, nrgy, 0.0
                x1 = x0
                f1 = F64(func.F(x1))
                calc = 1
            if not set_right and f3 < f2:
                x2 = x3
                f2 = F64(func.F(x2))
                calc = 2
        if f1 < f2:
            if calc == 2:
                f1 = F64(func.F(x1))
            nrgy, alpha = f1, x1
        else:
            if calc == 1:
                f2 = F64(func.F(x2))
            nrgy, alpha = f2, x2
        self.prev_val = -dfa * alpha
        return LS_S, float(nrgy), float(alpha)


class Brent(GoldenSection):
    """LineSearchBrent (src/ls.cpp:262-268, src/ls.h:525-706)."""

    def __init__(self, tol=float(np.sqrt(2.0 * EPS)), max_iter=5, zeps=EPS * 1.0e-3, brack=True):
        super().__init__(tol, max_iter, brack)
        self.zeps = zeps

    def min(self, func, nrgy, init_flag):
        cgold = 0.38196601
        st, dfa, max_a = _prep_checks(func)
        if st != LS_S:
            return st, nrgy, 0.0
        d = e = u = fu = F64(0.0)
        if self.brack:
            st, a, x, b, fa, fx, fb = bracket(func, dfa, max_a, 0.0, nrgy)
            if st != B_S:
                func.F_reset()
                return st, nrgy, 0.0
        else:
            a, fa = F64(0.0), F64(nrgy)
            b = F64(max_a)
            x = cgold * (b - a)
            fx = F64(func.F(x))
        w = v = x
        fw = fv = fx
        with np.errstate(all="ignore"):
            for _ in range(self.max_iter):
                xm = 0.5 * (a + b)
                tol1 = self.tol * abs(x) + self.zeps
                tol2 = 2.0 * tol1
                if abs(x - xm) <= (tol2 - 0.5 * (b - a)):
                    if u != x:
                        func.F(x)
                    self.prev_val = x
                    return LS_S, float(fx), float(x)
                if abs(e) > tol1:
                    r = (x - w) * (fx - fv)
                    q = (x - v) * (fx - fw)
                    p = (x - v) * q - (x - w) * r
                    q = 2.0 * (q - r)
                    if q > 0.0:
                        p = -p
                    q = abs(q)
                    etemp = e
                    e = d
                    if abs(p) >= abs(0.5 * q * etemp) or p <= q * (a - x) or p >= q * (b - x):
                        e = a - x if x >= xm else b - x
                        d = cgold * e
                    else:
                        d = p / q
                        u = x + d
                        if u - a < tol2 or b - u < tol2:
                            d = abs(tol1) if xm - x >= 0.0 else -abs(tol1)
                else:
                    e = a - x if x >= xm else b - x
                    d = cgold * e
                if abs(d) >= tol1:
                    u = x + d
                else:
                    u = x + abs(tol1) if d >= 0.0 else x - abs(tol1)
                fu = F64(func.F(u))
                if fu <= fx:
                    if u >= x:
                        a = x
                    else:
                        b = x
                    v, w, x = w, x, u
                    fv, fw, fx = fw, fx, fu
                else:
                    if u < x:
                        a = u
                    else:
                        b = u
                    if fu <= fw or w == x:
                        v, w, fv, fw = w, u, fw, fu
                    elif fu <= fv or v == x or v == w:
                        v, fv = u, fu
        if fa < fx:
            func.F_reset()
            return LS_MIN_ALPHA, nrgy, 0.0
        if u != x:
            func.F(x)
        self.prev_val = x
        return LS_S, float(fx), float(x)


def min_cg(system, max_nsteps, e_tol=1e-8, max_dx=1.0, ls=None):
    """MinCG::__run (src/min_cg.h:453-533). Returns a dict with steps,
    status, e0, e and tally rows (step, energy)."""
    ls = ls or BackTrack()
    hd = Handler(system, max_dx)
    ls.reset()
    e0 = hd.force_calc()
    e_curr = e0
    tally = [(0, e0)]
    err = MIN_F_MAX_ITER if max_nsteps == 0 else LS_S
    hd.h = hd.f.copy()
    f0_f0 = _dot(hd.f, hd.f)
    istep = 0
    while err == LS_S:
        if f0_f0 == 0.0:
            err = LS_F_GRAD0
            istep += 1
            continue
        hd.x0 = np.array(system.x, dtype=float)
        f0 = hd.f.copy()
        e_prev = e_curr
        hd.f_h = _dot(hd.f, hd.h)
        if hd.f_h < 0.0:
            hd.h = hd.f.copy()
            hd.f_h = f0_f0
        err, e_curr, alpha = ls.min(hd, e_curr, 1)
        if err != LS_S:
            continue
        hd.force_calc()
        tally.append((istep + 1, e_curr))
        if e_prev - e_curr < e_tol:
            err = MIN_S_TOLERANCE
        if istep + 1 == max_nsteps:
            err = MIN_F_MAX_ITER
        if err:
            istep += 1
            continue
        f_f = _dot(hd.f, hd.f)
        f_f0 = _dot(hd.f, f0)
        ratio = (f_f - f_f0) / f0_f0
        hd.h = ratio * hd.h + hd.f
        f0_f0 = f_f
        istep += 1
    return dict(steps=istep, status=err, e0=e0, e=e_curr, tally=np.array(tally))


def lbfgs_direction(f, s, y, rho, gamma):
    """The two-loop recursion of src/min_l-bfgs.h:150-162: s[0], y[0] the
    newest pair."""
    k = len(s)
    h = f.copy()
    alpha = np.zeros(k)
    for i in range(k):
        alpha[i] = -rho[i] * _dot(s[i], h)
        h = h + alpha[i] * y[i]
    h = h * gamma
    for i in range(k - 1, -1, -1):
        scl0 = -alpha[i] - rho[i] * _dot(y[i], h)
        h = h + scl0 * s[i]
    return h


def min_lbfgs(system, max_nsteps, m=2, e_tol=1e-8, max_dx=1.0, ls=None, trace=None):
    """MinLBFGS::__run (src/min_l-bfgs.h:114-217). ``trace``, when given, is
    called with (f, h, s, y, rho, gamma) after each direction is formed."""
    ls = ls or BackTrack()
    hd = Handler(system, max_dx)
    ls.reset()
    e0 = hd.force_calc()
    e_curr = e0
    tally = [(0, e0)]
    s, y = [], []
    rho = [0.0] * m
    gamma = 1.0
    k = 0
    err = MIN_F_MAX_ITER if max_nsteps == 0 else LS_S
    istep = 0
    while err == LS_S:
        hd.x0 = np.array(system.x, dtype=float)
        f0 = hd.f.copy()
        hd.h = lbfgs_direction(hd.f, s[:k], y[:k], rho[:k], gamma)
        if trace is not None:
            trace(hd.f.copy(), hd.h.copy(), [a.copy() for a in s[:k]], [a.copy() for a in y[:k]], list(rho[:k]),
                  gamma)
        e_prev = e_curr
        hd.f_h = _dot(hd.f, hd.h)
        if hd.f_h < 0.0:
            hd.h = hd.f.copy()
            k = 0
            hd.f_h = _dot(hd.f, hd.f)
        err, e_curr, alpha_m = ls.min(hd, e_curr, 0)
        if err != LS_S:
            continue
        hd.force_calc()
        tally.append((istep + 1, e_curr))
        if e_prev - e_curr < e_tol:
            err = MIN_S_TOLERANCE
        if istep + 1 == max_nsteps:
            err = MIN_F_MAX_ITER
        if err:
            istep += 1
            continue
        if m:
            if k != m:
                k += 1
            # cyclic_shift: the oldest of the first k slots takes the new pair
            s = s[:k - 1]
            y = y[:k - 1]
            for i in range(m - 1, 0, -1):
                rho[i] = rho[i - 1]
            s.insert(0, alpha_m * hd.h)
            y.insert(0, f0 - hd.f)
            inner0 = _dot(s[0], y[0])
            inner1 = _dot(y[0], y[0])
            gamma = inner0 / inner1
            rho[0] = 1.0 / inner0
        else:
            gamma = alpha_m * (_dot(hd.h, f0) - _dot(hd.h, hd.f)) / (
                _dot(hd.f, hd.f) + _dot(f0, f0) - 2.0 * _dot(hd.f, f0))
        istep += 1
    return dict(steps=istep, status=err, e0=e0, e=e_curr, tally=np.array(tally))
