"""Oracle restatement of the reference's variable-order BDF integrator of the
concentrations, DAEBDF (src/dae_bdf.cpp:103-700, src/dae_bdf.h:103-546),
over the oracle's DMD functions (c_d_calc, __J, DAEImplicit::newton).

Test infrastructure: only tests/ and the benchmark's CPU baseline use it.

The positions and widths stay fixed (the reference's min_error relaxation
between steps is the caller's: ``min_error`` is called where the reference
calls DAE::min_error_false). nreset = 0 means no periodic relaxation (the
reference takes (istep + 1) % nreset there, a division by zero).
"""
import numpy as np

MAX_Q = 5


def _l_calc(q, dt, t, l):
    """DAEBDFMath::l_calc (src/dae_bdf.h:209-226)."""
    l[0] = 1.0
    l[1:q + 1] = 0.0
    for i in range(q):
        k0 = 1.0 / (dt - t[i])
        for j in range(i + 1, 0, -1):
            l[j] += k0 * l[j - 1]


def _prep_A_bar_l(q, dt, t, dq, l, A_bar):
    """VC_y::prep_A_bar_l (src/dae_bdf.h:488-503)."""
    if dq == 1:
        _l_calc(q + 1, dt, t, l)
    elif dq == -1:
        _l_calc(q - 1, dt, t, l)
        p = 1.0
        for i in range(q - 1):
            p *= dt - t[i]
        for i in range(1, q):
            A_bar[i][q] -= p * l[i - 1]
    else:
        _l_calc(q, dt, t, l)


def _err_fac_calc(q, dt, t, lo, hi):
    """VC_y::err_fac_calc (src/dae_bdf.h:505-543): (err_fac, beta)."""
    beta_inv = 0.0
    for i in range(q):
        beta_inv += 1.0 / (dt - t[i])
    beta = 1.0 / beta_inv
    s_q = 0.0
    for i in range(q):
        s_q += 1.0 / (dt - t[i])
    err_fac = abs(1.0 / (s_q * (dt - t[q])))
    if q < MAX_Q:
        with np.errstate(divide="ignore", invalid="ignore"):  # t = 0 after a reset, as in the reference
            r = np.float64(-dt) / np.float64(t[q + 1])
            for i in range(1, q + 1):
                r *= (t[i] - dt) / np.float64(t[i])
        s_qq = s_q + 1.0 / (dt - t[q])
        hi[0] = abs(1.0 / (dt * (q + 2.0) * s_qq))
        hi[1] = -r
    if q > 1:
        p_q = 1.0
        for i in range(q):
            p_q *= dt - t[i]
        s_q_1 = s_q - 1.0 / (dt - t[q - 1])
        lo[0] = abs(1.0 / ((dt - t[q]) * s_q_1))
        lo[1] = p_q
    return err_fac, beta


class OBdf:
    """DAEBDF::run over an oracle_lib.ODMD system ``d`` (c [n][c_dim])."""

    def __init__(self, d, r_crd, zeta, masses, a_tol=1e-8, min_dt=np.finfo(float).eps, max_nsteps=1000,
                 max_nnewton_iters=5, max_ngmres_iters=5, nreset=0, ntally=0, min_error=None, x=None, alpha=None):
        self.d, self.r_crd, self.zeta, self.masses = d, r_crd, zeta, masses
        self.a_tol, self.min_dt = a_tol, min_dt
        self.max_nsteps, self.max_newton, self.max_gmres = max_nsteps, max_nnewton_iters, max_ngmres_iters
        self.nreset, self.ntally, self.min_error = nreset, ntally, min_error
        self.x, self.alpha = x, alpha  # the fixed positions and widths (for restoring c)
        # A[i][j]: binomial coefficients (src/dae_bdf.cpp:17-31)
        self.A = np.zeros((MAX_Q + 1, MAX_Q + 1))
        for i in range(MAX_Q + 1):
            self.A[0][i] = self.A[i][i] = 1.0
        for i in range(2, MAX_Q + 1):
            for j in range(1, i):
                self.A[j][i] = self.A[j][i - 1] + self.A[j - 1][i - 1]
        self.A_bar = np.zeros((MAX_Q + 1, MAX_Q + 1))
        self.t = np.zeros(MAX_Q + 1)
        self.l = np.zeros(MAX_Q + 1)
        self.lo, self.hi = np.zeros(2), np.zeros(2)
        self.counts = dict(nonlin_acc=0, nonlin_rej=0, intpol_acc=0, intpol_rej=0, integ_acc=0, integ_rej=0,
                           err_mins=0)
        self.tally = []

    # --- helpers over the oracle ---
    def _c(self):
        return self.d.state_c().ravel().copy()

    def _set_c(self, c):
        self.d.set_state(self.x, self.alpha, c.reshape(self.d.n, self.d.C))

    def _c_d(self):
        cd, _ = self.d.c_d(self.r_crd, self.zeta, self.masses)
        return cd.ravel()

    def _calc_err(self):
        """DAE::calc_err (src/dae.cpp:151-164) without box dofs: the norm of the
        gradient per unit c (calc_mu's f and f_alpha) per x and alpha dof at the
        current state."""
        g, ga = self.d.unit_grad()
        return np.sqrt((g * g).sum() + (ga * ga).sum()) / np.sqrt(g.size + ga.size)

    def _reset(self):
        """DAEBDF::reset (src/dae_bdf.cpp:560-604)."""
        cd = self._c_d()
        n_c = cd.size
        a_tol_sqrt = self.a_tol * np.sqrt(float(n_c))
        jc = self.d.J(self.r_crd, self.zeta, self.masses, cd.reshape(self.d.n, self.d.C)).ravel()
        norm = np.sqrt((jc * jc).sum()) / a_tol_sqrt
        c = self._c()
        max_dt = np.inf
        self.z = np.zeros((n_c, MAX_Q + 1))
        for i in range(n_c):
            if c[i] >= 0.0:
                max_dt = min(max_dt, self._get_max_dt(c[i], cd[i]))
                self.z[i, 0] = c[i]
                self.z[i, 1] = cd[i]
        self.q = 1
        with np.errstate(divide="ignore"):
            s = np.sqrt(2.0 / norm)
        self.dt = max(min(min(s, (self.t_fin - self.t_cur) * 0.001), max_dt), self.min_dt)
        self.t[:] = 0.0
        self.dy = np.zeros(n_c)

    @staticmethod
    def _get_max_dt(x, x_d):
        if x_d > 0.0:
            ans = (1.0 - x) / x_d
            while x + x_d * ans > 1.0:
                ans = np.nextafter(ans, -1.0)
            return ans
        if x_d < 0.0:
            ans = -x / x_d
            while x + x_d * ans < 0.0:
                ans = np.nextafter(ans, -1.0)
            return ans
        return np.inf

    def _interpolate(self):
        """DAEBDF::interpolate (src/dae_bdf.cpp:322-366)."""
        q, dt = self.q, self.dt
        for i in range(q + 1):
            p = 1.0
            for j in range(i, q + 1):
                self.A_bar[i][j] = self.A[i][j] * p
                p *= dt
        self.err_fac, self.beta = _err_fac_calc(q, dt, self.t, self.lo, self.hi)
        y0 = self.z[:, :q + 1] @ self.A_bar[0, :q + 1]
        if np.any((y0 < 0.0) | (y0 > 1.0)):
            self.counts["intpol_rej"] += 1
            return False
        self.y_0 = y0
        self.a = self.z[:, 1:q + 1] @ self.A_bar[1, 1:q + 1] - y0 * (1.0 / self.beta)
        self.counts["intpol_acc"] += 1
        return True

    def _newton(self):
        ok, it, _ = self.d.newton(self.r_crd, self.zeta, self.masses, self.beta,
                                  self.a.reshape(self.d.n, self.d.C), self.a_tol, self.err_fac, self.max_newton,
                                  self.max_gmres)
        if ok:
            if it:
                self.counts["nonlin_acc"] += 1
        else:
            self.counts["nonlin_rej"] += 1
        return ok

    def _shrink(self, r):
        if r * self.dt < self.min_dt:
            self.dt = self.min_dt
        elif r * self.dt > self.t_fin - self.t_cur - self.min_dt:
            self.dt = self.t_fin - self.t_cur - self.min_dt
        else:
            self.dt *= r

    def _fail_step(self, r):
        """newton_fail / integrate_fail (src/dae_bdf.cpp:660-700, 612-650)."""
        if self.t_fin - self.t_cur <= 2.0 * self.min_dt or self.dt == self.min_dt:
            if self.q > 1:
                self.q -= 1
        else:
            self._shrink(r)

    def _integrate(self):
        """DAEBDF::integrate (src/dae_bdf.cpp:254-301)."""
        c_0 = self._c()
        while True:
            if not self._interpolate():
                self._reset()  # interpolate_fail
                continue
            if not self._newton():
                self._fail_step(0.25)
                self._set_c(c_0)
                continue
            break
        c = self._c()
        n_c = c.size
        self.err = self.err_fac * np.sqrt(((c - self.y_0) ** 2).sum()) / (self.a_tol * np.sqrt(float(n_c)))
        if self.err < 1.0:
            self.counts["integ_acc"] += 1
            return True
        self._set_c(c_0)
        self.counts["integ_rej"] += 1
        return False

    def _prep_for_next(self):
        """DAEBDF::prep_for_next (src/dae_bdf.cpp:370-465)."""
        q = self.q
        c = self._c()
        a_tol_sqrt = self.a_tol * np.sqrt(float(c.size))
        eta = [0.0, np.power(0.5 / self.err, 1.0 / (q + 1.0)), 0.0]
        live = c >= 0.0
        with np.errstate(all="ignore"):
            if q < MAX_Q:
                tmp = (c - self.y_0 + self.hi[1] * self.dy)[live]
                norm = self.hi[0] * np.sqrt((tmp * tmp).sum()) / a_tol_sqrt
                eta[2] = np.power(0.5 / norm, 1.0 / (q + 2.0))
            if q > 1:
                tmp = (c - self.y_0 + self.lo[1] * self.z[:, q])[live]
                norm = self.lo[0] * np.sqrt((tmp * tmp).sum()) / a_tol_sqrt
                eta[0] = np.power(0.5 / norm, 1.0 / q)
        self.dy = c - self.y_0
        dt_r, self.dq = 1.0, 0
        if self.nconst_q > 2 + q and self.nconst_dt > 2 + q:
            if eta[0] > eta[1] and eta[0] > eta[2]:
                self.dq, dt_r = -1, eta[0]
            elif eta[2] > eta[0] and eta[2] > eta[1]:
                self.dq, dt_r = 1, eta[2]
            else:
                self.dq, dt_r = 0, eta[1]
        if 2.0 < dt_r:
            dt_r = 2.0
        elif dt_r <= 0.5:
            dt_r = 0.5
        left = self.t_fin - self.t_cur
        if dt_r * self.dt > left:
            new_dt = left
        elif left <= 2.0 * self.min_dt:
            new_dt = 2.0 * self.min_dt
        elif dt_r * self.dt < self.min_dt:
            new_dt = self.min_dt
        elif dt_r * self.dt >= left - self.min_dt:
            new_dt = left - self.min_dt
        else:
            new_dt = dt_r * self.dt
        self._update_z()
        for i in range(MAX_Q, 0, -1):
            self.t[i] = self.t[i - 1] - self.dt
        self.dt = new_dt
        self.q += self.dq

    def _update_z(self):
        """DAEBDF::update_z (src/dae_bdf.cpp:469-510) with DAEBDFMath::update_z(_inc)
        (src/dae_bdf.h:137-159): z_j <- sum_{k>=j} A_bar[j][k] z_k + dy l_j."""
        q, dq = self.q, self.dq
        _prep_A_bar_l(q, self.dt, self.t, dq, self.l, self.A_bar)
        dim = q if dq == 1 else q + dq
        if dq == 1:
            self.z[:, dim + 1] = self.l[dim + 1] * self.dy
        z = self.z
        for j in range(dim + 1):
            z[:, j] = z[:, j:dim + 1] @ self.A_bar[j, j:dim + 1] + self.dy * self.l[j]

    def run(self, t_tot):
        """DAEBDF::run (src/dae_bdf.cpp:103-250). Returns the number of steps."""
        self.t_cur, self.t_fin = 0.0, float(t_tot)
        if self.nreset and self.min_error:
            self._min_error()
        fe_err = prev_err = self._calc_err()
        self._reset()
        self.nconst_q = self.nconst_dt = 0
        if self.ntally:
            self.tally.append((0, self.t_cur, self.d.force()[3]))
        istep = 0
        while istep < self.max_nsteps and self.t_cur < self.t_fin:
            dt_prev, q_prev = self.dt, self.q
            while not self._integrate():
                self._fail_step(np.power(0.5 / self.err, 1.0 / float(self.q + 1)))
            self.nconst_q = 1 if q_prev != self.q else self.nconst_q + 1
            self.nconst_dt = 1 if dt_prev != self.dt else self.nconst_dt + 1
            dt_prev, q_prev = self.dt, self.q
            self.t_cur += dt_prev
            self._prep_for_next()
            if dt_prev != self.dt:
                self.nconst_dt = 0
            if q_prev != self.q:
                self.nconst_q = 0
            fe_err = self._calc_err()
            periodic = self.nreset and (istep + 1) % self.nreset == 0 and fe_err / self.a_tol > 1.0
            if self.nreset and self.min_error and (periodic or fe_err > 2.0 * prev_err):
                self._min_error()
                fe_err = self._calc_err()
                prev_err = fe_err
                self._reset()
                self.nconst_q = self.nconst_dt = 0
            istep += 1
            if self.ntally and istep % self.ntally == 0:
                self.tally.append((istep, self.t_cur, self.d.force()[3]))
        # the last step's line and relaxation (src/dae_bdf.cpp:203-217)
        if self.ntally and istep % self.ntally:
            self.tally.append((istep, self.t_cur, self.d.force()[3]))
        if self.nreset and self.min_error and istep % self.nreset:
            self._min_error()
        return istep

    def _min_error(self):
        """the caller's DAE::min_error_false; it returns the Newton steps it took,
        and a relaxation counts when there was one (src/dae.cpp:325)"""
        if self.min_error():
            self.counts["err_mins"] += 1
