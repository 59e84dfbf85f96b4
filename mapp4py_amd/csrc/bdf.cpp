// bdf.cpp -- DAEBDF on the host over the GPU DMD system.
#include "bdf.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace mapp {
namespace {

constexpr int kMaxQ = 5;

// DAEBDFMath::l_calc (src/dae_bdf.h:209-226): coefficients of the
// corrector polynomial from the past time offsets t
void l_calc(int q, double dt, const double* t, double* l)
{
    l[0] = 1.0;
    for (int i = 1; i < q + 1; ++i) l[i] = 0.0;
    for (int i = 0; i < q; ++i) {
        const double k0 = 1.0 / (dt - t[i]);
        for (int j = i + 1; j > 0; --j) l[j] += k0 * l[j - 1];
    }
}

class Bdf {
public:
    Bdf(DMDSystem& d, const double* r_crd, const double* zeta, const BdfParams& p)
        : d_(d), r_crd_(r_crd), zeta_(zeta), p_(p), nc_((size_t)d.nsites() * d.c_dim())
    {
        for (int i = 0; i <= kMaxQ; ++i) A_[0][i] = A_[i][i] = 1.0;  // binomial table (src/dae_bdf.cpp:17-31)
        for (int i = 2; i <= kMaxQ; ++i)
            for (int j = 1; j < i; ++j) A_[j][i] = A_[j][i - 1] + A_[j - 1][i - 1];
        c_.resize(nc_);
        c0_.resize(nc_);
        cd_.resize(nc_);
        jc_.resize(nc_);
        y0_.resize(nc_);
        a_.resize(nc_);
        dy_.assign(nc_, 0.0);
        z_.assign(nc_ * (kMaxQ + 1), 0.0);
        a_tol_sqrt_ = p.a_tol * std::sqrt((double)nc_);  // nc_dof = natms c_dim (src/ff_dmd.cpp:75)
    }

    BdfReport run(double t_tot, std::vector<double>* tally);

private:
    double* z(size_t i) { return z_.data() + i * (kMaxQ + 1); }
    void pull_c() { d_.get_state(nullptr, nullptr, c_.data()); }

    // DAE::calc_err without box dofs (src/dae.cpp:151-164): the gradient per
    // unit c that c_d_calc leaves (ForceFieldDMD::norm_sq of calc_mu's f and
    // f_alpha) at the state of the last c_d call
    double calc_err()
    {
        const double n = 3.0 * d_.nsites() + (double)nc_;
        return std::sqrt(d_.unit_grad_err_sq()) / std::sqrt(n);
    }

    void min_error()
    {
        // DAE::min_error_false (src/dae.cpp:273-330): MinCGDMD over x and alpha
        // at fixed c with e_tol 1e-9, max_dx 1, max_dalpha 0.1 and
        // LineSearchBrent's defaults (DAE::init, src/dae.cpp:115-118), then the
        // Newton stage; counted when the Newton stage took a step
        LineSearchParams brent;
        brent.kind = 2;
        d_.cg(p_.relax_max_iter, 1e-9, 1.0, 0.1, 0.0, &brent);
        double res = 0.0;
        if (d_.newton_xa(p_.a_tol, p_.max_nnewton_iters, p_.max_ngmres_iters, &res) > 0) ++rep_.err_mins;
    }

    static double get_max_dt(double x, double x_d)
    {
        if (x_d > 0.0) {
            double ans = (1.0 - x) / x_d;
            while (x + x_d * ans > 1.0) ans = std::nextafter(ans, -1.0);
            return ans;
        }
        if (x_d < 0.0) {
            double ans = -x / x_d;
            while (x + x_d * ans < 0.0) ans = std::nextafter(ans, -1.0);
            return ans;
        }
        return std::numeric_limits<double>::infinity();
    }

    // DAEBDF::reset (src/dae_bdf.cpp:560-604)
    void reset()
    {
        d_.c_d(r_crd_, zeta_, cd_.data());
        d_.J(cd_.data(), jc_.data());  // ForceFieldDMD::c_dd_norm
        double s = 0.0;
        for (size_t i = 0; i < nc_; ++i) s += jc_[i] * jc_[i];
        const double norm = std::sqrt(s) / a_tol_sqrt_;
        pull_c();
        double max_dt = std::numeric_limits<double>::infinity();
        std::fill(z_.begin(), z_.end(), 0.0);
        for (size_t i = 0; i < nc_; ++i)
            if (c_[i] >= 0.0) {
                max_dt = std::min(max_dt, get_max_dt(c_[i], cd_[i]));
                z(i)[0] = c_[i];
                z(i)[1] = cd_[i];
            }
        q_ = 1;
        dt_ = std::max(std::min(std::min(std::sqrt(2.0 / norm), (t_fin_ - t_cur_) * 0.001), max_dt), p_.min_dt);
        for (double& v : t_) v = 0.0;
        std::fill(dy_.begin(), dy_.end(), 0.0);
    }

    // VC_y::err_fac_calc (src/dae_bdf.h:505-543)
    double err_fac_calc()
    {
        const int q = q_;
        const double dt = dt_;
        const double* t = t_;
        double beta_inv = 0.0;
        for (int i = 0; i < q; ++i) beta_inv += 1.0 / (dt - t[i]);
        beta_ = 1.0 / beta_inv;
        double s_q = 0.0;
        for (int i = 0; i < q; ++i) s_q += 1.0 / (dt - t[i]);
        const double err_fac = std::fabs(1.0 / (s_q * (dt - t[q])));
        if (q < kMaxQ) {
            double r = -dt / t[q + 1];
            for (int i = 1; i < q + 1; ++i) r *= (t[i] - dt) / t[i];
            const double s_qq = s_q + 1.0 / (dt - t[q]);
            hi_[0] = std::fabs(1.0 / (dt * (q + 2.0) * s_qq));
            hi_[1] = -r;
        }
        if (q > 1) {
            double p_q = 1.0;
            for (int i = 0; i < q; ++i) p_q *= dt - t[i];
            const double s_q_1 = s_q - 1.0 / (dt - t[q - 1]);
            lo_[0] = std::fabs(1.0 / ((dt - t[q]) * s_q_1));
            lo_[1] = p_q;
        }
        return err_fac;
    }

    // DAEBDF::interpolate (src/dae_bdf.cpp:322-366): the predictor y_0 and
    // the history term a of the corrector equation
    bool interpolate()
    {
        const int q = q_;
        for (int i = 0; i < q + 1; ++i) {
            double s = 1.0;
            for (int j = i; j < q + 1; ++j) {
                Ab_[i][j] = A_[i][j] * s;
                s *= dt_;
            }
        }
        err_fac_ = err_fac_calc();
        const double beta_inv = 1.0 / beta_;
        for (size_t i = 0; i < nc_; ++i) {
            const double* zi = z(i);
            double y0 = 0.0;
            for (int k = 0; k < q + 1; ++k) y0 += Ab_[0][k] * zi[k];
            if (y0 < 0.0 || y0 > 1.0) {
                ++rep_.intpol_rej;
                return false;
            }
            double s = 0.0;
            for (int k = 1; k < q + 1; ++k) s += Ab_[1][k] * zi[k];
            y0_[i] = y0;
            a_[i] = s - y0 * beta_inv;
        }
        ++rep_.intpol_acc;
        return true;
    }

    bool newton()
    {
        int it = 0;
        const bool ok = d_.newton(r_crd_, zeta_, beta_, a_.data(), p_.a_tol, err_fac_, p_.max_nnewton_iters,
                                  p_.max_ngmres_iters, &it);
        if (ok) {
            if (it) ++rep_.nonlin_acc;
        } else {
            ++rep_.nonlin_rej;
        }
        return ok;
    }

    // newton_fail / integrate_fail (src/dae_bdf.cpp:612-700)
    void fail_step(double r)
    {
        if (t_fin_ - t_cur_ <= 2.0 * p_.min_dt || dt_ == p_.min_dt) {
            if (q_ > 1) --q_;
        } else if (r * dt_ < p_.min_dt) {
            dt_ = p_.min_dt;
        } else if (r * dt_ > t_fin_ - t_cur_ - p_.min_dt) {
            dt_ = t_fin_ - t_cur_ - p_.min_dt;
        } else {
            dt_ *= r;
        }
    }

    // DAEBDF::integrate (src/dae_bdf.cpp:254-301)
    bool integrate()
    {
        pull_c();
        c0_ = c_;
        while (true) {
            if (!interpolate()) {
                reset();  // interpolate_fail
                continue;
            }
            if (!newton()) {  // the step left c at c0
                fail_step(0.25);
                continue;
            }
            break;
        }
        pull_c();
        double s = 0.0;
        for (size_t i = 0; i < nc_; ++i) s += (c_[i] - y0_[i]) * (c_[i] - y0_[i]);
        err_ = err_fac_ * std::sqrt(s) / a_tol_sqrt_;
        if (err_ < 1.0) {
            ++rep_.integ_acc;
            return true;
        }
        d_.set_c(c0_.data());
        ++rep_.integ_rej;
        return false;
    }

    // DAEBDF::update_z (src/dae_bdf.cpp:469-510) with DAEBDFMath::update_z
    // and update_z_inc (src/dae_bdf.h:137-159)
    void update_z()
    {
        const int q = q_, dq = dq_;
        if (dq == 1) {
            l_calc(q + 1, dt_, t_, l_);
        } else if (dq == -1) {
            l_calc(q - 1, dt_, t_, l_);
            double p = 1.0;
            for (int i = 0; i < q - 1; ++i) p *= dt_ - t_[i];
            for (int i = 1; i < q; ++i) Ab_[i][q] -= p * l_[i - 1];
        } else {
            l_calc(q, dt_, t_, l_);
        }
        const int dim = dq == 1 ? q : q + dq;
        for (size_t i = 0; i < nc_; ++i) {
            double* zi = z(i);
            if (dq == 1) zi[dim + 1] = l_[dim + 1] * dy_[i];
            for (int j = 0; j < dim + 1; ++j) {
                double s = 0.0;
                for (int k = j; k < dim + 1; ++k) s += Ab_[j][k] * zi[k];
                zi[j] = s + dy_[i] * l_[j];
            }
        }
    }

    // DAEBDF::prep_for_next (src/dae_bdf.cpp:370-465): the next order and step
    void prep_for_next()
    {
        const int q = q_;
        pull_c();
        double eta[3] = {0.0, std::pow(0.5 / err_, 1.0 / (q + 1.0)), 0.0};
        if (q < kMaxQ) {
            double s = 0.0;
            for (size_t i = 0; i < nc_; ++i)
                if (c_[i] >= 0.0) {
                    const double v = c_[i] - y0_[i] + hi_[1] * dy_[i];
                    s += v * v;
                }
            const double norm = hi_[0] * std::sqrt(s) / a_tol_sqrt_;
            eta[2] = std::pow(0.5 / norm, 1.0 / (q + 2.0));
        }
        if (q > 1) {
            double s = 0.0;
            for (size_t i = 0; i < nc_; ++i)
                if (c_[i] >= 0.0) {
                    const double v = c_[i] - y0_[i] + lo_[1] * z(i)[q];
                    s += v * v;
                }
            const double norm = lo_[0] * std::sqrt(s) / a_tol_sqrt_;
            eta[0] = std::pow(0.5 / norm, 1.0 / q);
        }
        for (size_t i = 0; i < nc_; ++i) dy_[i] = c_[i] - y0_[i];
        double dt_r = 1.0;
        dq_ = 0;
        if (nconst_q_ > 2 + q && nconst_dt_ > 2 + q) {
            if (eta[0] > eta[1] && eta[0] > eta[2]) {
                dq_ = -1;
                dt_r = eta[0];
            } else if (eta[2] > eta[0] && eta[2] > eta[1]) {
                dq_ = 1;
                dt_r = eta[2];
            } else {
                dt_r = eta[1];
            }
        }
        if (2.0 < dt_r)
            dt_r = 2.0;
        else if (dt_r <= 0.5)
            dt_r = 0.5;
        const double left = t_fin_ - t_cur_;
        double new_dt;
        if (dt_r * dt_ > left)
            new_dt = left;
        else if (left <= 2.0 * p_.min_dt)
            new_dt = 2.0 * p_.min_dt;
        else if (dt_r * dt_ < p_.min_dt)
            new_dt = p_.min_dt;
        else if (dt_r * dt_ >= left - p_.min_dt)
            new_dt = left - p_.min_dt;
        else
            new_dt = dt_r * dt_;
        update_z();
        for (int i = kMaxQ; i > 0; --i) t_[i] = t_[i - 1] - dt_;
        dt_ = new_dt;
        q_ += dq_;
    }

    DMDSystem& d_;
    const double *r_crd_, *zeta_;
    BdfParams p_;
    size_t nc_;
    double A_[kMaxQ + 1][kMaxQ + 1] = {}, Ab_[kMaxQ + 1][kMaxQ + 1] = {};
    double t_[kMaxQ + 1] = {}, l_[kMaxQ + 2] = {}, lo_[2] = {}, hi_[2] = {};
    std::vector<double> c_, c0_, cd_, jc_, y0_, a_, dy_, z_;
    double a_tol_sqrt_ = 0.0, t_cur_ = 0.0, t_fin_ = 0.0, dt_ = 0.0, beta_ = 0.0, err_ = 0.0, err_fac_ = 0.0;
    int q_ = 1, dq_ = 0, nconst_q_ = 0, nconst_dt_ = 0;
    BdfReport rep_;
};

BdfReport Bdf::run(double t_tot, std::vector<double>* tally)
{
    t_cur_ = 0.0;
    t_fin_ = t_tot;
    if (p_.nreset) min_error();
    d_.c_d(r_crd_, zeta_, nullptr);
    double fe_err = calc_err(), prev_err = fe_err;
    reset();
    nconst_q_ = nconst_dt_ = 0;
    rep_.min_dt = std::numeric_limits<double>::infinity();
    if (tally && p_.ntally) {
        const DMDReport r = d_.derivative(false, nullptr);
        tally->insert(tally->end(), {0.0, t_cur_, r.fe});
    }
    int istep = 0;
    while (istep < p_.max_nsteps && t_cur_ < t_fin_) {
        double dt_prev = dt_;
        int q_prev = q_;
        while (!integrate()) fail_step(std::pow(0.5 / err_, 1.0 / (double)(q_ + 1)));
        rep_.max_q = std::max(rep_.max_q, q_);
        rep_.max_dt = std::max(rep_.max_dt, dt_);
        rep_.min_dt = std::min(rep_.min_dt, dt_);
        nconst_q_ = q_prev != q_ ? 1 : nconst_q_ + 1;
        nconst_dt_ = dt_prev != dt_ ? 1 : nconst_dt_ + 1;
        dt_prev = dt_;
        q_prev = q_;
        t_cur_ += dt_prev;
        prep_for_next();
        if (dt_prev != dt_) nconst_dt_ = 0;
        if (q_prev != q_) nconst_q_ = 0;
        fe_err = calc_err();
        const bool periodic = p_.nreset && (istep + 1) % p_.nreset == 0 && fe_err / p_.a_tol > 1.0;
        if (p_.nreset && (periodic || fe_err > 2.0 * prev_err)) {
            min_error();
            d_.c_d(r_crd_, zeta_, nullptr);
            fe_err = calc_err();
            prev_err = fe_err;
            reset();
            nconst_q_ = nconst_dt_ = 0;
        }
        ++istep;
        if (tally && p_.ntally && istep % p_.ntally == 0) {
            const DMDReport r = d_.derivative(false, nullptr);
            tally->insert(tally->end(), {(double)istep, t_cur_, r.fe});
        }
    }
    // the last step's line and relaxation (src/dae_bdf.cpp:203-217)
    if (tally && p_.ntally && istep % p_.ntally) {
        const DMDReport r = d_.derivative(false, nullptr);
        tally->insert(tally->end(), {(double)istep, t_cur_, r.fe});
    }
    if (p_.nreset && istep % p_.nreset) min_error();
    rep_.nsteps = istep;
    rep_.t = t_cur_;
    return rep_;
}

}  // namespace

BdfReport bdf_run(DMDSystem& d, const double* r_crd, const double* zeta, double t_tot, const BdfParams& p,
                  std::vector<double>* tally)
{
    MAPP_REQUIRE(t_tot > 0.0 && p.a_tol > 0.0 && p.min_dt > 0.0, ERR_ARG, "t_tot, a_tol and min_dt must be positive");
    MAPP_REQUIRE(p.max_nsteps >= 0 && p.nreset >= 0 && p.ntally >= 0, ERR_ARG, "counts must be >= 0");
    Bdf b(d, r_crd, zeta, p);
    return b.run(t_tot, tally);
}

}  // namespace mapp
