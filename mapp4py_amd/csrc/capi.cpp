// capi.cpp -- extern "C" boundary: status codes, thread-local error text,
// opaque handles over the C++ classes.
#include <cstring>
#include <memory>
#include <string>

#include "../../include/mapp_amd.h"
#include "bdf.hpp"
#include "cfg.hpp"
#include "dmd.hpp"
#include "gcmc.hpp"
#include "replica.hpp"
#include "surrogate.hpp"
#include "md.hpp"
#include "linesearch.hpp"
#include "rng.hpp"

namespace mapp {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace mapp

struct mapp_eam {
    std::shared_ptr<mapp::EamForceField> ff;
};
struct mapp_md {
    std::unique_ptr<mapp::MDSystem> sys;
};
struct mapp_dmd {
    std::unique_ptr<mapp::DMDSystem> sys;
};
struct mapp_surr {
    std::shared_ptr<mapp::Surrogate> sys;
};
struct mapp_cfg {
    mapp::CfgDMD data;
};
struct mapp_cfg_md {
    mapp::CfgMD data;
};
struct mapp_gcmc {
    std::unique_ptr<mapp::GcmcBatch> sys;
};

struct mapp_replica {
    std::unique_ptr<mapp::ReplicaBatch> sys;
};

namespace {

template <class F>
mapp_status guard(F&& f)
{
    try {
        mapp::set_last_error("");
        f();
        return MAPP_OK;
    } catch (const mapp::Error& e) {
        mapp::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        mapp::set_last_error("host out of memory");
        return MAPP_ERR_CAPACITY;
    } catch (const std::exception& e) {
        mapp::set_last_error(e.what());
        return MAPP_ERR_STATE;
    }
}

std::vector<std::string> names(const char* const* elems, int nelems)
{
    std::vector<std::string> v;
    for (int i = 0; elems && i < nelems; ++i) v.emplace_back(elems[i]);
    return v;
}

#define NEED(p) MAPP_REQUIRE((p) != nullptr, mapp::ERR_ARG, #p " must not be NULL")

}  // namespace

extern "C" {

const char* mapp_last_error(void) { return mapp::g_last_error.c_str(); }

const char* mapp_version(void) { return "mapp4py_amd 0.1 (gfx950)"; }

mapp_status mapp_eam_create_setfl(const char* path, const char* const* elems, int nelems, int dmd, mapp_eam** out)
{
    return guard([&] {
        NEED(path);
        NEED(out);
        auto t = mapp::read_dynamo(path, 0, names(elems, nelems), dmd != 0);
        *out = new mapp_eam{std::make_shared<mapp::EamForceField>(std::move(t))};
    });
}

mapp_status mapp_eam_create_fs(const char* path, const char* const* elems, int nelems, int dmd, mapp_eam** out)
{
    return guard([&] {
        NEED(path);
        NEED(out);
        auto t = mapp::read_dynamo(path, 1, names(elems, nelems), dmd != 0);
        *out = new mapp_eam{std::make_shared<mapp::EamForceField>(std::move(t))};
    });
}

mapp_status mapp_eam_create_funcfl(const char* const* paths, int nfiles, int dmd, mapp_eam** out)
{
    return guard([&] {
        NEED(paths);
        NEED(out);
        std::vector<std::string> p;
        for (int i = 0; i < nfiles; ++i) {
            NEED(paths[i]);
            p.emplace_back(paths[i]);
        }
        auto t = mapp::read_funcfl(p, dmd != 0);
        *out = new mapp_eam{std::make_shared<mapp::EamForceField>(std::move(t))};
    });
}

mapp_status mapp_eam_create_knots(int ne, int nr, int nrho, double dr, double drho, const double* F,
                                  const double* rho, const double* r_phi, int fs, int dmd, mapp_eam** out)
{
    return guard([&] {
        NEED(F);
        NEED(rho);
        NEED(r_phi);
        NEED(out);
        auto t = mapp::tables_from_knots(ne, nr, nrho, dr, drho, F, rho, r_phi, fs == 0, dmd != 0);
        *out = new mapp_eam{std::make_shared<mapp::EamForceField>(std::move(t))};
    });
}

void mapp_eam_destroy(mapp_eam* ff) { delete ff; }

mapp_status mapp_eam_cutoff(const mapp_eam* ff, int* ne, double* rc)
{
    return guard([&] {
        NEED(ff);
        const auto& t = ff->ff->tables();
        if (ne) *ne = t.ne;
        if (rc) std::memcpy(rc, t.rc.data(), sizeof(double) * t.rc.size());
    });
}

mapp_status mapp_eam_coefs(const mapp_eam* ff, int which, int a, int b, double* out)
{
    return guard([&] {
        NEED(ff);
        NEED(out);
        const auto& t = ff->ff->tables();
        MAPP_REQUIRE(a >= 0 && a < t.ne && b >= 0 && b < t.ne, mapp::ERR_ARG, "element index out of range");
        if (which == 0) std::memcpy(out, t.rphi_at(a, b), sizeof(double) * 4 * t.nr);
        else if (which == 1) std::memcpy(out, t.rho_at(a, b), sizeof(double) * 4 * t.nr);
        else if (which == 2) std::memcpy(out, t.F_at(a), sizeof(double) * t.nf() * t.nrho);
        else throw mapp::Error(mapp::ERR_ARG, "which must be 0, 1 or 2");
    });
}

mapp_status mapp_md_create(mapp_eam* ff, int natms, const double* H, const double* x, const uint8_t* elem,
                           double skin, mapp_md** out)
{
    return guard([&] {
        NEED(ff);
        NEED(H);
        NEED(x);
        NEED(elem);
        NEED(out);
        *out = new mapp_md{std::make_unique<mapp::MDSystem>(ff->ff, natms, H, x, elem, skin)};
    });
}

void mapp_md_destroy(mapp_md* md) { delete md; }

mapp_status mapp_md_set_x(mapp_md* md, const double* x)
{
    return guard([&] { NEED(md); NEED(x); md->sys->set_x(x); });
}
mapp_status mapp_md_get_x(mapp_md* md, double* x)
{
    return guard([&] { NEED(md); NEED(x); md->sys->get_x(x); });
}
mapp_status mapp_md_set_v(mapp_md* md, const double* v)
{
    return guard([&] { NEED(md); NEED(v); md->sys->set_v(v); });
}
mapp_status mapp_md_get_v(mapp_md* md, double* v)
{
    return guard([&] { NEED(md); NEED(v); md->sys->get_v(v); });
}

mapp_status mapp_md_set_box(mapp_md* md, const double* H, const double* x)
{
    return guard([&] {
        NEED(md);
        NEED(H);
        md->sys->set_box(H, x);
    });
}

mapp_status mapp_md_update(mapp_md* md, int* rebuilt)
{
    return guard([&] {
        NEED(md);
        const bool r = md->sys->update();
        if (rebuilt) *rebuilt = r ? 1 : 0;
    });
}

mapp_status mapp_md_force_calc(mapp_md* md, double* f, double* out7)
{
    return guard([&] {
        NEED(md);
        const mapp::MDReport r = md->sys->force_calc();
        if (out7) std::memcpy(out7, r.e, sizeof r.e);
        if (f) md->sys->get_f(f);
    });
}

mapp_status mapp_md_value(mapp_md* md, double* energy)
{
    return guard([&] {
        NEED(md);
        NEED(energy);
        *energy = md->sys->value();
    });
}

mapp_status mapp_md_value_noalloc(mapp_md* md, double* energy)
{
    return guard([&] {
        NEED(md);
        NEED(energy);
        *energy = md->sys->value_noalloc();
    });
}

mapp_status mapp_md_get_rho(mapp_md* md, double* rho)
{
    return guard([&] { NEED(md); NEED(rho); md->sys->get_rho(rho); });
}

mapp_status mapp_md_pairs_count(mapp_md* md, int64_t* npairs)
{
    return guard([&] { NEED(md); NEED(npairs); *npairs = md->sys->half_pairs_count(); });
}

mapp_status mapp_md_pairs(mapp_md* md, int32_t* i, int32_t* j, int8_t* img)
{
    return guard([&] { NEED(md); NEED(i); NEED(j); NEED(img); md->sys->half_pairs(i, j, img); });
}

mapp_status mapp_md_list_stats(mapp_md* md, int* max_neighbors, int64_t* full_entries, int* rebuilds)
{
    return guard([&] {
        NEED(md);
        int m = 0;
        int64_t s = 0;
        md->sys->list_stats(&m, &s);
        if (max_neighbors) *max_neighbors = m;
        if (full_entries) *full_entries = s;
        if (rebuilds) *rebuilds = md->sys->last_rebuilds();
    });
}

mapp_status mapp_md_nvt_run(mapp_md* md, const double* masses, double kB, double T, double dt, int nsteps,
                            double* out7, double* T_part, int* rebuilds)
{
    return guard([&] {
        NEED(md);
        NEED(masses);
        mapp::MDReport r{};
        const int nr = md->sys->nvt_run(masses, kB, T, dt, nsteps, &r, T_part);
        if (out7) std::memcpy(out7, r.e, sizeof r.e);
        if (rebuilds) *rebuilds = nr;
    });
}

mapp_status mapp_md_nst_run(mapp_md* md, const double* masses, double kB, double T, double dt, const double* S,
                            const double* tau, int nreset, int nsteps, double* out7, double* T_part, double* H,
                            double* hist, int* rebuilds)
{
    return guard([&] {
        NEED(md);
        NEED(masses);
        NEED(S);
        mapp::MDReport r{};
        const int nr = md->sys->nst_run(masses, kB, T, dt, S, tau, nreset, nsteps, &r, T_part, H, hist);
        if (out7) std::memcpy(out7, r.e, sizeof r.e);
        if (rebuilds) *rebuilds = nr;
    });
}

mapp_status mapp_md_muvt_run(mapp_md* md, const double* masses, double kB, double hP, double T, double dt, double mu,
                             int gas_elem, int seed, int nevery, int nattempts, int nsteps, double* out7,
                             double* T_part, int* counts, double* hist, int* builds)
{
    return guard([&] {
        NEED(md);
        NEED(masses);
        mapp::MDReport r{};
        const int nb = md->sys->muvt_run(masses, kB, hP, T, dt, mu, gas_elem, seed, nevery, nattempts, nsteps, &r,
                                         T_part, counts, hist);
        if (out7) std::memcpy(out7, r.e, sizeof r.e);
        if (builds) *builds = nb;
    });
}

mapp_status mapp_md_natms(mapp_md* md, int* natms)
{
    return guard([&] {
        NEED(md);
        NEED(natms);
        *natms = md->sys->natms();
    });
}

mapp_status mapp_md_get_elem(mapp_md* md, uint8_t* elem)
{
    return guard([&] {
        NEED(md);
        NEED(elem);
        md->sys->get_elem(elem);
    });
}

mapp_status mapp_md_minimize(mapp_md* md, const mapp_min_params* params, mapp_min_report* report, double* tally,
                             int max_tally_rows, int* tally_rows)
{
    return guard([&] {
        NEED(md);
        NEED(params);
        MAPP_REQUIRE(max_tally_rows >= 0 && (tally || max_tally_rows == 0), mapp::ERR_ARG, "invalid tally buffer");
        mapp::MinParams p;
        p.lbfgs = params->lbfgs;
        p.m = params->m;
        p.e_tol = params->e_tol;
        p.max_dx = params->max_dx;
        p.max_nsteps = params->max_nsteps;
        p.ls_c = params->ls_c;
        p.ls_rho = params->ls_rho;
        p.ls_min_alpha = params->ls_min_alpha;
        p.ls = params->ls;
        p.ls_max_iter = params->ls_max_iter;
        p.ls_brack = params->ls_brack;
        p.ls_tol = params->ls_tol;
        p.ls_zeps = params->ls_zeps;
        std::vector<double> rows;
        const mapp::MinReport r = md->sys->minimize(p, tally ? &rows : nullptr);
        if (report) {
            report->steps = r.steps;
            report->status = r.status;
            report->evals = r.evals;
            report->force_calcs = r.force_calcs;
            report->rebuilds = r.rebuilds;
            report->e0 = r.e0;
            report->e = r.e;
            report->f_norm = r.f_norm;
            report->f_max = r.f_max;
        }
        const int nrows = std::min((int)(rows.size() / 2), max_tally_rows);
        if (tally && nrows) std::memcpy(tally, rows.data(), sizeof(double) * 2 * nrows);
        if (tally_rows) *tally_rows = nrows;
    });
}

// the line-search fields of a mapp_min_params, checked: every search ends
// for parameters that pass (mapp::line_search_params_error)
static mapp::LineSearchParams line_search_params(const mapp_min_params& q)
{
    mapp::LineSearchParams lp;
    lp.kind = q.ls;
    lp.c = q.ls_c;
    lp.rho = q.ls_rho;
    lp.min_alpha = q.ls_min_alpha;
    lp.max_iter = q.ls_max_iter;
    lp.brack = q.ls_brack;
    lp.tol = q.ls_tol;
    lp.zeps = q.ls_zeps;
    const char* err = mapp::line_search_params_error(lp);
    MAPP_REQUIRE(err == nullptr, mapp::ERR_ARG, err ? err : "");
    return lp;
}

mapp_status mapp_line_search(const mapp_min_params* params, double dfa, double max_a, int init_flag,
                             double* prev_val, mapp_line_energy_fn energy, mapp_line_reset_fn reset, void* ctx,
                             double* energy_io, double* alpha, int* status)
{
    return guard([&] {
        NEED(params);
        NEED(prev_val);
        NEED(energy);
        NEED(reset);
        NEED(energy_io);
        NEED(alpha);
        NEED(status);
        const mapp::LineSearchParams lp = line_search_params(*params);
        auto f = [&](double t) { return energy(ctx, t); };
        auto r = [&]() { reset(ctx); };
        mapp::LineSearcher<decltype(f), decltype(r)> ls(lp, f, r);
        ls.prev_val = *prev_val;
        double e = *energy_io, a = 0.0;
        *status = ls.min(e, a, dfa, max_a, init_flag);
        *energy_io = e;
        *alpha = a;
        *prev_val = ls.prev_val;
    });
}

mapp_status mapp_md_profile(mapp_md* md, int enable)
{
    return guard([&] {
        NEED(md);
        auto& p = md->sys->profiler();
        p.reset();
        p.on = enable != 0;
    });
}

mapp_status mapp_md_profile_read(mapp_md* md, const char* kernel, double* total_ms, int64_t* launches)
{
    return guard([&] {
        NEED(md);
        NEED(kernel);
        auto& p = md->sys->profiler();
        p.flush();
        for (int k = 0; k < mapp::K_COUNT; ++k)
            if (std::strcmp(kernel, mapp::kernel_name(k)) == 0) {
                if (total_ms) *total_ms = p.total_ms[k];
                if (launches) *launches = p.launches[k];
                return;
            }
        throw mapp::Error(mapp::ERR_ARG, std::string("unknown kernel group ") + kernel);
    });
}

mapp_status mapp_md_set_batch(mapp_md* md, int rho_batch, int force_batch)
{
    return guard([&] { NEED(md); md->sys->set_batch(rho_batch, force_batch); });
}

mapp_status mapp_md_lds_window(mapp_md* md, int* rho_first_row, int* force_first_row)
{
    return guard([&] {
        NEED(md);
        NEED(rho_first_row);
        NEED(force_first_row);
        md->sys->lds_window(rho_first_row, force_first_row);
    });
}

mapp_status mapp_md_set_lds(mapp_md* md, int on, int* in_use)
{
    return guard([&] {
        NEED(md);
        const bool used = md->sys->set_lds(on != 0);
        if (in_use) *in_use = used ? 1 : 0;
    });
}

mapp_status mapp_dmd_create(mapp_eam* ff, int nsites, int c_dim, int nquad, const double* H, const double* x,
                            const uint8_t* elem, const double* alpha, const double* c, double skin, double kB,
                            double T, double hP, const double* masses, mapp_dmd** out)
{
    return guard([&] {
        NEED(ff); NEED(H); NEED(x); NEED(elem); NEED(alpha); NEED(c); NEED(masses); NEED(out);
        *out = new mapp_dmd{std::make_unique<mapp::DMDSystem>(ff->ff, nsites, c_dim, nquad, H, x, elem, alpha, c,
                                                              skin, kB, T, hP, masses)};
    });
}

void mapp_dmd_destroy(mapp_dmd* d) { delete d; }

mapp_status mapp_dmd_set_state(mapp_dmd* d, const double* x, const double* alpha, const double* c)
{
    return guard([&] { NEED(d); NEED(x); NEED(alpha); NEED(c); d->sys->set_state(x, alpha, c); });
}

mapp_status mapp_dmd_get_state(mapp_dmd* d, double* x, double* alpha, double* c)
{
    return guard([&] { NEED(d); d->sys->get_state(x, alpha, c); });
}

mapp_status mapp_dmd_update(mapp_dmd* d, int* rebuilt)
{
    return guard([&] {
        NEED(d);
        const bool r = d->sys->update();
        if (rebuilt) *rebuilt = r ? 1 : 0;
    });
}

mapp_status mapp_dmd_derivative(mapp_dmd* d, int with_c, const double* ext_mu, double* f, double* f_alpha,
                                double* mu, double* out8)
{
    return guard([&] {
        NEED(d);
        const mapp::DMDReport r = d->sys->derivative(with_c != 0, ext_mu);
        if (out8) {
            out8[0] = r.fe;
            for (int k = 0; k < 6; ++k) out8[1 + k] = r.FH[k];
            out8[7] = r.s;
        }
        if (f || f_alpha || mu) d->sys->get_forces(f, f_alpha, mu);
    });
}

mapp_status mapp_dmd_set_box(mapp_dmd* d, const double* H, const double* x)
{
    return guard([&] {
        NEED(d);
        NEED(H);
        d->sys->set_box(H, x);
    });
}

mapp_status mapp_dmd_force_calc_static(mapp_dmd* d, double* out8)
{
    return guard([&] {
        NEED(d);
        NEED(out8);
        const mapp::DMDReport r = d->sys->force_calc_static();
        out8[0] = r.fe;
        for (int k = 0; k < 6; ++k) out8[1 + k] = r.FH[k];
        out8[7] = r.s;
    });
}

mapp_status mapp_dmd_set_fp32(mapp_dmd* d, int mode)
{
    return guard([&] { NEED(d); d->sys->set_f32(mode); });
}

mapp_status mapp_dmd_knot_window(mapp_dmd* d, int* first_knot)
{
    return guard([&] {
        NEED(d);
        NEED(first_knot);
        *first_knot = d->sys->knots_in_lds() ? d->sys->knot_window_first() : -1;
    });
}

mapp_status mapp_dmd_set_batch(mapp_dmd* d, int batch)
{
    return guard([&] { NEED(d); d->sys->set_batch(batch); });
}

mapp_status mapp_dmd_c_d(mapp_dmd* d, const double* r_crd, const double* C, double* c_d, double* c_d_norm)
{
    return guard([&] {
        NEED(d);
        NEED(r_crd);
        NEED(C);
        const double v = d->sys->c_d(r_crd, C, c_d);
        if (c_d_norm) *c_d_norm = v;
    });
}

mapp_status mapp_dmd_J(mapp_dmd* d, const double* x, double* Ax)
{
    return guard([&] {
        NEED(d);
        NEED(x);
        NEED(Ax);
        d->sys->J(x, Ax);
    });
}

mapp_status mapp_dmd_newton(mapp_dmd* d, const double* r_crd, const double* C, double beta, const double* a,
                            double a_tol, double err_fac, int max_nnewton_iters, int max_ngmres_iters,
                            int* converged, int* iters)
{
    return guard([&] {
        NEED(d);
        NEED(r_crd);
        NEED(C);
        NEED(a);
        const bool ok = d->sys->newton(r_crd, C, beta, a, a_tol, err_fac, max_nnewton_iters, max_ngmres_iters, iters);
        if (converged) *converged = ok ? 1 : 0;
    });
}

mapp_status mapp_dmd_Jxa(mapp_dmd* d, const double* dx, const double* dalpha, double* Jdx, double* Jdalpha,
                         double* f, double* f_alpha)
{
    return guard([&] {
        NEED(d);
        NEED(dx);
        NEED(dalpha);
        NEED(Jdx);
        NEED(Jdalpha);
        d->sys->Jxa(dx, dalpha, Jdx, Jdalpha, f, f_alpha);
    });
}

mapp_status mapp_dmd_min_newton(mapp_dmd* d, double a_tol, int max_nnewton_iters, int max_ngmres_iters, int* steps,
                                double* res)
{
    return guard([&] {
        NEED(d);
        double r = 0.0;
        const int k = d->sys->newton_xa(a_tol, max_nnewton_iters, max_ngmres_iters, &r);
        if (steps) *steps = k;
        if (res) *res = r;
    });
}

mapp_status mapp_dmd_bdf_run(mapp_dmd* d, const double* r_crd, const double* C, double t_tot,
                             const mapp_bdf_params* params, mapp_bdf_report* report, double* tally,
                             int max_tally_rows, int* tally_rows)
{
    return guard([&] {
        NEED(d);
        NEED(r_crd);
        NEED(C);
        NEED(params);
        mapp::BdfParams p;
        p.a_tol = params->a_tol;
        p.min_dt = params->min_dt;
        p.max_nsteps = params->max_nsteps;
        p.max_nnewton_iters = params->max_nnewton_iters;
        p.max_ngmres_iters = params->max_ngmres_iters;
        p.nreset = params->nreset;
        p.ntally = params->ntally;
        p.relax_max_iter = params->relax_max_iter;
        std::vector<double> rows;
        const mapp::BdfReport r = mapp::bdf_run(*d->sys, r_crd, C, t_tot, p, &rows);
        const int nrows = (int)(rows.size() / 3);
        if (tally) std::memcpy(tally, rows.data(), sizeof(double) * 3 * std::min(nrows, std::max(max_tally_rows, 0)));
        if (tally_rows) *tally_rows = tally ? std::min(nrows, std::max(max_tally_rows, 0)) : 0;
        if (report) {
            report->nsteps = r.nsteps;
            report->max_q = r.max_q;
            report->t = r.t;
            report->max_dt = r.max_dt;
            report->min_dt = r.min_dt;
            report->nonlin_acc = r.nonlin_acc;
            report->nonlin_rej = r.nonlin_rej;
            report->intpol_acc = r.intpol_acc;
            report->intpol_rej = r.intpol_rej;
            report->integ_acc = r.integ_acc;
            report->integ_rej = r.integ_rej;
            report->err_mins = r.err_mins;
        }
    });
}

namespace {
std::vector<std::string> name_list(int n, const char* const* names)
{
    std::vector<std::string> out;
    for (int i = 0; i < n; ++i) {
        MAPP_REQUIRE(names[i] != nullptr, mapp::ERR_ARG, "null name");
        out.emplace_back(names[i]);
    }
    return out;
}
}  // namespace

mapp_status mapp_cfg_md_read(const char* path, mapp_cfg_md** out)
{
    return guard([&] {
        NEED(path);
        NEED(out);
        *out = nullptr;
        auto h = std::make_unique<mapp_cfg_md>();
        h->data = mapp::read_cfg_md(path);
        *out = h.release();
    });
}

mapp_status mapp_cfg_md_info(const mapp_cfg_md* cfg, int* natms, int* nelems, int* has_vel)
{
    return guard([&] {
        NEED(cfg);
        if (natms) *natms = cfg->data.natms;
        if (nelems) *nelems = cfg->data.nelems;
        if (has_vel) *has_vel = cfg->data.has_vel ? 1 : 0;
    });
}

mapp_status mapp_cfg_md_get(const mapp_cfg_md* cfg, double* H, double* x, double* v, uint8_t* elem,
                            double* masses)
{
    return guard([&] {
        NEED(cfg);
        const mapp::CfgMD& d = cfg->data;
        if (H) std::memcpy(H, d.H, sizeof d.H);
        if (x) std::memcpy(x, d.x.data(), sizeof(double) * d.x.size());
        if (v && d.has_vel) std::memcpy(v, d.v.data(), sizeof(double) * d.v.size());
        if (elem) std::memcpy(elem, d.elem.data(), d.elem.size());
        if (masses) std::memcpy(masses, d.masses.data(), sizeof(double) * d.masses.size());
    });
}

const char* mapp_cfg_md_element(const mapp_cfg_md* cfg, int i)
{
    if (!cfg || i < 0 || i >= cfg->data.nelems) return nullptr;
    return cfg->data.names[i].c_str();
}

void mapp_cfg_md_free(mapp_cfg_md* cfg) { delete cfg; }

mapp_status mapp_cfg_md_write(const char* path, int natms, const double* H, const double* x, const uint8_t* elem,
                              const double* v, int naux, const double* aux, const char* const* aux_names,
                              int nelems, const double* masses, const char* const* names)
{
    return guard([&] {
        NEED(path);
        NEED(H);
        NEED(masses);
        NEED(names);
        MAPP_REQUIRE(natms >= 0 && naux >= 0, mapp::ERR_ARG, "natms >= 0 and naux >= 0");
        if (natms) {
            NEED(x);
            NEED(elem);
            if (naux) NEED(aux);
        }
        if (naux) NEED(aux_names);
        mapp::write_cfg_md(path, natms, H, x, elem, v, naux, aux, name_list(naux, aux_names), nelems, masses,
                           name_list(nelems, names));
    });
}

mapp_status mapp_md_export_cfg(mapp_md* md, const char* path, int nelems, const double* masses,
                               const char* const* names, int with_vel)
{
    return guard([&] {
        NEED(md);
        NEED(path);
        NEED(masses);
        NEED(names);
        md->sys->export_cfg(path, nelems, masses, name_list(nelems, names), with_vel != 0);
    });
}

mapp_status mapp_cfg_dmd_read(const char* path, mapp_cfg** out)
{
    return guard([&] {
        NEED(path);
        NEED(out);
        *out = nullptr;
        auto h = std::make_unique<mapp_cfg>();
        h->data = mapp::read_cfg_dmd(path);
        *out = h.release();
    });
}

mapp_status mapp_cfg_dmd_info(const mapp_cfg* cfg, int* natms, int* c_dim, int* nelems)
{
    return guard([&] {
        NEED(cfg);
        if (natms) *natms = cfg->data.natms;
        if (c_dim) *c_dim = cfg->data.c_dim;
        if (nelems) *nelems = cfg->data.nelems;
    });
}

mapp_status mapp_cfg_dmd_get(const mapp_cfg* cfg, double* H, double* x, uint8_t* elem, double* alpha, double* c,
                             double* masses)
{
    return guard([&] {
        NEED(cfg);
        const mapp::CfgDMD& d = cfg->data;
        if (H) std::memcpy(H, d.H, sizeof d.H);
        if (x) std::memcpy(x, d.x.data(), sizeof(double) * d.x.size());
        if (elem) std::memcpy(elem, d.elem.data(), d.elem.size());
        if (alpha) std::memcpy(alpha, d.alpha.data(), sizeof(double) * d.alpha.size());
        if (c) std::memcpy(c, d.c.data(), sizeof(double) * d.c.size());
        if (masses) std::memcpy(masses, d.masses.data(), sizeof(double) * d.masses.size());
    });
}

const char* mapp_cfg_dmd_element(const mapp_cfg* cfg, int i)
{
    if (!cfg || i < 0 || i >= cfg->data.nelems) return nullptr;
    return cfg->data.names[i].c_str();
}

void mapp_cfg_dmd_free(mapp_cfg* cfg) { delete cfg; }

mapp_status mapp_cfg_dmd_write(const char* path, int natms, int c_dim, const double* H, const double* x,
                               const uint8_t* elem, const double* alpha, const double* c, int nelems,
                               const double* masses, const char* const* names)
{
    return guard([&] {
        NEED(path);
        NEED(H);
        NEED(masses);
        NEED(names);
        MAPP_REQUIRE(natms >= 0 && c_dim >= 1 && c_dim <= 4, mapp::ERR_ARG, "natms >= 0 and 1 <= c_dim <= 4");
        if (natms) {
            NEED(x);
            NEED(elem);
            NEED(alpha);
            NEED(c);
        }
        for (int i = 0; i < (natms * c_dim); ++i)
            MAPP_REQUIRE(elem[i] < nelems, mapp::ERR_ARG, "element index out of range");
        std::vector<std::string> nm;
        for (int i = 0; i < nelems; ++i) {
            NEED(names[i]);
            nm.emplace_back(names[i]);
        }
        mapp::write_cfg_dmd(path, natms, c_dim, H, x, elem, alpha, c, nelems, masses, nm);
    });
}

mapp_status mapp_dmd_export_cfg(mapp_dmd* d, const char* path, int nelems, const double* masses,
                                const char* const* names)
{
    return guard([&] {
        NEED(d);
        NEED(path);
        NEED(masses);
        NEED(names);
        std::vector<std::string> nm;
        for (int i = 0; i < nelems; ++i) {
            NEED(names[i]);
            nm.emplace_back(names[i]);
        }
        d->sys->export_cfg(path, nelems, masses, nm);
    });
}

mapp_status mapp_dmd_unit_gradient(mapp_dmd* d, double* g, double* g_alpha)
{
    return guard([&] {
        NEED(d);
        d->sys->unit_gradient(g, g_alpha);
    });
}

mapp_status mapp_dmd_J_entries(mapp_dmd* d, int64_t* live, int64_t* exchange)
{
    return guard([&] {
        NEED(d);
        NEED(live);
        NEED(exchange);
        d->sys->J_entries(live, exchange);
    });
}

mapp_status mapp_dmd_value(mapp_dmd* d, int with_c, const double* ext_mu, double* fe)
{
    return guard([&] { NEED(d); NEED(fe); *fe = d->sys->value(with_c != 0, ext_mu); });
}

mapp_status mapp_dmd_pairs_count(mapp_dmd* d, int64_t* npairs)
{
    return guard([&] { NEED(d); NEED(npairs); *npairs = d->sys->half_pairs_count(); });
}

mapp_status mapp_dmd_pairs(mapp_dmd* d, int32_t* i, int32_t* j, int8_t* img)
{
    return guard([&] { NEED(d); NEED(i); NEED(j); NEED(img); d->sys->half_pairs(i, j, img); });
}

mapp_status mapp_dmd_smear(mapp_dmd* d, int ei, int ej, int npts, const double* r, const double* alpha,
                           double* out)
{
    return guard([&] { NEED(d); NEED(r); NEED(alpha); NEED(out); d->sys->smear(ei, ej, npts, r, alpha, out); });
}

mapp_status mapp_dmd_min_cg(mapp_dmd* d, int max_iter, double e_tol, double max_dx, double max_dalpha,
                            double grad_tol, int* iters, int* status, double* out4)
{
    return guard([&] {
        NEED(d);
        const mapp::CGReport r = d->sys->cg(max_iter, e_tol, max_dx, max_dalpha, grad_tol);
        if (iters) *iters = r.iters;
        if (status) *status = r.status;
        if (out4) {
            out4[0] = r.fe0;
            out4[1] = r.fe;
            out4[2] = r.grad_norm;
            out4[3] = r.grad_max;
        }
    });
}

mapp_status mapp_dmd_min_lbfgs(mapp_dmd* d, int m, int max_iter, double e_tol, double max_dx, double max_dalpha,
                               double grad_tol, int* iters, int* status, double* out4)
{
    return guard([&] {
        NEED(d);
        const mapp::CGReport r = d->sys->lbfgs(m, max_iter, e_tol, max_dx, max_dalpha, grad_tol);
        if (iters) *iters = r.iters;
        if (status) *status = r.status;
        if (out4) {
            out4[0] = r.fe0;
            out4[1] = r.fe;
            out4[2] = r.grad_norm;
            out4[3] = r.grad_max;
        }
    });
}

mapp_status mapp_dmd_min_run(mapp_dmd* d, int lbfgs, int m, int max_iter, double e_tol, double max_dx,
                             double max_dalpha, double grad_tol, const mapp_min_params* ls, int* iters, int* status,
                             double* out4)
{
    return guard([&] {
        NEED(ls);
        const mapp::LineSearchParams lp = line_search_params(*ls);
        NEED(d);
        const mapp::CGReport r = lbfgs ? d->sys->lbfgs(m, max_iter, e_tol, max_dx, max_dalpha, grad_tol, &lp)
                                       : d->sys->cg(max_iter, e_tol, max_dx, max_dalpha, grad_tol, &lp);
        if (iters) *iters = r.iters;
        if (status) *status = r.status;
        if (out4) {
            out4[0] = r.fe0;
            out4[1] = r.fe;
            out4[2] = r.grad_norm;
            out4[3] = r.grad_max;
        }
    });
}

mapp_status mapp_dmd_profile(mapp_dmd* d, int enable)
{
    return guard([&] {
        NEED(d);
        auto& p = d->sys->profiler();
        p.reset();
        p.on = enable != 0;
    });
}

mapp_status mapp_dmd_profile_read(mapp_dmd* d, const char* kernel, double* total_ms, int64_t* launches)
{
    return guard([&] {
        NEED(d);
        NEED(kernel);
        auto& p = d->sys->profiler();
        p.flush();
        for (int k = 0; k < mapp::K_COUNT; ++k)
            if (std::strcmp(kernel, mapp::kernel_name(k)) == 0) {
                if (total_ms) *total_ms = p.total_ms[k];
                if (launches) *launches = p.launches[k];
                return;
            }
        throw mapp::Error(mapp::ERR_ARG, std::string("unknown kernel group ") + kernel);
    });
}

mapp_status mapp_gauss_hermite(int n, double* x, double* w)
{
    return guard([&] {
        NEED(x); NEED(w);
        MAPP_REQUIRE(n >= 1 && n <= mapp::kMaxQuad, mapp::ERR_ARG, "1 to 16 points");
        mapp::gauss_hermite(n, x, w);
    });
}

mapp_status mapp_create_T(int natms, const uint8_t* elem, const double* masses, double kB, double T, int seed,
                          double* v)
{
    return guard([&] {
        NEED(elem);
        NEED(masses);
        NEED(v);
        mapp::ParkMiller rng(seed);
        for (int i = 0; i < natms; ++i) {
            const double fac = std::sqrt(kB * T / masses[elem[i]]);
            for (int d = 0; d < 3; ++d) v[3 * (size_t)i + d] = rng.gaussian() * fac;
        }
    });
}


/* ---------------- batched grand-canonical exchange ---------------- */
mapp_status mapp_gcmc_create(mapp_eam* ff, int ncells, const int32_t* natms, const double* H, const double* x,
                             const uint8_t* elem, const int32_t* cap, int gas_elem, double mu, double T, double kB,
                             double hP, const double* masses, const int32_t* seeds, mapp_gcmc** out)
{
    return guard([&] {
        NEED(ff);
        NEED(natms);
        NEED(H);
        NEED(masses);
        NEED(seeds);
        NEED(out);
        *out = nullptr;
        int64_t tot = 0;
        for (int c = 0; c < ncells; ++c) tot += natms[c];
        MAPP_REQUIRE(tot == 0 || (x && elem), mapp::ERR_ARG, "x and elem must not be NULL");
        auto g = std::make_unique<mapp_gcmc>();
        g->sys = std::make_unique<mapp::GcmcBatch>(ff->ff, ncells, natms, H, x, elem, cap, gas_elem, mu, T, kB, hP,
                                                   masses, seeds);
        *out = g.release();
    });
}

void mapp_gcmc_destroy(mapp_gcmc* g) { delete g; }

mapp_status mapp_gcmc_attempts(mapp_gcmc* g, int nattempts, int8_t* decisions, double* dU)
{
    return guard([&] { NEED(g); g->sys->attempts(nattempts, decisions, dU); });
}

mapp_status mapp_gcmc_tune(mapp_gcmc* g, int nattempts, int32_t* waves, double* ms)
{
    return guard([&] {
        NEED(g);
        const mapp::GcmcTuneReport r = g->sys->tune(nattempts);
        for (int k = 0; k < mapp::kGcmcClasses; ++k)
            if (waves) waves[k] = r.waves[k];
        for (int k = 0; k <= mapp::kGcmcClasses; ++k)
            if (ms) ms[k] = r.ms[k];
    });
}

mapp_status mapp_gcmc_set_waves(mapp_gcmc* g, const int32_t* waves)
{
    return guard([&] {
        NEED(g);
        NEED(waves);
        int w[mapp::kGcmcClasses];
        for (int k = 0; k < mapp::kGcmcClasses; ++k) w[k] = waves[k];
        g->sys->set_waves(w);
    });
}

mapp_status mapp_gcmc_use_graph(mapp_gcmc* g, int on)
{
    return guard([&] { NEED(g); g->sys->use_graph(on != 0); });
}

mapp_status mapp_gcmc_xchng_energy(mapp_gcmc* g, int cell, int mode, const double* s, int igas, double* dU)
{
    return guard([&] {
        NEED(g);
        NEED(dU);
        *dU = g->sys->xchng_energy(cell, mode, s, igas);
    });
}

mapp_status mapp_gcmc_xchng_commit(mapp_gcmc* g, const double* v)
{
    return guard([&] {
        NEED(g);
        g->sys->xchng_commit(v);
    });
}

mapp_status mapp_gcmc_counts(mapp_gcmc* g, int32_t* natms, int32_t* ngas)
{
    return guard([&] { NEED(g); g->sys->counts(natms, ngas); });
}

mapp_status mapp_gcmc_get_cell(mapp_gcmc* g, int cell, double* x, uint8_t* elem, double* rho, double* F, double* v,
                               int32_t* rng_seed)
{
    return guard([&] { NEED(g); g->sys->get_cell(cell, x, elem, rho, F, v, rng_seed); });
}

mapp_status mapp_gcmc_profile(mapp_gcmc* g, int enable)
{
    return guard([&] {
        NEED(g);
        auto& p = g->sys->profiler();
        p.reset();
        p.on = enable != 0;
    });
}

mapp_status mapp_gcmc_profile_read(mapp_gcmc* g, const char* kernel, double* total_ms, int64_t* launches)
{
    return guard([&] {
        NEED(g);
        NEED(kernel);
        auto& p = g->sys->profiler();
        p.flush();
        for (int k = 0; k < mapp::K_COUNT; ++k)
            if (std::strcmp(kernel, mapp::kernel_name(k)) == 0) {
                if (total_ms) *total_ms = p.total_ms[k];
                if (launches) *launches = p.launches[k];
                return;
            }
        throw mapp::Error(mapp::ERR_ARG, std::string("unknown kernel group ") + kernel);
    });
}

/* ---------------- surrogate ---------------- */
mapp_status mapp_surr_create(int ne, const float* box4, int seed, mapp_surr** out)
{
    return guard([&] {
        NEED(box4);
        NEED(out);
        *out = nullptr;
        auto s = std::make_unique<mapp_surr>();
        s->sys = std::make_shared<mapp::Surrogate>(ne, mapp::SurrBox{box4[0], box4[1], box4[2], box4[3]}, seed);
        *out = s.release();
    });
}

void mapp_surr_destroy(mapp_surr* s) { delete s; }

mapp_status mapp_surr_train(mapp_surr* s, mapp_dmd* d, int nsamples, int steps, int batch, double lr,
                            double* loss2)
{
    return guard([&] {
        NEED(s);
        NEED(d);
        const mapp::SurrTrainReport r = s->sys->train(*d->sys, nsamples, steps, batch, lr);
        if (loss2) {
            loss2[0] = r.loss0;
            loss2[1] = r.loss;
        }
    });
}

mapp_status mapp_surr_loss_grad(mapp_surr* s, int net, int n, const float* x, const float* t, float* grad,
                                double* loss)
{
    return guard([&] {
        NEED(s);
        NEED(x);
        NEED(t);
        NEED(grad);
        const double l = s->sys->loss_grad(net, n, x, t, grad);
        if (loss) *loss = l;
    });
}

mapp_status mapp_surr_get_params(mapp_surr* s, int net, float* p)
{
    return guard([&] { NEED(s); NEED(p); s->sys->get_params(net, p); });
}

mapp_status mapp_surr_set_params(mapp_surr* s, int net, const float* p)
{
    return guard([&] { NEED(s); NEED(p); s->sys->set_params(net, p); });
}

mapp_status mapp_surr_get_norm(mapp_surr* s, int net, float* scale, float* shift)
{
    return guard([&] { NEED(s); NEED(scale); NEED(shift); s->sys->get_norm(net, scale, shift); });
}

mapp_status mapp_surr_set_norm(mapp_surr* s, int net, const float* scale, const float* shift)
{
    return guard([&] { NEED(s); NEED(scale); NEED(shift); s->sys->set_norm(net, scale, shift); });
}

mapp_status mapp_surr_infer(mapp_surr* s, int net, int n, const float* rin, float* out)
{
    return guard([&] { NEED(s); NEED(rin); NEED(out); s->sys->infer(net, n, rin, out); });
}

mapp_status mapp_dmd_use_surrogate(mapp_dmd* d, mapp_surr* s)
{
    return guard([&] { NEED(d); d->sys->use_surrogate(s ? s->sys : nullptr); });
}

mapp_status mapp_dmd_surrogate_coverage(mapp_dmd* d, int64_t* net_evals, int64_t* quadrature_slots)
{
    return guard([&] { NEED(d); d->sys->surrogate_coverage(net_evals, quadrature_slots); });
}

mapp_status mapp_replica_create(mapp_eam* ff, int nreplicas, const int32_t* natms, const double* H, const double* x,
                                const uint8_t* elem, double skin, mapp_replica** out)
{
    return guard([&] {
        NEED(ff);
        NEED(natms);
        NEED(H);
        NEED(x);
        NEED(elem);
        NEED(out);
        *out = nullptr;
        auto r = std::make_unique<mapp_replica>();
        r->sys = std::make_unique<mapp::ReplicaBatch>(ff->ff, nreplicas, natms, H, x, elem, skin);
        *out = r.release();
    });
}

void mapp_replica_destroy(mapp_replica* r) { delete r; }

mapp_status mapp_replica_min_cg(mapp_replica* r, int max_nsteps, double e_tol, double max_dx, double ls_c,
                                double ls_rho, double ls_min_alpha, int32_t* steps, int32_t* status, double* e0,
                                double* e, int32_t* evals)
{
    return guard([&] {
        NEED(r);
        r->sys->min_cg(max_nsteps, e_tol, max_dx, ls_c, ls_rho, ls_min_alpha, steps, status, e0, e, evals);
    });
}

mapp_status mapp_replica_tune(mapp_replica* r, int max_nsteps, double e_tol, double max_dx, double ls_c, double ls_rho,
                              double ls_min_alpha, int32_t* waves, double* ms)
{
    return guard([&] {
        NEED(r);
        const mapp::RepTuneReport t = r->sys->tune(max_nsteps, e_tol, max_dx, ls_c, ls_rho, ls_min_alpha);
        for (int k = 0; k < 5; ++k)
            if (waves) waves[k] = t.waves[k];
        for (int k = 0; k <= 5; ++k)
            if (ms) ms[k] = t.ms[k];
    });
}

mapp_status mapp_replica_set_waves(mapp_replica* r, const int32_t* waves)
{
    return guard([&] {
        NEED(r);
        NEED(waves);
        r->sys->set_waves(waves);
    });
}

mapp_status mapp_replica_use_graph(mapp_replica* r, int on)
{
    return guard([&] { NEED(r); r->sys->use_graph(on != 0); });
}

mapp_status mapp_replica_get_x(mapp_replica* r, int replica, double* x)
{
    return guard([&] {
        NEED(r);
        NEED(x);
        r->sys->get(replica, x);
    });
}

mapp_status mapp_replica_profile(mapp_replica* r, int enable)
{
    return guard([&] {
        NEED(r);
        auto& p = r->sys->profiler();
        p.reset();
        p.on = enable != 0;
    });
}

mapp_status mapp_replica_profile_read(mapp_replica* r, double* total_ms, int64_t* launches)
{
    return guard([&] {
        NEED(r);
        auto& p = r->sys->profiler();
        p.flush();
        if (total_ms) *total_ms = p.total_ms[mapp::K_REPLICA];
        if (launches) *launches = p.launches[mapp::K_REPLICA];
    });
}

}  // extern "C"
