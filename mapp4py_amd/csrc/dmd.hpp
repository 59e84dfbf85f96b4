// dmd.hpp -- device-resident diffusive-MD (DMD) system: Gaussian-smeared
// EAM by Gauss-Hermite quadrature over per-site components.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "linesearch.hpp"
#include "md.hpp"
#include "surrogate.hpp"

namespace mapp {

constexpr int kMaxQuad = 16;
constexpr int kMaxComp = 4;

// Kernel argument block of the smeared force field.
struct DmdArgs {
    int ne, nr, nrho, N, C;
    double dr_inv, drho_inv, rho_max, kBT, xmax;
    double xi[kMaxQuad], wi[kMaxQuad];
    double cut[kMaxElems * kMaxElems];  // r_c (no skin)
    double mu0[kMaxElems];
    const double4* tab;  // [ne*ne][nr][3]: r*phi[a][b], r*rho[a][b], r*rho[b][a] (FP64)
    const float4* tabf;  // the same rows rounded to FP32 (the FP32-row passes)
    const double* F5;    // [ne][nrho][5], quartic embedding
};

// FP32 knots of the LDS passes: the distinct table functions as (value,
// slope per unit offset) pairs [nf][nr], the distinct function of each
// (element pair, function) slot [nfn = ne*ne*3], and the knot window
// m_lo .. nr-1 (nwin knots per function) the workgroups hold in LDS
struct KnotArgs {
    const float2* glob;
    const int* fn;
    int nf, nr, nwin, m_lo, nfn;
};


// physicists' Gauss-Hermite rule (weight e^{-t^2}), n <= 16
void gauss_hermite(int n, double* x, double* w);

struct DMDReport {
    double fe;      // free energy (ext_mu term included when with_c)
    double FH[6];   // Voigt xx,xy,xz,yy,yz,zz
    double s;       // configurational entropy sum (units of kB)
};

// exchange parameters of the concentration rates: r_crd^2, the exchange
// width zeta and c_1 = sqrt(kT / 2m) / pi per element, beta = 1 / kT
struct CdArgs {
    double rsq_crd[kMaxElems], zeta[kMaxElems], c1[kMaxElems];
    double beta;
};

struct CGReport {
    int iters, status, rebuilds, evals;
    double fe0, fe, grad_norm, grad_max;
};

// Arrays of the Jacobian-product state, as the kernels take them
struct JState {
    uint32_t* idx;   // [maxn][ldl] partner component of a list entry
    double* r2;      // rho_{j->i}
    double* r1;      // rho_{i->j}
    double* t0;      // r phi + dE_i rho_{j->i} + dE_j rho_{i->j}
    int* cnt;        // [nc] list entries
    double* ddE;     // [nc] F''(rho)
    uint32_t* idx2;  // [maxn][ldl] second-list partner
    double *f0, *f1, *f2;  // exchange coefficients of each second-list entry
    double* S;       // [nc][3] their sums S0, S1, S2 per component
    int* cnt2;       // [nc]
};

class DMDSystem {
public:
    DMDSystem(std::shared_ptr<EamForceField> ff, int nsites, int c_dim, int nquad, const double* H, const double* x,
              const uint8_t* elem, const double* alpha, const double* c, double skin, double kB, double T,
              double hP, const double* masses);
    // list entries per batch of the quadrature passes (1 = one pair at a time)
    void set_batch(int b);
    // a new box and, unless x is null, new positions; rebuilds the grid and lists
    void set_box(const double* H, const double* x);
    // free energy, F_H and entropy of the state of the last c_d (no new passes)
    DMDReport force_calc_static();
    // table evaluation of the free energy and gradient (derivative, value,
    // the minimisers): 0 FP64 rows; 1 FP32 knots with the window in LDS (the
    // FP32 rows where the table is not a Hermite spline of its knots); 2 FP32
    // rows. FP32 modes keep FP64 moments and sums. The concentration-rate
    // path (c_d, J, the implicit and BDF steps) always reads the FP64 rows.
    void set_f32(int mode);
    int f32() const { return f32_; }
    // whether mode 1 holds the knots in LDS, and the window's first knot
    bool knots_in_lds() const { return knots_ok_; }
    int knot_window_first() const { return kn_.m_lo; }
    ~DMDSystem();

    int nsites() const { return n_; }
    int c_dim() const { return C_; }
    void set_state(const double* x, const double* alpha, const double* c);
    // c alone, in caller order
    void set_c(const double* c);
    // sum of f_x^2 and f_alpha^2 of the last evaluation (ForceFieldDMD::norm_sq)
    double grad_err_sq();
    // the gradient per unit c of the last c_d call (ForceFieldEAMDMD::calc_mu's
    // f and f_alpha): its squared norm, or the arrays in caller order
    double unit_grad_err_sq();
    void unit_gradient(double* g, double* g_alpha);
    void get_state(double* x, double* alpha, double* c);
    bool update();
    void rebuild();
    // ForceFieldDMD::new_derivative<with_c> / new_value<with_c>
    DMDReport derivative(bool with_c, const double* ext_mu);
    double value(bool with_c, const double* ext_mu);
    void get_forces(double* f, double* f_alpha, double* mu);
    // ExportCFGDMD: the current state as an extended CFG file (caller order)
    void export_cfg(const char* path, int nelems, const double* masses, const std::vector<std::string>& names);
    // ForceFieldDMD::c_d_calc: concentration rates over the second list
    // (same element, r < r_crd[e]) with exchange widths zeta[e]; c_d in caller
    // order [site][c_dim] when c_d_host is not null; returns max |c_d|
    double c_d(const double* r_crd, const double* zeta, double* c_d_host);
    // ForceFieldEAMDMD::__J: Ax = (d c_d / d c) x at the state and exchange
    // parameters of the last c_d call (an error if the state changed since);
    // x and Ax in caller order [site][c_dim]
    void J(const double* x, double* ax);
    // DAEImplicit::newton: solve c / beta + a - c_d(c) = 0 from the current c
    // by Newton steps with GMRES over J (at most max_gmres <= 8 products per
    // step); a in caller order. True on convergence (c holds the solution);
    // otherwise c is restored. iters = Newton steps as the reference counts them.
    bool newton(const double* r_crd, const double* zeta, double beta, const double* a, double a_tol,
                double err_fac, int max_newton, int max_gmres, int* iters);
    // ForceFieldEAMDMD::__prepJ_n_res + __J (src/ff_eam_dmd.cpp:364-831) at
    // the current state and fixed c: (jdx, jda) = -(d f / d(x, alpha)) (dx, da)
    // with f the gradient per unit c (calc_mu's f and f_alpha); caller order,
    // dx / jdx / f [site][3], da / jda / fa [site][c_dim]; f and fa optional
    void Jxa(const double* dx, const double* da, double* jdx, double* jda, double* f = nullptr,
             double* fa = nullptr);
    // the Newton stage of DAE::min_error_false (src/dae.cpp:294-325) over x
    // and alpha at fixed c (max_gmres <= 8); returns the steps taken and, in
    // res, |f| at the end
    int newton_xa(double a_tol, int max_newton, int max_gmres, double* res);
    // entries the last J state holds: list entries and second-list entries
    void J_entries(int64_t* live, int64_t* exchange);
    int64_t half_pairs_count();
    void half_pairs(int32_t* i, int32_t* j, int8_t* img);
    // MinCGDMD over x and alpha (c fixed)
    // ls: the line search and its parameters (nullptr: LineSearchBackTrack's defaults)
    CGReport cg(int max_iter, double e_tol, double max_dx, double max_dalpha, double grad_tol,
                const LineSearchParams* ls = nullptr);
    // MinLBFGSDMD::__run (src/min_l-bfgs_dmd.h:113-210) over x and alpha with
    // memory m and the back-tracking line search; report and stops as cg
    CGReport lbfgs(int m, int max_iter, double e_tol, double max_dx, double max_dalpha, double grad_tol,
                   const LineSearchParams* ls = nullptr);
    // smeared pair functions at given (r, alpha_ij): out[k*9 + q], q = rho_phi[3], drho_phi_dr[3], drho_phi_dalpha[3]
    void smear(int ei, int ej, int npts, const double* r, const double* a, double* out);
    // the same for device arrays and a table index tt = ei * ne + ej (synchronous)
    void smear_dev(int tt, int npts, const double* r, const double* a, double* out);
    Profiler& profiler() { return prof_; }
    // evaluations use the surrogate networks for pairs inside their training
    // box and the quadrature elsewhere; nullptr returns to the quadrature
    void use_surrogate(std::shared_ptr<Surrogate> s);
    // slots the last evaluation gave to the networks / to the quadrature
    void surrogate_coverage(int64_t* by_net, int64_t* by_quadrature);
    const DmdArgs& args() const { return A_; }
    hipStream_t stream() const { return stream_; }

private:
    void setup_grid();
    void build_list();
    void pass1(bool energy_only, int f32 = 0);
    void pass2(bool unit = false, int f32 = 0);
    void plan_knots();
    bool decide();
    void reduce_rows(const double* part, int nrows, int ncols, double* out);
    void J_prepare();
    void j_reserve();
    void J_dev(const double* x, double* ax);
    JState jstate();
    double xa_prepare();
    void Jxa_dev(const double* in, double* out);
    // device Krylov basis [m+1][nv], product vector and partial sums of one GMRES
    struct GmresWork {
        DevBuf<double> Q, v, red, part;
    };
    template <class Op>
    bool gmres(GmresWork& w, size_t nv, int m, Op&& A, const double* b, double tol, double* x, double& norm);

    std::shared_ptr<EamForceField> ff_;
    int n_, C_, nc_;  // sites, components per site, total components
    double skin_, max_cut_, scl_;
    Box box_;
    CellGrid grid_{};
    DmdArgs A_{};
    hipStream_t stream_ = nullptr;
    int nrebuild_ = 0;

    DevBuf<double4> x_, x_alt_, s_, x0_, s_tmp_;
    DevBuf<double> al_, al_alt_, c_, c_alt_, al0_;
    DevBuf<int> el_, el_alt_;
    DevBuf<int> id_, id_alt_, slot_of_;
    DevBuf<double4> f_;
    DevBuf<double> fa_, mu_, E_, dE_;
    DevBuf<double> cd_;              // concentration rates per component
    // Jacobian state: bumped by every change of x, alpha, c or the slot order
    int64_t ver_ = 0, jver_ = -1, jrp_ver_ = -1, jslot_ver_ = -1;
    CdArgs jP_{};
    DevBuf<double4> gu_;   // gradient per unit c of the last c_d, per site
    DevBuf<double> gau_;   // and per component (alpha)
    int64_t gu_ver_ = -1;
    int64_t static_ver_ = -1;        // state of the sums in static_ (the last c_d)
    DevBuf<double> static_;          // free energy, entropy and virial sums of the last c_d
    DevBuf<double> jmu_, jdE_, jddE_;
    DevBuf<uint32_t> jidx_, jidx2_;   // list entries / second list, slot-major
    DevBuf<double> jr2_, jr1_, jt0_, jf0_, jf1_, jf2_, jS_;
    DevBuf<int> jcnt_, jcnt2_;
    DevBuf<double> jx_, ja1_, jb0_, jax_, jxs_;
    // implicit-step vectors (slot order) and the Krylov basis [m+1][nc]
    DevBuf<double> nt_a_, nt_F_, nt_del_, nt_c0_, nt_red_;
    GmresWork nt_gm_;
    // second-derivative state: kXaQ values per list slot ([q][slot]), F'' and
    // fcoef per component, the change of F' of the last product, the
    // residual and the Newton step over (x, alpha) (flat, slot order)
    DevBuf<double> xa_slots_, xa_ddE_, xa_fc_, xa_dde_, xa_f_, xa_h_;
    size_t xa_lds_ = 0;
    int64_t xa_ver_ = -1;
    GmresWork xa_gm_;
    std::vector<double> masses_;
    std::vector<uint8_t> elem_host_;  // [site][c_dim] in caller order
    DevBuf<uint32_t> keys_, keys_alt_;
    DevBuf<int> vals_, vals_alt_;
    DevBuf<unsigned char> sort_tmp_;
    size_t sort_tmp_bytes_ = 0;
    DevBuf<int> cell_rank_, cell_start_;
    DevBuf<int> nn_;
    DevBuf<uint32_t> list_;
    int maxn_ = 0, ldl_ = 0;
    int batch_ = 4;                  // quadrature pass batch: 1 (per-pair loop), 2 or 4 list entries
    int f32_ = 0;                    // table mode of derivative / value (set_f32)
    // FP32 knots and their LDS window (plan_knots)
    DevBuf<float2> knots_;
    DevBuf<int> knot_fn_;
    KnotArgs kn_{};
    bool knots_ok_ = false;
    size_t knot_lds_ = 0;
    int knot_grid_ = 0;
    // their passes' packed neighbor records: (alpha, c), and (F', element)
    // from pass 1 of state ed_ver_
    DevBuf<double2> ac_, ed_;
    int64_t ed_ver_ = -1;
    DevBuf<int> flags_;
    DevBuf<double> part_, red_, red2_;
    DevBuf<double4> tab_;
    DevBuf<float4> tabf_;
    DevBuf<double> F5_;
    // CG work vectors (x part as double4 slots, alpha part per component)
    DevBuf<double4> cg_x0_, cg_h_, cg_f0_;
    DevBuf<double> cg_a0_, cg_ha_, cg_fa0_;
    Profiler prof_;
    void stage_surrogate();
    std::shared_ptr<Surrogate> surr_;
    DevBuf<float> s_y_;
    DevBuf<int> s_cnt_, s_off_;
    DevBuf<SurrRequest> s_req_;
    DevBuf<unsigned char> s_scan_tmp_;
    size_t s_scan_bytes_ = 0;
    int64_t s_net_ = 0, s_quad_ = 0;
};

}  // namespace mapp
