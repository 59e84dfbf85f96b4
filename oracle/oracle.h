/*
 * oracle.h -- FP64 CPU restatement of MAPP's EAM / DMD / GCMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness oracle and the
 * CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library (mapp4py_amd) never does.
 *
 * Every function restates the algorithm of the reference file:line given in
 * its comment; none of the reference's source is copied. All arithmetic is
 * IEEE double with floating-point contraction disabled (-ffp-contract=off)
 * so that the neighbor predicate reproduces the reference bit for bit.
 *
 * Parity pinning (what checks each part; see tests/test_oracle_*.py):
 *   splines       exact reproduction of cubics/quartics on a uniform grid,
 *                 knot interpolation, closed-form coefficient formulas
 *   setfl / fs    round trip through a generated file, shipped example tables
 *   neighbor list O(N^2) brute force over explicit periodic images
 *   EAM force     finite differences of the energy, Newton III, sum f = 0,
 *                 pair-only limit vs analytic pair sum, perfect lattice f = 0,
 *                 finite-strain derivative vs virial
 *   Gauss-Hermite closed form n=3, exactness on polynomials, sum w = sqrt(pi)
 *   DMD           closed-form Gaussian smearing, finite differences of the
 *                 free energy vs f_x, f_alpha, mu; alpha -> 0 gives MD
 *   Park-Miller   published check value (seed 1 -> 10000th = 1043618065)
 *   NVT / NHC     zero-force temperature history against an independent
 *                 integration of the chain's equations of motion; the
 *                 extended energy of that chain conserved with forces
 */
#ifndef MAPP_ORACLE_H
#define MAPP_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- splines (src/import_eam.cpp:62-115) ---------------- */
void o_interpolate4(size_t n, double (*s)[4]);
void o_interpolate5(size_t n, double (*s)[5]);

/* ---------------- tabulated EAM (src/import_eam.h:52-901) ---------------- */
typedef struct o_eam o_eam;
/* fmt: 0 = setfl (eam/alloy), 1 = fs (eam/fs). dmd != 0 builds the DMD
 * variant: quartic F ([5] coefficients) and r*rho tables
 * (src/ff_eam_dmd.cpp:80-94). Returns NULL and fills err on failure. */
o_eam* o_eam_read(const char* path, int fmt, int dmd, char* err, int errlen);
/* Builds an EAM object from knot arrays: F[ne][nrho], rho[ne][ne][nr]
 * (rho[i][j] = density function listed in element i's block at position j),
 * r_phi[ne][ne][nr] (symmetric). setfl_alias != 0 makes rho[i][j] alias
 * rho[i][0] as setfl does. */
o_eam* o_eam_from_knots(int ne, int nr, int nrho, double dr, double drho,
                        const double* F, const double* rho, const double* r_phi,
                        int setfl_alias, int dmd);
/* ImportEAM::funcfl (src/import_eam.h:68-344): one funcfl file per element */
o_eam* o_eam_read_funcfl(const char* const* paths, int nfiles, int dmd, char* err, int errlen);
void o_eam_free(o_eam*);
void o_eam_info(const o_eam*, int* ne, int* nr, int* nrho, double* dr, double* drho);
void o_eam_cut(const o_eam*, double* rc /* [ne*ne] */);
/* coefficient export for cross-checking: which 0=r_phi 1=rho 2=F */
void o_eam_coefs(const o_eam*, int which, int i, int j, double* out);

/* ---------------- Gauss-Hermite (src/xmath.cpp:440-589) ---------------- */
void o_quadrature_hg(int n, double* x, double* w);

/* ---------------- random numbers (src/random.cpp:7-68) ---------------- */
typedef struct { int seed; int gauss_chk; double reserved; } o_rng;
void o_rng_init(o_rng*, int seed);
double o_rng_uniform(o_rng*);
double o_rng_gaussian(o_rng*);

/* ---------------- MD system: ghosts, cells, half list, EAM ---------------- */
typedef struct o_md o_md;
/* DynamicMD::init (src/dynamic_md.cpp:34-52): wrap, ghosts, list, store x0.
 * H is the lower-triangular box, row-major H[3][3] (rows = box vectors). */
o_md* o_md_create(const o_eam* ff, int natms, const double* H, const double* x,
                  const uint8_t* elem, double skin);
void o_md_free(o_md*);
int o_md_natms(const o_md*);
int o_md_nghosts(const o_md*);
int64_t o_md_npairs(const o_md*);
/* half list as (i, owner j, image shift) triples in list order */
void o_md_pairs(const o_md*, int32_t* i, int32_t* j, int8_t* img /* [npairs][3] */);
void o_md_get_x(const o_md*, double* x);
void o_md_set_x(o_md*, const double* x);
/* DynamicMD::update<true> (src/dynamic_md.h:38-68). Returns 1 on rebuild. */
int o_md_update(o_md*);
/* force_calc (src/ff_md.cpp:125-144 + src/ff_eam.cpp:65-240).
 * out[0] = pe, out[1..6] = F_H Voigt (xx,xy,xz,yy,yz,zz). rho_out may be NULL. */
void o_md_force(o_md*, double* f, double* out7, double* rho_out);
/* value() (src/ff_md.cpp:148-155 + src/ff_eam.cpp:244-318) */
double o_md_energy(o_md*);
/* MDNVT::run (src/md_nvt.cpp:383-480) with NHC L=3, niters=1 on a system
 * whose velocities are v [N][3]; masses per element. Returns the number of
 * neighbor-list rebuilds. out7 as o_md_force after the last step. */
int o_md_nvt_run(o_md*, double* v, const double* masses, double kB, double T,
                 double dt, int nsteps, double* out7, double* T_part_out);
/* o_md_nvt_run with a record per step after its second thermostat call:
 * hist[nsteps][8] = T_part, potential energy, xi_1..xi_3 (the chain's
 * eta_d), and the integrals of xi_1..xi_3 over time, each thermostat call
 * adding its span dt/2 times the chain at its mid state. */
int o_md_nvt_run_hist(o_md*, double* v, const double* masses, double kB, double T,
                      double dt, int nsteps, double* out7, double* T_part_out, double* hist);
/* ThermostatNHC(dt, t_relax, 3, niters) (src/thermostat.cpp:19-104) called
 * ncalls times on a reduced temperature T_r that nothing else changes;
 * hist[ncalls][4] (may be NULL) = T_r, xi_1..xi_3 after each call. Returns
 * the product of the velocity rescales. */
double o_nhc_run(double T_r, int ndof, double dt, double t_relax, int niters, int ncalls, double* hist);
/* Test hook (0 = the reference's thermostat): plants a known mistake in the
 * thermostat, 1 ddt and ddt2 swapped, 2 the sign of the forward sweep's chain
 * coupling flipped, 3 N_f xi_1^2 in the reverse sweep too. */
extern int o_nhc_mutation;
/* Threads of the MD list build and force (1, the default, for the checker;
 * more only for the CPU baseline's all-cores timing). o_get_threads gives
 * the count in effect (1 without OpenMP). */
void o_set_threads(int n);
int o_get_threads(void);
/* MDMath::calc (src/md_nst.cpp:887-1002): E = exp(-dt2 (A + a I)),
 * F0 = dt2 phi(-dt2 (A + a I)), F1 = dt phi(dt A), phi(z) = (e^z - 1) / z,
 * for a lower-triangular A; all row-major [9] */
void o_nst_calc(double a, double dt, double dt2, const double* A, double* E, double* F0, double* F1);
/* MDNST::run (src/md_nst.cpp:376-460): see o_md.c; returns the rebuilds */
extern int o_nst_mutation;
int o_md_nst_run(o_md*, double* v, const double* masses, double kB, double T, double dt, const double* S9,
                 const double* tau9, int nreset, int nsteps, double* out7, double* T_part, double* H9, double* hist);
/* AtomsMD::create_T (src/atoms_md.cpp:130-160) */
void o_create_T(int natms, const uint8_t* elem, const double* masses, double kB,
                double T, int seed, double* v);

/* ---------------- DMD (src/ff_eam_dmd.cpp, src/neighbor_dmd.cpp) ---------------- */
typedef struct o_dmd o_dmd;
/* c_dim components per site; elem/alpha/c are [natms][c_dim]; c < 0 marks
 * an absent component. ff must have been built with dmd != 0. */
o_dmd* o_dmd_create(const o_eam* ff, int natms, int c_dim, int nquad,
                    const double* H, const double* x, const uint8_t* elem,
                    const double* alpha, const double* c, double skin,
                    double kB, double T, double hP, const double* masses);
void o_dmd_free(o_dmd*);
int64_t o_dmd_npairs(const o_dmd*);
void o_dmd_pairs(const o_dmd*, int32_t* i, int32_t* j, int8_t* img);
void o_dmd_set_state(o_dmd*, const double* x, const double* alpha, const double* c);
/* __force_calc (src/ff_eam_dmd.cpp:126-360) wrapped by derivative()
 * (src/ff_dmd.cpp:220-237). out: [0]=fe, [1..6]=F_H Voigt, [7]=s.
 * f [N][3], f_alpha [N*c_dim], mu [N*c_dim] (= f_c after __force_calc,
 * i.e. F(rho_i) + sum_j c_j (phi + rho dE_j)), E [N*c_dim] = F(rho). */
void o_dmd_force(o_dmd*, double* f, double* f_alpha, double* mu, double* out8);
/* value() (src/ff_dmd.cpp:223-231 + src/ff_eam_dmd.cpp:835-953) */
double o_dmd_energy(o_dmd*);
/* smeared pair quantities for one pair (r, alpha_ij): out[9] = rho_phi[3],
 * drho_phi_dr[3], drho_phi_dalpha[3] in the order phi(ei,ej), rho(ei,ej),
 * rho(ej,ei) (src/ff_eam_dmd.cpp:195-253). */
void o_dmd_pair_eval(const o_dmd*, int ei, int ej, double r, double alpha_ij, double* out9);
/* ForceFieldEAMDMD::calc_Q with theta (src/ff_eam_dmd.cpp:1611-1633) */
void o_dmd_calc_q(double gamma_i, double gamma_j, double mu_ji, double* Q, double* alpha_Q_sq, double* theta);
/* ForceFieldDMD::c_d_calc + ForceFieldEAMDMD::__c_d_calc (src/ff_dmd.cpp:449-456,
 * src/ff_eam_dmd.cpp:1359-1473): c_d [natms*c_dim]; returns max |c_d| */
double o_dmd_c_d(o_dmd*, const double* r_crd, const double* zeta, const double* masses, double* c_d);
/* ForceFieldEAMDMD::__J (src/ff_eam_dmd.cpp:1475-1575): Ax = (d c_d / d c) x */
void o_dmd_J(o_dmd*, const double* r_crd, const double* zeta, const double* masses, const double* x, double* Ax);
/* ForceFieldEAMDMD::calc_mu's f and f_alpha: the gradient per unit c */
void o_dmd_unit_grad(o_dmd*, double* g /* [natms][3] */, double* ga /* [natms*c_dim] */);
/* smear with the second derivatives (src/ff_eam_dmd.cpp:420-495): out[18] =
 * rho_phi[3], (1/r) d/dr [3], (1/a) d/da [3], ((1/r) d/dr)^2 [3],
 * (1/a) d/da (1/r) d/dr [3], ((1/a) d/da)^2 [3] */
void o_dmd_pair_eval2(const o_dmd*, int ei, int ej, double r, double alpha_ij, double* out18);
/* ForceFieldEAMDMD::__prepJ_n_res + __J (src/ff_eam_dmd.cpp:364-831):
 * (Jdx, Jda) = -(d f / d(x, alpha)) (dx, da) at fixed c, f the gradient per
 * unit c (o_dmd_unit_grad) */
void o_dmd_Jxa(o_dmd*, const double* dx, const double* da, double* Jdx, double* Jda);
/* the Newton stage of DAE::min_error_false (src/dae.cpp:294-325); returns the
 * steps taken, *res the final |f| */
int o_dmd_newton_xa(o_dmd*, double a_tol, int max_newton, int max_gmres, double* res);
extern int o_jxa_mutation;
/* the local positions [natms][3] and alpha [natms*c_dim] */
void o_dmd_get_xa(const o_dmd*, double* x, double* alpha);
/* the local concentrations [natms][c_dim] */
void o_dmd_get_c(const o_dmd*, double* c);
/* __GMRES::solve (src/gmres.h:40-60) on a dense row-major A [n][n]: returns 1
 * when |residual| < tol within m products; *norm = |x| */
int o_gmres_dense(int n, int m, const double* A, const double* b, double tol, double* x, double* norm);
/* DAEImplicit::newton (src/dae_imp.cpp:146-206): solve c / beta + a - c_d(c) = 0
 * from the current c; 1 on convergence (c holds the solution), else c restored */
int o_dmd_newton(o_dmd*, const double* r_crd, const double* zeta, const double* masses, double beta,
                 const double* a, double a_tol, double err_fac, int max_newton, int max_gmres, int* iters);

/* ---------------- GCMC (src/sgcmc.cpp, src/ff_eam.cpp:323-589) ---------------- */
/* One periodic cell under single-process SGCMC with EAM exchange energies.
 * cap bounds the atom count (inserted atoms included). */
typedef struct o_gcmc o_gcmc;
o_gcmc* o_gcmc_create(const o_eam* ff, int natms, const double* H, const double* x,
                      const uint8_t* elem, int gas_elem, double mu, double T, double kB,
                      double hP, const double* masses, int seed, int cap);
void o_gcmc_free(o_gcmc*);
/* nattempts SGCMC::attmpt calls. decisions[k]: 1 accepted insertion, -1
 * rejected insertion, 2 accepted deletion, -2 rejected deletion, 0 deletion
 * drawn with no gas atom (no further draws); dU[k] = delta_u (0 when 0).
 * Returns 0, or -1 when an insertion would exceed cap. */
int o_gcmc_attempts(o_gcmc*, int nattempts, int* decisions, double* dU);
/* the cell's atoms replaced by natms atoms (x Cartesian, wrapped on entry;
 * v NULL: zero), densities and F recomputed, the random stream kept (the
 * set-up of each GCMC::xchng); -1 if natms exceeds cap */
int o_gcmc_reload(o_gcmc*, int natms, const double* x, const uint8_t* elem, const double* v);
int o_gcmc_natms(const o_gcmc*);
int o_gcmc_ngas(const o_gcmc*);
int o_gcmc_rng_seed(const o_gcmc*);
/* out8: z_fac, vol, beta, lambda, sigma, cut, cut_s[0], cell_size[0];
 * iout4: max trial images, cells per dimension (3) */
void o_gcmc_info(const o_gcmc*, double* out8, int* iout4);
void o_gcmc_state(const o_gcmc*, double* x, double* s, uint8_t* elem, double* rho,
                  double* F, double* v);

/* MDMuVT::run (src/md_muvt.cpp:122-280): NVT with an SGCMC exchange of the
 * gas element after every nevery-th first half step. *md is replaced when the
 * atom count changes; v holds cap atoms; counts[4]: accepted insertions,
 * accepted deletions, insertion trials, deletion trials; hist [nsteps][3]:
 * T_part, potential energy, gas atoms (may be NULL). Returns the list builds,
 * -1 when an insertion would exceed cap. */
int o_md_muvt_run(o_md** md, double* v, int cap, const double* masses, double kB, double hP, double T, double dt,
                  double mu, int gas, int seed, int nevery, int nattempts, int nsteps, double* out7,
                  double* T_part, int* natms, int* counts, double* hist);

#ifdef __cplusplus
}
#endif
#endif
