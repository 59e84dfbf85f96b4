/*
 * mapp_amd.h -- C ABI of the MI355X (gfx950) implementation of MAPP's EAM /
 * DMD hot path.
 *
 * Conventions
 *   - Every call returns a status (0 = MAPP_OK). On failure mapp_last_error()
 *     returns a message for the calling thread. No exception crosses the ABI.
 *   - Arrays passed by pointer are host arrays owned by the caller, in the
 *     caller's atom order; x, v, f are [natms][3] doubles.
 *   - H is the 3x3 lower-triangular box matrix, row-major, rows = box vectors
 *     (the reference's atoms->H, src/atoms.cpp:187-194).
 *   - Handles are not thread safe; each owns one HIP stream.
 *   - Units are the caller's (the reference uses eV, Angstrom, amu).
 */
#ifndef MAPP_AMD_H
#define MAPP_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int mapp_status;
#define MAPP_OK 0
#define MAPP_ERR_ARG 1
#define MAPP_ERR_FILE 2
#define MAPP_ERR_HIP 3
#define MAPP_ERR_CAPACITY 4
#define MAPP_ERR_STATE 5

typedef struct mapp_eam mapp_eam; /* tabulated EAM tables resident on the GPU */
typedef struct mapp_md mapp_md;   /* atoms + neighbor list + dynamic state on the GPU */

/* Message of the last failed call on this thread ("" if none). */
const char* mapp_last_error(void);
/* Library version string. */
const char* mapp_version(void);

/* ---------------- force field construction ----------------
 * Replace ForceFieldEAM construction by sim.ff_eam_setfl / sim.ff_eam_fs
 * (src/ff_eam.cpp:594-801) over ImportEAM::setfl / fs (src/import_eam.h:348-901).
 * elems selects and orders the file's elements (NULL / 0: all, file order).
 * dmd != 0 builds the DMD variant (quartic F, r*rho tables) used by
 * ForceFieldEAMDMD (src/ff_eam_dmd.cpp:17-106). */
mapp_status mapp_eam_create_setfl(const char* path, const char* const* elems, int nelems, int dmd, mapp_eam** out);
mapp_status mapp_eam_create_fs(const char* path, const char* const* elems, int nelems, int dmd, mapp_eam** out);
/* ff_eam_funcfl (src/ff_eam.cpp:597-650) with ImportEAM::funcfl
 * (src/import_eam.h:68-344): one DYNAMO funcfl file per element, element i =
 * paths[i]; the files are resampled onto the largest spacing over the longest
 * range, r phi_ij = 27.2 * 0.529 Z_i Z_j. Files whose r table is shorter than
 * that grid are rejected (the reference reads past them). */
mapp_status mapp_eam_create_funcfl(const char* const* paths, int nfiles, int dmd, mapp_eam** out);
/* From knot arrays: F[ne][nrho], rho[ne][ne][nr] (rho[a][b] = density
 * function listed in element a's block at position b), r_phi[ne][ne][nr]
 * symmetric; fs == 0 uses rho[a][0] for every b as setfl does. */
mapp_status mapp_eam_create_knots(int ne, int nr, int nrho, double dr, double drho, const double* F,
                                  const double* rho, const double* r_phi, int fs, int dmd, mapp_eam** out);
void mapp_eam_destroy(mapp_eam* ff);
/* r_c[ne][ne] derived from the tables (src/import_eam.h:52-64) */
mapp_status mapp_eam_cutoff(const mapp_eam* ff, int* ne, double* rc);
/* Spline coefficients for cross-checking: which 0 = r_phi[a][b] ([nr][4]),
 * 1 = rho[a][b] ([nr][4]), 2 = F[a] ([nrho][4] or [5] for dmd). */
mapp_status mapp_eam_coefs(const mapp_eam* ff, int which, int a, int b, double* out);

/* ---------------- MD system ----------------
 * Replaces DynamicMD::init (src/dynamic_md.cpp:34-52): wraps the atoms into
 * the box, bins them into cells and builds the neighbor list with skin. */
mapp_status mapp_md_create(mapp_eam* ff, int natms, const double* H, const double* x, const uint8_t* elem,
                           double skin, mapp_md** out);
void mapp_md_destroy(mapp_md* md);
mapp_status mapp_md_set_x(mapp_md* md, const double* x);
mapp_status mapp_md_get_x(mapp_md* md, double* x);
mapp_status mapp_md_set_v(mapp_md* md, const double* v);
mapp_status mapp_md_get_v(mapp_md* md, double* v);
/* A new box, as the reference's callers change atoms->H (src/atoms.cpp:187-194:
 * H and B = H^-1) when they strain the system or integrate the box (md_nst,
 * the minimisers' H degrees of freedom), with the positions x in caller order
 * (NULL keeps the Cartesian positions). Rebuilds the cell grid, the image
 * shifts and the neighbor list; the next force_calc is that of the new box.
 * A box narrower than half the cut-off + skin leaves the system unchanged
 * (MAPP_ERR_CAPACITY). */
mapp_status mapp_md_set_box(mapp_md* md, const double* H, const double* x);
/* DynamicMD::update<true> (src/dynamic_md.h:38-68, decide at
 * src/dynamic_md.cpp:90-105): rebuilds when an atom moved more than skin/2. */
mapp_status mapp_md_update(mapp_md* md, int* rebuilt);
/* ForceFieldMD::force_calc (src/ff_md.cpp:125-144) with
 * ForceFieldEAM::__force_calc (src/ff_eam.cpp:65-240). out7[0] = energy,
 * out7[1..6] = F_H in Voigt order xx,xy,xz,yy,yz,zz (S_pe = -F_H / V).
 * f may be NULL. */
mapp_status mapp_md_force_calc(mapp_md* md, double* f, double* out7);
/* ForceFieldMD::value (src/ff_md.cpp:148-155) with
 * ForceFieldEAM::__energy_calc (src/ff_eam.cpp:244-318). */
mapp_status mapp_md_value(mapp_md* md, double* energy);
/* mapp_md_value without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, code that forbids hipMalloc). The same passes as mapp_md_value, the
 * same energy; their forces and densities go to a static scratch array that
 * the library holds on every GPU it runs on (24 MiB) instead of buffers that
 * mapp_md_value allocates on its first call. At most 1,048,576 atoms
 * (MAPP_ERR_CAPACITY beyond). Every call on one GPU uses that GPU's one
 * array: calls that overlap on one GPU (two threads, two handles) race. */
mapp_status mapp_md_value_noalloc(mapp_md* md, double* energy);
/* per-atom electron density of the last force_calc (the "rho" vector) */
mapp_status mapp_md_get_rho(mapp_md* md, double* rho);
/* Half list as NeighborMD::create_list stores it (src/neighbor_md.cpp:82-152):
 * triples (i, owner j, image shift) with i<j inside the box and every
 * periodic-image pair. */
mapp_status mapp_md_pairs_count(mapp_md* md, int64_t* npairs);
mapp_status mapp_md_pairs(mapp_md* md, int32_t* i, int32_t* j, int8_t* img);
mapp_status mapp_md_list_stats(mapp_md* md, int* max_neighbors, int64_t* full_entries, int* rebuilds);
/* MDNVT::run (src/md_nvt.cpp:383-480): velocity Verlet with a Nose-Hoover
 * chain (L=3, one iteration, t_relax = 100 dt), centre-of-mass drift removed
 * each step, neighbor rebuild on the skin criterion. masses[ne]. */
mapp_status mapp_md_nvt_run(mapp_md* md, const double* masses, double kB, double T, double dt, int nsteps,
                            double* out7, double* T_part, int* rebuilds);
/* MDNST::run (src/md_nst.cpp:376-460, pre_init and update_V_H at :291-351,
 * the exponential updates of update_x_d__x__x_d at :80-153 with MDMath::calc
 * at :843-1002): velocity Verlet under NHC(dt/2, 100 dt) with a lower-
 * triangular box velocity driven toward the target stress S [3][3]
 * (row-major, lower triangle read, NaN = a free component; at least one set)
 * under NHC(dt/2, 1000 dt), tau [3][3] the box relaxation times (NULL:
 * 1000 dt each). The reference box B_ref and volume are re-taken every nreset
 * steps (0: never). The box changes every step; the list is rebuilt only when
 * an atom has moved half the skin. masses per element; out7 = energy and F_H
 * of the last force call, T_part, H [3][3] the final box (any may be NULL),
 * hist [nsteps][14] (T_part, pe, H00, H10, H11, H20, H21, H22, then the
 * stress -F_H / vol as xx, yx, zx, yy, zy, zz after each step; may be NULL).
 * The handle keeps the new box. */
mapp_status mapp_md_nst_run(mapp_md* md, const double* masses, double kB, double T, double dt, const double* S,
                            const double* tau, int nreset, int nsteps, double* out7, double* T_part, double* H,
                            double* hist, int* rebuilds);
/* MDMuVT::run (src/md_muvt.cpp:122-280, the run without fixed degrees of
 * freedom): MDNVT's velocity Verlet under NHC(dt/2, 100 dt), where after every
 * nevery-th first half step the exchange gcmc.xchng(false, nattempts) takes
 * the place of the skin check: nattempts insertion/deletion attempts of
 * element gas_elem at chemical potential mu (eV), temperature T, with the
 * Planck constant hP in the table's units; the chain's Random(seed) stream
 * runs across the exchanges of one call. At one process PGCMC, which MDMuVT
 * uses, draws and decides as SGCMC (src/pgcmc.cpp:562-735,1163-1266,
 * src/sgcmc.cpp:251-444). Each exchange wraps the atoms into the box, leaves
 * them in the reference's order (an inserted atom appended, the last atom
 * moved into a deleted one's place; velocities travel with them, an inserted
 * atom's drawn by the chain) and rebuilds the list; ndof_part = 3 natms. The
 * system may hold at most 16,384 atoms. masses per element; out7 = energy and
 * F_H of the last force call; counts[4] = accepted insertions, accepted
 * deletions, insertion trials, deletion trials; hist [nsteps][2] = T_part and
 * gas atoms after each step (any may be NULL); builds = list builds
 * (rebuilds and exchanges). The handle keeps the new atoms: read their count
 * with mapp_md_natms, then positions, velocities and elements. */
mapp_status mapp_md_muvt_run(mapp_md* md, const double* masses, double kB, double hP, double T, double dt, double mu,
                             int gas_elem, int seed, int nevery, int nattempts, int nsteps, double* out7,
                             double* T_part, int* counts, double* hist, int* builds);
/* the handle's atom count (it changes under mapp_md_muvt_run) */
mapp_status mapp_md_natms(mapp_md* md, int* natms);
/* element index per atom, caller order (AtomsMD's elem vector) */
mapp_status mapp_md_get_elem(mapp_md* md, uint8_t* elem);
/* Energy minimisation of the positions: MinCG::__run (src/min_cg.h:453-533,
 * lbfgs = 0) or MinLBFGS::__run (src/min_l-bfgs.h:114-217, lbfgs = 1, memory
 * m) with MinMDHandler::ls_prep (src/min_cg.h:231-257: no step component
 * larger than max_dx) and one of the reference's line searches: ls = 0
 * LineSearchBackTrack::min (src/ls.h:711-769; c, rho, min_alpha as in
 * src/ls.cpp:456-461), ls = 1 LineSearchGoldenSection::min (src/ls.h:341-520)
 * or ls = 2 LineSearchBrent::min (src/ls.h:525-706), both over
 * LineSearch::bracket (src/ls.h:213-335) when ls_brack != 0, with tol,
 * max_iter and zeps as in src/ls.cpp:95-101,262-268. The box stays
 * fixed (H_dof all False) and displacements are per atom (affine = False).
 * status follows the reference's err_msgs order (src/min.cpp:7-25): 1 not
 * downhill, 2 zero gradient, 3 minimum alpha, 4 energy tolerance, 5 maximum
 * iterations, 8 bracketing found no lower energy. Rows (step, energy), the start and every accepted step, go to
 * tally [max_tally_rows][2]; *tally_rows counts the rows written. */
typedef struct {
    int lbfgs, m;
    double e_tol, max_dx;
    int max_nsteps;
    double ls_c, ls_rho, ls_min_alpha;
    int ls, ls_max_iter, ls_brack;
    double ls_tol, ls_zeps;
} mapp_min_params;
typedef struct {
    int steps, status, evals, force_calcs, rebuilds;
    double e0, e, f_norm, f_max;
} mapp_min_report;
mapp_status mapp_md_minimize(mapp_md* md, const mapp_min_params* params, mapp_min_report* report, double* tally,
                             int max_tally_rows, int* tally_rows);
/* One line search over a caller's energy along a line, as the minimisers
 * run it: LineSearch*::min (src/ls.h:213-769) entered with what
 * MinMDHandler::ls_prep (src/min_cg.h:231-257) leaves, dfa = -f.h and max_a.
 * energy(ctx, alpha) moves the caller's state to x0 + alpha h and returns its
 * energy; reset(ctx) moves it back to x0. params->ls, ls_c, ls_rho,
 * ls_min_alpha, ls_max_iter, ls_brack, ls_tol and ls_zeps choose and tune the
 * search (the other fields are not read); *prev_val carries the back-tracking
 * search's memory between calls (-1 after LineSearch::reset). On return
 * *status is the reference's code (0 success, 1-3 line-search failures,
 * 8 bracketing failure), *energy the accepted energy and *alpha the step. Runs
 * on the host only. */
typedef double (*mapp_line_energy_fn)(void* ctx, double alpha);
typedef void (*mapp_line_reset_fn)(void* ctx);
mapp_status mapp_line_search(const mapp_min_params* params, double dfa, double max_a, int init_flag,
                             double* prev_val, mapp_line_energy_fn energy, mapp_line_reset_fn reset, void* ctx,
                             double* energy_io, double* alpha, int* status);
/* Kernel timing with HIP events on the system's stream (no counterpart in
 * the reference, which has no profiler). enable != 0 starts recording and
 * clears the totals; read returns the summed duration and launch count of a
 * kernel group: "eam_rho", "eam_force", "build_list", "cell_sort",
 * "integrate", "eam_energy". */
mapp_status mapp_md_profile(mapp_md* md, int enable);
mapp_status mapp_md_profile_read(mapp_md* md, const char* kernel, double* total_ms, int64_t* launches);
/* Launch-shape knobs (no counterpart in the reference): neighbors resolved
 * per batch in the density pass (1, 2, 4, 8) and the force pass (1, 2, 4, 8;
 * the LDS force pass uses at most 4 in the integrator's steps, which sum no
 * energy or virial, and at most 2 otherwise). Any value gives the same
 * results to rounding. */
mapp_status mapp_md_set_batch(mapp_md* md, int rho_batch, int force_batch);
/* Where the passes of a one-element table read their spline rows (no
 * counterpart in the reference): on != 0 keeps a window of rows in LDS and
 * runs one persistent 1024-thread workgroup per CU (the default when the
 * device has the LDS); 0 reads the global tables. *in_use (may be NULL)
 * reports the mode in effect. */
mapp_status mapp_md_set_lds(mapp_md* md, int on, int* in_use);
/* First spline interval of the LDS row windows of the density and force
 * passes (rows from there to nr-2 are in LDS; -1 when the LDS passes are not
 * available). No counterpart in the reference. */
mapp_status mapp_md_lds_window(mapp_md* md, int* rho_first_row, int* force_first_row);
/* AtomsMD::create_T (src/atoms_md.cpp:130-160): Maxwell velocities from the
 * Park-Miller / Marsaglia stream of src/random.cpp:7-68 seeded with seed. */
mapp_status mapp_create_T(int natms, const uint8_t* elem, const double* masses, double kB, double T, int seed,
                          double* v);

/* ---------------- DMD (Gaussian-smeared EAM) ----------------
 * Replaces AtomsDMD + DynamicDMD::init (src/dynamic_dmd.cpp:34-57) with
 * ForceFieldEAMDMD (src/ff_eam_dmd.cpp:17-106). ff must be built with dmd=1.
 * Sites carry c_dim components: elem/alpha/c are [nsites][c_dim]; c < 0
 * marks an absent component (src/atoms_dmd.cpp:20-24). nquad Hermite points
 * (src/atoms_dmd.cpp:19). masses[ne] enter mu_0 = 1.5 kT (2 ln lambda - 1 -
 * ln pi), lambda = hP / sqrt(2 pi m kT) (src/ff_dmd.cpp:166-168). */
typedef struct mapp_dmd mapp_dmd;
mapp_status mapp_dmd_create(mapp_eam* ff, int nsites, int c_dim, int nquad, const double* H, const double* x,
                            const uint8_t* elem, const double* alpha, const double* c, double skin, double kB,
                            double T, double hP, const double* masses, mapp_dmd** out);
void mapp_dmd_destroy(mapp_dmd* d);
mapp_status mapp_dmd_set_state(mapp_dmd* d, const double* x, const double* alpha, const double* c);
mapp_status mapp_dmd_get_state(mapp_dmd* d, double* x, double* alpha, double* c);
/* A new box for the DMD system (atoms->H and B, src/atoms.cpp:187-194), with
 * the positions x in caller order (NULL keeps the Cartesian positions; alpha
 * and c are kept). Rebuilds the grid and the component lists. A box narrower
 * than half the list cut-off leaves the system unchanged (MAPP_ERR_CAPACITY). */
mapp_status mapp_dmd_set_box(mapp_dmd* d, const double* H, const double* x);
/* ForceFieldDMD::force_calc_static (src/ff_dmd.cpp:429-445) with
 * ForceFieldEAMDMD::__force_calc_static (src/ff_eam_dmd.cpp:1284-1340): the
 * free energy, F_H (Voigt) and entropy sum (out8 as mapp_dmd_derivative) of
 * the state the last mapp_dmd_c_d call evaluated, without new passes (as the
 * reference re-sums what calc_mu left); forces are not recomputed.
 * MAPP_ERR_STATE if x, alpha or c changed since that call. */
mapp_status mapp_dmd_force_calc_static(mapp_dmd* d, double* out8);
/* DynamicDMD::update (decide at src/dynamic_dmd.cpp:103-126) */
mapp_status mapp_dmd_update(mapp_dmd* d, int* rebuilt);
/* ForceFieldDMD::new_derivative<with_c> (src/ff_dmd.cpp:285-366) over
 * ForceFieldEAMDMD::__force_calc (src/ff_eam_dmd.cpp:126-360).
 * out8 = free energy, F_H Voigt[6], entropy sum S. f [nsites][3],
 * f_alpha / mu [nsites][c_dim] (any may be NULL). With with_c == 0, mu is
 * the raw f_c vector of __force_calc; with with_c != 0 it is the chemical
 * potential and ext_mu[ne] enters the free energy. */
mapp_status mapp_dmd_derivative(mapp_dmd* d, int with_c, const double* ext_mu, double* f, double* f_alpha,
                                double* mu, double* out8);
/* ForceFieldDMD::new_value<with_c> (src/ff_dmd.cpp:252-281) over
 * ForceFieldEAMDMD::__energy_calc (src/ff_eam_dmd.cpp:835-953) */
/* Launch-shape knob (no counterpart in the reference): list entries per batch
 * of the quadrature passes, 1 (one pair at a time), 2 or 4 (default). Any
 * value gives the same results to rounding. */
mapp_status mapp_dmd_set_batch(mapp_dmd* d, int batch);
/* Table evaluation of the quadrature passes behind mapp_dmd_derivative,
 * mapp_dmd_value and the DMD minimisers (ForceFieldEAMDMD::__force_calc /
 * __energy_calc, src/ff_eam_dmd.cpp:126-360,835-953; no such switch in the
 * reference, which is FP64 throughout). mode 0 (the default) reads the FP64
 * coefficient rows. mode 1 evaluates each pair spline in FP32 from (value,
 * slope) pairs of the knots of the table's distinct functions, the top
 * knots of which every workgroup holds in LDS; where a table function is not
 * the Hermite cubic of its knots, mode 1 reads the FP32 rows. mode 2 reads
 * FP32 copies of the coefficient rows. The FP32 modes keep the quadrature
 * moments and every sum in FP64 and apply to batches 2 and 4 (batch 1 stays
 * FP64). The concentration-rate path (mapp_dmd_c_d, mapp_dmd_J,
 * mapp_dmd_newton, mapp_dmd_bdf_run) always reads the FP64 rows. */
mapp_status mapp_dmd_set_fp32(mapp_dmd* d, int mode);

/* The first knot of table mode 1's LDS window (the knots below it are read
 * from device memory), or -1 where mode 1 reads the FP32 rows. No
 * counterpart in the reference: it reports this library's table layout. */
mapp_status mapp_dmd_knot_window(mapp_dmd* d, int* first_knot);
/* ForceFieldDMD::c_d_calc (src/ff_dmd.cpp:449-456) with
 * ForceFieldEAMDMD::__c_d_calc (src/ff_eam_dmd.cpp:1359-1473): the
 * concentration rates dc/dt of every component at the current state, over the
 * second list of NeighborDMD::create_2nd_list (src/neighbor_dmd.cpp:224-273:
 * same element, r < r_crd[e]). r_crd and C (the exchange width, zeta) are the
 * per-element parameters of ff_eam_setfl/fs (src/ff_eam_dmd.cpp:1734-1790).
 * c_d [natms][c_dim] in caller order (may be NULL: kept on the device);
 * c_d_norm = max |c_d| (may be NULL). */
mapp_status mapp_dmd_c_d(mapp_dmd* d, const double* r_crd, const double* C, double* c_d, double* c_d_norm);
/* ForceFieldEAMDMD::__J (src/ff_eam_dmd.cpp:1475-1575): Ax = (d c_d / d c) x,
 * the Jacobian of the concentration rates applied to x, at the state and
 * r_crd / C of the last mapp_dmd_c_d call (as __J reads the rho_phi, ddE and
 * chemical potentials c_d_calc left). x and Ax [natms][c_dim] in caller
 * order. MAPP_ERR_STATE if x, alpha or c changed since that call, or if it
 * ran through the surrogate networks. */
mapp_status mapp_dmd_J(mapp_dmd* d, const double* x, double* Ax);
/* DAEImplicit::newton (src/dae_imp.cpp:146-206) with update_c
 * (src/dae_imp.cpp:89-138) and __GMRES (src/gmres.h:40-60, src/gmres.cpp):
 * solve c / beta + a - c_d(c) = 0 for c from the current c (the implicit
 * equation of one BDF step: beta is the step's coefficient, a its history
 * term, [natms][c_dim] in caller order), Newton steps of GMRES over
 * mapp_dmd_J with at most max_ngmres_iters (1..8) products each. Tolerances
 * from a_tol sqrt(natms c_dim) and err_fac as DAEImplicit sets them.
 * *converged = 1: c holds the solution (read it with mapp_dmd_get_state);
 * 0: c is restored to its value at the call. *iters = Newton steps as the
 * reference counts them. An exactly zero residual gives no step (the
 * reference divides by zero there). */
mapp_status mapp_dmd_newton(mapp_dmd* d, const double* r_crd, const double* C, double beta, const double* a,
                            double a_tol, double err_fac, int max_nnewton_iters, int max_ngmres_iters,
                            int* converged, int* iters);
/* ForceFieldEAMDMD::__prepJ_n_res + __J (src/ff_eam_dmd.cpp:364-831), via
 * ForceFieldDMD::prepJ_n_res / Jnew (src/ff_dmd.cpp:495-525): at the current
 * state and fixed c, (Jdx, Jdalpha) = -(d f / d(x, alpha)) (dx, dalpha), f
 * the gradient per unit c that calc_mu leaves in f and f_alpha. dx, Jdx and f
 * [natms][3], dalpha, Jdalpha and f_alpha [natms][c_dim], caller order; f and
 * f_alpha may be NULL. FP64 table rows and second derivatives from the
 * Hermite moments xi^0..xi^4. */
mapp_status mapp_dmd_Jxa(mapp_dmd* d, const double* dx, const double* dalpha, double* Jdx, double* Jdalpha,
                         double* f, double* f_alpha);
/* The Newton stage of DAE::min_error_false (src/dae.cpp:294-325) at fixed c:
 * while |f| > a_tol sqrt(3 natms + natms c_dim) and fewer than
 * max_nnewton_iters steps, GMRES (at most max_ngmres_iters <= 8 products,
 * tolerance 0.005 a_tol sqrt(...), zero start) solves J h = f with the J of
 * mapp_dmd_Jxa; x and alpha take h whole unless an alpha would turn negative
 * (then half the largest fraction that keeps them non-negative,
 * MinDMDHandler::max_alpha_lcl_alpha); the list follows. *steps = Newton
 * steps taken, *res = |f| at the end (either may be NULL). */
mapp_status mapp_dmd_min_newton(mapp_dmd* d, double a_tol, int max_nnewton_iters, int max_ngmres_iters, int* steps,
                                double* res);
/* DAEBDF::run (src/dae_bdf.cpp:103-250, the coefficient algebra of
 * src/dae_bdf.h:103-546): variable-order (1..5) BDF integration of the
 * concentrations from the current state over t_tot, each step's implicit
 * equation solved by mapp_dmd_newton, order and step chosen as the reference
 * chooses them. Positions and widths stay fixed between relaxations; with
 * nreset > 0 they are relaxed at the start, every nreset steps when the
 * gradient error exceeds a_tol, whenever it doubles, and at the end when the
 * step count is not a multiple of nreset, by DAE::min_error_false
 * (src/dae.cpp:273-330): the GPU CG with LineSearchBrent, then the Newton
 * stage of mapp_dmd_min_newton with max_nnewton_iters and max_ngmres_iters
 * (err_mins counts the relaxations whose Newton stage took a step). nreset =
 * 0: no relaxation (the reference divides by nreset). With ntally > 0, rows
 * (step, time, free energy) at step 0, every ntally steps and at the last
 * step go to tally [max_tally_rows][3]; *tally_rows counts the rows written. */
typedef struct {
    double a_tol, min_dt;
    int max_nsteps, max_nnewton_iters, max_ngmres_iters, nreset, ntally, relax_max_iter;
} mapp_bdf_params;
typedef struct {
    int nsteps, max_q;
    double t, max_dt, min_dt;
    int nonlin_acc, nonlin_rej, intpol_acc, intpol_rej, integ_acc, integ_rej, err_mins;
} mapp_bdf_report;
mapp_status mapp_dmd_bdf_run(mapp_dmd* d, const double* r_crd, const double* C, double t_tot,
                             const mapp_bdf_params* params, mapp_bdf_report* report, double* tally,
                             int max_tally_rows, int* tally_rows);

/* CFG (AtomEye) files of MD systems. mapp_cfg_md_read replaces
 * ImportCFGMD::operator() (src/import_cfg.cpp:495-604) with
 * ImportCFG::read_header / read_body_std / read_body_ext
 * (src/import_cfg.cpp:32-480): standard files (rows of mass, symbol, s, s_d)
 * and extended ones (mass and symbol lines, rows of entry_count values),
 * positions wrapped into [0, 1) and returned Cartesian, elements numbered in
 * order of first appearance (the first mass of a symbol kept), velocities
 * v = s_d R H unless the file says .NO_VELOCITY. (has_vel 0), the box in
 * lower-triangular form times A. mapp_cfg_md_get fills H [3][3], x and v
 * [natms][3] (v ignored without velocities), elem [natms] and masses
 * [nelems]; any output may be NULL. */
typedef struct mapp_cfg_md mapp_cfg_md;
mapp_status mapp_cfg_md_read(const char* path, mapp_cfg_md** out);
mapp_status mapp_cfg_md_info(const mapp_cfg_md* cfg, int* natms, int* nelems, int* has_vel);
mapp_status mapp_cfg_md_get(const mapp_cfg_md* cfg, double* H, double* x, double* v, uint8_t* elem,
                            double* masses);
const char* mapp_cfg_md_element(const mapp_cfg_md* cfg, int i);
void mapp_cfg_md_free(mapp_cfg_md* cfg);
/* ExportCFGMD::write_header / write_body_sort (src/export_cfg_md.cpp:26-99)
 * for host arrays: atoms in the given order, positions as wrapped fractional
 * coordinates, velocities (v NULL: .NO_VELOCITY.) as s_d = v H^-1, then naux
 * auxiliary columns aux [natms][naux] named aux_names (the reference's
 * extra_vecs, one name per column); names holds nelems symbols. */
mapp_status mapp_cfg_md_write(const char* path, int natms, const double* H, const double* x, const uint8_t* elem,
                              const double* v, int naux, const double* aux, const char* const* aux_names,
                              int nelems, const double* masses, const char* const* names);
/* md.export_cfg (src/export_cfg_md.cpp:26-216) of a handle: its current
 * positions, box and elements, with velocities when with_vel is nonzero,
 * atoms in caller order (the reference's sort=True). */
mapp_status mapp_md_export_cfg(mapp_md* md, const char* path, int nelems, const double* masses,
                               const char* const* names, int with_vel);

/* Extended CFG (AtomEye) files of DMD systems. mapp_cfg_dmd_read replaces
 * ImportCFGDMD::operator() (src/import_cfg.cpp:669-854) with
 * ImportCFG::read_header / read_body_ext (src/import_cfg.cpp:32-356): sites
 * with fractional positions (wrapped into [0, 1)), one alpha and one c per
 * element (c = -1: absent), checked as the reference checks them, the
 * present elements of a site compacted into c_dim components (absent
 * components: element 0, c = -1, alpha = 0), the box in lower-triangular
 * form. mapp_cfg_dmd_get fills H [3][3] (rows = box vectors), x
 * [natms][3] (Cartesian), elem / alpha / c [natms][c_dim] and masses
 * [nelems]; mapp_cfg_dmd_element gives the symbol of element i (NULL if out
 * of range). */
typedef struct mapp_cfg mapp_cfg;
mapp_status mapp_cfg_dmd_read(const char* path, mapp_cfg** out);
mapp_status mapp_cfg_dmd_info(const mapp_cfg* cfg, int* natms, int* c_dim, int* nelems);
mapp_status mapp_cfg_dmd_get(const mapp_cfg* cfg, double* H, double* x, uint8_t* elem, double* alpha, double* c,
                             double* masses);
const char* mapp_cfg_dmd_element(const mapp_cfg* cfg, int i);
void mapp_cfg_dmd_free(mapp_cfg* cfg);
/* ExportCFGDMD::write_header / write_body (src/export_cfg_dmd.cpp:25-170) for
 * host arrays laid out as mapp_cfg_dmd_get gives them (x Cartesian, H lower
 * triangular); names holds nelems symbols. */
mapp_status mapp_cfg_dmd_write(const char* path, int natms, int c_dim, const double* H, const double* x,
                               const uint8_t* elem, const double* alpha, const double* c, int nelems,
                               const double* masses, const char* const* names);
/* ExportCFGDMD::write_header / write_body (src/export_cfg_dmd.cpp:25-170): the
 * current state of d as an extended CFG file, sites in caller order, with
 * the masses and symbols of the force field's nelems elements. */
mapp_status mapp_dmd_export_cfg(mapp_dmd* d, const char* path, int nelems, const double* masses,
                                const char* const* names);
/* The gradient per unit c that ForceFieldEAMDMD::calc_mu leaves in f and
 * f_alpha (src/ff_eam_dmd.cpp:1189-1278) at the state of the last
 * mapp_dmd_c_d call: g [natms][3], g_alpha [natms][c_dim] (either may be NULL);
 * it is what DAE::calc_err measures. MAPP_ERR_STATE if the state changed. */
mapp_status mapp_dmd_unit_gradient(mapp_dmd* d, double* g, double* g_alpha);
/* Diagnostic (no counterpart in the reference): the entries the state of the
 * last mapp_dmd_J call holds, list entries (those past the smeared cut hold
 * zeros) and second-list entries (exchange), summed over components. */
mapp_status mapp_dmd_J_entries(mapp_dmd* d, int64_t* entries, int64_t* exchange);
mapp_status mapp_dmd_value(mapp_dmd* d, int with_c, const double* ext_mu, double* fe);
/* NeighborDMD::create_list half list (src/neighbor_dmd.cpp:102-194) as
 * (component i, component j, image) triples in caller indices */
mapp_status mapp_dmd_pairs_count(mapp_dmd* d, int64_t* npairs);
mapp_status mapp_dmd_pairs(mapp_dmd* d, int32_t* i, int32_t* j, int8_t* img);
/* Smeared pair quantities at npts (r, alpha_ij) points for elements (ei, ej):
 * out[npts][9] = rho_phi[3], drho_phi_dr[3], drho_phi_dalpha[3] in the order
 * r*phi[ei][ej], r*rho[ei][ej], r*rho[ej][ei] (src/ff_eam_dmd.cpp:195-253). */
mapp_status mapp_dmd_smear(mapp_dmd* d, int ei, int ej, int npts, const double* r, const double* alpha,
                           double* out);
/* MinCGDMD over x and alpha with c fixed (src/min_cg_dmd.h:633-700) and the
 * backtracking line search (src/ls.h:711-770). Stops on an energy decrement
 * below e_tol (the reference's criterion) or when max |gradient component|
 * < grad_tol (grad_tol <= 0 disables). out4 = fe0, fe, |grad|_2, |grad|_max;
 * status: 1 zero gradient, 2 not downhill, 3 grad_tol reached, 4 line search
 * failed (minimum alpha reached, or for golden / Brent no bracket found),
 * 5 e_tol reached, 0 iteration limit; the same codes for every minimiser and
 * line search of mapp_dmd_min_lbfgs and mapp_dmd_min_run. */
mapp_status mapp_dmd_min_cg(mapp_dmd* d, int max_iter, double e_tol, double max_dx, double max_dalpha,
                            double grad_tol, int* iters, int* status, double* out4);
/* MinLBFGSDMD::__run over x and alpha with c fixed (src/min_l-bfgs_dmd.h:113-210,
 * memory m; MinDMDHandler::ls_prep at src/min_cg_dmd.h:432-458) and the
 * back-tracking line search with its defaults (src/ls.h:711-769). Stops,
 * out4 and status as mapp_dmd_min_cg. */
mapp_status mapp_dmd_min_lbfgs(mapp_dmd* d, int m, int max_iter, double e_tol, double max_dx, double max_dalpha,
                               double grad_tol, int* iters, int* status, double* out4);
/* mapp_dmd_min_cg (lbfgs = 0) or mapp_dmd_min_lbfgs (lbfgs = 1, memory m) with the
 * line search ls->ls and its parameters ls_c, ls_rho, ls_min_alpha, ls_max_iter,
 * ls_brack, ls_tol, ls_zeps as mapp_md_minimize reads them: the reference's
 * dmd.min_cg / dmd.min_lbfgs accept ls_bt, ls_golden or ls_brent
 * (src/min_cg_dmd.cpp:172, src/min_l-bfgs_dmd.cpp:76). */
mapp_status mapp_dmd_min_run(mapp_dmd* d, int lbfgs, int m, int max_iter, double e_tol, double max_dx,
                             double max_dalpha, double grad_tol, const mapp_min_params* ls, int* iters, int* status,
                             double* out4);
mapp_status mapp_dmd_profile(mapp_dmd* d, int enable);
mapp_status mapp_dmd_profile_read(mapp_dmd* d, const char* kernel, double* total_ms, int64_t* launches);
/* XMath::quadrature_hg (src/xmath.cpp:440-589): n-point physicists'
 * Gauss-Hermite nodes and weights, sum w = sqrt(pi) */
mapp_status mapp_gauss_hermite(int n, double* x, double* w);

/* ---------------- learned surrogate of the smeared functions ----------------
 * No counterpart in the reference: one MLP per ordered element pair (a, b),
 * 3 -> 64 -> 64 -> 6 with tanh, maps (r, alpha_a, alpha_b) to what
 * ForceFieldEAMDMD's quadrature computes for that pair
 * (src/ff_eam_dmd.cpp:195-253): (g_phi, g_rho_ab, dg_phi/dr, dg_rho/dr,
 * dg_phi/dalpha, dg_rho/dalpha), alpha = sqrt(alpha_a^2 + alpha_b^2).
 * box4 = r_lo, r_hi, alpha_lo, alpha_hi: the training box; pairs outside it
 * keep the quadrature. Training and inference run on BF16 MFMA with FP32
 * accumulation. Network index = a * ne + b. */
typedef struct mapp_surr mapp_surr;
mapp_status mapp_surr_create(int ne, const float* box4, int seed, mapp_surr** out);
void mapp_surr_destroy(mapp_surr* s);
/* nsamples quadrature samples per network from d's force field and rule,
 * then `steps` Adam steps of `batch` samples per network (cosine rate lr ->
 * lr/20). loss2 = mean squared error of the normalized outputs at the first
 * and the last step. */
mapp_status mapp_surr_train(mapp_surr* s, mapp_dmd* d, int nsamples, int steps, int batch, double lr,
                            double* loss2);
/* MSE loss and its gradient (flat parameter order W1[64][4] (r, alpha_a,
 * alpha_b, bias), W2[64][64], b2[64], W3[6][64], b3[6]) for normalized inputs
 * x[n][3] and targets t[n][6], n a multiple of 64, through the training kernel */
mapp_status mapp_surr_loss_grad(mapp_surr* s, int net, int n, const float* x, const float* t, float* grad,
                                double* loss);
mapp_status mapp_surr_get_params(mapp_surr* s, int net, float* p);
mapp_status mapp_surr_set_params(mapp_surr* s, int net, const float* p);
/* output de-normalization y = net * scale + shift, per output */
mapp_status mapp_surr_get_norm(mapp_surr* s, int net, float* scale, float* shift);
mapp_status mapp_surr_set_norm(mapp_surr* s, int net, const float* scale, const float* shift);
/* outputs [n][6] for raw inputs rin[n][3] = (r, alpha_a, alpha_b), inference kernel */
mapp_status mapp_surr_infer(mapp_surr* s, int net, int n, const float* rin, float* out);
/* DMD evaluations use s (NULL: the quadrature again). Kernel group "surr_infer"
 * of mapp_dmd_profile_read times the request staging and inference. */
mapp_status mapp_dmd_use_surrogate(mapp_dmd* d, mapp_surr* s);
/* network evaluations and quadrature-covered slots of the last evaluation */
mapp_status mapp_dmd_surrogate_coverage(mapp_dmd* d, int64_t* net_evals, int64_t* quadrature_slots);

/* ---------------- batched grand-canonical exchange (GCMC) ----------------
 * Many independent periodic cells, each running its own chain of SGCMC
 * insert/delete attempts of one gas element with EAM exchange energies:
 * SGCMC::xchng's attempt loop (src/sgcmc.cpp:145-230) with SGCMC::attmpt
 * (src/sgcmc.cpp:255-327) over ForceFieldEAM::init_xchng / pre_xchng_energy /
 * xchng_energy / post_xchng_energy (src/ff_eam.cpp:323-578); constants of
 * GCMC::init (src/gcmc.cpp:86-116). Cell c holds natms[c] atoms (x and elem
 * concatenated over cells, H[c] a row-major lower-triangular 3x3) and room
 * for cap[c] (cap NULL: natms + max(32, natms/8)); seeds[c] starts its
 * Park-Miller stream (src/random.cpp:7-68). masses[ne]; only the gas mass
 * enters (lambda, sigma). */
typedef struct mapp_gcmc mapp_gcmc;
mapp_status mapp_gcmc_create(mapp_eam* ff, int ncells, const int32_t* natms, const double* H, const double* x,
                             const uint8_t* elem, const int32_t* cap, int gas_elem, double mu, double T, double kB,
                             double hP, const double* masses, const int32_t* seeds, mapp_gcmc** out);
void mapp_gcmc_destroy(mapp_gcmc* g);
/* nattempts attempts in every cell, all cells in one persistent launch.
 * decisions/dU are [ncells][nattempts] (either may be NULL): 1 accepted
 * insertion, -1 rejected insertion, 2 accepted deletion, -2 rejected deletion,
 * 0 deletion drawn with no gas atom (src/sgcmc.cpp:280-281; dU 0). An
 * insertion accepted in a full cell stops that cell's chain and returns
 * MAPP_ERR_CAPACITY. */
mapp_status mapp_gcmc_attempts(mapp_gcmc* g, int nattempts, int8_t* decisions, double* dU);
/* Tile tuner (no counterpart in the reference): cells of the size classes
 * [0,256) [256,512) [512,1024) [1024,2048) [2048,inf) atoms run on one wave
 * below a threshold class and on four waves (one workgroup) from it up. Each
 * threshold 0..5 is timed with nattempts attempts on a scratch copy of the
 * whole batch and the fastest is kept; the chains are not advanced.
 * waves[5] reports the chosen shape per class, ms[6] the launch time of each
 * threshold. */
mapp_status mapp_gcmc_tune(mapp_gcmc* g, int nattempts, int32_t* waves, double* ms);
/* sets the per-class shape directly (1 or 4 waves per cell); results do not depend on it */
mapp_status mapp_gcmc_set_waves(mapp_gcmc* g, const int32_t* waves);
/* on != 0: the attempt launch is captured once in a HIP graph and replayed */
mapp_status mapp_gcmc_use_graph(mapp_gcmc* g, int on);
mapp_status mapp_gcmc_counts(mapp_gcmc* g, int32_t* natms, int32_t* ngas);
/* The exchange hooks of one trial, for a caller that keeps its own attempt
 * loop and random stream (SGCMC::attmpt, src/sgcmc.cpp:255-327, or PGCMC):
 * mapp_gcmc_xchng_energy is ForceFieldEAM::pre_xchng_energy + xchng_energy
 * (src/ff_eam.cpp:323-446) on cell `cell`: mode 1 an insertion of the gas
 * element at fractional s[3] (in [0, 1); the reference's s_buff), mode 2 the
 * deletion of the igas-th gas atom in atom order (the reference's igas);
 * *dU is the exchange energy in FP64. The trial is kept until the next
 * call. mapp_gcmc_xchng_commit applies it as accepted:
 * ForceFieldEAM::post_xchng_energy (src/ff_eam.cpp:451-483) with
 * SGCMC::ins_succ / del_succ (src/sgcmc.cpp:331-444) -- densities and
 * embedding energies of the touched atoms, the new atom appended with
 * velocity v[3] (the reference's gaussian() * sigma) or the last atom moved
 * into the deleted one's place. A rejected trial is simply not committed.
 * init_xchng (src/ff_eam.cpp:509-578) runs when the cells are created.
 * MAPP_ERR_STATE when there is no trial to commit (none made, already
 * committed, or mapp_gcmc_attempts ran since); MAPP_ERR_CAPACITY when an
 * insertion finds the cell full; MAPP_ERR_ARG when igas is not below the gas
 * count. */
mapp_status mapp_gcmc_xchng_energy(mapp_gcmc* g, int cell, int mode, const double* s, int igas, double* dU);
mapp_status mapp_gcmc_xchng_commit(mapp_gcmc* g, const double* v);
/* one cell's atoms in the reference's order (appended on insertion, last
 * atom moved into the hole on deletion): Cartesian x [n][3], elem, rho, F
 * (the exchange bookkeeping of ForceFieldEAM), v, and its stream's seed */
mapp_status mapp_gcmc_get_cell(mapp_gcmc* g, int cell, double* x, uint8_t* elem, double* rho, double* F, double* v,
                               int32_t* rng_seed);
mapp_status mapp_gcmc_profile(mapp_gcmc* g, int enable);
mapp_status mapp_gcmc_profile_read(mapp_gcmc* g, const char* kernel, double* total_ms, int64_t* launches);

/* ---------------- batched replica minimisation ---------------- */
/* Many small periodic cells ("replicas", each its own box H [9], atoms and
 * elements, concatenated in replica order), each minimised over its atom
 * positions by MinCG::__run (src/min_cg.h:453-533) with MinMDHandler's
 * ls_prep / F / F_reset (src/min_cg.h:231-278) and LineSearchBackTrack::min
 * (src/ls.h:711-769, init_flag 1), on the EAM energy and forces of
 * ForceFieldEAM (src/ff_eam.cpp:99-318), with the list rebuilt when an atom
 * has moved half the skin. The reference runs one system per MinCG; a batch
 * of replicas (the minimisations a replica driver issues one after another)
 * runs here in one persistent launch, one wave or one workgroup per replica.
 * FP64 throughout. Every box must be at least r_c + skin deep along each box
 * vector. */
typedef struct mapp_replica mapp_replica;
mapp_status mapp_replica_create(mapp_eam* ff, int nreplicas, const int32_t* natms, const double* H, const double* x,
                                const uint8_t* elem, double skin, mapp_replica** out);
void mapp_replica_destroy(mapp_replica* r);
/* every replica minimised from its current positions; per replica (any
 * may be NULL): steps taken, the reference's status (0..5 as MinCG's err;
 * 101 a list row over capacity, 102 an atom beyond two box lengths, both
 * reported as errors too), energies at the start and end, energy evaluations */
mapp_status mapp_replica_min_cg(mapp_replica* r, int max_nsteps, double e_tol, double max_dx, double ls_c,
                                double ls_rho, double ls_min_alpha, int32_t* steps, int32_t* status, double* e0,
                                double* e, int32_t* evals);
/* launch-shape tuner (no counterpart in the reference): one wave per replica
 * below a size-class threshold, a 256-thread workgroup from there up; every
 * threshold timed on a copy of the batch, the fastest kept. waves [5] per
 * class (atoms < 256, < 512, < 1024, < 2048, more), ms [6] per threshold */
mapp_status mapp_replica_tune(mapp_replica* r, int max_nsteps, double e_tol, double max_dx, double ls_c, double ls_rho,
                              double ls_min_alpha, int32_t* waves, double* ms);
mapp_status mapp_replica_set_waves(mapp_replica* r, const int32_t* waves);
/* capture the launch in a hipGraph and replay it on later calls with the same parameters */
mapp_status mapp_replica_use_graph(mapp_replica* r, int on);
/* positions of one replica after the last minimisation (not wrapped back into the box) */
mapp_status mapp_replica_get_x(mapp_replica* r, int replica, double* x);
mapp_status mapp_replica_profile(mapp_replica* r, int enable);
mapp_status mapp_replica_profile_read(mapp_replica* r, double* total_ms, int64_t* launches);

/* Diagnostic (no counterpart in the reference): copies `bytes` between two
 * device buffers `reps` times with a plain streaming kernel and reports the
 * achieved bandwidth (read + write bytes per second / 1e9). Used to calibrate
 * the HBM traffic counters and to measure the device's copy bandwidth. */
mapp_status mapp_diag_copy(size_t bytes, int reps, double* gbps);

#ifdef __cplusplus
}
#endif
#endif
