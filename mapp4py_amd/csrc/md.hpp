// md.hpp -- device-resident MD system: atoms, neighbor list, tabulated EAM.
#pragma once
#include <functional>
#include <memory>
#include <vector>

#include "box.hpp"
#include "common.hpp"
#include "tables.hpp"

namespace mapp {

constexpr int kMaxElems = 8;
constexpr int kIndexBits = 25;                 // neighbor entry: j in the low bits,
constexpr uint32_t kIndexMask = (1u << kIndexBits) - 1u;  // image code in the high 7
constexpr int kZeroImage = 62;                  // code of the (0,0,0) shift

// Kernel argument block describing the EAM tables on the device.
struct EamArgs {
    int ne, nr, nrho;
    double dr_inv, drho_inv, rho_max;
    float dr_invf;
    double cut_sq[kMaxElems * kMaxElems];
    const float4* tabA;   // [ne*ne][nr]: density at i due to j (rho[tj][ti])
    const float4* tabB;   // [ne*ne][nr][3]: (rho'[tj][ti] d1..d3, r*phi c0), (r*phi c1..c3, 0), (rho'[ti][tj] d1..d3, 0)
    const float4* tabC;   // one element only: [nr][2] = (rho' d1..d3, r*phi c0), (r*phi c1..c3, 0)
    const double* Ftab;   // [ne][nrho][4]
};

// ForceFieldEAM with its tables uploaded once (pair splines in FP32, the
// embedding function in FP64).
class EamForceField {
public:
    explicit EamForceField(EamTables t);
    const EamTables& tables() const { return t_; }
    EamArgs args() const;
private:
    EamTables t_;
    DevBuf<float4> tabA_, tabB_, tabC_;
    DevBuf<double> Ftab_;
};

// Box, cell grid and search radius of one neighbor build.
struct CellGrid {
    int nc[3];
    int R[3];
    int ncells;
    double cut_s[3];
};

// 16-byte position records: coordinates as 32-bit integers modulo 2^32 on a
// grid of q (a power of two, small enough that every listed separation stays
// below 2^31 q), the element (density pass) or F'(rho) (force pass of
// single-element tables) in the fourth word. Differences of two records, less
// the integer image shift n . cH, are exact integers, so separations carry
// about 1e-8 A.
struct Rec16Frame {
    double q, inv_q;
    unsigned cH[9];  // rows of H on the grid, modulo 2^32: cH[3 * row + dim]
    double cut_sq_u[kMaxElems * kMaxElems];  // squared cut-offs in grid units (cut^2 / q^2)
    double knot_u;                           // spline knots per grid unit (q / dr)
    float inv_qf;                            // 1/q
    const uint4* shift;                      // [128] image code -> grid shift (x, y, z) modulo 2^32
};

// second velocity-Verlet half kick carried out by the force pass
struct KickArgs {
    double dt2;
    double mass[kMaxElems];
    int on;
};

struct MDReport {
    double e[7];   // energy, F_H Voigt (xx,xy,xz,yy,yz,zz)
};

// MinCG / MinLBFGS with LineSearchBackTrack, LineSearchGoldenSection or
// LineSearchBrent over the atom positions (no box degrees of freedom,
// affine = False)
struct MinParams {
    int lbfgs = 1;        // 0: conjugate gradients, 1: L-BFGS
    int m = 2;            // L-BFGS memory (pairs of s, y)
    double e_tol = 1e-8;
    double max_dx = 1.0;
    int max_nsteps = 1000;
    double ls_c = 0.4, ls_rho = 0.5, ls_min_alpha = 0.0;      // back-tracking (ls = 0)
    int ls = 0;                                               // 0 back-track, 1 golden section, 2 Brent
    int ls_max_iter = 5, ls_brack = 1;                         // golden section and Brent
    double ls_tol = 2.1073424255447017e-08, ls_zeps = 2.220446049250313e-19;
};

// status: 0 line search succeeded (not returned), 1 not a downhill direction,
// 2 zero gradient, 3 minimum step length reached, 4 energy tolerance
// reached, 5 maximum number of steps, 8 bracketing found no lower energy
// (the reference's err_msgs order)
struct MinReport {
    int steps = 0, status = 0, evals = 0, force_calcs = 0, rebuilds = 0;
    double e0 = 0.0, e = 0.0, f_norm = 0.0, f_max = 0.0;
};
constexpr int kMaxLbfgsM = 64;

// A grand-canonical exchange run by nvt_run after every nevery-th first half
// step in place of the skin check (MDMuVT::run); it may replace the atoms
// through reset_atoms.
struct ExchangeHook {
    int nevery = 0;
    std::function<void()> run;
};

// largest system of MDSystem::value_noalloc: its scratch rows take 24 bytes
// per atom (24 MiB per GPU)
constexpr int kNoAllocMaxAtoms = 1 << 20;

// AtomsMD + NeighborMD + DynamicMD + ForceFieldEAM state resident in HBM.
// Atoms are kept in cell order ("slots"); id[slot] is the caller's index.
class MDSystem {
public:
    MDSystem(std::shared_ptr<EamForceField> ff, int natms, const double* H, const double* x,
             const uint8_t* elem, double skin);
    ~MDSystem();

    int natms() const { return n_; }
    // a new box H (lower triangular, rows = box vectors) and, unless x is
    // null, new positions; rebuilds the grid and the list
    void set_box(const double* H, const double* x);
    // the atoms replaced by natms new ones in the same box (caller order; v
    // may be null: zero), then the cell sort and the list
    void reset_atoms(int natms, const double* x, const uint8_t* elem, const double* v);
    void get_elem(uint8_t* elem) const;        // caller order
    void set_x(const double* x_host);          // caller order
    void get_x(double* x_host);
    void set_v(const double* v_host);
    void get_v(double* v_host);
    void get_f(double* f_host);
    void get_rho(double* rho_host);
    // ExportCFGMD of the current state (caller order, with or without the
    // velocities); one name and mass per element of the force field
    void export_cfg(const char* path, int nelems, const double* masses, const std::vector<std::string>& names,
                    bool with_vel);
    // DynamicMD::update: returns true if the list was rebuilt
    bool update();
    void rebuild();
    // ForceFieldMD::force_calc -> energy + F_H
    MDReport force_calc();
    double value();
    // value() with the passes' forces and densities in the library's static
    // scratch rows (at most kNoAllocMaxAtoms atoms): no device allocation
    double value_noalloc();
    int64_t half_pairs_count();
    void half_pairs(int32_t* i, int32_t* j, int8_t* img);
    void list_stats(int* maxn, int64_t* nfull);
    // MDNVT::run; masses per element; returns rebuild count. hook: an
    // exchange after every hook->nevery-th first half step; hist_T: T_part
    // after each step
    int nvt_run(const double* masses, double kB, double T, double dt, int nsteps, MDReport* rep, double* T_part,
                const ExchangeHook* hook = nullptr, double* hist_T = nullptr);
    // MDMuVT::run (src/md_muvt.cpp:122-280): NVT with nattempts exchange
    // attempts of element gas at chemical potential mu after every nevery-th
    // first half step, the chain's random stream seeded once per run; counts:
    // accepted insertions, accepted deletions, insertion trials, deletion
    // trials; hist [nsteps][2]: T_part and gas atoms after each step; returns
    // the list builds (rebuilds and exchanges)
    int muvt_run(const double* masses, double kB, double hP, double T, double dt, double mu, int gas, int seed,
                 int nevery, int nattempts, int nsteps, MDReport* rep, double* T_part, int* counts, double* hist);
    // MDNST::run (src/md_nst.cpp:376-460): S9 the target stress (row-major,
    // lower triangle read, NaN = free), tau9 the box relaxation times (nullptr:
    // 1000 dt), B_ref and the reference volume re-taken every nreset steps (0:
    // never); H9 the final box, hist [nsteps][14] (T_part, pe, H00, H10, H11,
    // H20, H21, H22, S_pe xx yx zx yy zy zz) when not null; returns the list
    // rebuilds
    int nst_run(const double* masses, double kB, double T, double dt, const double* S9, const double* tau9,
                int nreset, int nsteps, MDReport* rep, double* T_part, double* H9, double* hist);
    // MinCG::run / MinLBFGS::run (src/min_cg.h:453-533, src/min_l-bfgs.h:114-217);
    // tally rows (step, energy): the start and every accepted step
    MinReport minimize(const MinParams& p, std::vector<double>* tally);
    hipStream_t stream() const { return stream_; }
    Profiler& profiler() { return prof_; }
    // launch-shape knobs picked by the tuner: neighbors resolved per batch
    void set_batch(int rho, int force);
    // one-element tables: spline rows in LDS (persistent passes) or read from
    // the global tables; returns whether the LDS passes are in use
    bool set_lds(bool on);
    bool lds() const { return use_lds_; }
    // first spline interval held in LDS by the density and force passes (-1: no LDS passes)
    void lds_window(int* rho_first, int* force_first) const
    {
        *rho_first = lds_ok_ ? rho_m_lo_ : -1;
        *force_first = lds_ok_ ? f_m_lo_ : -1;
    }
    int batch_rho() const { return batch_rho_; }
    int batch_force() const { return batch_force_; }
    void sync();
    int last_rebuilds() const { return nrebuild_; }

private:
    void setup_grid();
    void cell_topology();  // cell ranks and stencil order of grid_'s cell counts
    void refresh_frame();  // image shifts of the records for the current box
    void build_list();
    // records_fresh: rec_ already holds the current positions (written by the skin check or a rebuild)
    // the two EAM passes; f_out / rho_out (nullptr: f_ and rho_) receive the
    // forces and densities
    void force_passes(const KickArgs* kick, bool records_fresh = false, float4* f_out = nullptr,
                      double* rho_out = nullptr, bool ev = true);
    MDReport reduce_report(int nparts);
    void reduce_rows(const double* part, int nrows, int ncols, double* out);
    bool decide();

    std::shared_ptr<EamForceField> ff_;
    int n_;
    double skin_, max_cut_;
    Box box_;
    std::vector<uint8_t> elem_host_;  // caller order
    CellGrid grid_{};
    hipStream_t stream_ = nullptr;
    int nrebuild_ = 0;

    DevBuf<double4> x_, x_alt_;      // x,y,z,type
    DevBuf<double4> s_;              // fractional coordinates at the last build
    DevBuf<double4> x0_;             // positions at the last build
    DevBuf<double4> v_, v_alt_;      // velocities (w unused)
    DevBuf<int> id_, id_alt_;        // slot -> caller index
    DevBuf<int> slot_of_;            // caller index -> slot
    DevBuf<float4> f_;               // forces (w unused)
    DevBuf<double> rho_;
    DevBuf<float4> val_f_;           // forces and densities written by value()'s passes
    DevBuf<double> val_rho_;
    DevBuf<uint4> rec_;              // 16-byte position records (coordinates on the grid, element)
    DevBuf<uint4> xr_;               // 16-byte pass-2 records (coordinates, F'(rho)), one element
    DevBuf<float> dF_;               // F'(rho) per slot, several elements
    Rec16Frame r16_{};
    DevBuf<uint4> shift_;            // image shifts of r16_
    DevBuf<uint32_t> keys_, keys_alt_;
    DevBuf<int> vals_, vals_alt_;
    DevBuf<double4> s_tmp_;
    DevBuf<unsigned char> sort_tmp_;
    size_t sort_tmp_bytes_ = 0;
    DevBuf<int> cell_rank_, cell_start_;
    DevBuf<uint4> cell_order_;       // per cell: the 27 stencil cells in Morton-rank order (32 bytes)
    bool ordered_ = false;           // a 3x3x3 stencil of distinct cells: cell_order_ holds it
    int topo_[6] = {0, 0, 0, 0, 0, 0};  // cells and stencil reach per direction of cell_rank_ / cell_order_
    DevBuf<int> nn_;
    DevBuf<uint32_t> list_;          // 16-byte blocks: entries k..k+3 of slot i at block (k / 4) * ldl_ + i
    uint4* list4() const { return reinterpret_cast<uint4*>(list_.p); }
    int maxn_ = 0;
    int ldl_ = 0;                    // leading dimension of the list (slots)
    DevBuf<int> flags_;              // [0] overflow max count, [1] rebuild needed
    DevBuf<double> part_;            // per-block partial sums
    DevBuf<double> red_;             // reduced values
    DevBuf<double> red2_;            // first-level reduction rows
    std::vector<double> masses_;
    Profiler prof_;
    int batch_rho_ = 8, batch_force_ = 8;
    // persistent one-element passes: LDS row windows [m_lo, m_lo + nwin)
    bool lds_ok_ = false, use_lds_ = false;
    int lds_grid_ = 0;
    int rho_m_lo_ = 0, rho_nwin_ = 0, f_m_lo_ = 0, f_nwin_ = 0;
    size_t rho_lds_bytes_ = 0, f_lds_bytes_ = 0;
    int mvv_rows_ = 0;               // rows of kinetic partial sums written by the last force pass
    // minimiser vectors in slot order: line-search origin, direction, forces
    // at the origin, L-BFGS pairs; a spare for permutations
    DevBuf<double4> mn_x0_, mn_h_, mn_f0_, mn_tmp_;
    std::vector<std::unique_ptr<DevBuf<double4>>> mn_s_, mn_y_;
    DevBuf<double> mn_part_, mn_red_;
    // the minimiser's vectors follow their atoms into a rebuilt slot order
    void min_follow(int nvec, DevBuf<double4>** vecs, int nshift);
    void plan_lds();
};

}  // namespace mapp
