// gcmc.hpp -- batched grand-canonical exchange of one gas element over many
// independent periodic cells: every cell runs its own SGCMC Markov chain
// (insert/delete attempts with EAM exchange energies) on the device, all
// cells in one persistent launch.
#pragma once
#include <memory>
#include <vector>

#include "md.hpp"

namespace mapp {

constexpr int kGcmcMaxTrials = 64;   // periodic images of one trial point
constexpr int kGcmcClasses = 5;      // cell-size classes of the tile tuner
constexpr int kGcmcWaveCap = 4096;   // atoms a one-wave cell may hold (64 per lane)

// cell-size class boundaries (atoms at creation): [0,256) [256,512) [512,1024) [1024,2048) [2048,inf)
int gcmc_class_of(int natms);

// Per-cell state and constants, resident on the device. The random stream
// (seed, cached Gaussian) and the atom count live here between launches.
struct GcmcCell {
    int off, cap, n, ngas;
    int seed, gchk, status, cls;
    double gval;
    double h[6];          // packed lower triangle of H
    double lo[3], hi[3];  // padded box in s for trial images (gas cutoff)
    double lo2[3], hi2[3];  // padded box in s for the initial densities (largest cutoff)
    double zv;            // z_fac * vol
};

struct GcmcArgs {
    int ne, nr, nrho, gas;
    double dr_inv, drho_inv, rho_max, beta, sigma;
    double cut_sq[kMaxElems * kMaxElems];
    const double4* phi;  // [ne*ne][nr] coefficients of r*phi
    const double4* rho;  // [ne*ne][nr] rho[a][b]: density of an a atom at a b atom
    const double4* F;    // [ne][nrho]
};

// Device arrays of one batch (one copy for production, one for tuning).
struct GcmcState {
    DevBuf<GcmcCell> cells;
    DevBuf<double4> xe;   // x, y, z, element
    DevBuf<double4> s;    // fractional coordinates
    DevBuf<double4> v;
    DevBuf<double> rho, F, rho_x, F_x;
    DevBuf<uint8_t> el;
};

struct GcmcTuneReport {
    int waves[kGcmcClasses];
    double ms[kGcmcClasses + 1];  // whole-batch launch time for threshold class 0..5
};

class GcmcBatch {
public:
    GcmcBatch(std::shared_ptr<EamForceField> ff, int ncells, const int* natms, const double* H, const double* x,
              const uint8_t* elem, const int* cap, int gas, double mu, double T, double kB, double hP,
              const double* masses, const int* seeds);
    ~GcmcBatch();

    int ncells() const { return nc_; }
    // nattempts SGCMC attempts in every cell; decisions/dU are [ncells][nattempts] (either may be null)
    void attempts(int nattempts, int8_t* decisions, double* dU);
    GcmcTuneReport tune(int nattempts);
    void set_waves(const int* waves);  // per class: 1 or 4 waves per cell
    void use_graph(bool on);
    void counts(int* natms, int* ngas);
    void get_cell(int c, double* x, uint8_t* elem, double* rho, double* F, double* v, int* seed);
    // cell c's atoms replaced by n atoms (x Cartesian, wrapped into the box
    // as x2s_lcl does; v may be null: zero), their densities and F recomputed
    // (init_xchng); the cell's random stream continues
    void load_cell(int c, int n, const double* x, const uint8_t* elem, const double* v);
    // one trial of one cell, driven by the caller (pre_xchng_energy +
    // xchng_energy): mode 1 inserts at fractional s, mode 2 deletes the
    // igas-th gas atom; returns the exchange energy and keeps the trial
    double xchng_energy(int cell, int mode, const double* s, int igas);
    // the kept trial accepted (post_xchng_energy after ins_succ / del_succ)
    void xchng_commit(const double* v);
    Profiler& profiler() { return prof_; }

private:
    void split_cells(const int* waves, std::vector<int>& q4, std::vector<int>& q1) const;
    void build_queues();
    void launch(GcmcState& st, int nattempts, int8_t* dec, double* du, const int* q4, int n4, const int* q1, int n1);
    void copy_state(GcmcState& dst, const GcmcState& src);
    void check_status(GcmcState& st);

    std::shared_ptr<EamForceField> ff_;
    int nc_ = 0;
    size_t total_cap_ = 0;
    GcmcArgs A_{};
    hipStream_t stream_ = nullptr;
    std::vector<GcmcCell> host_cells_;
    std::vector<int> n0_;
    int waves_[kGcmcClasses] = {1, 1, 1, 4, 4};
    GcmcState st_, scratch_;
    DevBuf<double4> phi_, rho4_, F4_;
    DevBuf<int> queue_;   // [0] next four-wave cell, [1] next one-wave cell
    DevBuf<int> q4_, q1_;
    int n4_ = 0, n1_ = 0;
    bool queues_dirty_ = true;
    DevBuf<int8_t> dec_;
    DevBuf<double> du_;
    int grid_ = 0;
    bool graph_on_ = false;
    hipGraphExec_t graph_ = nullptr;
    int graph_k_ = -1;
    Profiler prof_;
    struct PendingTrial {
        int cell = -1, mode = 0, del_idx = -1;
        double rhoi = 0.0, x[3] = {0, 0, 0}, s[3] = {0, 0, 0};
        int64_t ver = -1;
    };
    PendingTrial pending_;
    int64_t ver_ = 0;  // bumped by every change of the cells
    DevBuf<double> xout_;
};

}  // namespace mapp
