// replica.hpp -- batched energy minimisation of many small periodic cells
// ("replicas"): every replica runs MinCG::run with LineSearchBackTrack over
// its atom positions, all replicas in one persistent launch.
#pragma once
#include <memory>
#include <vector>

#include "md.hpp"

namespace mapp {

constexpr int kRepMaxImages = 2;  // image shifts -2..2 per box vector (the 7-bit list code)

// Per-replica constants and results, resident on the device.
struct RepCell {
    int off;         // first atom in the batch arrays
    int n;           // atoms
    long long loff;  // first list entry: entries [k][i] at loff + k * n + i
    int maxn;        // list capacity per atom
    int cls;         // size class of the launch-shape tuner
    int status;      // MinCG's error code at the end; 100 + k for a failure of the replica's own checks
    int steps, evals, builds;
    double e0, e, fmax;
    double h[6], b[6];  // packed lower triangles of H and B
    double cs[3];       // extent in s of the (r_c + skin) sphere along each box vector
};

struct RepArgs {
    int ne, nr, nrho;
    double dr_inv, drho_inv, rho_max;
    double cut_sq[kMaxElems * kMaxElems];    // r_c^2
    double cutsk_sq[kMaxElems * kMaxElems];  // (r_c + skin)^2
    double skin_sq4;                         // skin^2 / 4 (DynamicMD::decide)
    const double4* phi;  // [ne*ne][nr] r*phi coefficients
    const double4* rho;  // [ne*ne][nr] rho[a][b]: density of an a atom at a b atom
    const double4* F;    // [ne][nrho]
    // MinCG + LineSearchBackTrack
    int max_nsteps;
    double e_tol, max_dx, ls_c, ls_rho, ls_min_alpha;
};

struct RepArrays {
    double4* x;   // x, y, z, element
    double4* x0;  // line-search origin
    double4* xb;  // positions at the last list build
    double4* h;   // search direction
    double4* f;   // forces
    double4* f0;  // forces at the line-search origin
    double* dF;   // F'(rho)
    uint32_t* list;
    int* nn;
};

struct RepTuneReport {
    int waves[5];
    double ms[6];
};

class ReplicaBatch {
public:
    // nrep replicas; natms[r] atoms each, H [nrep][9] (lower triangular, rows =
    // box vectors), x and elem concatenated in replica order
    ReplicaBatch(std::shared_ptr<EamForceField> ff, int nrep, const int* natms, const double* H, const double* x,
                 const uint8_t* elem, double skin);
    ~ReplicaBatch();
    int nreplicas() const { return nr_; }
    // MinCG::run (src/min_cg.h:453-533) with LineSearchBackTrack (src/ls.h:711-769)
    // on every replica from its current positions; per-replica results (any
    // may be null)
    void min_cg(int max_nsteps, double e_tol, double max_dx, double c, double rho, double min_alpha, int* steps,
                int* status, double* e0, double* e, int* evals);
    // launch shape per size class: 1 (one wave per replica) or 4 (a
    // workgroup per replica); tune() times every threshold on copies of the
    // batch and keeps the fastest
    void set_waves(const int* waves);
    RepTuneReport tune(int max_nsteps, double e_tol, double max_dx, double c, double rho, double min_alpha);
    void use_graph(bool on);
    void get(int r, double* x);  // positions of replica r (may lie outside the box by the displacements)
    Profiler& profiler() { return prof_; }

private:
    struct State {
        DevBuf<RepCell> cells;
        DevBuf<double4> x;
    };
    void launch(State& st, const RepArgs& A, const int* q4, int n4, const int* q1, int n1);
    void build_queues(const int* waves, DevBuf<int>& q4, int& n4, DevBuf<int>& q1, int& n1) const;
    RepArgs args(int max_nsteps, double e_tol, double max_dx, double c, double rho, double min_alpha) const;

    std::shared_ptr<EamForceField> ff_;
    int nr_ = 0;
    size_t natot_ = 0, ltot_ = 0;
    double skin_ = 0.0;
    RepArgs A0_{};
    hipStream_t stream_ = nullptr;
    std::vector<RepCell> host_cells_;
    int waves_[5] = {1, 1, 4, 4, 4};
    State st_, scratch_;
    DevBuf<double4> x0_, xb_, h_, f_, f0_;
    DevBuf<double> dF_;
    DevBuf<uint32_t> list_;
    DevBuf<int> nn_, queue_, q4_, q1_;
    int n4_ = 0, n1_ = 0;
    bool queues_dirty_ = true;
    DevBuf<double4> phi_, rho4_, F4_;
    int grid_ = 0;
    bool graph_on_ = false;
    hipGraphExec_t graph_ = nullptr;
    RepArgs graph_args_{};
    Profiler prof_;
};

}  // namespace mapp
