// md_muvt.hip -- MDMuVT: NVT dynamics of an MD system whose gas atoms are
// exchanged with a reservoir at chemical potential mu every nevery steps
// (src/md_muvt.cpp:122-280). The dynamics are MDSystem::nvt_run's fused
// kernels; the exchange runs the single-process grand-canonical chain of the
// batched exchange kernel on a one-cell batch that holds the system's atoms
// in the reference's order, then hands the new atoms back to the MD system.
//
// At one process PGCMC, which MDMuVT constructs, draws, decides and orders
// the atoms exactly as SGCMC does: one concurrent attempt per round, the same
// Random(seed) stream, insertion points drawn in [0, 1) per dimension, the
// same acceptance expressions and the same add/del order
// (src/pgcmc.cpp:562-735,1163-1266 against src/sgcmc.cpp:251-444). The chain
// is created once per run (PGCMC gcmc(...); gcmc.init()) and keeps its random
// stream across exchanges.
#include <algorithm>
#include <vector>

#include "gcmc.hpp"
#include "md.hpp"

namespace mapp {

int MDSystem::muvt_run(const double* masses, double kB, double hP, double T, double dt, double mu, int gas, int seed,
                       int nevery, int nattempts, int nsteps, MDReport* rep, double* T_part, int* counts, double* hist)
{
    const EamTables& t = ff_->tables();
    MAPP_REQUIRE(gas >= 0 && gas < t.ne, ERR_ARG, "gas element index out of range");
    MAPP_REQUIRE(nevery >= 1 && nattempts >= 0 && nsteps >= 0, ERR_ARG,
                 "nevery must be positive, nattempts and nsteps non-negative");
    MAPP_REQUIRE(hP > 0.0, ERR_ARG, "the Planck constant must be positive");
    // the chain's cell holds the whole system, inserted atoms included
    const int cap = 4 * kGcmcWaveCap;
    MAPP_REQUIRE(n_ <= cap, ERR_CAPACITY,
                 "the exchange holds at most " + std::to_string(cap) + " atoms (this system has " +
                     std::to_string(n_) + ")");
    double H9[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) H9[3 * i + j] = box_.H[i][j];
    std::vector<double> xh((size_t)3 * cap), vh((size_t)3 * cap);
    std::vector<uint8_t> eh(cap);
    get_x(xh.data());
    int n0 = n_, cp = cap, sd = seed;
    GcmcBatch gb(ff_, 1, &n0, H9, xh.data(), elem_host_.data(), &cp, gas, mu, T, kB, hP, masses, &sd);
    int cnt[4] = {0, 0, 0, 0};
    std::vector<int8_t> dec(std::max(nattempts, 1));
    int ngas0 = 0;
    for (int i = 0; i < n_; ++i) ngas0 += elem_host_[i] == gas;
    std::vector<int> ngas_after;  // gas atoms after each exchange

    ExchangeHook hook;
    hook.nevery = nevery;
    hook.run = [&]() {
        // GCMC::xchng: the current atoms (after the first half step) in the
        // caller's order, wrapped into the box by the load
        get_x(xh.data());
        get_v(vh.data());
        gb.load_cell(0, n_, xh.data(), elem_host_.data(), vh.data());
        gb.attempts(nattempts, nattempts ? dec.data() : nullptr, nullptr);
        for (int k = 0; k < nattempts; ++k) {
            const int d = dec[k];
            if (d == 1) ++cnt[0];
            if (d == 2) ++cnt[1];
            if (d == 1 || d == -1) ++cnt[2]; else ++cnt[3];
        }
        int na = 0, ng = 0;
        gb.counts(&na, &ng);
        MAPP_REQUIRE(na >= 1, ERR_STATE, "the exchange removed every atom");
        gb.get_cell(0, xh.data(), eh.data(), nullptr, nullptr, vh.data(), nullptr);
        // DynamicMD::fin_xchng: the new atoms, the cell sort and the list
        reset_atoms(na, xh.data(), eh.data(), vh.data());
        ngas_after.push_back(ng);
    };
    std::vector<double> hT(std::max(nsteps, 1));
    const int builds = nvt_run(masses, kB, T, dt, nsteps, rep, T_part, &hook, hT.data());
    if (counts)
        for (int k = 0; k < 4; ++k) counts[k] = cnt[k];
    if (hist)
        for (int s = 0; s < nsteps; ++s) {
            // the exchanges run after steps nevery - 1, 2 nevery - 1, ...
            const int m = (s + 1) / nevery;
            hist[2 * (size_t)s] = hT[s];
            hist[2 * (size_t)s + 1] = m == 0 ? ngas0 : ngas_after[m - 1];
        }
    return builds;
}

}  // namespace mapp
