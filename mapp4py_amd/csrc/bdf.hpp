// bdf.hpp -- variable-order BDF integration of the DMD concentrations
// (DAEBDF): the order and step control on the host, every evaluation (rates,
// Jacobian products, the implicit Newton-GMRES step, free energy, CG
// relaxation) on the GPU through DMDSystem.
#pragma once
#include <vector>

#include "dmd.hpp"

namespace mapp {

struct BdfParams {
    double a_tol = 1e-8;
    double min_dt = 2.220446049250313e-16;
    int max_nsteps = 1000;
    int max_nnewton_iters = 5;
    int max_ngmres_iters = 5;
    int nreset = 0;  // 0: no periodic relaxation of x and alpha
    int ntally = 0;  // record (step, time, free energy) every ntally steps and at step 0
    int relax_max_iter = 1000;  // CG iterations of one relaxation
};

struct BdfReport {
    int nsteps = 0, max_q = 1;
    double t = 0.0, max_dt = 0.0, min_dt = 0.0;
    int nonlin_acc = 0, nonlin_rej = 0, intpol_acc = 0, intpol_rej = 0, integ_acc = 0, integ_rej = 0, err_mins = 0;
};

// DAEBDF::run (src/dae_bdf.cpp:103-250) from the current state to t_tot;
// tally rows (step, time, free energy) are appended to *tally when given
BdfReport bdf_run(DMDSystem& d, const double* r_crd, const double* zeta, double t_tot, const BdfParams& p,
                  std::vector<double>* tally);

}  // namespace mapp
