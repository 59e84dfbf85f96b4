// gcmc_harness.cpp -- C entry points over mapp::GcmcBatch built against the
// CPU SIMT emulation (test infrastructure). The force field comes straight
// from the table reader; nothing is uploaded to a device.
#include <cstdio>
#include "gcmc.hpp"
namespace mapp {
EamForceField::EamForceField(EamTables t) : t_(std::move(t)) {}
void set_last_error(const std::string&) {}
}  // namespace mapp
using mapp::GcmcBatch;
extern "C" {
void* emu_gcmc_create(const char* path, int ncells, const int* natms, const double* H, const double* x,
                      const uint8_t* elem, const int* cap, int gas, double mu, double T, double kB, double hP,
                      const double* masses, const int* seeds)
{
    try {
        auto ff = std::make_shared<mapp::EamForceField>(mapp::read_dynamo(path, 0, {}, false));
        return new GcmcBatch(ff, ncells, natms, H, x, elem, cap, gas, mu, T, kB, hP, masses, seeds);
    } catch (const std::exception& e) {
        fprintf(stderr, "emu create: %s\n", e.what());
        return nullptr;
    }
}
void emu_gcmc_free(void* g) { delete (GcmcBatch*)g; }
int emu_gcmc_attempts(void* g, int k, int8_t* dec, double* du)
{
    try {
        ((GcmcBatch*)g)->attempts(k, dec, du);
        return 0;
    } catch (const mapp::Error& e) {
        fprintf(stderr, "emu attempts: %s\n", e.what());
        return e.code;
    }
}
void emu_gcmc_set_waves(void* g, const int* w) { ((GcmcBatch*)g)->set_waves(w); }
void emu_gcmc_use_graph(void* g, int on) { ((GcmcBatch*)g)->use_graph(on != 0); }
void emu_gcmc_counts(void* g, int* n, int* ng) { ((GcmcBatch*)g)->counts(n, ng); }
void emu_gcmc_get_cell(void* g, int c, double* x, uint8_t* e, double* rho, double* F, double* v, int* seed)
{
    ((GcmcBatch*)g)->get_cell(c, x, e, rho, F, v, seed);
}
int emu_gcmc_xchng_energy(void* g, int cell, int mode, const double* s, int igas, double* du)
{
    try {
        *du = ((GcmcBatch*)g)->xchng_energy(cell, mode, s, igas);
        return 0;
    } catch (const mapp::Error& e) {
        fprintf(stderr, "emu xchng_energy: %s\n", e.what());
        return e.code;
    }
}
int emu_gcmc_xchng_commit(void* g, const double* v)
{
    try {
        ((GcmcBatch*)g)->xchng_commit(v);
        return 0;
    } catch (const mapp::Error& e) {
        fprintf(stderr, "emu xchng_commit: %s\n", e.what());
        return e.code;
    }
}
void emu_gcmc_tune(void* g, int k, int* w)
{
    auto r = ((GcmcBatch*)g)->tune(k);
    for (int i = 0; i < mapp::kGcmcClasses; ++i) w[i] = r.waves[i];
}
}
