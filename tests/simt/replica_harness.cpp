// replica_harness.cpp -- C entry points over mapp::ReplicaBatch built against
// the CPU SIMT emulation (test infrastructure). The force field comes
// straight from the table reader; nothing is uploaded to a device.
#include <cstdio>
#include "replica.hpp"
namespace mapp {
EamForceField::EamForceField(EamTables t) : t_(std::move(t)) {}
void set_last_error(const std::string&) {}
}  // namespace mapp
using mapp::ReplicaBatch;
extern "C" {
void* emu_replica_create(const char* path, int nrep, const int* natms, const double* H, const double* x,
                         const uint8_t* elem, double skin)
{
    try {
        auto ff = std::make_shared<mapp::EamForceField>(mapp::read_dynamo(path, 0, {}, false));
        return new ReplicaBatch(ff, nrep, natms, H, x, elem, skin);
    } catch (const std::exception& e) {
        fprintf(stderr, "emu create: %s\n", e.what());
        return nullptr;
    }
}
void emu_replica_free(void* r) { delete (ReplicaBatch*)r; }
int emu_replica_min_cg(void* r, int nsteps, double e_tol, double max_dx, int* steps, int* status, double* e0,
                       double* e)
{
    try {
        ((ReplicaBatch*)r)->min_cg(nsteps, e_tol, max_dx, 0.4, 0.5, 0.0, steps, status, e0, e, nullptr);
        return 0;
    } catch (const mapp::Error& e) {
        fprintf(stderr, "emu min_cg: %s\n", e.what());
        return e.code;
    }
}
void emu_replica_set_waves(void* r, const int* w) { ((ReplicaBatch*)r)->set_waves(w); }
void emu_replica_use_graph(void* r, int on) { ((ReplicaBatch*)r)->use_graph(on != 0); }
void emu_replica_get_x(void* r, int k, double* x) { ((ReplicaBatch*)r)->get(k, x); }
}
