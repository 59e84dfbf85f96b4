// cfg.hpp -- CFG (AtomEye) files: standard and extended files of MD
// systems, and extended files of DMD systems (sites with a fractional
// position, one alpha and one c per element).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mapp {

// A DMD configuration as the importer leaves it: the present elements of
// each site compacted into c_dim components (absent: element 0, c = -1,
// alpha = 0), positions Cartesian, H lower triangular with rows = box vectors.
struct CfgDMD {
    int natms = 0, nelems = 0, c_dim = 0;
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<double> x;      // [natms][3]
    std::vector<uint8_t> elem;  // [natms][c_dim]
    std::vector<double> alpha;  // [natms][c_dim]
    std::vector<double> c;      // [natms][c_dim]
    std::vector<double> masses;
    std::vector<std::string> names;
};

// An MD configuration as ImportCFGMD leaves it: positions Cartesian, H lower
// triangular with rows = box vectors, one element per atom (numbered in the
// order of first appearance), velocities Cartesian (v = s_d R H) when the
// file carries them.
struct CfgMD {
    int natms = 0, nelems = 0;
    bool has_vel = false;
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<double> x;      // [natms][3]
    std::vector<double> v;      // [natms][3] when has_vel
    std::vector<uint8_t> elem;  // [natms]
    std::vector<double> masses;
    std::vector<std::string> names;
};

// ImportCFGMD::operator() (src/import_cfg.cpp:495-604) over
// ImportCFG::read_header / read_body_ext / read_body_std
// (src/import_cfg.cpp:32-480)
CfgMD read_cfg_md(const char* path);

// ExportCFGMD::write_header / write_body_sort (src/export_cfg_md.cpp:26-99):
// atoms in the given order, fractional positions wrapped into [0, 1),
// velocities (v may be null: .NO_VELOCITY.) as v B, then naux auxiliary
// columns aux[natms][naux] named aux_names
void write_cfg_md(const char* path, int natms, const double* H, const double* x, const uint8_t* elem,
                  const double* v, int naux, const double* aux, const std::vector<std::string>& aux_names,
                  int nelems, const double* masses, const std::vector<std::string>& names);

// ImportCFGDMD::operator() (src/import_cfg.cpp:669-854) over
// ImportCFG::read_header / read_body_ext (src/import_cfg.cpp:32-330)
CfgDMD read_cfg_dmd(const char* path);

// ExportCFGDMD::write_header / write_body (src/export_cfg_dmd.cpp:25-170):
// sites in the given order, components expanded back to one alpha and one c
// per element
void write_cfg_dmd(const char* path, int natms, int c_dim, const double* H, const double* x, const uint8_t* elem,
                   const double* alpha, const double* c, int nelems, const double* masses,
                   const std::vector<std::string>& names);

}  // namespace mapp
